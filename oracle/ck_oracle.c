/*
 * ck_oracle.c -- CPU restatement of the reference cuckoo Monte Carlo trial.
 *
 * TEST INFRASTRUCTURE ONLY: the checker, never the thing measured or shipped.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * What pins each function (tests/test_oracle.py holds the checks):
 *   ock_sbox            GF(2^8) inverse + affine map; FIPS-197 S(0)=63, S(53)=ed,
 *                       and bijectivity.
 *   ock_aes_bytewise    FIPS-197 Appendix C.1 known answer; equal to the AES-NI
 *                       path (a hardware implementation) on random blocks; equal
 *                       to the reference build in oracle/_ref.
 *   ock_expand_key      FIPS-197 Appendix A.1 key schedule (2b7e1516... vector).
 *   ock_batch_key       host_tests.cpp:37-46 golden vectors (seed 123, id 0 and 7).
 *   ock_splitmix        published SplitMix64 outputs (seed 1234567, Vigna's test
 *                       vector) and the reference build.
 *   ock_parameters      host_tests.cpp:158-159 (23 / 32 buckets), closed forms of
 *                       the bgh25 rule, and the reference build over a pnum sweep.
 *   ock_bucket_*        bijection of the 8-round Feistel network on [0,2^2h)
 *                       (brute force on tiny domains), output < buckets, cached ==
 *                       direct, and the reference build.
 *   ock_trial_run       cpu_golden.hpp (5 bgh25 / 214 drrt18 failures in 32768
 *                       trials), smoke_expected.json, cpu_tests.cpp:30-39 affine
 *                       vectors, table invariants (each placed ordinal exactly
 *                       once), and the reference build's per-trial tables.
 *   extension (k!=3, stash>0, explicit buckets/kick, pnum>65534): no reference
 *                       counterpart.  Pinned by reduction (k=3, stash=0, derived
 *                       buckets/kick equals the reference bit-for-bit), by table
 *                       invariants and by brute-force hand traces at tiny sizes.
 */
#include "ck_oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#if defined(__x86_64__)
#include <immintrin.h>
#include <wmmintrin.h>
#endif

/* ---------------------------------------------------------------- AES */

static uint8_t g_sbox[256];
static pthread_once_t g_sbox_once = PTHREAD_ONCE_INIT;

static uint8_t gf_mul(uint8_t a, uint8_t b) {
    uint8_t r = 0;
    while (b) {
        if (b & 1) r ^= a;
        a = (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
        b >>= 1;
    }
    return r;
}
static uint8_t rotl8(uint8_t x, int s) { return (uint8_t)((x << s) | (x >> (8 - s))); }

/* FIPS-197 section 5.1.1: multiplicative inverse in GF(2^8) followed by the
 * affine transformation b ^ rotl(b,1..4) ^ 0x63. */
static void sbox_init(void) {
    for (unsigned x = 0; x < 256; ++x) {
        uint8_t inv = 0;
        if (x) {
            /* x^254 = x^-1 */
            uint8_t r = 1, base = (uint8_t)x;
            unsigned e = 254;
            while (e) {
                if (e & 1) r = gf_mul(r, base);
                base = gf_mul(base, base);
                e >>= 1;
            }
            inv = r;
        }
        g_sbox[x] = (uint8_t)(inv ^ rotl8(inv, 1) ^ rotl8(inv, 2) ^ rotl8(inv, 3) ^
                              rotl8(inv, 4) ^ 0x63);
    }
}
static const uint8_t* sbox(void) {
    pthread_once(&g_sbox_once, sbox_init);
    return g_sbox;
}
void ock_sbox(uint8_t out[256]) { memcpy(out, sbox(), 256); }

static uint8_t xtime(uint8_t x) { return (uint8_t)((x << 1) ^ ((x & 0x80) ? 0x1b : 0)); }

static uint32_t sub_word(uint32_t x) {
    const uint8_t* s = sbox();
    return ((uint32_t)s[x >> 24] << 24) | ((uint32_t)s[(x >> 16) & 255] << 16) |
           ((uint32_t)s[(x >> 8) & 255] << 8) | s[x & 255];
}

/* Key schedule on four big-endian words (cuckoo_bench.cu:136-149: the device
 * derives the per-trial key from AES output words directly). */
void ock_expand_words(const uint32_t words[4], ock_key* k) {
    for (int i = 0; i < 4; ++i) k->w[i] = words[i];
    uint8_t rcon = 1;
    for (int i = 4; i < 44; ++i) {
        uint32_t v = k->w[i - 1];
        if ((i & 3) == 0) {
            v = sub_word((v << 8) | (v >> 24)) ^ ((uint32_t)rcon << 24);
            rcon = xtime(rcon);
        }
        k->w[i] = k->w[i - 4] ^ v;
    }
}

/* expand_key(lo, hi) (cuckoo_bench.cu:86-99): the 16 key bytes are lo then hi
 * in little-endian order; round-key words are read big-endian from them. */
void ock_expand_key(uint64_t lo, uint64_t hi, ock_key* k) {
    uint8_t b[16];
    for (int i = 0; i < 8; ++i) {
        b[i] = (uint8_t)(lo >> (8 * i));
        b[8 + i] = (uint8_t)(hi >> (8 * i));
    }
    uint32_t words[4];
    for (int i = 0; i < 4; ++i)
        words[i] = ((uint32_t)b[4 * i] << 24) | ((uint32_t)b[4 * i + 1] << 16) |
                   ((uint32_t)b[4 * i + 2] << 8) | b[4 * i + 3];
    ock_expand_words(words, k);
}

/* Byte-oriented FIPS-197 cipher (host_aes128, cuckoo_bench.cu:360-392):
 * input and output are four big-endian words. */
void ock_aes_bytewise(const ock_key* key, const uint32_t in[4], uint32_t out[4]) {
    const uint8_t* s = sbox();
    uint8_t st[16];
    for (int j = 0; j < 4; ++j)
        for (int q = 0; q < 4; ++q)
            st[4 * j + q] = (uint8_t)(in[j] >> (24 - 8 * q)) ^ (uint8_t)(key->w[j] >> (24 - 8 * q));
    for (int round = 1; round <= 10; ++round) {
        uint8_t t[16];
        /* SubBytes + ShiftRows: row r of column c comes from column c+r. */
        for (int c = 0; c < 4; ++c)
            for (int r = 0; r < 4; ++r) t[4 * c + r] = s[st[4 * ((c + r) & 3) + r]];
        if (round != 10) {
            for (int c = 0; c < 4; ++c) {
                const uint8_t a0 = t[4 * c], a1 = t[4 * c + 1], a2 = t[4 * c + 2], a3 = t[4 * c + 3];
                st[4 * c + 0] = (uint8_t)(xtime(a0) ^ xtime(a1) ^ a1 ^ a2 ^ a3);
                st[4 * c + 1] = (uint8_t)(a0 ^ xtime(a1) ^ xtime(a2) ^ a2 ^ a3);
                st[4 * c + 2] = (uint8_t)(a0 ^ a1 ^ xtime(a2) ^ xtime(a3) ^ a3);
                st[4 * c + 3] = (uint8_t)(xtime(a0) ^ a0 ^ a1 ^ a2 ^ xtime(a3));
            }
        } else {
            memcpy(st, t, 16);
        }
        for (int j = 0; j < 4; ++j)
            for (int q = 0; q < 4; ++q) st[4 * j + q] ^= (uint8_t)(key->w[4 * round + j] >> (24 - 8 * q));
    }
    for (int j = 0; j < 4; ++j)
        out[j] = ((uint32_t)st[4 * j] << 24) | ((uint32_t)st[4 * j + 1] << 16) |
                 ((uint32_t)st[4 * j + 2] << 8) | st[4 * j + 3];
}

#if defined(__x86_64__)
static int g_aesni = -1;
int ock_has_aesni(void) {
    if (g_aesni < 0) g_aesni = __builtin_cpu_supports("aes") ? 1 : 0;
    return g_aesni;
}
/* AES-NI form of the same cipher (cpu_trials.hpp:19-43): the 16-byte state in
 * memory order is the big-endian bytes of the four input words. */
__attribute__((target("aes,sse4.1"))) static __m128i be_words(const uint32_t w[4]) {
    return _mm_set_epi32((int)__builtin_bswap32(w[3]), (int)__builtin_bswap32(w[2]),
                         (int)__builtin_bswap32(w[1]), (int)__builtin_bswap32(w[0]));
}
__attribute__((target("aes,sse4.1"))) static void aesni_block(const ock_key* k, const uint32_t in[4],
                                                             uint32_t out[4]) {
    __m128i st = _mm_xor_si128(be_words(in), be_words(k->w));
    for (int r = 1; r < 10; ++r) st = _mm_aesenc_si128(st, be_words(k->w + 4 * r));
    st = _mm_aesenclast_si128(st, be_words(k->w + 40));
    uint32_t little[4];
    _mm_storeu_si128((__m128i*)little, st);
    for (int i = 0; i < 4; ++i) out[i] = __builtin_bswap32(little[i]);
}
/* Round-key cache for many blocks under one key. */
typedef struct { __m128i rk[11]; } aesni_sched;
__attribute__((target("aes,sse4.1"))) static void aesni_load(const ock_key* k, aesni_sched* s) {
    for (int r = 0; r < 11; ++r) s->rk[r] = be_words(k->w + 4 * r);
}
/* Eight independent blocks in flight: the last word of AES(c0, r, half, right+i). */
__attribute__((target("aes,sse4.1"))) static void aesni_last8(const aesni_sched* s, uint32_t c0, uint32_t c1,
                                                             uint32_t c2, uint32_t right0, uint32_t out[8]) {
    __m128i st[8];
    for (int i = 0; i < 8; ++i) {
        const uint32_t w[4] = {c0, c1, c2, right0 + (uint32_t)i};
        st[i] = _mm_xor_si128(be_words(w), s->rk[0]);
    }
    for (int r = 1; r < 10; ++r)
        for (int i = 0; i < 8; ++i) st[i] = _mm_aesenc_si128(st[i], s->rk[r]);
    for (int i = 0; i < 8; ++i) {
        st[i] = _mm_aesenclast_si128(st[i], s->rk[10]);
        out[i] = __builtin_bswap32((uint32_t)_mm_extract_epi32(st[i], 3));
    }
}
#else
int ock_has_aesni(void) { return 0; }
#endif

void ock_aes(const ock_key* k, const uint32_t in[4], uint32_t out[4]) {
#if defined(__x86_64__)
    if (ock_has_aesni()) {
        aesni_block(k, in, out);
        return;
    }
#endif
    ock_aes_bytewise(k, in, out);
}

/* batch_key (cuckoo_bench.cu:393-404). */
void ock_batch_key(uint64_t seed, uint64_t batch, ock_key* out) {
    ock_key master;
    ock_expand_key(seed, seed ^ 0x9e3779b97f4a7c15ULL, &master);
    const uint32_t in[4] = {0x42415443u, (uint32_t)(batch >> 32), (uint32_t)batch, 0x50524731u};
    uint32_t x[4];
    ock_aes_bytewise(&master, in, x);
    const uint64_t prg_lo = ((uint64_t)x[0] << 32) | x[1];
    const uint64_t prg_hi = ((uint64_t)x[2] << 32) | x[3];
    ock_expand_key(seed ^ prg_lo, prg_hi ^ (seed ^ 0x243f6a8885a308d3ULL), out);
}

/* SplitMix64 (cuckoo_core.hpp:110-114). */
uint64_t ock_splitmix(uint64_t* state) {
    *state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* --------------------------------------------------------- parameters */

unsigned ock_bit_width(uint64_t v) {
    unsigned bits = 0;
    for (; v; v >>= 1) ++bits;
    return bits;
}

/* original_log2_u64 (cuckoo_bench.cu:488-498): a float-rounded atanh series;
 * the DRRT18 bucket count depends on its exact rounding. */
float ock_original_log2(uint64_t n) {
    if (n == 0) return -INFINITY;
    const unsigned e = ock_bit_width(n) - 1;
    const double x = (double)n / (double)(1ULL << e);
    if (x >= 2.0) return (float)(e + 1);
    const double z = (x - 1.0) / (x + 1.0), z2 = z * z;
    double term = z, sum = 0.0;
    for (int k = 1; k <= 23; k += 2) {
        sum += term / k;
        term *= z2;
    }
    return (float)((double)e + 2.0 * sum * 1.4426950408889634074);
}

int ock_parameters(uint32_t mode, uint32_t pnum, uint32_t ell_i, uint32_t k, uint32_t buckets,
                   uint32_t kick, uint32_t stash, uint32_t input_mode, ock_params* x) {
    memset(x, 0, sizeof *x);
    if (mode != OCK_MODE_BGH25 && mode != OCK_MODE_DRRT18) return -1;
    if (k < 2 || k > 4 || pnum == 0 || ell_i > 61 || input_mode > 1) return -2;
    if (ell_i < 32 && (1ULL << ell_i) < pnum) return -3;
    uint32_t b, kk;
    if (mode == OCK_MODE_BGH25) {
        const uint64_t overhead = pnum < 30 ? 100 : 50;
        b = (uint32_t)((uint64_t)pnum * (100 + overhead) / 100);
        kk = 3 * pnum;
    } else {
        const float lf = ock_original_log2(pnum);
        const float f = (float)((double)pnum * ((double)(170.0f + lf) / 123.5));
        b = (uint32_t)ceil((double)f);
        kk = 100;
    }
    if (buckets != OCK_DERIVED) b = buckets;
    if (kick != OCK_DERIVED) kk = kick;
    if (b == 0) return -4;
    x->mode = mode;
    x->k = k;
    x->pnum = pnum;
    x->buckets = b;
    x->kick = kk;
    x->stash = stash;
    x->ell_i = ell_i;
    x->input_mode = input_mode;
    const uint64_t range = 1ULL << ell_i;
    const uint64_t source = range * k;
    x->bucket_size = source / b + (source % b != 0);
    x->ell_d = (uint32_t)ceil(ock_original_log2(x->bucket_size));
    if (b > UINT64_MAX / x->bucket_size) return -5;
    x->domain = (uint64_t)b * x->bucket_size;
    unsigned bits = ock_bit_width(x->domain - 1);
    if (bits < 2) bits = 2;
    if (bits & 1u) ++bits;
    x->half_bits = bits / 2;
    x->mask = (x->half_bits >= 64) ? ~0ULL : ((1ULL << x->half_bits) - 1ULL);
    x->range_mask = range - 1ULL;
    return 0;
}

/* ------------------------------------------------------- trial pieces */

/* Per-trial derivation (cuckoo_bench.cu:222-238, cpu_trials.hpp:71-80). */
void ock_trial_setup(const ock_key* batch, uint32_t offset, const ock_params* p, ock_trial* t) {
    uint32_t w[4];
    const uint32_t kin[4] = {0x4b455931u, 0, offset, 0x50525031u};
    ock_aes(batch, kin, w);
    ock_expand_words(w, &t->key);
    const uint32_t win[4] = {0x57414c4bu, 0, offset, 0x524e4731u};
    ock_aes(batch, win, w);
    t->walk = ((uint64_t)w[0] << 32) | w[1];
    t->shift = (((uint64_t)w[2] << 32) | w[3]) & p->range_mask;
    uint64_t input_state = t->shift ^ 0xd1b54a32d192ed03ULL;
    t->step = (ock_splitmix(&input_state) | 1ULL) & p->range_mask;
    t->randomness = ock_splitmix(&t->walk);
}

/* trial_element (cuckoo_bench.cu:189-193). */
uint64_t ock_element(const ock_trial* t, const ock_params* p, uint32_t ordinal) {
    if (p->input_mode == 0) return ordinal;
    return (t->shift + (uint64_t)ordinal * t->step) & p->range_mask;
}

static uint32_t bucket_of(const ock_params* p, uint64_t y) {
    if (p->domain <= 0xffffffffULL) return (uint32_t)y / (uint32_t)p->bucket_size;
    return (uint32_t)(y / p->bucket_size);
}

/* Direct Feistel PRP with cycle walking (cuckoo_bench.cu:161-184,
 * cpu_trials.hpp:56-68); y = k*element + which generalises 3*element+which. */
uint32_t ock_bucket_direct(const ock_trial* t, const ock_params* p, uint32_t ordinal, uint32_t which) {
    uint64_t y = ock_element(t, p, ordinal) * p->k + which;
    do {
        uint64_t left = y >> p->half_bits, right = y & p->mask;
        for (uint32_t r = 0; r < 8; ++r) {
            const uint32_t in[4] = {0x43455350u, r, p->half_bits, (uint32_t)right};
            uint32_t h[4];
            ock_aes(&t->key, in, h);
            const uint64_t next = (left ^ ((uint64_t)h[3] & p->mask)) & p->mask;
            left = right;
            right = next;
        }
        y = (left << p->half_bits) | right;
    } while (y >= p->domain);
    return bucket_of(p, y);
}

/* Round cache: rounds[r<<half | right] = low half_bits of the round function
 * (gpu_shared.cuh:68-77, cpu_trials.hpp:139-142).  Needs half_bits <= 16. */
void ock_build_rounds(const ock_trial* t, const ock_params* p, uint16_t* rounds) {
    const uint32_t width = 1u << p->half_bits;
#if defined(__x86_64__)
    if (ock_has_aesni() && width >= 8) {
        aesni_sched s;
        aesni_load(&t->key, &s);
        for (uint32_t r = 0; r < 8; ++r)
            for (uint32_t right = 0; right < width; right += 8) {
                uint32_t h[8];
                aesni_last8(&s, 0x43455350u, r, p->half_bits, right, h);
                for (int i = 0; i < 8; ++i) rounds[(r << p->half_bits) + right + i] = (uint16_t)(h[i] & p->mask);
            }
        return;
    }
#endif
    for (uint32_t r = 0; r < 8; ++r)
        for (uint32_t right = 0; right < width; ++right) {
            const uint32_t in[4] = {0x43455350u, r, p->half_bits, right};
            uint32_t h[4];
            ock_aes(&t->key, in, h);
            rounds[(r << p->half_bits) + right] = (uint16_t)(h[3] & p->mask);
        }
}

/* CachedBucketsT::operator() (cuckoo_core.hpp:56-71). */
uint32_t ock_bucket_cached(const ock_trial* t, const ock_params* p, const uint16_t* rounds,
                           uint32_t ordinal, uint32_t which) {
    uint64_t y = ock_element(t, p, ordinal) * p->k + which;
    const unsigned h = p->half_bits;
    do {
        uint64_t left = y >> h, right = y & p->mask;
        for (unsigned r = 0; r < 8; ++r) {
            const uint64_t next = left ^ (uint64_t)rounds[((size_t)r << h) + right];
            left = right;
            right = next & p->mask;
        }
        y = (left << h) | right;
    } while (y >= p->domain);
    return bucket_of(p, y);
}

void ock_bucket_map(const ock_key* batch, uint32_t offset, const ock_params* p, uint32_t* map) {
    ock_trial t;
    ock_trial_setup(batch, offset, p, &t);
    uint16_t* rounds = NULL;
    if (p->half_bits <= 16) {
        rounds = (uint16_t*)malloc(sizeof(uint16_t) * ((size_t)8 << p->half_bits));
        ock_build_rounds(&t, p, rounds);
    }
    for (uint32_t o = 0; o < p->pnum; ++o)
        for (uint32_t w = 0; w < p->k; ++w)
            map[(size_t)o * p->k + w] = rounds ? ock_bucket_cached(&t, p, rounds, o, w)
                                               : ock_bucket_direct(&t, p, o, w);
    free(rounds);
}

/* ------------------------------------------------------------ insertion */

#define EMPTY 0xffffffffu
static inline uint32_t pack(uint32_t ordinal, uint32_t choice) { return ordinal | (choice << 30); }
static inline uint32_t ord_of(uint32_t e) { return e & 0x3fffffffu; }
static inline uint32_t choice_of(uint32_t e) { return e >> 30; }

typedef struct {
    const ock_trial* t;
    const ock_params* p;
    const uint16_t* rounds;
} bucket_fn;
static inline uint32_t eval(const bucket_fn* f, uint32_t ordinal, uint32_t which) {
    return f->rounds ? ock_bucket_cached(f->t, f->p, f->rounds, ordinal, which)
                     : ock_bucket_direct(f->t, f->p, ordinal, which);
}

/* insert_core (cuckoo_core.hpp:117-161), generalised to k candidates and a
 * stash.  Wherever the reference returns "failed", the item in hand goes to
 * the stash while it has room; with stash == 0 and k == 3 this is exactly the
 * reference. */
static int insert_trial(const ock_params* p, const bucket_fn* f, uint32_t* bins, uint64_t walk,
                        uint64_t randomness, ock_stats* st) {
    const uint32_t k = p->k;
    uint32_t stash_used = 0;
    for (uint32_t b = 0; b < p->buckets; ++b) bins[b] = EMPTY;
    for (uint32_t item = 0; item < p->pnum; ++item) {
        if (p->mode == OCK_MODE_BGH25) {
            uint32_t cur = item, collisions = 0;
            for (;;) {
                if (randomness < k) randomness = ock_splitmix(&walk);
                const uint32_t which = (uint32_t)(randomness % k);
                randomness /= k;
                const uint32_t target = eval(f, cur, which);
                st->evals++;
                st->probes++;
                const uint32_t old = bins[target];
                if (old == EMPTY) {
                    bins[target] = pack(cur, which);
                    st->placed++;
                    break;
                }
                if (++collisions > p->kick) {
                    if (stash_used < p->stash) {
                        ++stash_used;
                        st->stashed++;
                        break;
                    }
                    return 1;
                }
                bins[target] = pack(cur, which);
                cur = ord_of(old);
            }
        } else {
            const uint32_t initial = eval(f, item, 0);
            st->evals++;
            st->probes++;
            if (bins[initial] == EMPTY) {
                bins[initial] = pack(item, 0);
                st->placed++;
                continue;
            }
            int inserted = 0;
            for (uint32_t c = 1; c < k; ++c) {
                const uint32_t target = eval(f, item, c);
                st->evals++;
                st->probes++;
                if (bins[target] == EMPTY) {
                    bins[target] = pack(item, c);
                    st->placed++;
                    inserted = 1;
                    break;
                }
            }
            if (inserted) continue;
            uint32_t cur = item, choice = 0, target = initial, kicks = 0;
            for (;;) {
                const uint32_t displaced = bins[target];
                st->probes++;
                if (displaced == EMPTY) {
                    bins[target] = pack(cur, choice);
                    st->placed++;
                    break;
                }
                if (kicks >= p->kick) {
                    if (stash_used < p->stash) {
                        ++stash_used;
                        st->stashed++;
                        break;
                    }
                    return 1;
                }
                bins[target] = pack(cur, choice);
                cur = ord_of(displaced);
                choice = choice_of(displaced);
                const uint32_t previous = target;
                int found = 0;
                for (uint32_t s = 0; s < k; ++s) {
                    choice = (choice + 1) % k;
                    const uint32_t next = eval(f, cur, choice);
                    st->evals++;
                    if (next != previous) {
                        target = next;
                        found = 1;
                        break;
                    }
                }
                if (!found) {
                    if (stash_used < p->stash) {
                        ++stash_used;
                        st->stashed++;
                        break;
                    }
                    return 1;
                }
                ++kicks;
            }
        }
    }
    return 0;
}

int ock_trial_run(const ock_key* batch, uint32_t offset, const ock_params* p, uint32_t* bins,
                  uint16_t* rounds, ock_stats* st) {
    ock_trial t;
    ock_trial_setup(batch, offset, p, &t);
    if (rounds) ock_build_rounds(&t, p, rounds);
    const bucket_fn f = {&t, p, rounds};
    ock_stats local = {0, 0, 0, 0};
    const int failed = insert_trial(p, &f, bins, t.walk, t.randomness, &local);
    if (st) {
        st->probes += local.probes;
        st->evals += local.evals;
        st->stashed += local.stashed;
        st->placed += local.placed;
    }
    return failed;
}

/* ------------------------------------------------------ threaded range */

typedef struct {
    const ock_key* batch;
    const ock_params* p;
    uint32_t first, count;
    int cached;
    uint8_t* flags;
    uint32_t* probes;
    volatile uint32_t next;
    pthread_mutex_t lock;
    ock_stats totals;
} range_job;

static void* range_worker(void* arg) {
    range_job* j = (range_job*)arg;
    const ock_params* p = j->p;
    uint32_t* bins = (uint32_t*)malloc(sizeof(uint32_t) * p->buckets);
    uint16_t* rounds = NULL;
    if (j->cached && p->half_bits <= 16) rounds = (uint16_t*)malloc(sizeof(uint16_t) * ((size_t)8 << p->half_bits));
    ock_stats mine = {0, 0, 0, 0};
    for (;;) {
        const uint32_t start = __atomic_fetch_add(&j->next, 16u, __ATOMIC_RELAXED);
        if (start >= j->count) break;
        const uint32_t end = start + 16u < j->count ? start + 16u : j->count;
        for (uint32_t i = start; i < end; ++i) {
            ock_stats st = {0, 0, 0, 0};
            j->flags[i] = (uint8_t)ock_trial_run(j->batch, j->first + i, p, bins, rounds, &st);
            if (j->probes) j->probes[i] = (uint32_t)st.probes;
            mine.probes += st.probes;
            mine.evals += st.evals;
            mine.stashed += st.stashed;
            mine.placed += st.placed;
        }
    }
    pthread_mutex_lock(&j->lock);
    j->totals.probes += mine.probes;
    j->totals.evals += mine.evals;
    j->totals.stashed += mine.stashed;
    j->totals.placed += mine.placed;
    pthread_mutex_unlock(&j->lock);
    free(bins);
    free(rounds);
    return NULL;
}

uint32_t ock_run_range(const ock_key* batch, uint32_t first, uint32_t count, const ock_params* p,
                       int threads, int cached, uint32_t* failed, uint32_t* probes, ock_stats* totals) {
    if (threads < 1) threads = 1;
    range_job j;
    memset(&j, 0, sizeof j);
    j.batch = batch;
    j.p = p;
    j.first = first;
    j.count = count;
    j.cached = cached;
    j.flags = (uint8_t*)calloc(count ? count : 1, 1);
    j.probes = probes;
    j.next = 0;
    pthread_mutex_init(&j.lock, NULL);
    pthread_t* tids = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 1; i < threads; ++i) pthread_create(&tids[i], NULL, range_worker, &j);
    range_worker(&j);
    for (int i = 1; i < threads; ++i) pthread_join(tids[i], NULL);
    free(tids);
    pthread_mutex_destroy(&j.lock);
    uint32_t n = 0;
    for (uint32_t i = 0; i < count; ++i)
        if (j.flags[i]) failed[n++] = first + i;
    free(j.flags);
    if (totals) *totals = j.totals;
    return n;
}
