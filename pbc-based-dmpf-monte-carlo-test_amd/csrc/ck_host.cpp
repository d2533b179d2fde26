// ck_host.cpp -- host half of the C ABI: experiment shapes, the AES S-box and
// T-table, the host AES-128 and the per-batch key derivation.
#include "ck_internal.hpp"

#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>

namespace ck {

namespace {
uint8_t gmul(uint8_t a, uint8_t b) {
    uint8_t r = 0;
    for (; b; b >>= 1) {
        if (b & 1) r ^= a;
        a = static_cast<uint8_t>((a << 1) ^ ((a & 0x80) ? 0x1b : 0));
    }
    return r;
}

struct Tables {
    uint8_t sbox[256];
    uint32_t te0[256];
    Tables() {
        // FIPS-197 5.1.1: inverse in GF(2^8), then the affine map with 0x63.
        for (unsigned x = 0; x < 256; ++x) {
            uint8_t inv = 0;
            for (unsigned y = 1; y < 256 && x; ++y)
                if (gmul(static_cast<uint8_t>(x), static_cast<uint8_t>(y)) == 1) {
                    inv = static_cast<uint8_t>(y);
                    break;
                }
            uint8_t s = inv;
            for (int r = 1; r <= 4; ++r) s ^= static_cast<uint8_t>((inv << r) | (inv >> (8 - r)));
            sbox[x] = static_cast<uint8_t>(s ^ 0x63);
        }
        // Column of MixColumns applied to S(x): bytes {2s, s, s, 3s}, MSB first.
        for (unsigned x = 0; x < 256; ++x) {
            const uint32_t s = sbox[x], s2 = gmul(static_cast<uint8_t>(s), 2), s3 = s2 ^ s;
            te0[x] = (s2 << 24) | (s << 16) | (s << 8) | s3;
        }
    }
};
const Tables& tables() {
    static const Tables t;
    return t;
}
thread_local std::string g_error;
}  // namespace

const uint8_t* sbox() { return tables().sbox; }
const uint32_t* te0() { return tables().te0; }

void set_error(const std::string& e) { g_error = e; }
const char* last_error() { return g_error.c_str(); }

static uint32_t sub_word(uint32_t x) {
    const uint8_t* s = sbox();
    return (uint32_t(s[x >> 24]) << 24) | (uint32_t(s[(x >> 16) & 255]) << 16) |
           (uint32_t(s[(x >> 8) & 255]) << 8) | s[x & 255];
}

void expand_words(const uint32_t w4[4], uint32_t rk[44]) {
    for (int i = 0; i < 4; ++i) rk[i] = w4[i];
    uint32_t rcon = 1;
    for (int i = 4; i < 44; ++i) {
        uint32_t v = rk[i - 1];
        if ((i & 3) == 0) {
            v = sub_word((v << 8) | (v >> 24)) ^ (rcon << 24);
            rcon = ((rcon << 1) ^ ((rcon & 0x80) ? 0x1b : 0)) & 0xff;
        }
        rk[i] = rk[i - 4] ^ v;
    }
}

// expand_key(lo, hi) (cuckoo_bench.cu:86-99): key bytes are lo then hi, each
// little-endian; the schedule reads them as big-endian words.
void expand_key(uint64_t lo, uint64_t hi, uint32_t rk[44]) {
    uint32_t w[4];
    for (int i = 0; i < 2; ++i) {
        const uint64_t half = i ? hi : lo;
        w[2 * i] = __builtin_bswap32(static_cast<uint32_t>(half));
        w[2 * i + 1] = __builtin_bswap32(static_cast<uint32_t>(half >> 32));
    }
    expand_words(w, rk);
}

static inline uint32_t ror(uint32_t x, int s) { return (x >> s) | (x << (32 - s)); }

// T-table AES-128 on big-endian words (the device uses the same formulation).
void host_aes(const uint32_t rk[44], const uint32_t in[4], uint32_t out[4]) {
    const uint32_t* t = te0();
    const uint8_t* sb = sbox();
    uint32_t s[4];
    for (int j = 0; j < 4; ++j) s[j] = in[j] ^ rk[j];
    for (int r = 1; r < 10; ++r) {
        uint32_t n[4];
        for (int j = 0; j < 4; ++j)
            n[j] = t[s[j] >> 24] ^ ror(t[(s[(j + 1) & 3] >> 16) & 255], 8) ^
                   ror(t[(s[(j + 2) & 3] >> 8) & 255], 16) ^ ror(t[s[(j + 3) & 3] & 255], 24) ^ rk[4 * r + j];
        std::memcpy(s, n, sizeof s);
    }
    for (int j = 0; j < 4; ++j)
        out[j] = ((uint32_t(sb[s[j] >> 24]) << 24) | (uint32_t(sb[(s[(j + 1) & 3] >> 16) & 255]) << 16) |
                  (uint32_t(sb[(s[(j + 2) & 3] >> 8) & 255]) << 8) | sb[s[(j + 3) & 3] & 255]) ^
                 rk[40 + j];
}

// batch_key (cuckoo_bench.cu:393-404): AES(master, "BATC", id_hi, id_lo, "PRG1")
// under master = expand_key(seed, seed ^ golden ratio), folded with the seed.
void batch_key(uint64_t seed, uint64_t id, uint32_t rk[44]) {
    uint32_t master[44];
    expand_key(seed, seed ^ 0x9e3779b97f4a7c15ULL, master);
    const uint32_t in[4] = {0x42415443u, static_cast<uint32_t>(id >> 32), static_cast<uint32_t>(id), 0x50524731u};
    uint32_t x[4];
    host_aes(master, in, x);
    const uint64_t prg_lo = (uint64_t(x[0]) << 32) | x[1];
    const uint64_t prg_hi = (uint64_t(x[2]) << 32) | x[3];
    expand_key(seed ^ prg_lo, prg_hi ^ (seed ^ 0x243f6a8885a308d3ULL), rk);
}

static unsigned bit_width(uint64_t v) {
    unsigned b = 0;
    for (; v; v >>= 1) ++b;
    return b;
}

// original_log2_u64 (cuckoo_bench.cu:488-498): float-rounded atanh series.
float original_log2(uint64_t n) {
    if (n == 0) return -std::numeric_limits<float>::infinity();
    const unsigned e = bit_width(n) - 1;
    const double x = double(n) / double(1ULL << e);
    if (x >= 2.0) return float(e + 1);
    const double z = (x - 1.0) / (x + 1.0), z2 = z * z;
    double term = z, sum = 0.0;
    for (int k = 1; k <= 23; k += 2) {
        sum += term / k;
        term *= z2;
    }
    return float(double(e) + 2.0 * sum * 1.4426950408889634074);
}

int derive_params(uint32_t mode, uint32_t pnum, uint32_t ell_i, uint32_t k, uint32_t buckets, uint32_t kick,
                  uint32_t stash, uint32_t input_mode, ck_params* x) {
    std::memset(x, 0, sizeof *x);
    if (mode != CK_MODE_BGH25 && mode != CK_MODE_DRRT18) return fail(CK_EINVAL, "mode must be 1 (bgh25) or 2 (drrt18)");
    if (k < 2 || k > 4) return fail(CK_EINVAL, "k must be 2..4");
    if (pnum == 0 || pnum > (1u << 30)) return fail(CK_EINVAL, "pnum must be 1..2^30");
    if (ell_i > 61) return fail(CK_EINVAL, "ell_i must be <= 61");
    if (ell_i < 32 && (uint64_t(1) << ell_i) < pnum) return fail(CK_EINVAL, "2^ell_i must be >= pnum");
    if (input_mode > 1) return fail(CK_EINVAL, "input_mode must be 0 (sequential) or 1 (affine)");
    uint32_t b, kk;
    if (mode == CK_MODE_BGH25) {  // cuckoo_bench.cu:502-505
        const uint64_t overhead = pnum < 30 ? 100 : 50;
        b = static_cast<uint32_t>(uint64_t(pnum) * (100 + overhead) / 100);
        kk = 3 * pnum;
    } else {  // cuckoo_bench.cu:507-511, float rounding kept
        const float lf = original_log2(pnum);
        const float f = static_cast<float>(double(pnum) * (double(170.0f + lf) / 123.5));
        b = static_cast<uint32_t>(std::ceil(double(f)));
        kk = 100;
    }
    if (buckets != CK_DERIVED) b = buckets;
    if (kick != CK_DERIVED) kk = kick;
    if (b == 0) return fail(CK_EINVAL, "buckets must be positive");
    x->mode = mode;
    x->k = k;
    x->pnum = pnum;
    x->buckets = b;
    x->kick = kk;
    x->stash = stash;
    x->ell_i = ell_i;
    x->input_mode = input_mode;
    const uint64_t source = (uint64_t(1) << ell_i) * k;
    x->bucket_size = source / b + (source % b != 0);
    x->ell_d = static_cast<uint32_t>(std::ceil(original_log2(x->bucket_size)));
    if (b > std::numeric_limits<uint64_t>::max() / x->bucket_size) return fail(CK_EINVAL, "PRP domain overflow");
    x->domain = uint64_t(b) * x->bucket_size;
    unsigned bits = std::max(bit_width(x->domain - 1), 2u);
    if (bits & 1u) ++bits;
    x->half_bits = bits / 2;
    x->mask = x->half_bits >= 64 ? ~0ULL : ((1ULL << x->half_bits) - 1);
    x->range_mask = (uint64_t(1) << ell_i) - 1;
    return CK_OK;
}

int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

int check_params(const ck_params* p) {
    if (!p) return fail(CK_EINVAL, "null params");
    ck_params x;
    const int rc = derive_params(p->mode, p->pnum, p->ell_i, p->k, p->buckets, p->kick, p->stash, p->input_mode, &x);
    if (rc) return rc;
    if (std::memcmp(&x, p, sizeof x) != 0) return fail(CK_EINVAL, "params are not the output of ck_params_derive");
    return CK_OK;
}

}  // namespace ck

extern "C" {
int ck_params_derive(uint32_t mode, uint32_t pnum, uint32_t ell_i, uint32_t k, uint32_t buckets, uint32_t kick,
                     uint32_t stash, uint32_t input_mode, ck_params* out) {
    if (!out) return ck::fail(CK_EINVAL, "null output");
    return ck::derive_params(mode, pnum, ell_i, k, buckets, kick, stash, input_mode, out);
}
void ck_batch_key(uint64_t seed, uint64_t batch_id, uint32_t round_keys[44]) { ck::batch_key(seed, batch_id, round_keys); }
void ck_host_aes(const uint32_t round_keys[44], const uint32_t in[4], uint32_t out[4]) {
    ck::host_aes(round_keys, in, out);
}
const char* ck_last_error(void) { return ck::last_error(); }

int ck_trial_inputs_of(const ck_params* p, const uint32_t batch_round_keys[44], uint32_t offset,
                       ck_trial_inputs* out) {
    const int rc = ck::check_params(p);
    if (rc) return rc;
    if (!batch_round_keys || !out) return ck::fail(CK_EINVAL, "null argument");
    const auto splitmix = [](uint64_t& state) {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    uint32_t w[4];
    const uint32_t key_block[4] = {0x4b455931u, 0u, offset, 0x50525031u};  // "KEY1", 0, offset, "PRP1"
    ck::host_aes(batch_round_keys, key_block, w);
    ck::expand_words(w, out->prp_key);
    const uint32_t walk_block[4] = {0x57414c4bu, 0u, offset, 0x524e4731u};  // "WALK", 0, offset, "RNG1"
    ck::host_aes(batch_round_keys, walk_block, w);
    out->walk = (uint64_t(w[0]) << 32) | w[1];
    out->shift = ((uint64_t(w[2]) << 32) | w[3]) & p->range_mask;
    uint64_t state = out->shift ^ 0xd1b54a32d192ed03ull;
    out->step = (splitmix(state) | 1ull) & p->range_mask;
    out->randomness = splitmix(out->walk);
    return CK_OK;
}

int ck_round_cache(const ck_params* p, const uint32_t prp_key[44], uint16_t* rounds) {
    const int rc = ck::check_params(p);
    if (rc) return rc;
    if (!prp_key || !rounds) return ck::fail(CK_EINVAL, "null argument");
    if (p->half_bits > 16) return ck::fail(CK_EINVAL, "a round cache needs half_bits <= 16");
    const uint32_t half = p->half_bits, n = 1u << half;
    for (uint32_t r = 0; r < 8; ++r)
        for (uint32_t right = 0; right < n; ++right) {
            const uint32_t in[4] = {0x43455350u, r, half, right};  // "CESP", round, half, right
            uint32_t o[4];
            ck::host_aes(prp_key, in, o);
            rounds[(r << half) | right] = static_cast<uint16_t>(o[3] & p->mask);
        }
    return CK_OK;
}
}
