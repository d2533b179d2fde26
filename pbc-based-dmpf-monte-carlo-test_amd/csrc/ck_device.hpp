// ck_device.hpp -- device building blocks shared by the HIP kernels:
// AES-128 over an LDS T-table, the AES key schedule, SplitMix64, the Feistel
// PRP bucket function (round-cache and direct forms), exact division by the
// bucket size, and the ballot/popcount failure append.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ck {

// Register budgets that let a build block and chain waves share a CU (the HBM
// tier overlaps the build of pass i+1 with the chains of pass i).  A SIMD has
// 512 VGPRs per lane; a 16-wave build block puts 4 waves on each SIMD, so at
// 96 VGPRs a wave it leaves 128 for chain waves (20-38 VGPRs each).  At 128
// VGPRs the build block fills the register file, no chain wave can share its
// CU, and the overlap collapses (measured: 1.7x slower passes at n = 2^14).
#define CK_BUILD_REGS __attribute__((amdgpu_waves_per_eu(5)))

constexpr uint32_t EMPTY = 0xffffffffu;
constexpr uint32_t ORD_MASK = 0x3fffffffu;  // table entry = ordinal | choice << 30

// Kernel-argument form of ck_params plus the host-precomputed divider.
struct Shape {
    uint32_t k, pnum, buckets, kick, stash, half, input_mode, mode;
    uint64_t bucket_size, domain, mask, range_mask;
    uint64_t div_magic;  // ceil(2^64 / bucket_size) for the 32-bit quotient; 0 when bucket_size == 1
    uint32_t dbits;      // bits per bucket index in a packed candidate descriptor
    uint32_t pad;
};
struct RoundKeys {
    uint32_t w[44];
};

__device__ __forceinline__ uint32_t rotr(uint32_t x, uint32_t s) { return __builtin_amdgcn_alignbit(x, x, s); }
// a ^ b ^ c in one VALU op (v_bitop3_b32, truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}


// T0 in LDS, replicated 2^shift times: entry x of copy c sits at
// (x << shift) | c.  With shift 5 a lane L reading copy L%32 touches bank L%32
// whatever the byte index (no bank conflicts); shift 0 is the plain 1 KB table.
// Lookups form the LDS byte address (x << (shift + 2)) + (base + copy * 4)
// with one shift-add and read through an LDS-space pointer, so a lookup costs
// a byte extract, that shift-add and the read.
typedef const __attribute__((address_space(3))) uint32_t* lds_u32_cptr;
struct LdsTab {
    uint32_t lane4b;  // LDS byte address of this lane's copy of entry 0
    uint32_t shift2;  // shift + 2
    __device__ __forceinline__ uint32_t t(uint32_t x) const {
        return *reinterpret_cast<lds_u32_cptr>(static_cast<uintptr_t>((x << shift2) + lane4b));
    }
    __device__ __forceinline__ uint32_t s(uint32_t x) const { return (t(x) >> 16) & 255u; }
};
__device__ __forceinline__ LdsTab make_tab(const uint32_t* lds, uint32_t shift) {
    const uint32_t base = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((lds_u32_cptr)lds));
    return LdsTab{base + ((threadIdx.x & ((1u << shift) - 1u)) << 2), shift + 2};
}

// FIPS-197 AES-128 encryption in the T-table form (Daemen & Rijmen): rounds
// 1..9 are T0[b0]^ror8(T0[b1])^ror16(T0[b2])^ror24(T0[b3])^rk, the last round
// substitutes bytes only.  Input and output are big-endian column words.
__device__ __forceinline__ void aes_enc(const LdsTab& T, const uint32_t* rk, uint32_t a, uint32_t b, uint32_t c,
                                        uint32_t d, uint32_t out[4]) {
    uint32_t s0 = a ^ rk[0], s1 = b ^ rk[1], s2 = c ^ rk[2], s3 = d ^ rk[3];
#pragma unroll
    for (int r = 1; r < 10; ++r) {
        const uint32_t n0 = xor3(xor3(T.t(s0 >> 24), rotr(T.t((s1 >> 16) & 255), 8), rotr(T.t((s2 >> 8) & 255), 16)),
                                 rotr(T.t(s3 & 255), 24), rk[4 * r]);
        const uint32_t n1 = xor3(xor3(T.t(s1 >> 24), rotr(T.t((s2 >> 16) & 255), 8), rotr(T.t((s3 >> 8) & 255), 16)),
                                 rotr(T.t(s0 & 255), 24), rk[4 * r + 1]);
        const uint32_t n2 = xor3(xor3(T.t(s2 >> 24), rotr(T.t((s3 >> 16) & 255), 8), rotr(T.t((s0 >> 8) & 255), 16)),
                                 rotr(T.t(s1 & 255), 24), rk[4 * r + 2]);
        const uint32_t n3 = xor3(xor3(T.t(s3 >> 24), rotr(T.t((s0 >> 16) & 255), 8), rotr(T.t((s1 >> 8) & 255), 16)),
                                 rotr(T.t(s2 & 255), 24), rk[4 * r + 3]);
        s0 = n0;
        s1 = n1;
        s2 = n2;
        s3 = n3;
    }
    out[0] = ((T.s(s0 >> 24) << 24) | (T.s((s1 >> 16) & 255) << 16) | (T.s((s2 >> 8) & 255) << 8) | T.s(s3 & 255)) ^
             rk[40];
    out[1] = ((T.s(s1 >> 24) << 24) | (T.s((s2 >> 16) & 255) << 16) | (T.s((s3 >> 8) & 255) << 8) | T.s(s0 & 255)) ^
             rk[41];
    out[2] = ((T.s(s2 >> 24) << 24) | (T.s((s3 >> 16) & 255) << 16) | (T.s((s0 >> 8) & 255) << 8) | T.s(s1 & 255)) ^
             rk[42];
    out[3] = ((T.s(s3 >> 24) << 24) | (T.s((s0 >> 16) & 255) << 16) | (T.s((s1 >> 8) & 255) << 8) | T.s(s2 & 255)) ^
             rk[43];
}

// Only the last output word (the Feistel round function uses word 3).
__device__ __forceinline__ uint32_t aes_last(const LdsTab& T, const uint32_t* rk, uint32_t a, uint32_t b, uint32_t c,
                                             uint32_t d) {
    uint32_t o[4];
    aes_enc(T, rk, a, b, c, d, o);
    return o[3];
}

// FIPS-197 5.2 key expansion from four big-endian words (the device derives
// each trial's PRP key this way, cuckoo_bench.cu:136-149).
__device__ __forceinline__ void expand_key_dev(const LdsTab& T, const uint32_t w[4], uint32_t rk[44]) {
    rk[0] = w[0];
    rk[1] = w[1];
    rk[2] = w[2];
    rk[3] = w[3];
    uint32_t rcon = 1;
#pragma unroll
    for (int i = 4; i < 44; ++i) {
        uint32_t v = rk[i - 1];
        if ((i & 3) == 0) {
            v = ((T.s((v >> 16) & 255) << 24) | (T.s((v >> 8) & 255) << 16) | (T.s(v & 255) << 8) | T.s(v >> 24)) ^
                (rcon << 24);
            rcon = ((rcon << 1) ^ ((rcon & 0x80) ? 0x1b : 0)) & 0xff;
        }
        rk[i] = rk[i - 4] ^ v;
    }
}

// SplitMix64 (Steele, Lea & Flood), the walk generator (cuckoo_core.hpp:110-114).
__device__ __forceinline__ uint64_t splitmix(uint64_t& state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Exact y / d for 32-bit y with M = ceil(2^64/d) (Lemire, Kaser & Kurz,
// "Faster remainder by direct computation", 2019): floor(M*y / 2^64).
__device__ __forceinline__ uint32_t div_bucket(const Shape& S, uint32_t y) {
    if (S.div_magic == 0) return y;  // bucket_size == 1
    const uint32_t lo = static_cast<uint32_t>(S.div_magic), hi = static_cast<uint32_t>(S.div_magic >> 32);
    const uint64_t mid = uint64_t(hi) * y + __umulhi(lo, y);
    return static_cast<uint32_t>(mid >> 32);
}

// Per-trial derivation (cuckoo_bench.cu:222-238): PRP key = schedule of
// AES_batch("KEY1",0,offset,"PRP1"); (walk, affine shift) from
// AES_batch("WALK",0,offset,"RNG1"); step and the first walk draw by SplitMix64.
struct TrialSeed {
    uint32_t rk[44];
    uint64_t walk, rnd, shift, step;
};
__device__ __forceinline__ void derive_trial(const LdsTab& T, const RoundKeys& bk, const Shape& S, uint32_t offset,
                                             TrialSeed& ts) {
    uint32_t w[4];
    aes_enc(T, bk.w, 0x4b455931u, 0u, offset, 0x50525031u, w);
    expand_key_dev(T, w, ts.rk);
    aes_enc(T, bk.w, 0x57414c4bu, 0u, offset, 0x524e4731u, w);
    ts.walk = (uint64_t(w[0]) << 32) | w[1];
    ts.shift = ((uint64_t(w[2]) << 32) | w[3]) & S.range_mask;
    uint64_t input_state = ts.shift ^ 0xd1b54a32d192ed03ull;
    ts.step = (splitmix(input_state) | 1ull) & S.range_mask;
    ts.rnd = splitmix(ts.walk);
}

__device__ __forceinline__ uint64_t element_of(const Shape& S, uint64_t shift, uint64_t step, uint32_t ordinal) {
    return S.input_mode == 0 ? uint64_t(ordinal) : (shift + uint64_t(ordinal) * step) & S.range_mask;
}

// N independent AES-128 blocks advanced round by round together (more loads in
// flight per lane); returns the last output word of each block.
template <int N>
__device__ __forceinline__ void aes_last_n(const LdsTab& T, const uint32_t* rk, const uint32_t (&in)[N][4],
                                           uint32_t (&out)[N]) {
    uint32_t s[N][4];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) s[i][j] = in[i][j] ^ rk[j];
#pragma unroll
    for (int r = 1; r < 10; ++r) {
        uint32_t n[N][4];
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                n[i][j] = xor3(xor3(T.t(s[i][j] >> 24), rotr(T.t((s[i][(j + 1) & 3] >> 16) & 255), 8),
                                    rotr(T.t((s[i][(j + 2) & 3] >> 8) & 255), 16)),
                               rotr(T.t(s[i][(j + 3) & 3] & 255), 24), rk[4 * r + j]);
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) s[i][j] = n[i][j];
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
        out[i] = ((T.s(s[i][3] >> 24) << 24) | (T.s((s[i][0] >> 16) & 255) << 16) | (T.s((s[i][1] >> 8) & 255) << 8) |
                  T.s(s[i][2] & 255)) ^
                 rk[43];
}

// Round-cache blocks share all but the last word: the block of entry
// (r, x) is ("CESP", r, half, x) with x < 2^half <= 2^16, so after the first
// AddRoundKey only bytes 2 and 3 of column 3 depend on x.  Round 1 then
// depends on x in two T-table terms only (column 0 through byte 3, column 1
// through byte 2; columns 2 and 3 are fixed), and round 2 in the eight terms
// that read columns 0 and 1.  CacheRound holds, for one (key, r, half), the x-free
// part of each: round-1 columns 0 and 1 without their x term, and the round-2
// columns without their terms from columns 0 and 1 (round key included).
struct CacheRound {
    uint32_t c0, c1, d0, d1, d2, d3;
    uint32_t k2, k3;  // bytes 2 and 3 of rk[3], the key bytes that x meets
};
__device__ __forceinline__ CacheRound cache_round(const LdsTab& T, const uint32_t* rk, uint32_t r, uint32_t half) {
    const uint32_t s0 = 0x43455350u ^ rk[0], s1 = r ^ rk[1], s2 = half ^ rk[2], s3 = rk[3];
    CacheRound c;
    c.c0 = xor3(T.t(s0 >> 24), rotr(T.t((s1 >> 16) & 255), 8), rotr(T.t((s2 >> 8) & 255), 16)) ^ rk[4];
    c.c1 = xor3(T.t(s1 >> 24), rotr(T.t((s2 >> 16) & 255), 8), rotr(T.t(s0 & 255), 24)) ^ rk[5];
    const uint32_t n2 = xor3(xor3(T.t(s2 >> 24), rotr(T.t((s3 >> 16) & 255), 8), rotr(T.t((s0 >> 8) & 255), 16)),
                             rotr(T.t(s1 & 255), 24), rk[6]);
    const uint32_t n3 = xor3(xor3(T.t(s3 >> 24), rotr(T.t((s0 >> 16) & 255), 8), rotr(T.t((s1 >> 8) & 255), 16)),
                             rotr(T.t(s2 & 255), 24), rk[7]);
    c.d0 = xor3(rotr(T.t((n2 >> 8) & 255), 16), rotr(T.t(n3 & 255), 24), rk[8]);
    c.d1 = xor3(rotr(T.t((n2 >> 16) & 255), 8), rotr(T.t((n3 >> 8) & 255), 16), rk[9]);
    c.d2 = xor3(T.t(n2 >> 24), rotr(T.t((n3 >> 16) & 255), 8), rk[10]);
    c.d3 = xor3(T.t(n3 >> 24), rotr(T.t(n2 & 255), 24), rk[11]);
    c.k2 = (s3 >> 8) & 255;
    c.k3 = s3 & 255;
    return c;
}

// The last output word of N round-cache blocks (r, x[i]) sharing r: rounds 1
// and 2 from CacheRound, then rounds 3..10 as in aes_last_n.  Equal to
// aes_last(T, rk, "CESP", r, half, x[i]) for x[i] < 2^16.
template <int N>
__device__ __forceinline__ void aes_cache_n(const LdsTab& T, const uint32_t* rk, const CacheRound& c,
                                            const uint32_t (&x)[N], uint32_t (&out)[N]) {
    uint32_t s[N][4];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const uint32_t n0 = c.c0 ^ rotr(T.t(c.k3 ^ (x[i] & 255)), 24);
        const uint32_t n1 = c.c1 ^ rotr(T.t(c.k2 ^ ((x[i] >> 8) & 255)), 16);
        s[i][0] = xor3(c.d0, T.t(n0 >> 24), rotr(T.t((n1 >> 16) & 255), 8));
        s[i][1] = xor3(c.d1, T.t(n1 >> 24), rotr(T.t(n0 & 255), 24));
        s[i][2] = xor3(c.d2, rotr(T.t((n0 >> 8) & 255), 16), rotr(T.t(n1 & 255), 24));
        s[i][3] = xor3(c.d3, rotr(T.t((n0 >> 16) & 255), 8), rotr(T.t((n1 >> 8) & 255), 16));
    }
#pragma unroll
    for (int r = 3; r < 10; ++r) {
        uint32_t n[N][4];
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                n[i][j] = xor3(xor3(T.t(s[i][j] >> 24), rotr(T.t((s[i][(j + 1) & 3] >> 16) & 255), 8),
                                    rotr(T.t((s[i][(j + 2) & 3] >> 8) & 255), 16)),
                               rotr(T.t(s[i][(j + 3) & 3] & 255), 24), rk[4 * r + j]);
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) s[i][j] = n[i][j];
    }
#pragma unroll
    for (int i = 0; i < N; ++i)
        out[i] = ((T.s(s[i][3] >> 24) << 24) | (T.s((s[i][0] >> 16) & 255) << 16) | (T.s((s[i][1] >> 8) & 255) << 8) |
                  T.s(s[i][2] & 255)) ^
                 rk[43];
}

// The k candidates of one element through the cached Feistel network in
// lockstep: each pass applies the 8 rounds to every candidate (the k dependent
// LDS chains overlap) and keeps the result only for candidates still outside
// the domain, so each candidate ends after exactly the passes its own cycle
// walk takes.  Same values as k evaluations of cuckoo::CachedBuckets
// (include/cuckoo_trial.hpp) over the same cache.
template <int K>
__device__ __forceinline__ void feistel_cached_k(const uint16_t* cache, const Shape& S, uint64_t elem,
                                                 uint32_t (&out)[K]) {
    const uint32_t h = S.half, m = static_cast<uint32_t>(S.mask);
    uint32_t y[K];
#pragma unroll
    for (int w = 0; w < K; ++w) y[w] = static_cast<uint32_t>(elem * K + w);
    uint32_t done = 0;
    do {
        uint32_t L[K], R[K];
#pragma unroll
        for (int w = 0; w < K; ++w) {
            L[w] = y[w] >> h;
            R[w] = y[w] & m;
        }
#pragma unroll
        for (uint32_t r = 0; r < 8; ++r)
#pragma unroll
            for (int w = 0; w < K; ++w) {
                const uint32_t nx = L[w] ^ cache[(r << h) + R[w]];
                L[w] = R[w];
                R[w] = nx & m;
            }
#pragma unroll
        for (int w = 0; w < K; ++w) {
            if (!((done >> w) & 1u)) {
                y[w] = (L[w] << h) | R[w];
                if (uint64_t(y[w]) < S.domain) done |= 1u << w;
            }
        }
    } while (done != (1u << K) - 1u);
#pragma unroll
    for (int w = 0; w < K; ++w) out[w] = div_bucket(S, y[w]);
}

// Feistel PRP evaluating AES directly (any half): the round function is the
// last word of AES_key("CESP", round, half, right) (cuckoo_bench.cu:161-184).
__device__ __noinline__ uint32_t feistel_direct(const LdsTab& T, const uint32_t* rk, const Shape& S, uint64_t elem,
                                                uint32_t k, uint32_t which) {
    uint64_t y = elem * k + which;
    do {
        uint64_t L = y >> S.half, R = y & S.mask;
#pragma unroll 1
        for (uint32_t r = 0; r < 8; ++r) {
            const uint32_t f = aes_last(T, rk, 0x43455350u, r, S.half, static_cast<uint32_t>(R));
            const uint64_t nx = (L ^ (uint64_t(f) & S.mask)) & S.mask;
            L = R;
            R = nx;
        }
        y = (L << S.half) | R;
    } while (y >= S.domain);
    if (S.domain <= 0xffffffffull) return div_bucket(S, static_cast<uint32_t>(y));
    return static_cast<uint32_t>(y / S.bucket_size);
}

// which = r % K, r /= K for the 64-bit walk word (cuckoo_core.hpp:126-127).
template <int K>
__device__ __forceinline__ uint32_t take_choice(uint64_t& r) {
    if constexpr (K == 2) {
        const uint32_t w = static_cast<uint32_t>(r) & 1u;
        r >>= 1;
        return w;
    } else if constexpr (K == 4) {
        const uint32_t w = static_cast<uint32_t>(r) & 3u;
        r >>= 2;
        return w;
    } else {
        const uint64_t q = __umul64hi(r, 0xAAAAAAAAAAAAAAABull) >> 1;  // floor(r/3)
        const uint32_t w = static_cast<uint32_t>(r - q * 3);
        r = q;
        return w;
    }
}

// Wave-local LDS visibility: every lane's earlier LDS writes are seen by the
// wave's later LDS reads.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The chain kernels' block order.  The dispatcher hands workgroup i to XCD
// i mod 8, and each XCD translates addresses through its own TLBs.  xcd_block
// turns the dispatch index b of n blocks into a block index such that each XCD
// runs one contiguous range of blocks (XCDs 0 .. n mod 8 - 1 get one block more
// than the rest), so the trial slots of one XCD -- and their tables, laid out
// in slot order -- span one eighth of the pass's memory instead of all of it.
// It is a bijection on [0, n).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
    const uint32_t q = n >> 3, r = n & 7u, x = b & 7u, j = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
}

// Failure counting: one 64-bit ballot per wave, one atomic per wave with a
// failure, each failing lane writes at base + its rank among failing lanes.
__device__ __forceinline__ void append_failures(bool failed, uint32_t offset, uint32_t* __restrict__ count,
                                                uint32_t* __restrict__ offsets) {
    const uint64_t m = __ballot(failed);
    if (m == 0) return;
    const int leader = __builtin_ctzll(m);
    const uint32_t lane = __lane_id();
    uint32_t base = 0;
    if (lane == static_cast<uint32_t>(leader)) base = atomicAdd(count, static_cast<uint32_t>(__popcll(m)));
    base = __shfl(base, leader);
    if (failed) {
        const uint32_t rank =
            __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
        offsets[base + rank] = offset;
    }
}

// Sum of a per-lane count over the wave, added with one 64-bit atomic.
__device__ __forceinline__ void add_wave_total(uint32_t v, unsigned long long* __restrict__ total) {
    uint64_t s = v;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d);
    if (__lane_id() == 0 && s) atomicAdd(total, static_cast<unsigned long long>(s));
}

}  // namespace ck
