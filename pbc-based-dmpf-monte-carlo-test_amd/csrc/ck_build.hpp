// ck_build.hpp -- the per-trial build steps shared by the single-shape kernels
// and the mixed-job engine: wave-uniform trial derivation, the wave's round
// cache, and the bucket-map row store.
#pragma once
#include "ck_device.hpp"
#include "cuckoo_mc.h"

namespace ck {

// Kernel-argument shape from validated params (host).
inline Shape make_shape(const ck_params& p) {
    Shape s{};
    s.k = p.k;
    s.pnum = p.pnum;
    s.buckets = p.buckets;
    s.kick = p.kick;
    s.stash = p.stash;
    s.half = p.half_bits;
    s.input_mode = p.input_mode;
    s.mode = p.mode;
    s.bucket_size = p.bucket_size;
    s.domain = p.domain;
    s.mask = p.mask;
    s.range_mask = p.range_mask;
    s.div_magic = p.bucket_size <= 1 ? 0 : (~0ull / p.bucket_size) + 1;
    uint32_t bb = 1;
    while (bb < 32 && (uint64_t(1) << bb) < p.buckets) ++bb;
    s.dbits = bb;
    return s;
}


// Stages T0 (2^shift copies) from `src` (constant or global memory) into LDS;
// returns the lane's view.
__device__ __forceinline__ LdsTab stage_tab_from(const uint32_t* src, uint32_t* lds, uint32_t shift) {
    const uint32_t n = 256u << shift;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = src[i >> shift];
    __syncthreads();
    return make_tab(lds, shift);
}

__device__ __forceinline__ void derive_uniform(const LdsTab& T, const RoundKeys& bk, const Shape& S, uint32_t offset,
                                               TrialSeed& ts) {
    derive_trial(T, bk, S, offset, ts);
#pragma unroll
    for (int i = 0; i < 44; ++i) ts.rk[i] = __builtin_amdgcn_readfirstlane(ts.rk[i]);
}

// The wave's round cache: cache[r<<half | right] = F_r(right) & mask.  For
// half >= 8 the rounds r are taken in turn: the x-free part of AES rounds 1
// and 2 is computed once per r (cache_round) and each lane encrypts four
// entries of round r at a time from there (126 of the block's 148 table
// lookups).  Smaller caches encrypt four entries at a time, one per quarter
// of the cache (entries is a multiple of 256 for half >= 5).
__device__ __forceinline__ void build_cache(const LdsTab& T, const TrialSeed& ts, const Shape& S, uint16_t* cache,
                                            uint32_t lane) {
    const uint32_t entries = 8u << S.half, rmask = (1u << S.half) - 1u;
    if (S.half >= 8) {
        const uint32_t per_round = 1u << S.half;
        for (uint32_t r = 0; r < 8; ++r) {
            const CacheRound cr = cache_round(T, ts.rk, r, S.half);
            uint16_t* row = cache + (r << S.half);
            for (uint32_t x0 = lane; x0 < per_round; x0 += 256) {
                const uint32_t x[4] = {x0, x0 + 64, x0 + 128, x0 + 192};
                uint32_t out[4];
                aes_cache_n<4>(T, ts.rk, cr, x, out);
#pragma unroll
                for (int i = 0; i < 4; ++i) row[x[i]] = static_cast<uint16_t>(out[i] & S.mask);
            }
        }
    } else if ((entries & 255u) == 0) {
        const uint32_t q = entries >> 2;
        for (uint32_t e = lane; e < q; e += 64) {
            uint32_t in[4][4], out[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t x = e + i * q;
                in[i][0] = 0x43455350u;
                in[i][1] = x >> S.half;
                in[i][2] = S.half;
                in[i][3] = x & rmask;
            }
            aes_last_n<4>(T, ts.rk, in, out);
#pragma unroll
            for (int i = 0; i < 4; ++i) cache[e + i * q] = static_cast<uint16_t>(out[i] & S.mask);
        }
    } else {
        for (uint32_t e = lane; e < entries; e += 64)
            cache[e] = static_cast<uint16_t>(aes_last(T, ts.rk, 0x43455350u, e >> S.half, S.half, e & rmask) & S.mask);
    }
    wave_lds_sync();
}

// The build's outputs -- bucket-map rows and the cleared table -- are written
// once and read by the chain kernel only after the whole build, so they are
// stored non-temporally: streamed out without displacing the lines that the
// chains of the other pass, running at the same time, are reading and writing
// back.  Measured at n = 2^14 with the build overlapping the chains: chain
// 99.3 -> 96.1 ms per 65,536-trial pass; a second, ordinary clear of the table
// costs 2.8 ms of chain time.

// The bucket map row of ordinal o: k entries, or one packed descriptor.
template <int K, class M>
__device__ __forceinline__ void store_candidates(M* m, const Shape& S, uint32_t o, const uint32_t (&c)[K]) {
    if constexpr (sizeof(M) == 8) {
        uint64_t d = 0;
#pragma unroll
        for (int w = 0; w < K; ++w) d |= uint64_t(c[w]) << (w * S.dbits);
        __builtin_nontemporal_store(d, &m[o]);
    } else {
#pragma unroll
        for (int w = 0; w < K; ++w) __builtin_nontemporal_store(static_cast<M>(c[w]), &m[o * K + w]);
    }
}

// Fills a table of `bytes` (a multiple of 16) with all-ones (empty), the
// wave's 64 lanes storing 16 bytes each per step.
__device__ __forceinline__ void clear_table(char* table, size_t bytes, uint32_t lane) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4* t = reinterpret_cast<u32x4*>(table);
    const u32x4 e = {EMPTY, EMPTY, EMPTY, EMPTY};
    for (size_t i = lane; i < bytes / 16; i += 64) __builtin_nontemporal_store(e, &t[i]);
}

}  // namespace ck
