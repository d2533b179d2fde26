// ck_kernels.hip -- gfx950 kernels of the cuckoo Monte Carlo path and the
// C-ABI entry points that launch them.
//
// Work per trial is split in two stages:
//   build : one wave per trial.  Derives the PRP key and walk, builds the
//           8*2^half round cache in LDS, evaluates bucket(o, w) for every
//           ordinal and candidate (the bucket map), clears the trial's table.
//           Full-lane integer work: AES T-table rounds and cached Feistel.
//   chain : one lane per trial (HBM tier).  Runs the serial insertion over the
//           HBM-resident map and table; one probe = one random table read (+
//           one map read when an evicted item moves).  Latency-bound, so the
//           launch keeps tens of thousands of chains in flight.
// Failures are appended with one ballot and one atomic per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "ck_device.hpp"
#include "ck_internal.hpp"

namespace ck {

__constant__ uint32_t c_te0[256];

// ------------------------------------------------------------------ kernels

// Stages the replicated T0 into LDS; returns the lane's view.
__device__ __forceinline__ LdsTab stage_tab(uint32_t* lds) {
    for (uint32_t i = threadIdx.x; i < 256u * TE_REPL; i += blockDim.x) lds[i] = c_te0[i / TE_REPL];
    __syncthreads();
    return LdsTab{lds, threadIdx.x & (TE_REPL - 1)};
}

// build: one wave per trial slot.  CACHED selects the round-cache form
// (half <= 12); otherwise every bucket evaluation runs AES directly.
template <int K, bool CACHED>
__global__ __launch_bounds__(256) void k_build(Shape S, RoundKeys bk, uint32_t first, uint32_t count,
                                               uint32_t* __restrict__ map, size_t map_stride,
                                               uint32_t* __restrict__ bins, size_t bins_stride,
                                               uint64_t* __restrict__ walks, uint64_t* __restrict__ rnds,
                                               uint32_t* __restrict__ key_out) {
    extern __shared__ uint32_t lds[];
    const LdsTab T = stage_tab(lds);
    const uint32_t waves = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t slot = blockIdx.x * waves + wave;
    if (slot >= count) return;  // wave-uniform; no block barrier follows

    TrialSeed ts;
    derive_trial(T, bk, S, first + slot, ts);
#pragma unroll
    for (int i = 0; i < 44; ++i) ts.rk[i] = __builtin_amdgcn_readfirstlane(ts.rk[i]);
    if (key_out && lane < 44) {
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 44; ++i) v = (lane == uint32_t(i)) ? ts.rk[i] : v;
        key_out[slot * 44 + lane] = v;
    }

    uint32_t* m = map + slot * map_stride;
    if constexpr (CACHED) {
        uint16_t* cache = reinterpret_cast<uint16_t*>(lds + 256 * TE_REPL) + (size_t(wave) << (S.half + 3));
        const uint32_t entries = 8u << S.half, rmask = (1u << S.half) - 1u;
        for (uint32_t e = lane; e < entries; e += 64)
            cache[e] = static_cast<uint16_t>(aes_last(T, ts.rk, 0x43455350u, e >> S.half, S.half, e & rmask) & S.mask);
        wave_lds_sync();
        for (uint32_t o = lane; o < S.pnum; o += 64) {
            const uint64_t el = element_of(S, ts.shift, ts.step, o);
#pragma unroll
            for (int w = 0; w < K; ++w) m[o * K + w] = feistel_cached(cache, S, el, K, w);
        }
    } else {
        for (uint32_t o = lane; o < S.pnum; o += 64) {
            const uint64_t el = element_of(S, ts.shift, ts.step, o);
            for (int w = 0; w < K; ++w) m[o * K + w] = feistel_direct(T, ts.rk, S, el, K, w);
        }
    }

    uint4* t4 = reinterpret_cast<uint4*>(bins + slot * bins_stride);
    const uint4 e4 = make_uint4(EMPTY, EMPTY, EMPTY, EMPTY);
    for (uint32_t i = lane; i < bins_stride / 4; i += 64) t4[i] = e4;
    if (lane == 0) {
        walks[slot] = ts.walk;
        rnds[slot] = ts.rnd;
    }
}

// BGH25 random-walk insertion (cuckoo_core.hpp:123-133) with the k/stash
// extension: fail (or stash the item in hand) once collisions exceed kick.
template <int K>
__device__ __forceinline__ bool chain_bgh(const Shape& S, const uint32_t* __restrict__ m, uint32_t* __restrict__ t,
                                          uint64_t walk, uint64_t rnd, uint32_t& probes) {
    uint32_t item = 0, cur = 0, coll = 0, stash_left = S.stash, pr = 0;
    bool failed = false;
    while (item < S.pnum) {
        if (rnd < K) rnd = splitmix(walk);
        const uint32_t which = take_choice<K>(rnd);
        const uint32_t b = m[cur * K + which];
        const uint32_t old = t[b];
        ++pr;
        if (old == EMPTY) {
            t[b] = cur | (which << 30);
            cur = ++item;
            coll = 0;
            continue;
        }
        if (++coll > S.kick) {
            if (stash_left) {
                --stash_left;
                cur = ++item;
                coll = 0;
                continue;
            }
            failed = true;
            break;
        }
        t[b] = cur | (which << 30);
        cur = old & ORD_MASK;
    }
    probes = pr;
    return failed;
}

template <int K>
__device__ __forceinline__ uint32_t pick(const uint32_t (&c)[K], uint32_t i) {
    uint32_t v = c[0];
#pragma unroll
    for (int w = 1; w < K; ++w) v = (i == uint32_t(w)) ? c[w] : v;
    return v;
}

// DRRT18 deterministic insertion (cuckoo_core.hpp:134-158) with the k/stash
// extension: try candidates 0..k-1, then evict from candidate 0; the displaced
// item cycles its stored choice to the first bucket != the one it left.
template <int K>
__device__ __forceinline__ bool chain_drrt(const Shape& S, const uint32_t* __restrict__ m, uint32_t* __restrict__ t,
                                           uint32_t& probes) {
    uint32_t pr = 0, stash_left = S.stash;
    bool failed = false;
    for (uint32_t item = 0; item < S.pnum && !failed; ++item) {
        uint32_t c[K];
#pragma unroll
        for (int w = 0; w < K; ++w) c[w] = m[item * K + w];
        bool placed = false;
#pragma unroll
        for (int w = 0; w < K; ++w) {
            if (placed) break;
            ++pr;
            if (t[c[w]] == EMPTY) {
                t[c[w]] = item | (uint32_t(w) << 30);
                placed = true;
            }
        }
        if (placed) continue;
        uint32_t cur = item, choice = 0, target = c[0], kicks = 0;
        for (;;) {
            const uint32_t disp = t[target];
            ++pr;
            if (disp == EMPTY) {
                t[target] = cur | (choice << 30);
                break;
            }
            if (kicks >= S.kick) {
                if (stash_left) {
                    --stash_left;
                    break;
                }
                failed = true;
                break;
            }
            t[target] = cur | (choice << 30);
            cur = disp & ORD_MASK;
            choice = disp >> 30;
            const uint32_t prev = target;
            uint32_t cc[K];
#pragma unroll
            for (int w = 0; w < K; ++w) cc[w] = m[cur * K + w];
            bool found = false;
#pragma unroll
            for (int s = 0; s < K; ++s) {
                if (found) break;
                choice = (choice + 1 == K) ? 0 : choice + 1;
                const uint32_t nb = pick<K>(cc, choice);
                if (nb != prev) {
                    target = nb;
                    found = true;
                }
            }
            if (!found) {
                if (stash_left) {
                    --stash_left;
                    break;
                }
                failed = true;
                break;
            }
            ++kicks;
        }
    }
    probes = pr;
    return failed;
}

// chain (HBM tier): one lane per trial slot.
template <int K, bool BGH>
__global__ __launch_bounds__(256) void k_chain(Shape S, uint32_t first, uint32_t count,
                                               const uint32_t* __restrict__ map, size_t map_stride,
                                               uint32_t* __restrict__ bins, size_t bins_stride,
                                               const uint64_t* __restrict__ walks, const uint64_t* __restrict__ rnds,
                                               uint32_t* __restrict__ fail_count, uint32_t* __restrict__ fail_offsets,
                                               uint32_t* __restrict__ probes_out,
                                               unsigned long long* __restrict__ probe_total) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    bool failed = false;
    uint32_t pr = 0;
    if (slot < count) {
        const uint32_t* m = map + slot * map_stride;
        uint32_t* t = bins + slot * bins_stride;
        if constexpr (BGH)
            failed = chain_bgh<K>(S, m, t, walks[slot], rnds[slot], pr);
        else
            failed = chain_drrt<K>(S, m, t, pr);
        if (probes_out) probes_out[slot] = pr;
    }
    append_failures(failed, first + slot, fail_count, fail_offsets);
    add_wave_total(pr, probe_total);
}

__global__ void k_zero(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// FIPS-197 C.1 through the device key schedule and T-table AES on every lane
// of one wave (aes_kat / cached_aes_kat, cuckoo_bench.cu:322-354).
__global__ __launch_bounds__(64) void k_aes_kat(uint32_t* out) {
    extern __shared__ uint32_t lds[];
    const LdsTab T = stage_tab(lds);
    const uint32_t key[4] = {0x00010203u, 0x04050607u, 0x08090a0bu, 0x0c0d0e0fu};
    uint32_t rk[44], o[4];
    expand_key_dev(T, key, rk);
    aes_enc(T, rk, 0x00112233u, 0x44556677u, 0x8899aabbu, 0xccddeeffu, o);
    for (int i = 0; i < 4; ++i) out[threadIdx.x * 4 + i] = o[i];
}

// ---------------------------------------------------------------- host side

#define CKH(expr)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(CK_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct Workspace {
    void* base = nullptr;
    size_t bytes = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    Workspace ws;
    uint32_t tier = CK_TIER_AUTO;
    uint32_t chunk_cap = 0;
    size_t budget = size_t(48) << 30;  // HBM budget for per-slot maps and tables
    size_t lds_max = 64 * 1024;        // dynamic LDS per block the device allows
};

static Shape make_shape(const ck_params& p) {
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
    return s;
}

static int check_params(const ck_params* p) {
    if (!p) return fail(CK_EINVAL, "null params");
    ck_params x;
    const int rc = derive_params(p->mode, p->pnum, p->ell_i, p->k, p->buckets, p->kick, p->stash, p->input_mode, &x);
    if (rc) return rc;
    if (std::memcmp(&x, p, sizeof x) != 0) return fail(CK_EINVAL, "params are not the output of ck_params_derive");
    return CK_OK;
}

// Largest round cache a build wave keeps in LDS next to the 32 KB T-table.
static constexpr uint32_t kMaxCacheHalf = 12;

struct BuildPlan {
    bool cached;
    uint32_t waves;  // trials per build block
    size_t lds;
};
static BuildPlan plan_build(const ck_params& p, size_t lds_max) {
    BuildPlan b{};
    const size_t tab = 256 * TE_REPL * sizeof(uint32_t);
    const size_t cache = (size_t(8) << p.half_bits) * sizeof(uint16_t);
    b.cached = p.half_bits <= kMaxCacheHalf && tab + cache <= lds_max;
    if (b.cached) {
        b.waves = static_cast<uint32_t>(std::min<size_t>(4, (lds_max - tab) / cache));
        b.lds = tab + b.waves * cache;
    } else {
        b.waves = 4;
        b.lds = tab;
    }
    return b;
}

struct Layout {
    size_t map_stride, bins_stride;  // u32 entries per slot
    size_t off_map, off_bins, off_walk, off_rnd, off_probes, off_fail, off_count, off_key, total;
};
static Layout layout_for(const ck_params& p, uint32_t chunk) {
    Layout L{};
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    L.map_stride = (size_t(p.pnum) * p.k + 63) & ~size_t(63);
    L.bins_stride = (size_t(p.buckets) + 63) & ~size_t(63);
    size_t o = 0;
    L.off_map = o;
    o = align(o + L.map_stride * chunk * 4);
    L.off_bins = o;
    o = align(o + L.bins_stride * chunk * 4);
    L.off_walk = o;
    o = align(o + size_t(chunk) * 8);
    L.off_rnd = o;
    o = align(o + size_t(chunk) * 8);
    L.off_probes = o;
    o = align(o + size_t(chunk) * 4);
    L.off_fail = o;
    o = align(o + size_t(chunk) * 4);
    L.off_count = o;
    o = align(o + 64);
    L.off_key = o;
    o = align(o + size_t(chunk) * 44 * 4);
    L.total = o;
    return L;
}

static uint32_t default_chunk(const Ctx& c, const ck_params& p, uint32_t count) {
    const size_t per_slot = ((size_t(p.pnum) * p.k + 63) & ~size_t(63)) * 4 + ((size_t(p.buckets) + 63) & ~size_t(63)) * 4 +
                            16 + 8 + 176;
    size_t fit = c.budget / per_slot;
    uint32_t target = 1u << 17;  // lanes in flight: 2048 waves, 8 per CU
    if (c.chunk_cap) target = c.chunk_cap;
    uint32_t chunk = static_cast<uint32_t>(std::min<size_t>({fit, size_t(target), size_t(count)}));
    return std::max<uint32_t>(chunk, 1);
}

static int ensure_ws(Ctx& c, size_t bytes) {
    if (c.ws.bytes >= bytes) return CK_OK;
    if (c.ws.base) CKH(hipFree(c.ws.base));
    c.ws.base = nullptr;
    c.ws.bytes = 0;
    CKH(hipMalloc(&c.ws.base, bytes));
    c.ws.bytes = bytes;
    return CK_OK;
}

template <int K>
static void launch_build(const BuildPlan& bp, const Shape& S, const RoundKeys& bk, uint32_t first, uint32_t count,
                         uint32_t* map, size_t map_stride, uint32_t* bins, size_t bins_stride, uint64_t* walks,
                         uint64_t* rnds, uint32_t* keys, hipStream_t st) {
    const uint32_t blocks = (count + bp.waves - 1) / bp.waves;
    if (bp.cached)
        hipLaunchKernelGGL((k_build<K, true>), dim3(blocks), dim3(64 * bp.waves), bp.lds, st, S, bk, first, count, map,
                           map_stride, bins, bins_stride, walks, rnds, keys);
    else
        hipLaunchKernelGGL((k_build<K, false>), dim3(blocks), dim3(64 * bp.waves), bp.lds, st, S, bk, first, count, map,
                           map_stride, bins, bins_stride, walks, rnds, keys);
}

template <int K>
static void launch_chain(bool bgh, const Shape& S, uint32_t first, uint32_t count, const uint32_t* map,
                         size_t map_stride, uint32_t* bins, size_t bins_stride, const uint64_t* walks,
                         const uint64_t* rnds, uint32_t* fc, uint32_t* fo, uint32_t* probes,
                         unsigned long long* ptotal, hipStream_t st) {
    const uint32_t blocks = (count + 255) / 256;
    if (bgh)
        hipLaunchKernelGGL((k_chain<K, true>), dim3(blocks), dim3(256), 0, st, S, first, count, map, map_stride, bins,
                           bins_stride, walks, rnds, fc, fo, probes, ptotal);
    else
        hipLaunchKernelGGL((k_chain<K, false>), dim3(blocks), dim3(256), 0, st, S, first, count, map, map_stride, bins,
                           bins_stride, walks, rnds, fc, fo, probes, ptotal);
}

static int set_lds_limits(size_t lds_max) {
    // Opt every build variant into the device's full LDS (needed above 64 KB).
    const int lim = static_cast<int>(lds_max);
#define OPT(K)                                                                                             \
    CKH(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_build<K, true>),                              \
                            hipFuncAttributeMaxDynamicSharedMemorySize, lim));                            \
    CKH(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_build<K, false>),                             \
                            hipFuncAttributeMaxDynamicSharedMemorySize, lim));
    OPT(2) OPT(3) OPT(4)
#undef OPT
    return CK_OK;
}

// One pass over trials [first, first+count) of the batch: build + chain per
// chunk; failing offsets are appended to out (unsorted).
static int run_hbm(Ctx& c, const ck_params& p, const uint32_t rk[44], uint32_t first, uint32_t count,
                   std::vector<uint32_t>& out, uint32_t* probes_host, uint32_t* tables_host, uint32_t* keys_host,
                   uint64_t* walks_host, uint64_t* rnds_host, uint32_t* map_host, ck_stats* st) {
    const Shape S = make_shape(p);
    RoundKeys bk;
    std::memcpy(bk.w, rk, sizeof bk.w);
    const uint32_t chunk = default_chunk(c, p, count);
    const Layout L = layout_for(p, chunk);
    int rc = ensure_ws(c, L.total);
    if (rc) return rc;
    char* base = static_cast<char*>(c.ws.base);
    auto* map = reinterpret_cast<uint32_t*>(base + L.off_map);
    auto* bins = reinterpret_cast<uint32_t*>(base + L.off_bins);
    auto* walks = reinterpret_cast<uint64_t*>(base + L.off_walk);
    auto* rnds = reinterpret_cast<uint64_t*>(base + L.off_rnd);
    auto* probes = reinterpret_cast<uint32_t*>(base + L.off_probes);
    auto* fo = reinterpret_cast<uint32_t*>(base + L.off_fail);
    auto* fc = reinterpret_cast<uint32_t*>(base + L.off_count);
    auto* keys = reinterpret_cast<uint32_t*>(base + L.off_key);
    const BuildPlan bp = plan_build(p, c.lds_max);
    const bool bgh = p.mode == CK_MODE_BGH25;
    const bool want_probes = probes_host != nullptr;
    std::vector<uint32_t> chunk_fail(chunk);
    if (st) {
        st->tier = CK_TIER_HBM;
        st->chunk = chunk;
    }
    for (uint32_t done = 0; done < count;) {
        const uint32_t n = std::min(chunk, count - done);
        const uint32_t at = first + done;
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, c.stream, fc, 16u);
        CKH(hipEventRecord(c.ev[0], c.stream));
        switch (p.k) {
            case 2: launch_build<2>(bp, S, bk, at, n, map, L.map_stride, bins, L.bins_stride, walks, rnds, keys_host ? keys : nullptr, c.stream); break;
            case 3: launch_build<3>(bp, S, bk, at, n, map, L.map_stride, bins, L.bins_stride, walks, rnds, keys_host ? keys : nullptr, c.stream); break;
            default: launch_build<4>(bp, S, bk, at, n, map, L.map_stride, bins, L.bins_stride, walks, rnds, keys_host ? keys : nullptr, c.stream); break;
        }
        CKH(hipGetLastError());
        CKH(hipEventRecord(c.ev[1], c.stream));
        auto* ptot = reinterpret_cast<unsigned long long*>(fc + 2);
        uint32_t* pr_out = want_probes ? probes : nullptr;
        switch (p.k) {
            case 2: launch_chain<2>(bgh, S, at, n, map, L.map_stride, bins, L.bins_stride, walks, rnds, fc, fo, pr_out, ptot, c.stream); break;
            case 3: launch_chain<3>(bgh, S, at, n, map, L.map_stride, bins, L.bins_stride, walks, rnds, fc, fo, pr_out, ptot, c.stream); break;
            default: launch_chain<4>(bgh, S, at, n, map, L.map_stride, bins, L.bins_stride, walks, rnds, fc, fo, pr_out, ptot, c.stream); break;
        }
        CKH(hipGetLastError());
        CKH(hipEventRecord(c.ev[2], c.stream));
        uint32_t counters[4] = {0, 0, 0, 0};  // failures, pad, probe total (u64)
        CKH(hipMemcpyAsync(counters, fc, sizeof counters, hipMemcpyDeviceToHost, c.stream));
        CKH(hipStreamSynchronize(c.stream));
        const uint32_t nf = counters[0];
        if (st) st->probes += (uint64_t(counters[3]) << 32) | counters[2];
        if (nf > n) return fail(CK_ERANGE, "device failure count out of range");
        if (nf) {
            CKH(hipMemcpy(chunk_fail.data(), fo, size_t(nf) * 4, hipMemcpyDeviceToHost));
            out.insert(out.end(), chunk_fail.begin(), chunk_fail.begin() + nf);
        }
        if (probes_host) CKH(hipMemcpy(probes_host + done, probes, size_t(n) * 4, hipMemcpyDeviceToHost));
        if (tables_host)
            CKH(hipMemcpy2D(tables_host + size_t(done) * p.buckets, size_t(p.buckets) * 4, bins, L.bins_stride * 4,
                            size_t(p.buckets) * 4, n, hipMemcpyDeviceToHost));
        if (keys_host) CKH(hipMemcpy(keys_host + size_t(done) * 44, keys, size_t(n) * 44 * 4, hipMemcpyDeviceToHost));
        if (walks_host) CKH(hipMemcpy(walks_host + done, walks, size_t(n) * 8, hipMemcpyDeviceToHost));
        if (rnds_host) CKH(hipMemcpy(rnds_host + done, rnds, size_t(n) * 8, hipMemcpyDeviceToHost));
        if (map_host)
            CKH(hipMemcpy2D(map_host + size_t(done) * p.pnum * p.k, size_t(p.pnum) * p.k * 4, map, L.map_stride * 4,
                            size_t(p.pnum) * p.k * 4, n, hipMemcpyDeviceToHost));
        if (st) {
            float b = 0, ch = 0;
            CKH(hipEventElapsedTime(&b, c.ev[0], c.ev[1]));
            CKH(hipEventElapsedTime(&ch, c.ev[1], c.ev[2]));
            st->build_ms += b;
            st->chain_ms += ch;
            st->launches += 3;
        }
        done += n;
    }
    return CK_OK;
}

}  // namespace ck

using namespace ck;

struct ck_ctx {
    Ctx c;
};

extern "C" {

int ck_ctx_create(int device, ck_ctx** out) {
    if (!out) return fail(CK_EINVAL, "null output");
    *out = nullptr;
    int n = 0;
    CKH(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(CK_EINVAL, "no such HIP device");
    CKH(hipSetDevice(device));
    auto* ctx = new ck_ctx();
    ctx->c.device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&ctx->c.ev[i]);
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_te0), te0(), 256 * sizeof(uint32_t));
    if (e != hipSuccess) {
        ck_ctx_destroy(ctx);
        return fail(CK_EDEVICE, std::string("context setup: ") + hipGetErrorString(e));
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeSharedMemPerBlockOptin, device) == hipSuccess && lds > 0)
        ctx->c.lds_max = static_cast<size_t>(lds);
    else if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        ctx->c.lds_max = static_cast<size_t>(lds);
    const int rc = set_lds_limits(ctx->c.lds_max);
    if (rc) {
        ck_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return CK_OK;
}

void ck_ctx_destroy(ck_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.ws.base) (void)hipFree(ctx->c.ws.base);
    for (auto& ev : ctx->c.ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

int ck_ctx_set_tier(ck_ctx* ctx, uint32_t tier) {
    if (!ctx || tier > CK_TIER_LDS) return fail(CK_EINVAL, "bad tier");
    ctx->c.tier = tier;
    return CK_OK;
}

int ck_ctx_set_chunk(ck_ctx* ctx, uint32_t trials) {
    if (!ctx) return fail(CK_EINVAL, "null ctx");
    ctx->c.chunk_cap = trials;
    return CK_OK;
}

int ck_aes_self_test(ck_ctx* ctx) {
    if (!ctx) return fail(CK_EINVAL, "null ctx");
    CKH(hipSetDevice(ctx->c.device));
    uint32_t* d = nullptr;
    CKH(hipMalloc(&d, 64 * 4 * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_aes_kat, dim3(1), dim3(64), 256 * TE_REPL * 4, ctx->c.stream, d);
    std::vector<uint32_t> h(64 * 4);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, h.size() * 4, hipMemcpyDeviceToHost, ctx->c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->c.stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CK_EDEVICE, std::string("AES self test: ") + hipGetErrorString(e));
    const uint32_t want[4] = {0x69c4e0d8u, 0x6a7b0430u, 0xd8cdb780u, 0x70b4c55au};
    for (int lane = 0; lane < 64; ++lane)
        if (std::memcmp(h.data() + lane * 4, want, sizeof want))
            return fail(CK_ESELFTEST, "device AES-128 FIPS-197 known-answer test failed");
    return CK_OK;
}

static int launch_common(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t first, uint32_t count,
                         uint32_t* failed_offsets, uint32_t* n_failed, uint32_t* probes, uint32_t* tables,
                         ck_stats* stats) {
    if (!ctx || !rk || !n_failed || (count && !failed_offsets)) return fail(CK_EINVAL, "null argument");
    int rc = check_params(p);
    if (rc) return rc;
    if (uint64_t(first) + count > (uint64_t(1) << 32)) return fail(CK_EINVAL, "trial range exceeds 32-bit offsets");
    CKH(hipSetDevice(ctx->c.device));
    if (stats) std::memset(stats, 0, sizeof *stats);
    std::vector<uint32_t> fails;
    if (count) {
        rc = run_hbm(ctx->c, *p, rk, first, count, fails, probes, tables, nullptr, nullptr, nullptr, nullptr, stats);
        if (rc) return rc;
    }
    std::sort(fails.begin(), fails.end());
    if (std::adjacent_find(fails.begin(), fails.end()) != fails.end())
        return fail(CK_ERANGE, "duplicate failure offsets from the device");
    std::copy(fails.begin(), fails.end(), failed_offsets);
    *n_failed = static_cast<uint32_t>(fails.size());
    if (stats) {
        stats->trials = count;
        stats->failed = fails.size();
    }
    return CK_OK;
}

int ck_launch(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t first, uint32_t count,
              uint32_t* failed_offsets, uint32_t* n_failed, ck_stats* stats) {
    return launch_common(ctx, p, rk, first, count, failed_offsets, n_failed, nullptr, nullptr, stats);
}

int ck_launch_debug(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t first, uint32_t count,
                    uint32_t* failed_offsets, uint32_t* n_failed, uint32_t* probes, uint32_t* tables,
                    ck_stats* stats) {
    return launch_common(ctx, p, rk, first, count, failed_offsets, n_failed, probes, tables, stats);
}

int ck_trial_map(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t offset, uint32_t key_words[44],
                 uint64_t* walk, uint64_t* randomness, uint32_t* map) {
    if (!ctx || !rk || !key_words || !walk || !randomness || !map) return fail(CK_EINVAL, "null argument");
    int rc = check_params(p);
    if (rc) return rc;
    CKH(hipSetDevice(ctx->c.device));
    std::vector<uint32_t> fails;
    const uint32_t saved = ctx->c.chunk_cap;
    ctx->c.chunk_cap = 1;
    rc = run_hbm(ctx->c, *p, rk, offset, 1, fails, nullptr, nullptr, key_words, walk, randomness, map, nullptr);
    ctx->c.chunk_cap = saved;
    return rc;
}

}  // extern "C"
