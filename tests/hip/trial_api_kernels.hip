// trial_api_kernels.hip -- test-only HIP library: the trial API of
// include/cuckoo_trial.hpp instantiated in a gfx950 kernel with caller-side
// Bins and Buckets, one lane per trial -- what a caller that brings its own
// table and bucket function writes.  tests/test_trial_api.py drives it and
// compares outcomes and final tables with the reference's.
//
//   combination 0: StridedBins (u32) + CachedBuckets over the trial's round cache
//   combination 1: StridedBins (u32) + PrecomputedBuckets over its bucket map
//   combination 2: SharedBins (u16) + PrecomputedBuckets
//   combination 3: SharedBins (u16) + CachedBuckets
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cuckoo_trial.hpp"

namespace {

struct TrialShape {
    uint32_t bgh, pnum, kick, buckets, half_bits, sequential;
    uint64_t domain, bucket_size, mask, range_mask;
};
struct TrialIn {
    uint64_t walk, randomness, shift, step;
};

template <bool BGH, class Bins, class Buckets>
__device__ bool one_trial(const TrialShape& s, Bins bins, Buckets buckets, const TrialIn& in, uint32_t& probes) {
    return cuckoo::insert_core_k<BGH, 3>(s.pnum, s.kick, 0u, bins, buckets, in.walk, in.randomness, probes);
}

template <class Bins, class Buckets>
__device__ bool run(const TrialShape& s, Bins bins, Buckets buckets, const TrialIn& in, uint32_t& probes) {
    return s.bgh ? one_trial<true>(s, bins, buckets, in, probes) : one_trial<false>(s, bins, buckets, in, probes);
}

__global__ __launch_bounds__(64) void k_trials(uint32_t combo, TrialShape s, uint32_t count,
                                               const TrialIn* __restrict__ in, const uint16_t* __restrict__ rounds,
                                               const uint32_t* __restrict__ maps, uint32_t* __restrict__ t32,
                                               uint16_t* __restrict__ t16, uint32_t* __restrict__ failed,
                                               uint32_t* __restrict__ probes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const cuckoo::CachedBuckets<uint16_t, 3> cached{rounds + (size_t(i) << (s.half_bits + 3)), 1, s.domain,
                                                    s.bucket_size, s.mask, s.range_mask, in[i].shift, in[i].step,
                                                    s.half_bits, s.sequential != 0};
    const cuckoo::PrecomputedBuckets<uint32_t, 3> mapped{maps + size_t(i) * s.pnum * 3};
    const cuckoo::StridedBins wide{t32 + size_t(i) * s.buckets, 1};
    const cuckoo::SharedBins narrow{t16 + size_t(i) * s.buckets, 1};
    uint32_t pr = 0;
    bool f = false;
    switch (combo) {
        case 0: f = run(s, wide, cached, in[i], pr); break;
        case 1: f = run(s, wide, mapped, in[i], pr); break;
        case 2: f = run(s, narrow, mapped, in[i], pr); break;
        default: f = run(s, narrow, cached, in[i], pr); break;
    }
    failed[i] = f ? 1u : 0u;
    probes[i] = pr;
}

}  // namespace

// Runs `count` trials of one shape under `combo`; tables_out receives
// count x buckets u32 entries in the reference's form (ordinal | choice << 16,
// empty all-ones).  Returns 0, or the HIP error code.
extern "C" int trial_api_device_run(uint32_t combo, uint32_t bgh, uint32_t pnum, uint32_t kick, uint32_t buckets,
                                    uint32_t half_bits, uint32_t sequential, uint64_t domain, uint64_t bucket_size,
                                    uint64_t mask, uint64_t range_mask, uint32_t count, const uint64_t* inputs,
                                    const uint16_t* rounds, const uint32_t* maps, uint32_t* tables_out,
                                    uint32_t* failed_out, uint32_t* probes_out) {
    if (combo > 3 || half_bits > 16 || count == 0 || (combo >= 2 && pnum >= (1u << 14))) return -1;
    const TrialShape s{bgh, pnum, kick, buckets, half_bits, sequential, domain, bucket_size, mask, range_mask};
    const size_t rounds_n = size_t(count) << (half_bits + 3), maps_n = size_t(count) * pnum * 3,
                 tab_n = size_t(count) * buckets;
    TrialIn* d_in = nullptr;
    uint16_t *d_rounds = nullptr, *d_t16 = nullptr;
    uint32_t *d_maps = nullptr, *d_t32 = nullptr, *d_failed = nullptr, *d_probes = nullptr;
    hipError_t e = hipMalloc(&d_in, count * sizeof(TrialIn));
    if (e == hipSuccess) e = hipMalloc(&d_rounds, rounds_n * 2);
    if (e == hipSuccess) e = hipMalloc(&d_maps, maps_n * 4);
    if (e == hipSuccess) e = hipMalloc(&d_t32, tab_n * 4);
    if (e == hipSuccess) e = hipMalloc(&d_t16, tab_n * 2);
    if (e == hipSuccess) e = hipMalloc(&d_failed, count * 4);
    if (e == hipSuccess) e = hipMalloc(&d_probes, count * 4);
    if (e == hipSuccess) e = hipMemcpy(d_in, inputs, count * sizeof(TrialIn), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_rounds, rounds, rounds_n * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_maps, maps, maps_n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d_t32, 0xff, tab_n * 4);
    if (e == hipSuccess) e = hipMemset(d_t16, 0xff, tab_n * 2);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_trials, dim3((count + 63) / 64), dim3(64), 0, 0, combo, s, count, d_in, d_rounds,
                           d_maps, d_t32, d_t16, d_failed, d_probes);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && combo >= 2) {
        // Widen the u16 entries (ordinal | choice << 14) to the u32 form on the host.
        uint16_t* h16 = new uint16_t[tab_n];
        e = hipMemcpy(h16, d_t16, tab_n * 2, hipMemcpyDeviceToHost);
        for (size_t j = 0; e == hipSuccess && j < tab_n; ++j)
            tables_out[j] = h16[j] == 0xffffu ? 0xffffffffu
                                              : (cuckoo::SharedBins::ordinal(h16[j]) |
                                                 (cuckoo::SharedBins::choice(h16[j]) << 16));
        delete[] h16;
    } else if (e == hipSuccess) {
        e = hipMemcpy(tables_out, d_t32, tab_n * 4, hipMemcpyDeviceToHost);
    }
    if (e == hipSuccess) e = hipMemcpy(failed_out, d_failed, count * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(probes_out, d_probes, count * 4, hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    (void)hipFree(d_rounds);
    (void)hipFree(d_maps);
    (void)hipFree(d_t32);
    (void)hipFree(d_t16);
    (void)hipFree(d_failed);
    (void)hipFree(d_probes);
    return e == hipSuccess ? 0 : static_cast<int>(e);
}
