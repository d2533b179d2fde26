// ck_randbench.hip -- the achievable rate of the insertion chain's memory
// access pattern on this GPU, without the chain's logic: every lane walks its
// own table (by default 65536 lanes x 192 KB, the tables of one headline
// pass), and each step reads one random 8-byte slot of it and writes the slot
// back changed, the next slot depending on the value read (a dependent chain,
// as in cuckoo insertion).  Prints steps/s for the read+write form and for a
// read-only form; bench.py divides the chain's probes/s by the first.
//
//   ck_randbench [lanes] [table_bytes] [steps]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

__device__ __forceinline__ uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <bool WRITE>
__global__ __launch_bounds__(256) void k_walk(uint64_t* __restrict__ buf, uint64_t slots_per_lane, uint32_t lanes,
                                              uint32_t steps, unsigned long long* __restrict__ sink) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= lanes) return;
    uint64_t* t = buf + uint64_t(lane) * slots_per_lane;
    uint64_t state = mix(lane + 1);
    uint64_t acc = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        const uint64_t i = state % slots_per_lane;
        const uint64_t v = t[i];
        if constexpr (WRITE) t[i] = v + state;
        acc += v;
        state = mix(state ^ v);
    }
    if (acc == 0x5eed) sink[0] = acc;  // keeps the loads live
}

#define CHECK(x)                                                                              \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) {                                                               \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                      \
            return 1;                                                                         \
        }                                                                                     \
    } while (0)

}  // namespace

int main(int argc, char** argv) {
    const uint32_t lanes = argc > 1 ? uint32_t(std::strtoul(argv[1], nullptr, 10)) : 65536u;
    const uint64_t table = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 196608ull;
    const uint32_t steps = argc > 3 ? uint32_t(std::strtoul(argv[3], nullptr, 10)) : 32768u;
    if (lanes == 0 || table < 8 || steps == 0 || (table & 7)) {
        std::fprintf(stderr, "usage: ck_randbench [lanes] [table_bytes, multiple of 8] [steps]\n");
        return 2;
    }
    const uint64_t slots = table / 8;
    uint64_t* buf = nullptr;
    unsigned long long* sink = nullptr;
    CHECK(hipMalloc(&buf, uint64_t(lanes) * table));
    CHECK(hipMalloc(&sink, sizeof(unsigned long long)));
    CHECK(hipMemset(buf, 0, uint64_t(lanes) * table));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    const dim3 grid((lanes + 255) / 256), block(256);
    for (int form = 0; form < 2; ++form) {
        for (int rep = 0; rep < 2; ++rep) {  // the first run warms the TLB and clocks
            CHECK(hipEventRecord(a));
            if (form == 0)
                hipLaunchKernelGGL(k_walk<true>, grid, block, 0, 0, buf, slots, lanes, steps, sink);
            else
                hipLaunchKernelGGL(k_walk<false>, grid, block, 0, 0, buf, slots, lanes, steps, sink);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(b));
            CHECK(hipEventSynchronize(b));
        }
        float ms = 0;
        CHECK(hipEventElapsedTime(&ms, a, b));
        const double rate = double(lanes) * steps / (ms / 1e3);
        std::printf("%s lanes=%u table_bytes=%llu steps=%u ms=%.3f steps_per_s=%.6e\n",
                    form == 0 ? "read_write" : "read_only", lanes, (unsigned long long)table, steps, ms, rate);
    }
    CHECK(hipFree(buf));
    CHECK(hipFree(sink));
    return 0;
}
