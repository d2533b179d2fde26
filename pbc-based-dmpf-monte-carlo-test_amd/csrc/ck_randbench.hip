// ck_randbench.hip -- the rate the memory system gives the insertion chain's
// access pattern, with nothing else on the dependent path.
//
// Every lane owns a table of 8-byte slots (by default 192 KB, the descriptor
// table of one n = 2^14 trial) filled with random words.  A step reads one
// slot and, in the read+write form, writes it back changed; the next slot is
// mulhi32(low word of the value read ^ ctr, slots) with ctr advanced by a
// constant every step.  So the dependent path is load -> one XOR -> one
// 32-bit multiply-high -> address -> load, which is no longer than the chain's
// (load -> shift/mask of a bucket field -> address -> load), and ctr keeps the
// walk from settling into a short cycle of cached slots (a fixed random map of
// 24,576 slots would cycle within ~200 steps).  Tables are 256-byte aligned
// and disjoint, as the chain's are.
//
//   ck_randbench [--search K] [table_bytes] [steps] [lanes ...]
//
// For each lane count it prints one line per form:
//   read_write lanes=L table_bytes=B steps=S ms=T steps_per_s=R
//   read_only  ...
//   read_only_w4 ...   (four independent walks per lane, S steps in all)
// The rate belongs to the memory the tables land in (the chain's passes run
// 15 % faster in some allocations than in others), so with --search K it then
// allocates K buffers for 65,536 lanes, all held at once, runs both forms in
// each and prints
//   search_read_write buffer=i lanes=65536 ... steps_per_s=R
//   search_read_only  ...
//   search_read_only_w4 ...
// bench.py takes the best read-only rate over all lines (either form) as the
// ceiling that it sets the chain's DRAM request rate against.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

constexpr uint32_t kGolden32 = 0x9E3779B9u;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Fills words [0, n) with mix64(index + salt): random table contents.
__global__ __launch_bounds__(256) void k_fill(uint64_t* __restrict__ buf, uint64_t n, uint64_t salt) {
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        buf[i] = mix64(i + salt);
}

// W independent walks per lane over the lane's table, steps / W steps each
// (W = 1: the chain's single dependent path; W = 4: four loads in flight per
// lane, the read-only rate with latency taken out of the way).
template <bool WRITE, int W>
__global__ __launch_bounds__(256) void k_walk(uint64_t* __restrict__ buf, uint64_t pitch, uint32_t slots,
                                              uint32_t lanes, uint32_t steps, unsigned long long* __restrict__ sink) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= lanes) return;
    uint64_t* t = buf + uint64_t(lane) * pitch;
    uint32_t ctr[W], i[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
        ctr[w] = (lane * W + w) * kGolden32;
        i[w] = __umulhi(static_cast<uint32_t>(mix64(uint64_t(lane) * W + w)), slots);
    }
    uint64_t acc = 0;
    for (uint32_t s = 0; s < steps / W; ++s) {
        uint64_t v[W];
#pragma unroll
        for (int w = 0; w < W; ++w) v[w] = t[i[w]];
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if constexpr (WRITE) t[i[w]] = v[w] ^ ctr[w];
            acc += v[w];
            i[w] = __umulhi(static_cast<uint32_t>(v[w]) ^ ctr[w], slots);
            ctr[w] += kGolden32;
        }
    }
    if (acc == 0x5eedull) sink[0] = acc;  // keeps the loads live
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

// The three forms at `lanes` lanes over buf; prints one line per form with `tag`.
static int measure(const char* tag, int buffer, uint64_t* buf, uint64_t pitch, uint32_t slots, uint32_t lanes,
                   uint64_t table, uint32_t steps, unsigned long long* sink, hipEvent_t a, hipEvent_t b) {
    const dim3 grid((lanes + 255) / 256), block(256);
    static const char* const kForm[3] = {"read_write", "read_only", "read_only_w4"};
    for (int form = 0; form < 3; ++form) {
        float ms = 0;
        for (int rep = 0; rep < 2; ++rep) {  // the first run warms the TLB and clocks
            CHECK(hipEventRecord(a));
            if (form == 0)
                hipLaunchKernelGGL((k_walk<true, 1>), grid, block, 0, 0, buf, pitch, slots, lanes, steps, sink);
            else if (form == 1)
                hipLaunchKernelGGL((k_walk<false, 1>), grid, block, 0, 0, buf, pitch, slots, lanes, steps, sink);
            else
                hipLaunchKernelGGL((k_walk<false, 4>), grid, block, 0, 0, buf, pitch, slots, lanes, steps, sink);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(b));
            CHECK(hipEventSynchronize(b));
            CHECK(hipEventElapsedTime(&ms, a, b));
        }
        const uint32_t done = form == 2 ? steps / 4 * 4 : steps;
        const double rate = double(lanes) * done / (ms / 1e3);
        if (buffer >= 0)
            std::printf("%s%s buffer=%d lanes=%u table_bytes=%llu steps=%u ms=%.3f steps_per_s=%.6e\n", tag,
                        kForm[form], buffer, lanes, (unsigned long long)table, done, ms, rate);
        else
            std::printf("%s%s lanes=%u table_bytes=%llu steps=%u ms=%.3f steps_per_s=%.6e\n", tag, kForm[form],
                        lanes, (unsigned long long)table, done, ms, rate);
        std::fflush(stdout);
    }
    return 0;
}

int main(int argc, char** argv) {
    int arg = 1;
    uint32_t search = 0;
    if (argc > 2 && std::string(argv[1]) == "--search") {
        search = uint32_t(std::strtoul(argv[2], nullptr, 10));
        arg = 3;
    }
    const uint64_t table = argc > arg ? std::strtoull(argv[arg], nullptr, 10) : 196608ull;
    const uint32_t steps = argc > arg + 1 ? uint32_t(std::strtoul(argv[arg + 1], nullptr, 10)) : 16384u;
    std::vector<uint32_t> lane_counts;
    for (int a = arg + 2; a < argc; ++a) lane_counts.push_back(uint32_t(std::strtoul(argv[a], nullptr, 10)));
    if (lane_counts.empty()) lane_counts = {32768u, 65536u, 131072u, 262144u};
    uint32_t max_lanes = 0;
    for (uint32_t l : lane_counts) max_lanes = l > max_lanes ? l : max_lanes;
    if (table < 8 || (table & 7) || table / 8 > 0xffffffffull || steps == 0 || max_lanes == 0 ||
        max_lanes > (1u << 22) || search > 64) {
        std::fprintf(stderr,
                     "usage: ck_randbench [--search K <= 64] [table_bytes, multiple of 8] [steps] [lanes <= 2^22 ...]\n");
        return 2;
    }
    const uint32_t slots = static_cast<uint32_t>(table / 8);
    const uint64_t pitch = ((table + 255) & ~uint64_t(255)) / 8;  // words per lane table
    const uint64_t words = pitch * max_lanes;
    size_t free_b = 0, total_b = 0;
    CHECK(hipMemGetInfo(&free_b, &total_b));
    if (words * 8 > free_b / 10 * 9) {
        std::fprintf(stderr, "tables need %.1f GiB, more than 90%% of the free %.1f GiB\n", words * 8 / 1073741824.0,
                     free_b / 1073741824.0);
        return 2;
    }
    uint64_t* buf = nullptr;
    unsigned long long* sink = nullptr;
    CHECK(hipMalloc(&buf, words * 8));
    CHECK(hipMalloc(&sink, sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_fill, dim3(8192), dim3(256), 0, 0, buf, words, 0x243F6A8885A308D3ull);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    for (uint32_t lanes : lane_counts)
        if (measure("", -1, buf, pitch, slots, lanes, table, steps, sink, a, b)) return 1;
    CHECK(hipFree(buf));
    if (search) {
        // Candidate buffers for 65,536 lanes, all held at once (as the library's
        // pass-buffer search holds its candidates), as many as fit 90 % of the
        // free memory.
        const uint32_t lanes = 65536;
        const uint64_t bytes = pitch * lanes * 8;
        CHECK(hipMemGetInfo(&free_b, &total_b));
        const uint64_t fit = (free_b / 10 * 9) / bytes;
        std::vector<uint64_t*> cand;
        for (uint32_t i = 0; i < search && i < fit; ++i) {
            uint64_t* c = nullptr;
            CHECK(hipMalloc(&c, bytes));
            cand.push_back(c);
        }
        for (size_t i = 0; i < cand.size(); ++i) {
            hipLaunchKernelGGL(k_fill, dim3(8192), dim3(256), 0, 0, cand[i], pitch * lanes, 0x243F6A8885A308D3ull);
            CHECK(hipGetLastError());
            if (measure("search_", int(i), cand[i], pitch, slots, lanes, table, steps, sink, a, b)) return 1;
        }
        for (uint64_t* c : cand) CHECK(hipFree(c));
    }
    CHECK(hipFree(sink));
    return 0;
}
