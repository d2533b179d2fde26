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
//   read_write_w4 ...  (the same with the write-back)
//   read_write_7of8 ... (one walk per lane, seven steps in eight write back)
// The rate belongs to the memory the tables land in (the chain's passes run
// 15 % faster in some allocations than in others), so with --search K it then
// allocates K buffers for 65,536 lanes, all held at once, runs both forms in
// each and prints
//   search_read_write buffer=i lanes=65536 ... steps_per_s=R
//   search_read_only  ...
//   search_read_only_w4 ...
// With --regions it then builds one range for 65,536 lanes out of three
// pieces from the device's three memory regions (as the library places its
// cuckoo tables) and prints
//   regions_found=R piece_bytes=P
//   regions_read_write lanes=65536 ... steps_per_s=R    (and the other forms)
// bench.py takes the best read-only rate over all lines (either form) as the
// ceiling that it sets the chain's DRAM request rate against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ck_regions.hpp"

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

// W independent walks per lane over the lane's table, steps / W steps each;
// WRITE 0: read only, 1: every step writes its slot back changed, 2: seven
// steps in eight do
// (W = 1: the chain's single dependent path; W = 4: four loads in flight per
// lane, the read-only rate with latency taken out of the way).
template <int WRITE, int W>
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
            if constexpr (WRITE == 1) t[i[w]] = v[w] ^ ctr[w];
            if constexpr (WRITE == 2)  // seven steps in eight write back (the chain writes back 0.86 per probe)
                if (v[w] & 7) t[i[w]] = v[w] ^ ctr[w];
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

// Two pieces, the first `half` lanes over tables of `slots` words at a, the
// rest at b: the read + write walk, whose rate tells whether a and b lie in
// one memory region.
__global__ __launch_bounds__(256) void k_pair(uint64_t* __restrict__ a, uint64_t* __restrict__ b, uint32_t half,
                                              uint32_t slots, uint32_t steps, unsigned long long* __restrict__ sink) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= 2 * half) return;
    uint64_t* t = (lane < half ? a : b) + uint64_t(lane < half ? lane : lane - half) * slots;
    uint32_t ctr = lane * kGolden32;
    uint32_t i = __umulhi(static_cast<uint32_t>(mix64(lane)), slots);
    uint64_t acc = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        const uint64_t v = t[i];
        t[i] = v ^ ctr;
        acc += v;
        i = __umulhi(static_cast<uint32_t>(v) ^ ctr, slots);
        ctr += kGolden32;
    }
    if (acc == 0x5eedull) sink[0] = acc;
}

}  // namespace

// A range of `bytes` made of three virtual-memory pieces from three memory
// regions (two where only two show), found as the library's table placement
// finds them: pieces allocated in batches, each paired with piece 0 and, if
// apart from it, with the piece most clearly apart (pair walk with 24 KB
// tables, 768 MiB of each piece), until every region holds one.
struct Spread {
    char* va = nullptr;
    size_t piece = 0;
    int regions = 0;
    hipMemGenericAllocationHandle_t h[3] = {};
};
static int spread_range(size_t bytes, unsigned long long* sink, hipEvent_t a, hipEvent_t b, Spread* out) {
    const size_t align = size_t(2) << 20;
    const size_t P = ((bytes + 2) / 3 + align - 1) & ~(align - 1);
    if (P < (size_t(1) << 30)) {
        std::fprintf(stderr, "--regions needs pieces of 1 GiB or more\n");
        return 2;
    }
    size_t free_b = 0, total_b = 0;
    CHECK(hipMemGetInfo(&free_b, &total_b));
    const size_t maxp = std::min<size_t>(96, free_b / 10 * 9 / P);
    hipMemAllocationProp prop{};
    prop.type = hipMemAllocationTypePinned;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = 0;
    hipMemAccessDesc acc{};
    acc.location = prop.location;
    acc.flags = hipMemAccessFlagsProtReadWrite;
    char* pool = nullptr;
    CHECK(hipMemAddressReserve(reinterpret_cast<void**>(&pool), maxp * P, 0, nullptr, 0));
    std::vector<hipMemGenericAllocationHandle_t> h;
    std::vector<float> t0, t1;
    std::vector<int> region;
    size_t ref1 = 0;
    auto pair = [&](size_t x, size_t y, float* ms) {
        for (int rep = 0; rep < 2; ++rep) {
            CHECK(hipEventRecord(a));
            hipLaunchKernelGGL(k_pair, dim3(256), dim3(256), 0, 0, reinterpret_cast<uint64_t*>(pool + x * P),
                               reinterpret_cast<uint64_t*>(pool + y * P), 32768u, 3072u, 1024u, sink);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(b));
            CHECK(hipEventSynchronize(b));
            CHECK(hipEventElapsedTime(ms, a, b));
        }
        return 0;
    };
    int pick[3] = {0, -1, -1};
    while (h.size() < maxp && pick[2] < 0) {
        const size_t from = h.size();
        for (int k = 0; k < 4 && h.size() < maxp; ++k) {
            hipMemGenericAllocationHandle_t x{};
            if (hipMemCreate(&x, P, &prop, 0) != hipSuccess) break;
            CHECK(hipMemMap(pool + h.size() * P, P, 0, x, 0));
            CHECK(hipMemSetAccess(pool + h.size() * P, P, &acc, 1));
            h.push_back(x);
        }
        if (h.size() == from) break;
        t0.resize(h.size(), 0.f);
        t1.resize(h.size(), 0.f);
        region.resize(h.size(), 0);
        for (size_t j = std::max<size_t>(from, 1); j < h.size(); ++j)
            if (pair(0, j, &t0[j])) return 1;
        float f0 = 0, s0 = 0, f1 = 0, s1 = 0;
        std::vector<size_t> all;
        for (size_t j = 1; j < h.size(); ++j) all.push_back(j);
        if (!ck::split_modes(t0, all, 1.12f, &f0, &s0)) continue;
        std::vector<size_t> others;
        for (size_t j = 1; j < h.size(); ++j) {
            region[j] = t0[j] >= s0 / 1.06f ? 0 : t0[j] <= f0 * 1.06f ? 1 : -1;
            if (region[j] == 1) others.push_back(j);
        }
        if (!ref1 && !others.empty()) {
            ref1 = others[0];
            for (size_t j : others)
                if (t0[j] < t0[ref1]) ref1 = j;
        }
        std::vector<size_t> others1;
        for (size_t j : others)
            if (j != ref1) {
                if (t1[j] == 0 && pair(ref1, j, &t1[j])) return 1;
                others1.push_back(j);
            }
        if (ck::split_modes(t1, others1, 1.12f, &f1, &s1))
            for (size_t j : others)
                if (j != ref1) region[j] = t1[j] >= s1 / 1.06f ? 1 : t1[j] <= f1 * 1.06f ? 2 : -1;
        pick[1] = pick[2] = -1;
        for (size_t j = 1; j < h.size(); ++j)
            if (region[j] > 0 && pick[region[j]] < 0) pick[region[j]] = int(j);
    }
    out->regions = pick[1] < 0 ? 1 : pick[2] < 0 ? 2 : 3;
    if (pick[1] < 0) pick[1] = 1;  // one region seen: any pieces
    if (pick[2] < 0) {              // two: the second piece of the first region
        pick[2] = -1;
        for (size_t j = 1; j < h.size() && pick[2] < 0; ++j)
            if (region[j] == 0) pick[2] = int(j);
        if (pick[2] < 0) pick[2] = 2;
    }
    if (h.size() < 3) {
        std::fprintf(stderr, "--regions: fewer than three pieces\n");
        return 2;
    }
    CHECK(hipMemAddressReserve(reinterpret_cast<void**>(&out->va), 3 * P, 0, nullptr, 0));
    for (int r = 0; r < 3; ++r) {
        CHECK(hipMemMap(out->va + r * P, P, 0, h[pick[r]], 0));
        out->h[r] = h[pick[r]];
    }
    CHECK(hipMemSetAccess(out->va, 3 * P, &acc, 1));
    out->piece = P;
    CHECK(hipDeviceSynchronize());
    for (size_t j = 0; j < h.size(); ++j) CHECK(hipMemUnmap(pool + j * P, P));
    CHECK(hipMemAddressFree(pool, maxp * P));
    for (size_t j = 0; j < h.size(); ++j)
        if (int(j) != pick[0] && int(j) != pick[1] && int(j) != pick[2]) CHECK(hipMemRelease(h[j]));
    return 0;
}

// The five forms at `lanes` lanes over buf; prints one line per form with `tag`.
static int measure(const char* tag, int buffer, uint64_t* buf, uint64_t pitch, uint32_t slots, uint32_t lanes,
                   uint64_t table, uint32_t steps, unsigned long long* sink, hipEvent_t a, hipEvent_t b) {
    const dim3 grid((lanes + 255) / 256), block(256);
    static const char* const kForm[5] = {"read_write", "read_only", "read_only_w4", "read_write_w4",
                                         "read_write_7of8"};
    for (int form = 0; form < 5; ++form) {
        float ms = 0;
        for (int rep = 0; rep < 2; ++rep) {  // the first run warms the TLB and clocks
            CHECK(hipEventRecord(a));
            if (form == 0)
                hipLaunchKernelGGL((k_walk<1, 1>), grid, block, 0, 0, buf, pitch, slots, lanes, steps, sink);
            else if (form == 1)
                hipLaunchKernelGGL((k_walk<0, 1>), grid, block, 0, 0, buf, pitch, slots, lanes, steps, sink);
            else if (form == 2)
                hipLaunchKernelGGL((k_walk<0, 4>), grid, block, 0, 0, buf, pitch, slots, lanes, steps, sink);
            else if (form == 3)
                hipLaunchKernelGGL((k_walk<1, 4>), grid, block, 0, 0, buf, pitch, slots, lanes, steps, sink);
            else
                hipLaunchKernelGGL((k_walk<2, 1>), grid, block, 0, 0, buf, pitch, slots, lanes, steps, sink);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(b));
            CHECK(hipEventSynchronize(b));
            CHECK(hipEventElapsedTime(&ms, a, b));
        }
        const uint32_t done = form == 2 || form == 3 ? steps / 4 * 4 : steps;
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
    bool regions = false;
    for (;;) {
        if (argc > arg + 1 && std::string(argv[arg]) == "--search") {
            search = uint32_t(std::strtoul(argv[arg + 1], nullptr, 10));
            arg += 2;
        } else if (argc > arg && std::string(argv[arg]) == "--regions") {
            regions = true;
            arg += 1;
        } else {
            break;
        }
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
                     "usage: ck_randbench [--search K <= 64] [--regions] [table_bytes, multiple of 8] [steps] "
                     "[lanes <= 2^22 ...]\n");
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
        // fallback search holds its candidates), as many as fit 90 % of the
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
    if (regions) {
        const uint32_t lanes = 65536;
        Spread sp;
        if (int rc = spread_range(pitch * lanes * 8, sink, a, b, &sp)) return rc;
        std::printf("regions_found=%d piece_bytes=%zu\n", sp.regions, sp.piece);
        hipLaunchKernelGGL(k_fill, dim3(8192), dim3(256), 0, 0, reinterpret_cast<uint64_t*>(sp.va), pitch * lanes,
                           0x243F6A8885A308D3ull);
        CHECK(hipGetLastError());
        if (measure("regions_", -1, reinterpret_cast<uint64_t*>(sp.va), pitch, slots, lanes, table, steps, sink, a, b))
            return 1;
        CHECK(hipDeviceSynchronize());
        for (int r = 0; r < 3; ++r) CHECK(hipMemUnmap(sp.va + r * sp.piece, sp.piece));
        for (int r = 0; r < 3; ++r) CHECK(hipMemRelease(sp.h[r]));
        CHECK(hipMemAddressFree(sp.va, 3 * sp.piece));
    }
    CHECK(hipFree(sink));
    return 0;
}
