// ck_kernels.hip -- gfx950 kernels of the cuckoo Monte Carlo path and the
// C-ABI entry points that launch them.
//
// Each trial is (1) a full-lane build -- PRP key and walk derivation, the
// 8*2^half round cache, the bucket map bucket(o, w) for every ordinal and
// candidate, an empty table -- and (2) the serial insertion chain over that map
// and table.  Two placements ("tiers"):
//   HBM tier : k_build (one wave per trial) writes map and table to HBM, then
//              k_chain runs one lane per trial; one probe is one random slot
//              read (+ a map read when an evicted item moves).  The chains are
//              latency-bound, so a launch keeps ~10^5 of them in flight.
//   LDS tier : k_lds, one wave per trial with map, table and round cache in
//              the wave's LDS region; the wave builds with all 64 lanes, then
//              lane 0 runs the chain on chip.  Waves claim trials from an atomic
//              cursor, so build-phase and chain-phase waves overlap on a CU.
// Failures are appended with one ballot and one atomic per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <mutex>
#include <cstring>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "ck_build.hpp"
#include "ck_chain.hpp"
#include "ck_device.hpp"
#include "ck_internal.hpp"
#include "ck_lane.hpp"
#include "ck_regions.hpp"

namespace ck {

__constant__ uint32_t c_te0[256];

// Counter block at the start of the per-launch scratch (u32 words):
//   [0] failures  [2..3] probe total (u64)  [4] LDS-tier work cursor
constexpr uint32_t CNT_FAIL = 0, CNT_PROBES = 2, CNT_CURSOR = 4, CNT_WORDS = 16;

// Where a pass's per-slot tables lie.  n = 1: one range, slot s at base[0] +
// s * bytes.  n = 3: three ranges, one per memory region, slot s in range
// s mod 3 at (s / 3) * bytes, so that neighbouring slots -- the lanes of a
// wave, and whichever slots are still running at a pass's end -- spread their
// random accesses over the three regions.
struct TabRef {
    char* base[3];
    uint32_t n;
    __host__ __device__ char* at(size_t s, size_t bytes) const {
        return n == 1 ? base[0] + s * bytes : base[s % 3] + (s / 3) * bytes;
    }
};

// ------------------------------------------------------------------ kernels

// Stages T0 (2^shift copies) into LDS; returns the lane's view.
__device__ __forceinline__ LdsTab stage_tab(uint32_t* lds, uint32_t shift) {
    const uint32_t n = 256u << shift;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = c_te0[i >> shift];
    __syncthreads();
    return make_tab(lds, shift);
}

// build (HBM tier): one wave per trial slot.  CACHED selects the round-cache
// form (half <= 12); otherwise every bucket evaluation runs AES directly.
// M is the map entry width (u16 when buckets <= 65536), or u64 for the
// descriptor layout (one packed candidate list per ordinal).  With `clear`, the
// slot's table (table_bytes, a multiple of 16) is filled with all-ones (empty);
// without, the chain's pass tag makes the previous passes' entries read as
// empty.  The map rows cover ordinals [item0, item1), row 0 being item0; the
// walk and random word are written with the first window (item0 = 0).
template <int K, bool CACHED, class M>
__global__ __launch_bounds__(1024) CK_BUILD_REGS void k_build(Shape S, RoundKeys bk, uint32_t first, uint32_t count,
                                               uint32_t tab_shift, M* __restrict__ map, size_t map_stride,
                                               TabRef tables, size_t table_bytes,
                                               uint64_t* __restrict__ walks, uint64_t* __restrict__ rnds,
                                               uint32_t* __restrict__ key_out, uint32_t clear, uint32_t item0,
                                               uint32_t item1) {
    extern __shared__ uint32_t lds[];
    const LdsTab T = stage_tab(lds, tab_shift);
    const uint32_t waves = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t slot = blockIdx.x * waves + wave;
    if (slot >= count) return;  // wave-uniform; no block barrier follows
    TrialSeed ts;
    derive_uniform(T, bk, S, first + slot, ts);
    if (key_out && lane < 44) {
        uint32_t v = 0;
#pragma unroll
        for (int i = 0; i < 44; ++i) v = (lane == uint32_t(i)) ? ts.rk[i] : v;
        key_out[slot * 44 + lane] = v;
    }

    M* m = map + slot * map_stride;
    if constexpr (CACHED) {
        uint16_t* cache = reinterpret_cast<uint16_t*>(lds + (256u << tab_shift)) + (size_t(wave) << (S.half + 3));
        build_cache(T, ts, S, cache, lane);
        for (uint32_t o = item0 + lane; o < item1; o += 64) {
            uint32_t c[K];
            feistel_cached_k<K>(cache, S, element_of(S, ts.shift, ts.step, o), c);
            store_candidates<K>(m, S, o - item0, c);
        }
    } else {
        for (uint32_t o = item0 + lane; o < item1; o += 64) {
            const uint64_t el = element_of(S, ts.shift, ts.step, o);
            uint32_t c[K];
            for (int w = 0; w < K; ++w) c[w] = feistel_direct(T, ts.rk, S, el, K, w);
            store_candidates<K>(m, S, o - item0, c);
        }
    }

    if (clear) clear_table(tables.at(slot, table_bytes), table_bytes, lane);
    if (lane == 0 && item0 == 0) {
        walks[slot] = ts.walk;
        rnds[slot] = ts.rnd;
    }
}

// chain (HBM tier): one lane per trial slot; map and table widths M and E
// (M = u64 selects the descriptor layout, whose entries carry `tag`).  XCHG
// selects one-round-trip probes (atomic exchange / compare-and-swap; needs 32-
// or 64-bit entries; always on cleared tables, tag 0).
template <int K, bool BGH, class M, class E, bool XCHG>
__global__ __launch_bounds__(256) void k_chain(Shape S, uint32_t first, uint32_t count,
                                               const M* __restrict__ map, size_t map_stride,
                                               TabRef tables, size_t table_bytes,
                                               const uint64_t* __restrict__ walks, const uint64_t* __restrict__ rnds,
                                               uint32_t* __restrict__ counters, uint32_t* __restrict__ fail_offsets,
                                               uint32_t* __restrict__ probes_out, uint64_t tag) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    bool failed = false;
    uint32_t pr = 0;
    if (slot < count) {
        const M* m = map + slot * map_stride;
        E* t = reinterpret_cast<E*>(tables.at(slot, table_bytes));
        if constexpr (sizeof(M) == 8) {
            const DescCodec D{S.dbits, K * S.dbits, (1ull << S.dbits) - 1, (1ull << (K * S.dbits)) - 1,
                              K * S.dbits + 2, XCHG ? 0ull : tag};
            if constexpr (XCHG) {
                const DescView<K> v{D, m};
                if constexpr (BGH)
                    failed = chain_bgh_xchg<K>(S, v, t, walks[slot], rnds[slot], pr);
                else
                    failed = chain_drrt_xchg<K>(S, v, t, pr);
            } else if constexpr (BGH) {
                failed = chain_bgh_desc<K>(S, D, m, t, walks[slot], rnds[slot], pr);
            } else {
                failed = chain_drrt_desc<K>(S, D, m, t, pr);
            }
        } else if constexpr (XCHG) {
            const MapView<K, M, E> v{m};
            if constexpr (BGH)
                failed = chain_bgh_xchg<K>(S, v, t, walks[slot], rnds[slot], pr);
            else
                failed = chain_drrt_xchg<K>(S, v, t, pr);
        } else {
            if constexpr (BGH)
                failed = chain_bgh<K>(S, m, t, walks[slot], rnds[slot], pr);
            else
                failed = chain_drrt<K>(S, m, t, pr);
        }
        if (probes_out) probes_out[slot] = pr;
    }
    append_failures(failed, first + slot, counters + CNT_FAIL, fail_offsets);
    add_wave_total(pr, reinterpret_cast<unsigned long long*>(counters + CNT_PROBES));
}

// Debug: per-slot digest of the final table (slot_digest over occupied buckets).
template <int K, class M, class E>
__global__ __launch_bounds__(256) void k_digest(Shape S, uint32_t count, const M* __restrict__ map, size_t map_stride,
                                                TabRef tables, size_t table_bytes,
                                                unsigned long long* __restrict__ out) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= count) return;
    const M* m = map + slot * map_stride;
    const E* t = reinterpret_cast<const E*>(tables.at(slot, table_bytes));
    uint64_t sum = 0;
    for (uint32_t b = 0; b < S.buckets; ++b) {
        const E e = t[b];
        uint64_t h = 0;
        uint32_t choice;
        if constexpr (sizeof(E) == 8) {
            if (e == EMPTY64) continue;
            const uint32_t cs = K * S.dbits;
            choice = static_cast<uint32_t>(e >> cs) & 3u;
#pragma unroll
            for (int w = 0; w < K; ++w) h += cand_hash_term(w, static_cast<uint32_t>((e >> (w * S.dbits)) & ((1ull << S.dbits) - 1)));
        } else {
            if (e == empty_entry<E>()) continue;
            const uint32_t o = entry_ordinal<E>(e);
            choice = entry_choice<E>(e);
#pragma unroll
            for (int w = 0; w < K; ++w) h += cand_hash_term(w, m[o * K + w]);
        }
        sum += slot_digest(b, h, choice);
    }
    out[slot] = sum;
}

// chain (HBM tier, residual layout): M is the candidate stream (u64 packed
// descriptor or u32 map rows), E the residual table entry (u32, u64, E44 or E48).  The
// launch inserts items [item0, item1), whose candidates the stream holds.
// With `state` (a trial split into windows of items), the chain resumes from
// and saves, per slot: the walk and random word, state = stash left | failed
// << 31, and probes_out = probes so far; a failed trial is skipped by later
// windows.  Without it, item0 = 0 and item1 = pnum.  With `xcd_order` the
// blocks take their slots in XCD order (xcd_block): measured at n = 2^20, where
// a windowed pass's tables span about 270 GiB, 8,184 -> 8,355 trials/s (two
// interleaved pairs); at n = 2^14 (6.4 GB of tables) it was 0.2 % slower, so
// only windowed passes use it.
template <int K, bool BGH, class M, class E>
__global__ __launch_bounds__(256) void k_chain_res(Shape S, uint32_t first, uint32_t count,
                                                   const M* __restrict__ map, size_t map_stride,
                                                   TabRef tables, size_t table_bytes,
                                                   uint64_t* __restrict__ walks, uint64_t* __restrict__ rnds,
                                                   uint32_t* __restrict__ counters, uint32_t* __restrict__ fail_offsets,
                                                   uint32_t* __restrict__ probes_out, uint32_t item0, uint32_t item1,
                                                   uint32_t* __restrict__ state, uint32_t xcd_order) {
    const uint32_t slot = (xcd_order ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x) * blockDim.x + threadIdx.x;
    bool failed = false;
    uint32_t pr = 0;
    if (slot < count) {
        uint32_t flags = S.stash, before = 0;
        if (state && item0) {
            flags = state[slot];
            before = probes_out[slot];
        }
        if (!(flags >> 31)) {
            const ResCodec<K, typename ResWord<E>::type> R{S.dbits, (K - 1) * S.dbits, (1ull << S.dbits) - 1};
            const CandStream<K, M> st{map + slot * map_stride, S.dbits, (1ull << S.dbits) - 1};
            const auto t = res_table<E>(tables.at(slot, table_bytes));
            const ItemSpan span{item0, item1};
            uint32_t stash_left = flags;
            if constexpr (BGH) {
                uint64_t walk = walks[slot], rnd = rnds[slot];
                failed = chain_bgh_res<K>(S, R, st, t, span, walk, rnd, stash_left, pr);
                if (state) {
                    walks[slot] = walk;
                    rnds[slot] = rnd;
                }
            } else {
                failed = chain_drrt_res<K>(S, R, st, t, span, stash_left, pr);
            }
            if (state) state[slot] = stash_left | (failed ? 1u << 31 : 0u);
        }
        if (probes_out) probes_out[slot] = before + pr;
    }
    append_failures(failed, first + slot, counters + CNT_FAIL, fail_offsets);
    add_wave_total(pr, reinterpret_cast<unsigned long long*>(counters + CNT_PROBES));
}

// Debug: per-slot digest of a residual-layout table.
template <int K, class E>
__global__ __launch_bounds__(256) void k_digest_res(Shape S, uint32_t count, TabRef tables,
                                                    size_t table_bytes, unsigned long long* __restrict__ out) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= count) return;
    using W = typename ResWord<E>::type;
    const ResCodec<K, W> R{S.dbits, (K - 1) * S.dbits, (1ull << S.dbits) - 1};
    const auto t = res_table<E>(tables.at(slot, table_bytes));
    uint64_t sum = 0;
    for (uint32_t b = 0; b < S.buckets; ++b) {
        const W e = t[b];
        if (e == static_cast<W>(~W(0))) continue;
        Cands c = {};
        const uint32_t choice = R.dec(e, b, c);
        uint64_t h = 0;
#pragma unroll
        for (int w = 0; w < K; ++w) h += cand_hash_term(w, c[w]);
        sum += slot_digest(b, h, choice);
    }
    out[slot] = sum;
}

// chain (lane tier): one trial per lane, each of the block's `lanes` lanes
// (lanes <= 64, one wave per block) holding its whole residual table and the
// ring of its candidate stream in LDS.  Block LDS: [tables: lanes x buckets
// entries E, entry b of lane j at element b*lanes + j, rounded up to 16 bytes |
// rings: kLaneRing u64 words per lane, word i of lane j at i*lanes + j].  The
// stream holds one packed u64 per item (bits per bucket index S.dbits), rows of
// map_stride words per slot.  With `digests`, each lane also writes the digest
// of its final table (k_digest_res over the LDS table).
template <int K, bool BGH, class E>
__global__ __launch_bounds__(64) void k_chain_lane(Shape S, uint32_t first, uint32_t count, uint32_t lanes,
                                                   const uint64_t* __restrict__ map, size_t map_stride,
                                                   const uint64_t* __restrict__ walks, const uint64_t* __restrict__ rnds,
                                                   uint32_t* __restrict__ counters, uint32_t* __restrict__ fail_offsets,
                                                   uint32_t* __restrict__ probes_out,
                                                   unsigned long long* __restrict__ digests) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(3))) u32x4* lds_u32x4p;
    extern __shared__ uint32_t lds[];
    const uint32_t j = threadIdx.x;
    const uint32_t slot = blockIdx.x * lanes + j;
    const bool mine = j < lanes && slot < count;
    const uint32_t base = lds_offset(lds);
    const uint32_t tbytes = (lanes * S.buckets * uint32_t(sizeof(E)) + 15u) & ~15u;
    const u32x4 empty4 = {EMPTY, EMPTY, EMPTY, EMPTY};
    for (uint32_t i = j; i < tbytes / 16; i += 64)
        *reinterpret_cast<lds_u32x4p>(static_cast<uintptr_t>(base + i * 16)) = empty4;
    wave_lds_sync();

    const LaneTable<E> t{base + j * uint32_t(sizeof(E)), lanes * uint32_t(sizeof(E))};
    LaneFeed<K> f;
    f.row = (typename LaneFeed<K>::grow)(map + (mine ? size_t(slot) * map_stride : 0));
    f.ring = base + tbytes + j * 8u;
    f.ring_stride = lanes * 8u;
    f.bb = S.dbits;
    f.bmask = (1ull << S.dbits) - 1;
    if (mine) {
        f.issue(0);
        f.land();
        f.issue(kLaneChunk);
        f.land();
    }
    const ResCodec<K, E> R{S.dbits, (K - 1) * S.dbits, (1ull << S.dbits) - 1};
    uint32_t pr = 0;
    bool failed;
    if constexpr (BGH) {
        uint64_t walk = mine ? walks[slot] : 0, rnd = mine ? rnds[slot] : 0;
        // Wait for these two loads here: inside the chain loop a wait for a
        // load issued before the loop would also wait for the refill loads in
        // flight (the vector memory counter is in order), every iteration.
        asm volatile("" ::"v"(walk), "v"(rnd));
        failed = chain_bgh_lane<K, E>(S, R, f, t, mine, walk, rnd, pr);
    } else
        failed = chain_drrt_lane<K, E>(S, R, f, t, mine, pr);
    if (mine && probes_out) probes_out[slot] = pr;
    if (mine && digests) {
        uint64_t sum = 0;
        for (uint32_t b = 0; b < S.buckets; ++b) {
            const E e = t.get(b);
            if (e == static_cast<E>(~E(0))) continue;
            Cands c = {};
            const uint32_t choice = R.dec(e, b, c);
            uint64_t h = 0;
#pragma unroll
            for (int w = 0; w < K; ++w) h += cand_hash_term(w, c[w]);
            sum += slot_digest(b, h, choice);
        }
        digests[slot] = sum;
    }
    append_failures(failed, first + slot, counters + CNT_FAIL, fail_offsets);
    add_wave_total(pr, reinterpret_cast<unsigned long long*>(counters + CNT_PROBES));
}

// Table placement: the insertion chains' access pattern without their logic
// -- each of `lanes` lanes walks its own table of `slots` 8-byte words (lane
// l's at tables.at(l, pitch * 8)), reading a word and writing it back changed,
// the next index taken from the value read.  The contents are whatever the
// memory holds (every index is below `slots`); `sink` only keeps the loads
// live.
__global__ __launch_bounds__(256) void k_place_walk(TabRef tables, uint64_t pitch, uint32_t slots, uint32_t lanes,
                                                    uint32_t steps, unsigned long long* __restrict__ sink) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= lanes) return;
    uint64_t* t = reinterpret_cast<uint64_t*>(tables.at(lane, pitch * 8));
    uint32_t ctr = lane * 0x9E3779B9u;
    uint32_t i = __umulhi(lane * 0x85EBCA6Bu + 0x27D4EB2Fu, slots);
    uint64_t acc = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        const uint64_t v = t[i];
        t[i] = v ^ ctr;
        acc += v;
        i = __umulhi(static_cast<uint32_t>(v) ^ ctr, slots);
        ctr += 0x9E3779B9u;
    }
    if (acc == 0x5eedull) sink[0] = acc;
}

// Region placement: the same walk with 2 x `half` lanes, the first half over
// tables of `slots` words at a, the second half at b (half x slots words at
// each), so that its rate tells whether a and b lie in one memory region.
__global__ __launch_bounds__(256) void k_place_pair(uint64_t* __restrict__ a, uint64_t* __restrict__ b, uint32_t half,
                                                    uint32_t slots, uint32_t steps,
                                                    unsigned long long* __restrict__ sink) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= 2 * half) return;
    uint64_t* t = (lane < half ? a : b) + uint64_t(lane < half ? lane : lane - half) * slots;
    uint32_t ctr = lane * 0x9E3779B9u;
    uint32_t i = __umulhi(lane * 0x85EBCA6Bu + 0x27D4EB2Fu, slots);
    uint64_t acc = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        const uint64_t v = t[i];
        t[i] = v ^ ctr;
        acc += v;
        i = __umulhi(static_cast<uint32_t>(v) ^ ctr, slots);
        ctr += 0x9E3779B9u;
    }
    if (acc == 0x5eedull) sink[0] = acc;
}

// LDS tier: a persistent wave per trial at a time.  Wave region layout:
// [round cache 8<<half u16][map pnum*K M, unless ROUND][table buckets E].
// The chain is the trial API's insert_core_k (include/cuckoo_trial.hpp) on
// lane 0, its buckets read from the map (PrecomputedBuckets) or, with ROUND,
// evaluated over the round cache at every probe (CachedBuckets: no map, so a
// trial needs less LDS).
template <int K, bool BGH, class M, class E, bool ROUND>
__global__ __launch_bounds__(512) void k_lds(Shape S, RoundKeys bk, uint32_t first, uint32_t count, uint32_t tab_shift,
                                             uint32_t wave_bytes, uint32_t cache_bytes, uint32_t map_bytes,
                                             uint32_t* __restrict__ counters, uint32_t* __restrict__ fail_offsets,
                                             uint32_t* __restrict__ probes_out, uint32_t* __restrict__ tables_out,
                                             unsigned long long* __restrict__ digests_out) {
    extern __shared__ uint32_t lds[];
    const LdsTab T = stage_tab(lds, tab_shift);
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    char* region = reinterpret_cast<char*>(lds + (256u << tab_shift)) + size_t(wave) * wave_bytes;
    uint16_t* cache = reinterpret_cast<uint16_t*>(region);
    M* map = reinterpret_cast<M*>(region + cache_bytes);
    E* table = reinterpret_cast<E*>(region + cache_bytes + map_bytes);
    const cuckoo::PackedBins<E, EntryCodec<E>::SHIFT> bins{table, 1};
    // Every iteration starts and ends wave-converged: the only lane-divergent
    // region is the chain itself.  Single-lane side effects (claiming a slot,
    // publishing the outcome) are written as whole-wave operations on
    // wave-uniform values, so the compiler has no lane-0 branch to thread from
    // the end of one iteration into the start of the next.
    // The cursor is monotone, so no wave can claim more than count + 1 times.
    for (uint32_t iter = 0; iter <= count; ++iter) {
        const uint32_t claimed = atomicAdd(counters + CNT_CURSOR, lane == 0 ? 1u : 0u);
        const uint32_t slot = __builtin_amdgcn_readfirstlane(claimed);
        if (slot >= count) break;
        TrialSeed ts;
        derive_uniform(T, bk, S, first + slot, ts);
        build_cache(T, ts, S, cache, lane);
        const cuckoo::CachedBuckets<uint16_t, K> feistel{cache, 1, S.domain, S.bucket_size, S.mask, S.range_mask,
                                                         ts.shift, ts.step, S.half, S.input_mode == 0};
        if constexpr (!ROUND) {
            for (uint32_t o = lane; o < S.pnum; o += 64) {
                uint32_t c[K];
                feistel_cached_k<K>(cache, S, element_of(S, ts.shift, ts.step, o), c);
#pragma unroll
                for (int w = 0; w < K; ++w) map[o * K + w] = static_cast<M>(c[w]);
            }
        }
        for (uint32_t i = lane; i < S.buckets; i += 64) table[i] = empty_entry<E>();
        wave_lds_sync();
        bool failed_lane = false;
        uint32_t pr_lane = 0;
        if (lane == 0) {
            if constexpr (ROUND)
                failed_lane = cuckoo::insert_core_k<BGH, K>(S.pnum, S.kick, S.stash, bins, feistel, ts.walk, ts.rnd,
                                                            pr_lane);
            else
                failed_lane = cuckoo::insert_core_k<BGH, K>(S.pnum, S.kick, S.stash, bins,
                                                            cuckoo::PrecomputedBuckets<M, K>{map}, ts.walk, ts.rnd,
                                                            pr_lane);
        }
        wave_lds_sync();
        const uint32_t failed = __builtin_amdgcn_readfirstlane(failed_lane ? 1u : 0u);
        const uint32_t pr = __builtin_amdgcn_readfirstlane(pr_lane);
        if (tables_out) {
            uint32_t* out = tables_out + size_t(slot) * S.buckets;
            for (uint32_t i = lane; i < S.buckets; i += 64) {
                const E e = table[i];
                out[i] = e == empty_entry<E>() ? EMPTY : (entry_ordinal<E>(e) | (entry_choice<E>(e) << 30));
            }
        }
        if (digests_out) {
            uint64_t part = 0;
            for (uint32_t b = lane; b < S.buckets; b += 64) {
                const E e = table[b];
                if (e == empty_entry<E>()) continue;
                const uint32_t o = entry_ordinal<E>(e);
                uint64_t h = 0;
#pragma unroll
                for (int w = 0; w < K; ++w)
                    h += cand_hash_term(w, ROUND ? feistel(o, uint32_t(w)) : uint32_t(map[o * K + w]));
                part += slot_digest(b, h, entry_choice<E>(e));
            }
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) part += __shfl_xor(part, d);
            digests_out[slot] = part;  // same value from every lane
        }
        if (probes_out) probes_out[slot] = pr;  // same value from every lane
        if (failed) {
            const uint32_t at = __builtin_amdgcn_readfirstlane(atomicAdd(counters + CNT_FAIL, lane == 0 ? 1u : 0u));
            fail_offsets[at] = first + slot;  // same value from every lane
        }
        atomicAdd(reinterpret_cast<unsigned long long*>(counters + CNT_PROBES),
                  static_cast<unsigned long long>(lane == 0 ? pr : 0u));
    }
}

__global__ void k_zero(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// The first `words` words of each of `rows` rows `stride` words apart.
__global__ void k_zero_rows(uint32_t* p, size_t stride, uint32_t words, uint64_t rows) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < rows * words) p[(i / words) * stride + i % words] = 0;
}

// FIPS-197 C.1 through the device key schedule and T-table AES on every lane
// of one wave, with the replicated (shift 5) and plain (shift 0) tables
// (aes_kat / cached_aes_kat, cuckoo_bench.cu:322-354).
__global__ __launch_bounds__(64) void k_aes_kat(uint32_t* out, uint32_t tab_shift) {
    extern __shared__ uint32_t lds[];
    const LdsTab T = stage_tab(lds, tab_shift);
    const uint32_t key[4] = {0x00010203u, 0x04050607u, 0x08090a0bu, 0x0c0d0e0fu};
    uint32_t rk[44], o[4];
    expand_key_dev(T, key, rk);
    aes_enc(T, rk, 0x00112233u, 0x44556677u, 0x8899aabbu, 0xccddeeffu, o);
    for (int i = 0; i < 4; ++i) out[threadIdx.x * 4 + i] = o[i];
}

// ---------------------------------------------------------------- host side

#define CKH(expr)                                                                                              \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess) return fail(CK_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

struct Workspace {
    void* base = nullptr;
    size_t bytes = 0;
};

// The tables of one pass buffer, held apart from it by a placed plan: one
// hipMalloc (pieces empty), or `pieces` physical allocations of `piece` bytes
// each mapped one after another at `base` (virtual-memory API).
struct TableRange {
    void* base = nullptr;
    size_t bytes = 0;
    size_t piece = 0;
    std::vector<hipMemGenericAllocationHandle_t> pieces;
};

// The workspace of ck_launch_static: a file-scope device array, so that
// launch allocates nothing.  Every GPU the library runs on holds its own copy
// from load until the process ends.  Host code never uses the name `scratch`
// as an address (on the host it only names the symbol); static_scratch_of()
// resolves this device's address once, with the device current, and kernels
// receive that pointer.
static __device__ __align__(256) unsigned char scratch[CK_STATIC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices] = {};

static int static_scratch_of(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return fail(CK_EINVAL, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* addr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !addr)
            return fail(CK_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[device] = addr;
    }
    *out = g_scratch_addr[device];
    return CK_OK;
}

// Workspaces that an instantiated graph's kernel nodes have pointed into.  On
// ROCm 7.2 the runtime keeps such an allocation after hipFree, even once the
// graph instances are destroyed, the streams synchronised and destroyed and
// the graph and default memory pools trimmed (measured: each graph-mode
// context left its 40 GiB workspace behind, cumulatively).  The library
// therefore never frees them: it keeps them per device and hands them to the
// next context on that device that needs a workspace they can hold.
static std::mutex g_retired_mu;
static std::vector<Workspace> g_retired[kMaxDevices];

static void retire_ws(int device, const Workspace& w) {
    std::lock_guard<std::mutex> lock(g_retired_mu);
    g_retired[device].push_back(w);
}
// The smallest retired workspace of at least `bytes`, removed from the pool.
static bool take_retired(int device, size_t bytes, Workspace& out) {
    std::lock_guard<std::mutex> lock(g_retired_mu);
    auto& pool = g_retired[device];
    auto best = pool.end();
    for (auto it = pool.begin(); it != pool.end(); ++it)
        if (it->bytes >= bytes && (best == pool.end() || it->bytes < best->bytes)) best = it;
    if (best == pool.end()) return false;
    out = *best;
    pool.erase(best);
    return true;
}
static size_t largest_retired(int device) {
    std::lock_guard<std::mutex> lock(g_retired_mu);
    size_t m = 0;
    for (const auto& w : g_retired[device]) m = std::max(m, w.bytes);
    return m;
}

// Where a kernel goes: onto a stream, into a graph as a new node after `deps`,
// or into an existing node of an instantiated graph (new argument values; the
// kernel and its launch shape stay).  All three pass the arguments as one
// array of pointers to values of the kernel's own parameter types.
struct Dispatch {
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    const hipGraphNode_t* deps = nullptr;
    size_t ndeps = 0;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t node = nullptr;  // added node (graph), or the node to rewrite (exec)
    hipError_t err = hipSuccess;

    template <class... P, class... A>
    void operator()(void (*k)(P...), dim3 grid, dim3 block, size_t lds, A&&... args) {
        std::tuple<std::decay_t<P>...> vals(std::forward<A>(args)...);
        put(reinterpret_cast<void*>(k), grid, block, lds, vals, std::index_sequence_for<P...>{});
    }

   private:
    template <class T, size_t... I>
    void put(void* fn, dim3 grid, dim3 block, size_t lds, T& vals, std::index_sequence<I...>) {
        void* ptrs[] = {static_cast<void*>(&std::get<I>(vals))...};
        hipError_t e;
        if (!graph && !exec) {
            e = hipLaunchKernel(fn, grid, block, ptrs, lds, stream);
        } else {
            hipKernelNodeParams kp{};
            kp.func = fn;
            kp.gridDim = grid;
            kp.blockDim = block;
            kp.sharedMemBytes = static_cast<unsigned>(lds);
            kp.kernelParams = ptrs;
            kp.extra = nullptr;
            e = graph ? hipGraphAddKernelNode(&node, graph, deps, ndeps, &kp)
                      : hipGraphExecKernelNodeSetParams(exec, node, &kp);
        }
        if (e != hipSuccess && err == hipSuccess) err = e;
    }
};

// An HBM-tier launch as a graph (graph mode): every pass's build, counter
// clear and chain as graph nodes with the stream order's dependencies,
// replayed for the same shape and trial count with new batch keys and
// offsets.  Each pass has its own counters and failure list in `aux`, so
// nothing is read back until the whole graph has run.
struct HbmGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    ck_params p{};
    uint32_t count = 0, layout = 0, chunk = 0, nbuf = 0;
    void* ws = nullptr;
    uint64_t last_use = 0;
    std::vector<hipGraphNode_t> build, chain;
    uint32_t* aux = nullptr;  // per pass: CNT_WORDS counters, then up to `chunk` failure offsets
    size_t aux_words = 0;
    void drop() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (g) (void)hipGraphDestroy(g);
        exec = nullptr;
        g = nullptr;
        build.clear();
        chain.clear();
        ws = nullptr;
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;        // chains (HBM tier) and fused kernels (LDS tier)
    hipStream_t build_stream = nullptr;  // builds of the next HBM pass, overlapping the chains
    hipEvent_t ev[4] = {};
    hipEvent_t ev_build[2][2] = {};  // [buffer][start, end]
    hipEvent_t ev_chain[2][2] = {};
    Workspace ws;
    // Table placement (see place_buffers): the tables of the two pass buffers
    // are held apart from them, spread over the device's memory regions (or,
    // where that cannot be done, the two fastest of place_cands candidate
    // allocations); ws then holds the rest of both pass buffers.
    TableRange tabs[2];
    size_t placed_for = 0;           // table-region size tabs were placed for (0: not placed)
    uint32_t place_cands = 0;        // on/off, and candidates per fallback search (0: no placement)
    std::vector<float> place_ms;     // last placement: walk ms per candidate or piece, in allocation order
    int place_kept[2] = {-1, -1};    // candidates (or first pieces) kept for pass buffers 0 and 1
    std::vector<int32_t> place_region;   // region placement: each piece's region (empty: search)
    int32_t place_used[6] = {-1, -1, -1, -1, -1, -1};  // pieces mapped as buffer 0's, then buffer 1's tables
    uint32_t tier = CK_TIER_AUTO;
    uint32_t layout = CK_LAYOUT_AUTO;
    uint32_t chunk_cap = 0;
    uint32_t window = 0;  // items per residual-layout window (0 = automatic)
    uint32_t prp = CK_PRP_AUTO;          // round cache or direct AES for bucket evaluations
    uint32_t lds_variant = CK_LDS_AUTO;  // LDS tier: bucket map or round cache per probe
    size_t budget_fixed = 0;             // workspace budget set by the caller (0: from the free memory)
    void* static_scratch = nullptr;      // this device's address of `scratch` (ck_launch_static)
    bool ws_static = false;              // the workspace is `scratch`: never grown, freed or reused across calls
    bool ws_graphed = false;             // an instantiated graph has pointed into the workspace: retire, never free
    size_t budget = size_t(96) << 30;  // HBM budget for the double-buffered maps and tables (set per call)
    size_t lds_block = 64 * 1024;      // dynamic LDS per block the device allows
    size_t lds_cu = 64 * 1024;         // LDS per CU
    uint32_t cus = 1;
    // Descriptor tables reused without a clear: the plan of the tables in the
    // workspace, and per pass buffer whether the tables of its first `slots`
    // slots hold only entries of that plan written with tags up to `tag`.
    uint64_t table_plan = 0;
    struct {
        bool valid;
        uint32_t slots;
        uint64_t tag;
    } table_tags[2] = {};
    bool graph_on = false;  // HBM-tier launches run as replayed graphs
    static constexpr int kGraphs = 4;  // graphs kept per context, least recently used replaced first
    HbmGraph graphs[kGraphs];
    uint64_t graph_clock = 0;
};

// Round caches above 2^12 entries per round do not pay for themselves in LDS.
static constexpr uint32_t kMaxCacheHalf = 12;
static constexpr uint32_t kRepl = 5;  // T0 copies = 2^5 for conflict-free lookups
static size_t tab_bytes(uint32_t shift) { return size_t(256) << shift << 2; }
static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Entry widths: a u16 map holds bucket indices < 65536; a u16 table entry holds
// ordinal | choice << 14, which needs ordinals < 2^14 and never equals the
// all-ones empty entry (choice 3 with ordinal 16383 would).
static bool map16_ok(const ck_params& p) { return p.buckets <= 65536; }
static bool table16_ok(const ck_params& p) { return p.pnum < 16384 || (p.pnum == 16384 && p.k <= 3); }

// Trials (waves) per build block: up to 16, as LDS allows next to the T-table,
// so a CU keeps enough waves to hide the LDS latency of the AES rounds.
static constexpr uint32_t kBuildWaves = 16;

struct BuildPlan {
    bool cached;
    uint32_t waves;  // trials per build block
    size_t lds;
};
static bool cache_possible(const ck_params& p, size_t lds_block) {
    return p.half_bits <= kMaxCacheHalf && tab_bytes(kRepl) + (size_t(8) << p.half_bits) * sizeof(uint16_t) <= lds_block;
}
static BuildPlan plan_build(const ck_params& p, size_t lds_block, uint32_t prp) {
    BuildPlan b{};
    const size_t tab = tab_bytes(kRepl);
    const size_t cache = (size_t(8) << p.half_bits) * sizeof(uint16_t);
    b.cached = prp != CK_PRP_DIRECT && cache_possible(p, lds_block);
    if (b.cached) {
        b.waves = static_cast<uint32_t>(std::min<size_t>(kBuildWaves, (lds_block - tab) / cache));
        b.lds = tab + b.waves * cache;
    } else {
        b.waves = 4;
        b.lds = tab;
    }
    return b;
}

struct LdsPlan {
    bool ok = false, map16 = false, bins16 = false, round = false;
    uint32_t waves = 0, tab_shift = 0, blocks_per_cu = 0;
    size_t cache_b = 0, map_b = 0, bins_b = 0, wave_b = 0, lds = 0;
};
static LdsPlan plan_lds(const ck_params& p, size_t lds_block, size_t lds_cu, bool round) {
    LdsPlan L;
    L.round = round;
    if (p.half_bits > kMaxCacheHalf) return L;
    auto a16 = [](size_t x) { return (x + 15) & ~size_t(15); };
    L.map16 = map16_ok(p);
    L.bins16 = table16_ok(p);
    L.cache_b = a16((size_t(8) << p.half_bits) * 2);
    L.map_b = round ? 0 : a16(size_t(p.pnum) * p.k * (L.map16 ? 2 : 4));
    L.bins_b = a16(size_t(p.buckets) * (L.bins16 ? 2 : 4));
    L.wave_b = L.cache_b + L.map_b + L.bins_b;
    uint32_t best = 0;
    for (uint32_t shift : {kRepl, 0u}) {
        for (uint32_t w = 1; w <= 8; ++w) {
            const size_t blk = tab_bytes(shift) + w * L.wave_b;
            if (blk > lds_block) break;
            const uint32_t per_cu = static_cast<uint32_t>(lds_cu / blk);
            const uint32_t waves_cu = std::min<uint32_t>(per_cu * w, 32);
            if (waves_cu > best) {
                best = waves_cu;
                L.waves = w;
                L.tab_shift = shift;
                L.blocks_per_cu = per_cu;
                L.lds = blk;
            }
        }
    }
    L.ok = best > 0;
    return L;
}


// Descriptor layout fits when k packed bucket indices and a 2-bit choice leave
// the top bit of a u64 clear (so all-ones stays the empty entry).
static bool desc_ok(const ck_params& p) { return p.k * make_shape(p).dbits + 2 <= 63; }

// Residual layout: (k-1)*bits + 2 must stay below an all-ones entry -- fewer
// bits than the entry width, or exactly the width with a choice below 3.
static bool res_fits(const ck_params& p, uint32_t width) {
    const uint32_t need = (p.k - 1) * make_shape(p).dbits + 2;
    return need < width || (need == width && p.k <= 3);
}
static bool res_ok(const ck_params& p) { return res_fits(p, 64); }
static bool res_e32(const ck_params& p) { return res_fits(p, 32); }
static bool res_e44(const ck_params& p) { return res_fits(p, 44); }
static bool res_e48(const ck_params& p) { return res_fits(p, 48); }
static bool res_stream64(const ck_params& p) { return p.k * make_shape(p).dbits <= 64; }

// Lane tier: residual entries of at most 32 bits (16 when they fit) and a u64
// candidate stream.  A block is one wave with `lanes` of its lanes running
// trials; the planner takes the lanes per block that put the most trials on a
// CU (blocks of lanes x (buckets entries + kLaneRing ring words) bytes).
struct LanePlan {
    bool ok = false, e16 = false;
    uint32_t lanes = 0, blocks_per_cu = 0, lanes_per_cu = 0;
    size_t lds = 0;
};
static bool lane_fits(const ck_params& p) { return res_e32(p) && res_stream64(p); }
static size_t lane_block_bytes(const ck_params& p, uint32_t lanes, size_t eb) {
    return ((size_t(lanes) * p.buckets * eb + 15) & ~size_t(15)) + size_t(lanes) * kLaneRing * 8;
}
static LanePlan plan_lane(const ck_params& p, size_t lds_block, size_t lds_cu) {
    LanePlan L;
    if (!lane_fits(p)) return L;
    L.e16 = res_fits(p, 16);
    const size_t eb = L.e16 ? 2 : 4;
    for (uint32_t w = 64; w >= 1; --w) {
        const size_t blk = lane_block_bytes(p, w, eb);
        if (blk > lds_block) continue;
        const uint32_t per_cu = static_cast<uint32_t>(std::min<size_t>(lds_cu / blk, 32));  // 32 waves per CU
        if (per_cu * w > L.lanes_per_cu) {
            L.lanes_per_cu = per_cu * w;
            L.lanes = w;
            L.blocks_per_cu = per_cu;
            L.lds = blk;
        }
    }
    L.ok = L.lanes_per_cu > 0;
    return L;
}
// Tier of a launch: auto is the HBM tier, measured faster than both on-chip
// tiers on every shape tried.
//  - LDS tier (a wave per trial; n = 1024 and n = 2^14, both modes): an LDS
//    chain probe costs two dependent LDS round trips on one lane while the
//    HBM tier keeps ~10^5 chains in flight.
//  - Lane tier (bgh25, k = 2 and 3, n = 256 .. 4096, m = 1.3n .. 1.5n, one
//    65,536-trial pass without overlap): its chains ran at 26.3 G probes/s
//    at n = 256 (80 lanes per CU) down to 5.9 G at n = 2048 (12 lanes per
//    CU), against 25-41 G for the HBM tier, whose small tables stay in the
//    L2 and the MALL; with passes overlapped it lost 0.13-0.78x end to end,
//    because its chain blocks hold the LDS that the build blocks need.  A
//    lane-tier iteration (one probe per lane) took 580-780 ns against
//    ~1.3 us for an HBM-tier probe at low load: too few lanes fit on chip.
// The autotuner measures both on-chip tiers and keeps the fastest placement.
static uint32_t choose_tier(const Ctx& c, const ck_params&, bool = false) {
    return c.tier != CK_TIER_AUTO ? c.tier : CK_TIER_HBM;
}

// HBM-tier pass buffer: [counters | map C x map_bytes (x2 with windows) |
// walk C | rnd C | probes C | fail C | keys C x 44 | digests C | state C |
// tables C x table_bytes], twice (double-buffered).  In the descriptor layout the
// "map" is one u64 candidate descriptor per ordinal and the table entries are
// u64.  With item windows (residual layout) the map holds the candidates of
// `window` items, in two halves: the build of window j+1 fills one while the
// chain of window j reads the other.
// Internal layout value of the lane tier: residual entries, tables in LDS.
static constexpr uint32_t kLayoutLane = 100;
struct HbmPlan {
    bool desc, map16, table16, xchg;
    bool lane;           // lane tier: no tables in HBM, k_chain_lane
    LanePlan lp;
    bool res, stream64, e32, e44, e48;  // residual layout: u64 candidate stream, u32 / 44-bit / 48-bit entries
    uint32_t window;     // items per map (pnum unless windowed)
    uint32_t nwin;       // windows per trial
    size_t map_stride;   // map entries per slot
    size_t map_bytes;    // per slot
    size_t table_bytes;  // per slot
    size_t off_cnt, off_map, off_tables, off_walk, off_rnd, off_probes, off_fail, off_key, off_digest, off_state,
        buffer;
};
static HbmPlan plan_hbm(const ck_params& p, uint32_t chunk, uint32_t layout, uint32_t window) {
    HbmPlan h{};
    h.window = window;
    h.nwin = (p.pnum + window - 1) / window;
    const bool desc = layout == CK_LAYOUT_DESC || layout == CK_LAYOUT_DESC_PLAIN;
    h.desc = desc;
    h.xchg = layout == CK_LAYOUT_DESC || layout == CK_LAYOUT_MAP_XCHG;
    h.lane = layout == kLayoutLane;
    h.res = layout == CK_LAYOUT_RESIDUAL || h.lane;
    h.stream64 = h.res && res_stream64(p);
    h.e32 = h.res && res_e32(p);
    h.e44 = h.res && !h.e32 && res_e44(p);
    h.e48 = h.res && !h.e32 && !h.e44 && res_e48(p);
    h.map16 = !desc && !h.res && map16_ok(p);
    h.table16 = !desc && !h.res && !h.xchg && table16_ok(p);
    const bool m64 = desc || h.stream64;
    const size_t mw = m64 ? 8 : h.map16 ? 2 : 4;
    const size_t tw = h.res ? (h.e32 ? 4 : h.e48 ? 6 : 8) : desc ? 8 : h.table16 ? 2 : 4;
    h.map_bytes = align256(size_t(window) * (m64 ? 1 : p.k) * mw);
    h.map_stride = h.map_bytes / mw;
    h.table_bytes = h.lane ? 0
                   : h.e44 ? align256(Entries44::bytes(p.buckets))
                           : align256(size_t(p.buckets) * tw + (h.e48 ? 8 : 0));  // 48-bit reads reach 2 bytes past
    size_t o = 0;
    h.off_cnt = o;
    o = align256(o + CNT_WORDS * 4);
    h.off_map = o;
    o = align256(o + h.map_bytes * chunk * (h.nwin > 1 ? 2 : 1));
    h.off_walk = o;
    o = align256(o + size_t(chunk) * 8);
    h.off_rnd = o;
    o = align256(o + size_t(chunk) * 8);
    h.off_probes = o;
    o = align256(o + size_t(chunk) * 4);
    h.off_fail = o;
    o = align256(o + size_t(chunk) * 4);
    h.off_key = o;
    o = align256(o + size_t(chunk) * 44 * 4);
    h.off_digest = o;
    o = align256(o + size_t(chunk) * 8);
    h.off_state = o;
    o = align256(o + (h.nwin > 1 ? size_t(chunk) * 4 : 0));
    h.off_tables = o;  // last, so that a searched placement can hold the tables apart
    o = align256(o + h.table_bytes * chunk);
    h.buffer = o;
    return h;
}

// Below this many chains per pass the exchange probes win (measured at
// n = 2^20: 2048 chains, 595 vs 487 trials/s); above it load + store wins.
static constexpr uint32_t kXchgBelow = 1u << 14;

// Trials per pass: 65536 chains in flight saturate the random-access path
// (measured: 32768 slower, 131072 and more no faster).
static constexpr uint32_t kHbmChunk = 1u << 16;
static size_t slot_bytes(const ck_params& p, uint32_t layout, uint32_t window) {
    const HbmPlan one = plan_hbm(p, 1, layout, window);
    return one.map_bytes * (one.nwin > 1 ? 2 : 1) + one.table_bytes + 8 + 8 + 4 + 4 + 176 + (one.nwin > 1 ? 4 : 0);
}
// Pass buffers: two, so the build of pass i+1 runs under the chain of pass i,
// unless memory caps a pass below its target.  Then one buffer of twice the
// slots: a memory-capped shape (n = 2^20: 19.3 MB per slot) is bound by the
// latency of its chains, and its build is a few percent of the chain time.
static uint32_t hbm_buffers(const Ctx& c, const ck_params& p, uint32_t count, uint32_t layout) {
    const uint32_t target = c.chunk_cap ? c.chunk_cap : kHbmChunk;
    return c.budget / (2 * slot_bytes(p, layout, p.pnum)) < std::min<size_t>(target, count) ? 1 : 2;
}
static uint32_t hbm_chunk(const Ctx& c, const ck_params& p, uint32_t count, uint32_t layout, uint32_t nbuf,
                          uint32_t window) {
    const size_t fit = std::max<size_t>(1, c.budget / (nbuf * slot_bytes(p, layout, window)));
    const uint32_t target = c.chunk_cap ? c.chunk_cap : kHbmChunk;
    return static_cast<uint32_t>(std::max<size_t>(1, std::min<size_t>({fit, size_t(target), size_t(count)})));
}

// Item windows (residual layout).  When memory caps a single-buffered pass
// below its target, each slot keeps the candidates of 2 x pnum / kWindows items
// instead of pnum, and a pass runs as kWindows build + chain launches, each
// chain stopping at an item boundary.  At n = 2^20 the candidate stream (one
// u64 per item, 8.4 MB) is half of a slot; the memory that windows free holds
// more slots, and these chains are latency-bound, so more chains in flight is
// more throughput.  Measured at n = 2^20 with u64 entries (trials/s): no
// windows 3,425 (15,137 chains per pass); 8 windows 5,346; 16: 5,813; 32: 6,089
// (25,552 chains, the same 4.1 s of chain per pass).  With 48-bit entries:
// 32 windows 6,960 (33,556 chains, 4.75 s of chain per pass); 64: 7,146 (8,051
// once each 48-bit write became one 8-byte store).  With 44-bit entries and 64
// windows: 37,614 chains, 8,206-8,231 against 8,051-8,061 for 48-bit entries
// (interleaved runs).  Each window's build derives the trial and its round
// cache again, under the previous window's chain.
static constexpr uint32_t kWindows = 64;
static uint32_t stream_window(const Ctx& c, const ck_params& p, uint32_t count, uint32_t layout, bool allowed) {
    if (!allowed || layout != CK_LAYOUT_RESIDUAL) return p.pnum;
    if (c.window) return std::min(c.window, p.pnum);
    if (hbm_buffers(c, p, count, layout) != 1) return p.pnum;
    const uint32_t target = std::min(c.chunk_cap ? c.chunk_cap : kHbmChunk, count);
    if (hbm_chunk(c, p, count, layout, 1, p.pnum) >= target) return p.pnum;
    return (p.pnum + kWindows - 1) / kWindows;
}

// HBM budget of the pass buffers: 95 % of the free device memory (the current
// workspace counts as free, since growing it frees it first), at most 272 GiB
// of the 288.  Large shapes (n = 2^20: 19.3 MB per trial slot) are
// latency-bound, so every extra slot is an extra chain in flight.
static constexpr size_t kBudgetCap = size_t(272) << 30;
static size_t hbm_budget(const Ctx& c) {
    if (c.budget_fixed) return c.budget_fixed;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return size_t(96) << 30;
    const size_t own = std::max(c.ws.bytes + c.tabs[0].bytes + c.tabs[1].bytes, largest_retired(c.device));  // memory a new workspace can use
    return std::min(kBudgetCap, (free_b + own) / 20 * 19);
}

// LDS-tier per-launch scratch: [counters | probes C | fail C | digests C | tables C x buckets].
struct LdsScratch {
    size_t off_cnt, off_probes, off_fail, off_digest, off_tables, total;
};
static LdsScratch lds_scratch(const ck_params& p, uint32_t chunk, bool tables) {
    LdsScratch L{};
    size_t o = 0;
    L.off_cnt = o;
    o = align256(o + CNT_WORDS * 4);
    L.off_probes = o;
    o = align256(o + size_t(chunk) * 4);
    L.off_fail = o;
    o = align256(o + size_t(chunk) * 4);
    L.off_digest = o;
    o = align256(o + size_t(chunk) * 8);
    L.off_tables = o;
    o = align256(o + (tables ? size_t(chunk) * p.buckets * 4 : 0));
    L.total = o;
    return L;
}
static uint32_t lds_chunk(const Ctx& c, uint32_t count, bool tables, const ck_params& p) {
    uint32_t target = c.chunk_cap ? c.chunk_cap : (1u << 20);
    if (tables) target = std::min<uint32_t>(target, static_cast<uint32_t>(std::max<size_t>(1, (size_t(1) << 30) / (size_t(p.buckets) * 4))));
    return std::max<uint32_t>(1, std::min(target, count));
}

static void forget_tables(Ctx& c) {
    c.table_plan = 0;
    c.table_tags[0] = c.table_tags[1] = {};
}

// Frees (or, after graph use, retires) the workspace and a placed second pass
// buffer.  Graphs whose kernel nodes point into them are destroyed first (they
// could never be replayed against a new workspace).
static int free_tables(Ctx& c, TableRange& t) {
    if (!t.base) return CK_OK;
    if (t.pieces.empty()) {
        if (c.ws_graphed)
            retire_ws(c.device, Workspace{t.base, t.bytes});
        else
            CKH(hipFree(t.base));
    } else {
        // Virtual-memory pieces: unmapped and released once idle, graph use or
        // not (the runtime returns them to the device, unlike hipMalloc memory
        // that graph nodes pointed into).  Graphs pointing here are gone by now.
        CKH(hipDeviceSynchronize());
        for (size_t i = 0; i < t.pieces.size(); ++i) CKH(hipMemUnmap(static_cast<char*>(t.base) + i * t.piece, t.piece));
        for (hipMemGenericAllocationHandle_t h : t.pieces) CKH(hipMemRelease(h));
        CKH(hipMemAddressFree(t.base, t.bytes));
    }
    t = TableRange{};
    return CK_OK;
}

static int release_ws(Ctx& c) {
    forget_tables(c);
    for (HbmGraph& g : c.graphs)
        if (g.exec && g.ws == c.ws.base) g.drop();
    if (c.ws.base) {
        if (c.ws_graphed)
            retire_ws(c.device, c.ws);
        else
            CKH(hipFree(c.ws.base));
        c.ws = Workspace{};
    }
    for (TableRange& t : c.tabs)
        if (int rc = free_tables(c, t)) return rc;
    c.ws_graphed = false;
    c.placed_for = 0;
    return CK_OK;
}

// Grows the workspace (one allocation).  CK_ENOMEM when the device cannot
// hold `bytes`.
static int ensure_ws(Ctx& c, size_t bytes) {
    if (c.ws.bytes >= bytes) return CK_OK;
    if (c.ws_static)
        return fail(CK_EINVAL, "this launch needs a " + std::to_string(bytes >> 20) +
                                   " MiB workspace; ck_launch_static's array holds " +
                                   std::to_string(c.ws.bytes >> 20) + " MiB (one trial of this shape does not fit)");
    if (int rc = release_ws(c)) return rc;
    Workspace reuse;
    if (take_retired(c.device, bytes, reuse)) {
        c.ws = reuse;
        c.ws_graphed = true;
        return CK_OK;
    }
    const hipError_t e = hipMalloc(&c.ws.base, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        c.ws.base = nullptr;
        return fail(CK_ENOMEM, "device memory: cannot allocate a " + std::to_string(bytes >> 20) + " MiB workspace");
    }
    if (e != hipSuccess) return fail(CK_EDEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    c.ws.bytes = bytes;
    return CK_OK;
}

struct PassArgs {
    Shape S;
    RoundKeys bk;
    uint32_t first, count;
    char* map;
    size_t map_stride;
    TabRef tables;
    size_t table_bytes;
    uint64_t* walks;
    uint64_t* rnds;
    uint32_t* cnt;
    uint32_t* fail;
    uint32_t* probes;  // null unless requested
    uint32_t* keys;    // null unless requested
    unsigned long long* digests;  // null unless requested
    uint32_t clear;    // the build clears the tables
    uint64_t tag;      // descriptor-entry tag of the pass
    uint32_t item0, item1;  // items whose candidates the map holds (residual windows; else 0, pnum)
    uint32_t* state;   // residual windows: per-slot chain state; else null
};

template <int K, class M>
static void launch_build_t(const BuildPlan& bp, const PassArgs& a, Dispatch& st) {
    const uint32_t blocks = (a.count + bp.waves - 1) / bp.waves;
    M* map = reinterpret_cast<M*>(a.map);
    if (bp.cached)
        st(&k_build<K, true, M>, dim3(blocks), dim3(64 * bp.waves), bp.lds, a.S, a.bk, a.first, a.count, kRepl, map,
           a.map_stride, a.tables, a.table_bytes, a.walks, a.rnds, a.keys, a.clear, a.item0, a.item1);
    else
        st(&k_build<K, false, M>, dim3(blocks), dim3(64 * bp.waves), bp.lds, a.S, a.bk, a.first, a.count, kRepl, map,
           a.map_stride, a.tables, a.table_bytes, a.walks, a.rnds, a.keys, a.clear, a.item0, a.item1);
}
template <int K>
static void launch_build_k(const BuildPlan& bp, const HbmPlan& h, const PassArgs& a, Dispatch& st) {
    if (h.desc || h.stream64)
        launch_build_t<K, uint64_t>(bp, a, st);
    else if (h.map16)
        launch_build_t<K, uint16_t>(bp, a, st);
    else
        launch_build_t<K, uint32_t>(bp, a, st);
}

template <int K, bool BGH, class M, class E, bool XCHG>
static void launch_chain_t(const PassArgs& a, Dispatch& st) {
    const uint32_t blocks = (a.count + 255) / 256;
    st(&k_chain<K, BGH, M, E, XCHG>, dim3(blocks), dim3(256), 0, a.S, a.first, a.count,
       reinterpret_cast<const M*>(a.map), a.map_stride, a.tables, a.table_bytes, a.walks, a.rnds, a.cnt, a.fail,
       a.probes, a.tag);
}
template <int K, class M, class E>
static void launch_digest_t(const PassArgs& a, Dispatch& st) {
    st(&k_digest<K, M, E>, dim3((a.count + 255) / 256), dim3(256), 0, a.S, a.count, reinterpret_cast<const M*>(a.map),
       a.map_stride, a.tables, a.table_bytes, a.digests);
}
template <int K>
static void launch_digest_k(const HbmPlan& h, const PassArgs& a, Dispatch& st) {
    if (h.lane)
        return;  // k_chain_lane wrote the digests from its LDS tables
    if (h.res && h.e32)
        st(&k_digest_res<K, uint32_t>, dim3((a.count + 255) / 256), dim3(256), 0, a.S, a.count, a.tables,
           a.table_bytes, a.digests);
    else if (h.res && h.e44)
        st(&k_digest_res<K, E44>, dim3((a.count + 255) / 256), dim3(256), 0, a.S, a.count, a.tables,
           a.table_bytes, a.digests);
    else if (h.res && h.e48)
        st(&k_digest_res<K, E48>, dim3((a.count + 255) / 256), dim3(256), 0, a.S, a.count, a.tables,
           a.table_bytes, a.digests);
    else if (h.res)
        st(&k_digest_res<K, uint64_t>, dim3((a.count + 255) / 256), dim3(256), 0, a.S, a.count, a.tables,
           a.table_bytes, a.digests);
    else if (h.desc)
        launch_digest_t<K, uint64_t, uint64_t>(a, st);
    else if (h.map16 && h.table16)
        launch_digest_t<K, uint16_t, uint16_t>(a, st);
    else if (h.map16)
        launch_digest_t<K, uint16_t, uint32_t>(a, st);
    else if (h.table16)
        launch_digest_t<K, uint32_t, uint16_t>(a, st);
    else
        launch_digest_t<K, uint32_t, uint32_t>(a, st);
}

template <int K, bool BGH, class M, class E>
static void launch_chain_res_t(const PassArgs& a, Dispatch& st) {
    const uint32_t blocks = (a.count + 255) / 256;
    st(&k_chain_res<K, BGH, M, E>, dim3(blocks), dim3(256), 0, a.S, a.first, a.count,
       reinterpret_cast<const M*>(a.map), a.map_stride, a.tables, a.table_bytes, a.walks, a.rnds, a.cnt, a.fail,
       a.probes, a.item0, a.item1, a.state, a.state ? 1u : 0u);
}
template <int K, bool BGH, class E>
static void launch_chain_lane_t(const HbmPlan& h, const PassArgs& a, Dispatch& st) {
    const uint32_t lanes = h.lp.lanes;
    st(&k_chain_lane<K, BGH, E>, dim3((a.count + lanes - 1) / lanes), dim3(64), h.lp.lds, a.S, a.first, a.count, lanes,
       reinterpret_cast<const uint64_t*>(a.map), a.map_stride, static_cast<const uint64_t*>(a.walks),
       static_cast<const uint64_t*>(a.rnds), a.cnt, a.fail, a.probes, a.digests);
}
template <int K, bool BGH>
static void launch_chain_kb(const HbmPlan& h, const PassArgs& a, Dispatch& st) {
    if (h.lane && h.lp.e16)
        launch_chain_lane_t<K, BGH, uint16_t>(h, a, st);
    else if (h.lane)
        launch_chain_lane_t<K, BGH, uint32_t>(h, a, st);
    else if (h.res && h.stream64 && h.e32)
        launch_chain_res_t<K, BGH, uint64_t, uint32_t>(a, st);
    else if (h.res && h.stream64 && h.e44)
        launch_chain_res_t<K, BGH, uint64_t, E44>(a, st);
    else if (h.res && h.stream64 && h.e48)
        launch_chain_res_t<K, BGH, uint64_t, E48>(a, st);
    else if (h.res && h.stream64)
        launch_chain_res_t<K, BGH, uint64_t, uint64_t>(a, st);
    else if (h.res && h.e32)
        launch_chain_res_t<K, BGH, uint32_t, uint32_t>(a, st);
    else if (h.res && h.e44)
        launch_chain_res_t<K, BGH, uint32_t, E44>(a, st);
    else if (h.res && h.e48)
        launch_chain_res_t<K, BGH, uint32_t, E48>(a, st);
    else if (h.res)
        launch_chain_res_t<K, BGH, uint32_t, uint64_t>(a, st);
    else if (h.desc && h.xchg)
        launch_chain_t<K, BGH, uint64_t, uint64_t, true>(a, st);
    else if (h.desc)
        launch_chain_t<K, BGH, uint64_t, uint64_t, false>(a, st);
    else if (h.xchg && h.map16)
        launch_chain_t<K, BGH, uint16_t, uint32_t, true>(a, st);
    else if (h.xchg)
        launch_chain_t<K, BGH, uint32_t, uint32_t, true>(a, st);
    else if (h.map16 && h.table16)
        launch_chain_t<K, BGH, uint16_t, uint16_t, false>(a, st);
    else if (h.map16)
        launch_chain_t<K, BGH, uint16_t, uint32_t, false>(a, st);
    else if (h.table16)
        launch_chain_t<K, BGH, uint32_t, uint16_t, false>(a, st);
    else
        launch_chain_t<K, BGH, uint32_t, uint32_t, false>(a, st);
}
template <int K>
static void launch_chain_k(bool bgh, const HbmPlan& h, const PassArgs& a, Dispatch& st) {
    if (bgh)
        launch_chain_kb<K, true>(h, a, st);
    else
        launch_chain_kb<K, false>(h, a, st);
}
static void launch_build(uint32_t k, const BuildPlan& bp, const HbmPlan& h, const PassArgs& a, Dispatch& st) {
    switch (k) {
        case 2: launch_build_k<2>(bp, h, a, st); break;
        case 3: launch_build_k<3>(bp, h, a, st); break;
        default: launch_build_k<4>(bp, h, a, st); break;
    }
}
static void launch_chain(uint32_t k, bool bgh, const HbmPlan& h, const PassArgs& a, Dispatch& st) {
    switch (k) {
        case 2: launch_chain_k<2>(bgh, h, a, st); break;
        case 3: launch_chain_k<3>(bgh, h, a, st); break;
        default: launch_chain_k<4>(bgh, h, a, st); break;
    }
}

template <int K, bool BGH>
static void launch_lds_kb(const LdsPlan& L, uint32_t blocks, const Shape& S, const RoundKeys& bk, uint32_t first,
                          uint32_t count, uint32_t* cnt, uint32_t* fo, uint32_t* probes, uint32_t* tables,
                          unsigned long long* digests, hipStream_t st) {
    const dim3 grid(blocks), block(64 * L.waves);
    const uint32_t wb = static_cast<uint32_t>(L.wave_b), cb = static_cast<uint32_t>(L.cache_b),
                   mb = static_cast<uint32_t>(L.map_b);
#define LDS_ARGS S, bk, first, count, L.tab_shift, wb, cb, mb, cnt, fo, probes, tables, digests
    if (L.round && L.bins16)
        hipLaunchKernelGGL((k_lds<K, BGH, uint16_t, uint16_t, true>), grid, block, L.lds, st, LDS_ARGS);
    else if (L.round)
        hipLaunchKernelGGL((k_lds<K, BGH, uint16_t, uint32_t, true>), grid, block, L.lds, st, LDS_ARGS);
    else if (L.map16 && L.bins16)
        hipLaunchKernelGGL((k_lds<K, BGH, uint16_t, uint16_t, false>), grid, block, L.lds, st, LDS_ARGS);
    else if (L.map16)
        hipLaunchKernelGGL((k_lds<K, BGH, uint16_t, uint32_t, false>), grid, block, L.lds, st, LDS_ARGS);
    else if (L.bins16)
        hipLaunchKernelGGL((k_lds<K, BGH, uint32_t, uint16_t, false>), grid, block, L.lds, st, LDS_ARGS);
    else
        hipLaunchKernelGGL((k_lds<K, BGH, uint32_t, uint32_t, false>), grid, block, L.lds, st, LDS_ARGS);
#undef LDS_ARGS
}
template <int K>
static void launch_lds(bool bgh, const LdsPlan& L, uint32_t blocks, const Shape& S, const RoundKeys& bk,
                       uint32_t first, uint32_t count, uint32_t* cnt, uint32_t* fo, uint32_t* probes,
                       uint32_t* tables, unsigned long long* digests, hipStream_t st) {
    if (bgh)
        launch_lds_kb<K, true>(L, blocks, S, bk, first, count, cnt, fo, probes, tables, digests, st);
    else
        launch_lds_kb<K, false>(L, blocks, S, bk, first, count, cnt, fo, probes, tables, digests, st);
}

template <class Fn>
static int for_each_lds_kernel(Fn&& fn) {
#define ONE(K, B, M, E, R)                                                           \
    {                                                                                \
        const int rc = fn(reinterpret_cast<const void*>(&k_lds<K, B, M, E, R>));     \
        if (rc) return rc;                                                           \
    }
#define ALL_ME(K, B)                                                                                   \
    ONE(K, B, uint16_t, uint16_t, false) ONE(K, B, uint16_t, uint32_t, false) ONE(K, B, uint32_t, uint16_t, false) \
    ONE(K, B, uint32_t, uint32_t, false) ONE(K, B, uint16_t, uint16_t, true) ONE(K, B, uint16_t, uint32_t, true)
    ALL_ME(2, true) ALL_ME(2, false) ALL_ME(3, true) ALL_ME(3, false) ALL_ME(4, true) ALL_ME(4, false)
#undef ALL_ME
#undef ONE
    return CK_OK;
}

static int set_lds_limits(size_t lds_block) {
    const int lim = static_cast<int>(lds_block);
    auto opt = [&](const void* f) -> int {
        CKH(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lim));
        return CK_OK;
    };
    int rc;
#define B1(K, C)                                                                                          \
    reinterpret_cast<const void*>(&k_build<K, C, uint16_t>), reinterpret_cast<const void*>(&k_build<K, C, uint32_t>), \
        reinterpret_cast<const void*>(&k_build<K, C, uint64_t>)
    for (const void* f : {B1(2, true), B1(2, false), B1(3, true), B1(3, false), B1(4, true), B1(4, false)})
        if ((rc = opt(f))) return rc;
#undef B1
#define L1(K, B)                                                                   \
    reinterpret_cast<const void*>(&k_chain_lane<K, B, uint16_t>), reinterpret_cast<const void*>(&k_chain_lane<K, B, uint32_t>)
    for (const void* f : {L1(2, true), L1(2, false), L1(3, true), L1(3, false), L1(4, true), L1(4, false)})
        if ((rc = opt(f))) return rc;
#undef L1
    return for_each_lds_kernel(opt);
}

struct HostOut {
    std::vector<uint32_t>* fails;
    uint32_t* probes;   // per trial, may be null
    uint32_t* tables;   // per trial x buckets (ordinal | choice << 30), may be null
    uint32_t* keys;     // HBM tier debug: per trial 44 words
    uint64_t* walks;
    uint64_t* rnds;
    uint32_t* map;      // HBM tier debug: per trial pnum*k
    uint64_t* digests;  // per trial table digest, may be null
};

static int take_counters(const uint32_t* cnt, const uint32_t* fo, uint32_t n, HostOut& out, ck_stats* st) {
    uint32_t counters[4] = {0, 0, 0, 0};
    CKH(hipMemcpy(counters, cnt, sizeof counters, hipMemcpyDeviceToHost));
    const uint32_t nf = counters[CNT_FAIL];
    if (nf > n) return fail(CK_ERANGE, "device failure count out of range");
    if (nf) {
        const size_t at = out.fails->size();
        out.fails->resize(at + nf);
        CKH(hipMemcpy(out.fails->data() + at, fo, size_t(nf) * 4, hipMemcpyDeviceToHost));
    }
    if (st) st->probes += (uint64_t(counters[CNT_PROBES + 1]) << 32) | counters[CNT_PROBES];
    return CK_OK;
}

// Copies n rows of `width` entries (u16 or u32, row pitch `pitch` bytes) into
// u32 rows of `width` entries, widening table entries to ordinal | choice << 30.
static int copy_rows(uint32_t* dst, const char* src, size_t pitch, size_t width, uint32_t n, bool src16, bool entries) {
    if (!src16) {
        CKH(hipMemcpy2D(dst, width * 4, src, pitch, width * 4, n, hipMemcpyDeviceToHost));
        return CK_OK;
    }
    std::vector<uint16_t> tmp(width * n);
    CKH(hipMemcpy2D(tmp.data(), width * 2, src, pitch, width * 2, n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < tmp.size(); ++i) {
        const uint32_t v = tmp[i];
        dst[i] = !entries ? v : v == 0xffffu ? EMPTY : ((v & 0x3fffu) | ((v >> 14) << 30));
    }
    return CK_OK;
}

// Table reuse without a clear (descriptor layout with load + store probes).
// Entries carry the tag of the pass that wrote them in the bits above the
// choice, and entries of any other tag read as empty.  A buffer's tables
// therefore need clearing only on first use, when the plan of the tables in
// the workspace changes (position, size, tag position), when the tag field is
// used up (all-ones, the cleared value, is never a tag), and when a pass uses
// more slots than the buffer's last clear covered (a clear covers the slots of
// its own pass only).  A clear writes 192 KB per trial at n = 2^14, traffic
// that slows the other pass's chains.
static uint64_t table_plan(const Ctx& c, const ck_params& p, const HbmPlan& h, uint32_t nbuf, bool taggable) {
    uint64_t x = reinterpret_cast<uintptr_t>(c.ws.base);
    for (uint64_t v : {uint64_t(reinterpret_cast<uintptr_t>(c.tabs[0].base)),
                       uint64_t(reinterpret_cast<uintptr_t>(c.tabs[1].base)), uint64_t(h.buffer), uint64_t(h.off_tables), uint64_t(h.table_bytes), uint64_t(nbuf),
                       uint64_t(p.k) * make_shape(p).dbits, uint64_t(taggable)}) {
        x = (x ^ v) * 0x9E3779B97F4A7C15ull;
        x ^= x >> 29;
    }
    return x | 1;  // 0 means no plan
}
static void pass_tags(Ctx& c, uint32_t b, bool taggable, uint32_t tag_bits, PassArgs& a) {
    auto& T = c.table_tags[b];
    const uint64_t max_tag = (uint64_t(1) << std::min<uint32_t>(tag_bits, 63)) - 2;
    if (taggable && T.valid && a.count <= T.slots && T.tag < max_tag) {
        a.clear = 0;
        a.tag = ++T.tag;
        return;
    }
    a.clear = 1;
    a.tag = 0;
    T.valid = taggable;
    T.slots = a.count;
    T.tag = 0;
}

// Graph mode of the HBM tier.  The first call for a (shape, trial count,
// layout, pass size, workspace) builds a graph of every pass -- build and chain
// nodes, with the dependencies the two streams give in the eager form (builds
// in order; chains in order; a chain after its build; a build after the chain
// that last used its buffer) -- and later calls rewrite only the batch key and
// offsets of the build and chain nodes before replaying it.  Every pass writes
// its own counters and failure list, read back once the whole graph has run;
// one kernel ahead of the graph, on the stream it is launched on, clears the
// counters of every pass.  (As graph nodes of their own, without dependencies,
// the clears were spread over the executor's queues, and a chain whose clear
// sat behind the next pass's build on the build queue waited for that build:
// measured at the headline, every other build ran alone, 0.90x the eager
// rate.)  Device time is reported as chain_ms for the whole replay (build_ms
// stays 0: the kernels of a graph are not timed one by one).
static int run_hbm_graph(Ctx& c, const ck_params& p, uint32_t first, uint32_t count, uint32_t layout, uint32_t nbuf,
                         uint32_t chunk, const HbmPlan& h, const BuildPlan& bp, PassArgs (&args)[2], bool taggable,
                         uint32_t tag_bits, HostOut& out, ck_stats* st) {
    const uint32_t passes = (count + chunk - 1) / chunk;
    const size_t stride = (CNT_WORDS + size_t(chunk) + 63) & ~size_t(63);  // u32 words per pass, 256-byte aligned
    const bool bgh = p.mode == CK_MODE_BGH25;
    HbmGraph* hit = nullptr;
    HbmGraph* lru = &c.graphs[0];
    for (HbmGraph& x : c.graphs) {
        if (x.exec && x.ws == c.ws.base && x.count == count && x.layout == layout && x.chunk == chunk &&
            x.nbuf == nbuf && std::memcmp(&x.p, &p, sizeof p) == 0)
            hit = &x;
        if (x.last_use < lru->last_use) lru = &x;
    }
    const bool same = hit != nullptr;
    HbmGraph& G = same ? *hit : *lru;
    G.last_use = ++c.graph_clock;
    if (!same) {
        G.drop();
        if (G.aux_words < stride * passes) {
            if (G.aux) CKH(hipFree(G.aux));
            G.aux = nullptr;
            G.aux_words = 0;
            CKH(hipMalloc(&G.aux, stride * passes * 4));
            G.aux_words = stride * passes;
        }
    }
    // Pass i's arguments: its buffer's regions, its trial range, its counters.
    // Called once per pass, in pass order, for every build or replay.
    auto pass_args = [&](uint32_t i) {
        const uint32_t b = nbuf == 2 ? (i & 1) : 0;
        PassArgs a = args[b];
        a.first = first + i * chunk;
        a.count = std::min(chunk, count - i * chunk);
        a.cnt = G.aux + i * stride;
        a.fail = a.cnt + CNT_WORDS;
        pass_tags(c, b, taggable, tag_bits, a);
        return a;
    };
    if (!same) {
        CKH(hipGraphCreate(&G.g, 0));
        G.build.assign(passes, nullptr);
        G.chain.assign(passes, nullptr);
        for (uint32_t i = 0; i < passes; ++i) {
            const PassArgs a = pass_args(i);
            std::vector<hipGraphNode_t> deps;
            if (i) deps.push_back(G.build[i - 1]);
            if (i >= nbuf) deps.push_back(G.chain[i - nbuf]);
            Dispatch bd;
            bd.graph = G.g;
            bd.deps = deps.data();
            bd.ndeps = deps.size();
            launch_build(p.k, bp, h, a, bd);
            CKH(bd.err);
            G.build[i] = bd.node;
            std::vector<hipGraphNode_t> cdeps{G.build[i]};
            if (i) cdeps.push_back(G.chain[i - 1]);
            Dispatch cd;
            cd.graph = G.g;
            cd.deps = cdeps.data();
            cd.ndeps = cdeps.size();
            launch_chain(p.k, bgh, h, a, cd);
            CKH(cd.err);
            G.chain[i] = cd.node;
        }
        CKH(hipGraphInstantiate(&G.exec, G.g, nullptr, nullptr, 0));
        c.ws_graphed = true;
        G.p = p;
        G.ws = c.ws.base;
        G.count = count;
        G.layout = layout;
        G.chunk = chunk;
        G.nbuf = nbuf;
    } else {
        for (uint32_t i = 0; i < passes; ++i) {
            const PassArgs a = pass_args(i);
            Dispatch bd;
            bd.exec = G.exec;
            bd.node = G.build[i];
            launch_build(p.k, bp, h, a, bd);
            CKH(bd.err);
            Dispatch cd;
            cd.exec = G.exec;
            cd.node = G.chain[i];
            launch_chain(p.k, bgh, h, a, cd);
            CKH(cd.err);
        }
    }
    CKH(hipEventRecord(c.ev[0], c.build_stream));
    const uint64_t zero_words = uint64_t(passes) * CNT_WORDS;
    hipLaunchKernelGGL(k_zero_rows, dim3(static_cast<uint32_t>((zero_words + 255) / 256)), dim3(256), 0,
                       c.build_stream, G.aux, stride, CNT_WORDS, uint64_t(passes));
    CKH(hipGetLastError());
    CKH(hipGraphLaunch(G.exec, c.build_stream));
    CKH(hipEventRecord(c.ev[1], c.build_stream));
    CKH(hipEventSynchronize(c.ev[1]));
    std::vector<uint32_t> cnt(size_t(passes) * 4);
    CKH(hipMemcpy2D(cnt.data(), 16, G.aux, stride * 4, 16, passes, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < passes; ++i) {
        const uint32_t n = std::min(chunk, count - i * chunk), nf = cnt[i * 4 + CNT_FAIL];
        if (nf > n) return fail(CK_ERANGE, "device failure count out of range");
        if (nf) {
            const size_t at = out.fails->size();
            out.fails->resize(at + nf);
            CKH(hipMemcpy(out.fails->data() + at, G.aux + i * stride + CNT_WORDS, size_t(nf) * 4,
                          hipMemcpyDeviceToHost));
        }
        if (st) st->probes += (uint64_t(cnt[i * 4 + CNT_PROBES + 1]) << 32) | cnt[i * 4 + CNT_PROBES];
    }
    if (st) {
        float ms = 0;
        CKH(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
        st->chain_ms += ms;
        st->span_ms += ms;
        st->launches += 2 * passes + 1;
        st->graph = same ? 1 : 2;
    }
    return CK_OK;
}

// What an HBM-tier launch of `count` trials will use: the table layout, item
// windows, pass buffers, trials per pass, the workspace plan and the build.
struct HbmDecision {
    uint32_t layout = 0, window = 0, nbuf = 0, chunk = 0;
    bool graph = false;
    HbmPlan h{};
    BuildPlan bp{};
};
// `tables`: the caller wants final tables (ordinal entries); `graph`: the launch
// runs as a graph; `windows_ok`: item windows are allowed (no per-trial map,
// walk or random-word output).  Uses c.budget.
static int decide_hbm(const Ctx& c, const ck_params& p, uint32_t count, bool tables, bool graph, bool windows_ok,
                      HbmDecision& d, bool lane = false) {
    if (c.prp == CK_PRP_CACHE && !cache_possible(p, c.lds_block))
        return fail(CK_EINVAL, "a round cache needs half_bits <= 12 (this shape evaluates AES directly)");
    if (lane) {
        // Lane tier: the HBM tier's build and candidate stream, the tables in LDS.
        const LanePlan L = plan_lane(p, c.lds_block, c.lds_cu);
        if (!L.ok)
            return fail(CK_EINVAL, "shape does not fit the lane tier (residual entries above 32 bits, or one "
                                   "trial's table and ring above the LDS of a block)");
        if (tables) return fail(CK_EINVAL, "the lane tier keeps no ordinal tables; use the HBM map layout for tables");
        d.layout = kLayoutLane;
        d.graph = graph;
        d.window = p.pnum;
        d.nbuf = hbm_buffers(c, p, count, kLayoutLane);
        d.chunk = hbm_chunk(c, p, count, kLayoutLane, d.nbuf, p.pnum);
        d.h = plan_hbm(p, d.chunk, kLayoutLane, p.pnum);
        d.h.lp = L;
        d.bp = plan_build(p, c.lds_block, c.prp);
        return CK_OK;
    }
    // Auto: descriptor entries with load + store probes where they fit (fastest
    // with tens of thousands of chains in flight, and reusable without a clear);
    // otherwise residual entries, still one random read per probe (measured
    // against the bucket map at n = 40000, k = 4, m = 50000: 164K vs 141K
    // trials/s; at n = 2^20, memory-capped: 2416 vs 2170); otherwise ordinal
    // entries, with exchange probes when few chains are in flight (one round
    // trip per probe matters when latency, not throughput, bounds the pass).
    // Residual entries in 32 bits come first where they fit: half the descriptor's
    // table, and measured faster at n = 2^14 (chain 86.6 vs 94.5 ms per
    // 65,536-trial pass, 746K vs 685K trials/s, three interleaved pairs).
    uint32_t layout = c.layout;
    if (layout == CK_LAYOUT_AUTO) {
        if (!tables && res_e32(p) && res_stream64(p))
            layout = CK_LAYOUT_RESIDUAL;
        else if (!tables && desc_ok(p))
            layout = CK_LAYOUT_DESC_PLAIN;
        else if (!tables && res_ok(p))
            layout = CK_LAYOUT_RESIDUAL;
        else
            layout = hbm_chunk(c, p, count, CK_LAYOUT_MAP_XCHG, hbm_buffers(c, p, count, CK_LAYOUT_MAP_XCHG),
                               p.pnum) < kXchgBelow
                         ? CK_LAYOUT_MAP_XCHG
                         : CK_LAYOUT_MAP;
    }
    const bool desc = layout == CK_LAYOUT_DESC || layout == CK_LAYOUT_DESC_PLAIN;
    if (desc && !desc_ok(p)) return fail(CK_EINVAL, "descriptor layout needs k*bits(buckets-1) + 2 <= 63");
    if (desc && tables) return fail(CK_EINVAL, "the descriptor layout does not keep ordinals; use the map layout for tables");
    if (layout == CK_LAYOUT_RESIDUAL && !res_ok(p))
        return fail(CK_EINVAL, "residual layout needs (k-1)*bits(buckets-1) + 2 <= 63");
    if (layout == CK_LAYOUT_RESIDUAL && tables)
        return fail(CK_EINVAL, "the residual layout does not keep ordinals; use the map layout for tables");
    d.layout = layout;
    d.graph = graph;
    d.window = stream_window(c, p, count, layout, !graph && windows_ok);
    d.nbuf = d.window < p.pnum ? 1 : hbm_buffers(c, p, count, layout);
    d.chunk = hbm_chunk(c, p, count, layout, d.nbuf, d.window);
    d.h = plan_hbm(p, d.chunk, layout, d.window);
    d.bp = plan_build(p, c.lds_block, c.prp);
    return CK_OK;
}

// Table placement.  The chain rate of a pass belongs to the memory its tables
// land in.  The device's memory falls into three regions of about 96 GB
// (found by pairing 1.5-2 GiB allocations: a dependent random read-write walk
// with half its lanes in each of two allocations runs at 22.7 G steps/s when
// both lie in one region and at 29.1 G when they lie in two, and every
// allocation pairs slow with one third of the others), and random accesses
// confined to one region run slower than accesses spread over several: 65,536
// lanes walking 96 KB tables made 20.5 G steps/s in one region, 24.7 G split
// 2:1 over two, 27.3 G split evenly over three.  Allocations fill the regions
// in turn, so an ordinary buffer of a few GB lies in one region (the chains'
// 93-95 ms per pass at n = 2^14), and one that happens to straddle a boundary
// runs fast (73-79 ms): neither the contiguous-allocation flag nor any
// allocation size changes that.
//
// A double-buffered plan (placeable) therefore holds the tables of each pass
// buffer apart from it.  place_regions allocates pieces of a third of a pass's
// tables (virtual-memory API; kRegionMinPiece at least, since the pair walk
// needs 768 MiB per piece to stay out of the caches -- small tables use the
// start of each piece), as many as the budget holds, sorts them into regions
// by pair walks (k_place_pair) and gives each pass buffer one piece from each
// of three regions (2:1 over two where only two show), every third slot's
// table in each piece (TabRef n = 3), so that the lanes of a wave and the
// slots still running at a pass's end spread over the regions too.  Small
// tables gain as much as large ones (n = 1024 to 8192: 15-23 % more trials/s).
// Where fewer than two regions show or the budget holds fewer than six
// pieces, place_buffers falls back to a search: up to place_cands candidate
// table allocations, k_place_walk (the chains' access pattern) over each, the
// two fastest kept.  The rest of both pass buffers (candidate maps, per-slot
// words) shares one allocation; spreading it too changed nothing.  Placement
// reruns only when a plan needs tables of another size, so it is for contexts
// that run many passes of one shape (ck_ctx_set_place_search turns it on; the
// CLI and the headline bench do).  Results do not depend on it: every pass
// computes the same trials anywhere.
// Placed plans: passes of at least kPlaceMinChains chains (a walk over fewer
// lanes does not load the random-access path) in buffers of at least 256 MiB.
static constexpr size_t kPlaceMinBuffer = size_t(256) << 20;
static constexpr uint32_t kPlaceWalkSteps = 4096;
static constexpr uint32_t kPlaceMinChains = 1u << 14;
static constexpr size_t kRegionMinPiece = size_t(1) << 30;
static constexpr size_t kRegionPieceAlign = size_t(2) << 20;
static constexpr uint32_t kRegionMaxPieces = 192;
static constexpr uint32_t kRegionPairLanes = 65536, kRegionPairSlots = 3072, kRegionPairSteps = 1024;
static constexpr float kRegionModes = 1.12f;  // slow / fast pair time at least this: two regions seen
static constexpr uint32_t kRegionCands = 6;   // candidate table ranges timed
static constexpr float kRegionNear = 1.06f;   // within this factor of a mode: in that mode
static constexpr uint32_t kRegionBatch = 8;   // pieces allocated between classifications

// tables_out: the launch returns the tables, which it copies as one range.
static bool placeable(const Ctx& c, const HbmDecision& d, bool tables_out) {
    return c.place_cands >= 2 && !c.ws_static && !tables_out && d.nbuf == 2 && d.h.nwin == 1 &&
           !d.h.lane && d.h.table_bytes > 0 && d.chunk >= kPlaceMinChains && d.h.buffer >= kPlaceMinBuffer &&
           c.budget >= 2 * d.h.off_tables + 3 * (d.h.buffer - d.h.off_tables);
}

// Event-timed launch of `launch` on stream s: once to warm the translations,
// once timed.
template <typename F>
static hipError_t timed_twice(hipStream_t s, hipEvent_t e0, hipEvent_t e1, F&& launch, float* ms) {
    for (int rep = 0; rep < 2; ++rep) {
        hipError_t e = hipEventRecord(e0, s);
        if (e == hipSuccess) {
            launch();
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

static void region_access(int device, hipMemAllocationProp* prop, hipMemAccessDesc* acc) {
    *prop = hipMemAllocationProp{};
    prop->type = hipMemAllocationTypePinned;
    prop->location.type = hipMemLocationTypeDevice;
    prop->location.id = device;
    *acc = hipMemAccessDesc{};
    acc->location = prop->location;
    acc->flags = hipMemAccessFlagsProtReadWrite;
}

// Pieces are allocated in batches and sorted as they come, and allocation
// stops once every region holds `want` of them: a device whose memory was
// freed in scattered pieces shows all three regions within a few batches (and
// makes each new piece wait for the driver to clear it), a fresh one only
// after the first two regions are used up.  t0: each piece paired with piece
// 0; t1: each piece of another region than piece 0's paired with ref1, the one
// most clearly apart from piece 0.  Pieces near the slow mode of t0 share
// piece 0's region; near the fast one they do not, and a second split among
// them in t1, where it shows, is a third region; pieces in between are left
// out (-1).  (Where the memory after the first region was handed out in
// pieces that each span two regions, no second split shows.)
int region_pool(int device, hipStream_t stream, hipEvent_t e0, hipEvent_t e1, unsigned long long* sink,
                size_t piece, size_t max_pieces, size_t want, RegionPool* out, std::string* err) {
    RegionPool& R = *out;
    R = RegionPool{};
    R.piece = piece;
    if (piece < kRegionMinPiece || piece % kRegionPieceAlign || max_pieces < 2) return CK_OK;
    hipMemAllocationProp prop;
    hipMemAccessDesc acc;
    region_access(device, &prop, &acc);
    if (hipMemAddressReserve(reinterpret_cast<void**>(&R.pool), max_pieces * piece, 0, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        R.pool = nullptr;
        return CK_OK;
    }
    R.reserved = max_pieces * piece;
    auto pair_ms = [&](size_t a, size_t b, float* ms) {
        return timed_twice(
            stream, e0, e1,
            [&]() {
                hipLaunchKernelGGL(k_place_pair, dim3(kRegionPairLanes / 256), dim3(256), 0, stream,
                                   reinterpret_cast<uint64_t*>(R.pool + a * piece),
                                   reinterpret_cast<uint64_t*>(R.pool + b * piece), kRegionPairLanes / 2,
                                   kRegionPairSlots, kRegionPairSteps, sink);
            },
            ms);
    };
    std::vector<float> t0, t1;
    size_t ref1 = 0;
    bool full = false;
    while (!full) {
        const size_t from = R.ready;
        for (size_t k = 0; k < kRegionBatch && R.ready < max_pieces; ++k) {
            hipMemGenericAllocationHandle_t x{};
            if (hipMemCreate(&x, piece, &prop, 0) != hipSuccess) {  // the device is full (or has no such API)
                (void)hipGetLastError();
                full = true;
                break;
            }
            R.h.push_back(x);
            hipError_t e = hipMemMap(R.pool + R.ready * piece, piece, 0, x, 0);
            if (e == hipSuccess) {
                ++R.mapped;
                e = hipMemSetAccess(R.pool + R.ready * piece, piece, &acc, 1);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                full = true;
                break;
            }
            ++R.ready;
        }
        if (R.ready == max_pieces) full = true;
        if (R.ready == from) break;
        const size_t n = R.ready;
        t0.resize(n, 0.f);
        t1.resize(n, 0.f);
        R.region.resize(n, 0);
        for (size_t j = std::max<size_t>(from, 1); j < n; ++j)
            if (hipError_t e = pair_ms(0, j, &t0[j])) {
                *err = std::string("pair walk: ") + hipGetErrorString(e);
                return CK_EDEVICE;
            }
        std::vector<size_t> all;
        for (size_t j = 1; j < n; ++j) all.push_back(j);
        float f0 = 0, s0 = 0;
        R.two = split_modes(t0, all, kRegionModes, &f0, &s0);
        std::vector<size_t> others;
        for (size_t j = 1; j < n; ++j) {
            R.region[j] = !R.two ? 0 : t0[j] >= s0 / kRegionNear ? 0 : t0[j] <= f0 * kRegionNear ? 1 : -1;
            if (R.region[j] == 1) others.push_back(j);
        }
        if (R.two && !ref1 && !others.empty()) {
            ref1 = others[0];
            for (size_t j : others)
                if (t0[j] < t0[ref1]) ref1 = j;
        }
        std::vector<size_t> others1;
        for (size_t j : others)
            if (j != ref1) {
                if (t1[j] == 0)
                    if (hipError_t e = pair_ms(ref1, j, &t1[j])) {
                        *err = std::string("pair walk: ") + hipGetErrorString(e);
                        return CK_EDEVICE;
                    }
                others1.push_back(j);
            }
        float f1 = 0, s1 = 0;
        if (ref1 && split_modes(t1, others1, kRegionModes, &f1, &s1))
            for (size_t j : others1) R.region[j] = t1[j] >= s1 / kRegionNear ? 1 : t1[j] <= f1 * kRegionNear ? 2 : -1;
        if (R.of(0).size() >= want && R.of(1).size() >= want && R.of(2).size() >= want) break;
    }
    return CK_OK;
}

int map_pieces(int device, const RegionPool& p, const std::vector<size_t>& pick, void** va, std::string* err) {
    *va = nullptr;
    hipMemAllocationProp prop;
    hipMemAccessDesc acc;
    region_access(device, &prop, &acc);
    const size_t bytes = pick.size() * p.piece;
    if (hipMemAddressReserve(va, bytes, 0, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        *va = nullptr;
        *err = "hipMemAddressReserve of " + std::to_string(bytes >> 20) + " MiB";
        return CK_EDEVICE;
    }
    size_t n = 0;
    while (n < pick.size() && hipMemMap(static_cast<char*>(*va) + n * p.piece, p.piece, 0, p.h[pick[n]], 0) == hipSuccess)
        ++n;
    if (n == pick.size() && hipMemSetAccess(*va, bytes, &acc, 1) == hipSuccess) return CK_OK;
    (void)hipGetLastError();
    for (size_t i = 0; i < n; ++i) (void)hipMemUnmap(static_cast<char*>(*va) + i * p.piece, p.piece);
    (void)hipMemAddressFree(*va, bytes);
    *va = nullptr;
    *err = "mapping pieces into a range";
    return CK_EDEVICE;
}

void unmap_pieces(void* va, size_t piece, size_t n) {
    if (!va) return;
    for (size_t i = 0; i < n; ++i) (void)hipMemUnmap(static_cast<char*>(va) + i * piece, piece);
    (void)hipMemAddressFree(va, n * piece);
}

void drop_pool(RegionPool& p, const std::vector<hipMemGenericAllocationHandle_t>& keep) {
    (void)hipDeviceSynchronize();
    for (size_t j = 0; j < p.mapped; ++j) (void)hipMemUnmap(p.pool + j * p.piece, p.piece);
    if (p.pool) (void)hipMemAddressFree(p.pool, p.reserved);
    for (hipMemGenericAllocationHandle_t x : p.h)
        if (std::find(keep.begin(), keep.end(), x) == keep.end()) (void)hipMemRelease(x);
    p = RegionPool{};
}

// Region placement of a plan's tables (see above): pieces holding every
// third slot's table (1 GiB at least), up to six candidate ranges of one piece
// per region (or, two regions seen, one piece of piece 0's region and two of
// the other, from either end of their allocation order), each timed with
// k_place_walk at the plan's own geometry; the two fastest become the table
// ranges.  *placed stays false, with nothing allocated beyond ws, when it does
// not apply: too few pieces in the budget, no virtual-memory support, fewer
// than two regions seen, or too few pieces in them.
static int place_regions(Ctx& c, const HbmDecision& d, bool* placed) {
    *placed = false;
    const size_t rest = d.h.off_tables, tabs = d.h.buffer - d.h.off_tables;
    // Each of a range's three pieces holds every third slot's table; pieces
    // are at least kRegionMinPiece, so that the pair walks can sort them.
    const size_t third = ((size_t(d.chunk) + 2) / 3) * d.h.table_bytes;
    const size_t P = std::max(kRegionMinPiece, (third + kRegionPieceAlign - 1) & ~(kRegionPieceAlign - 1));
    const size_t room = c.budget > 2 * rest ? c.budget - 2 * rest : 0;
    const size_t maxp = std::min<size_t>(kRegionMaxPieces, room / P);
    if (maxp < 6) return CK_OK;
    auto* sink = reinterpret_cast<unsigned long long*>(c.ws.base);  // the first pass's counters, cleared per pass
    RegionPool rp;
    std::string err;
    if (int rc = region_pool(c.device, c.stream, c.ev_chain[0][0], c.ev_chain[0][1], sink, P, maxp, kRegionCands,
                             &rp, &err)) {
        drop_pool(rp, {});
        return fail(rc, "table placement: " + err);
    }
    const std::vector<size_t> of[3] = {rp.of(0), rp.of(1), rp.of(2)};
    std::vector<std::array<size_t, 3>> cand;
    if (rp.two && of[1].size() >= 2 && of[2].size() >= 2) {
        const size_t k = std::min({size_t(kRegionCands), of[0].size(), of[1].size(), of[2].size()});
        for (size_t i = 0; i < k; ++i) cand.push_back({of[0][i], of[1][i], of[2][i]});
    } else if (rp.two) {
        std::vector<size_t> rest1 = of[1];
        rest1.insert(rest1.end(), of[2].begin(), of[2].end());
        std::sort(rest1.begin(), rest1.end());
        const size_t k = std::min({size_t(kRegionCands), of[0].size(), rest1.size() / 2});
        for (size_t i = 0; i < k; ++i) cand.push_back({of[0][i], rest1[i], rest1[rest1.size() - 1 - i]});
    }
    if (cand.size() < 2) {
        drop_pool(rp, {});
        return CK_OK;
    }
    std::vector<float> cms(cand.size(), 0.f);
    const uint64_t pitch = d.h.table_bytes / 8;
    int rc = CK_OK;
    for (size_t i = 0; i < cand.size() && rc == CK_OK; ++i) {
        void* test = nullptr;
        rc = map_pieces(c.device, rp, {cand[i][0], cand[i][1], cand[i][2]}, &test, &err);
        if (rc == CK_OK) {
            const hipError_t e = timed_twice(
                c.stream, c.ev_chain[0][0], c.ev_chain[0][1],
                [&]() {
                    char* t0 = static_cast<char*>(test);
                    hipLaunchKernelGGL(k_place_walk, dim3((d.chunk + 255) / 256), dim3(256), 0, c.stream,
                                       TabRef{{t0, t0 + P, t0 + 2 * P}, 3}, pitch, static_cast<uint32_t>(pitch),
                                       d.chunk, kPlaceWalkSteps, sink);
                },
                &cms[i]);
            if (e != hipSuccess) {
                rc = CK_EDEVICE;
                err = std::string("walk over a candidate range: ") + hipGetErrorString(e);
            }
            (void)hipDeviceSynchronize();
            unmap_pieces(test, P, 3);
        }
    }
    std::vector<int> order(cand.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = static_cast<int>(i);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cms[x] < cms[y]; });
    TableRange out[2];
    for (int b = 0; b < 2 && rc == CK_OK; ++b) {
        const std::array<size_t, 3>& pick = cand[order[b]];
        void* va = nullptr;
        rc = map_pieces(c.device, rp, {pick[0], pick[1], pick[2]}, &va, &err);
        if (rc) break;
        out[b].base = va;
        out[b].bytes = 3 * P;
        out[b].piece = P;
        for (size_t j : pick) out[b].pieces.push_back(rp.h[j]);
    }
    if (rc != CK_OK) {
        (void)hipDeviceSynchronize();
        for (TableRange& t : out) unmap_pieces(t.base, P, t.pieces.size());
        drop_pool(rp, {});
        return fail(rc, "table placement: " + err);
    }
    std::vector<hipMemGenericAllocationHandle_t> keep = out[0].pieces;
    keep.insert(keep.end(), out[1].pieces.begin(), out[1].pieces.end());
    c.place_region = rp.region;
    drop_pool(rp, keep);
    c.tabs[0] = std::move(out[0]);
    c.tabs[1] = std::move(out[1]);
    c.placed_for = tabs;
    c.place_ms = cms;
    for (int b = 0; b < 2; ++b)
        for (int r = 0; r < 3; ++r) c.place_used[3 * b + r] = static_cast<int32_t>(cand[order[b]][r]);
    c.place_kept[0] = order[0];
    c.place_kept[1] = order[1];
    forget_tables(c);
    *placed = true;
    return CK_OK;
}

static int place_buffers(Ctx& c, const HbmDecision& d) {
    const size_t rest = d.h.off_tables, tabs = d.h.buffer - d.h.off_tables;
    if (c.placed_for == tabs && c.ws.bytes >= 2 * rest && c.tabs[0].bytes >= tabs && c.tabs[1].bytes >= tabs)
        return CK_OK;
    if (int rc = release_ws(c)) return rc;
    hipError_t e = hipSuccess;
    Workspace reuse;
    if (take_retired(c.device, 2 * rest, reuse)) {  // one a graph-mode context left
        c.ws = reuse;
        c.ws_graphed = true;
    } else {
        e = hipMalloc(&c.ws.base, 2 * rest);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            c.ws.base = nullptr;
            return fail(CK_ENOMEM, "device memory: cannot allocate the pass buffers");
        }
        if (e != hipSuccess) return fail(CK_EDEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
        c.ws.bytes = 2 * rest;
    }
    bool placed = false;
    if (int rc = place_regions(c, d, &placed)) return rc;
    if (placed) return CK_OK;
    const size_t room = c.budget > 2 * rest ? c.budget - 2 * rest : 0;
    const uint32_t want = static_cast<uint32_t>(std::min<size_t>(c.place_cands, room / tabs));
    std::vector<void*> cand;
    for (uint32_t i = 0; i < want; ++i) {
        void* b = nullptr;
        e = hipMalloc(&b, tabs);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            break;
        }
        if (e != hipSuccess) {
            for (void* x : cand) (void)hipFree(x);
            return fail(CK_EDEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        cand.push_back(b);
    }
    if (cand.size() < 2) {
        for (void* x : cand) (void)hipFree(x);
        return fail(CK_ENOMEM, "device memory: cannot allocate two " + std::to_string(tabs >> 20) + " MiB table regions");
    }
    // Each candidate: the chains' access pattern over its table regions (one
    // lane per trial slot, a dependent random read and write of 8-byte words
    // of the lane's own table).
    const uint64_t pitch = d.h.table_bytes / 8;
    const uint32_t lanes = d.chunk;
    auto* sink = reinterpret_cast<unsigned long long*>(c.ws.base);
    std::vector<float> ms(cand.size(), 0.f);
    for (size_t i = 0; i < cand.size(); ++i) {
        const hipError_t e = timed_twice(
            c.stream, c.ev_chain[0][0], c.ev_chain[0][1],
            [&]() {
                hipLaunchKernelGGL(k_place_walk, dim3((lanes + 255) / 256), dim3(256), 0, c.stream,
                                   TabRef{{static_cast<char*>(cand[i]), nullptr, nullptr}, 1}, pitch,
                                   static_cast<uint32_t>(pitch), lanes, kPlaceWalkSteps, sink);
            },
            &ms[i]);
        if (e != hipSuccess) {
            (void)hipDeviceSynchronize();
            for (void* x : cand) (void)hipFree(x);
            return fail(CK_EDEVICE, std::string("walk over a candidate table allocation: ") + hipGetErrorString(e));
        }
    }
    std::vector<int> order(cand.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = static_cast<int>(i);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ms[x] < ms[y]; });
    for (size_t i = 2; i < order.size(); ++i) CKH(hipFree(cand[order[i]]));
    for (int b = 0; b < 2; ++b) {
        c.tabs[b] = TableRange{};
        c.tabs[b].base = cand[order[b]];
        c.tabs[b].bytes = tabs;
    }
    c.placed_for = tabs;
    c.place_ms = ms;
    c.place_region.clear();
    std::fill(c.place_used, c.place_used + 6, -1);
    c.place_kept[0] = order[0];
    c.place_kept[1] = order[1];
    forget_tables(c);
    return CK_OK;
}

// HBM tier: build + chain per pass of up to `chunk` trial slots.  Two buffers
// and two streams let the build of pass i+1 run while pass i chains.
static int run_hbm(Ctx& c, const ck_params& p, const uint32_t rk[44], uint32_t first, uint32_t count, HostOut& out,
                   ck_stats* st, bool lane = false) {
    c.budget = hbm_budget(c);
    // A workspace within 3 % of the budget is planned into as it is: growing it
    // would free and map hundreds of GiB again (seconds at the memory-capped
    // n = 2^20 shape) for a few more chains per pass.
    if (!c.budget_fixed && c.ws.bytes < c.budget && c.ws.bytes >= c.budget / 100 * 97) c.budget = c.ws.bytes;
    const bool graph_wanted =
        c.graph_on && !out.probes && !out.tables && !out.keys && !out.walks && !out.rnds && !out.map && !out.digests;
    HbmDecision d;
    int rc = CK_OK;
    bool placed = false;
    // A workspace that does not fit (another context on the device took memory
    // since the budget was read) is planned again at half the budget.
    for (int attempt = 0;; ++attempt) {
        rc = decide_hbm(c, p, count, out.tables != nullptr, graph_wanted, !out.map && !out.walks && !out.rnds, d,
                        lane);
        if (rc) return rc;
        placed = placeable(c, d, out.tables != nullptr);
        rc = placed ? place_buffers(c, d) : ensure_ws(c, d.nbuf * d.h.buffer);
        if (rc != CK_ENOMEM || attempt == 8 || d.chunk == 1) break;
        c.budget /= 2;
    }
    if (rc) return rc;
    const uint32_t layout = d.layout, window = d.window, nbuf = d.nbuf, chunk = d.chunk;
    const bool graph = d.graph;
    const HbmPlan& h = d.h;
    const BuildPlan& bp = d.bp;
    (void)window;
    const bool bgh = p.mode == CK_MODE_BGH25;
    PassArgs args[2];
    for (uint32_t b = 0; b < nbuf; ++b) {
        char* base = static_cast<char*>(c.ws.base) + b * (placed ? h.off_tables : h.buffer);
        PassArgs& a = args[b];
        a.S = make_shape(p);
        std::memcpy(a.bk.w, rk, sizeof a.bk.w);
        a.map = base + h.off_map;
        a.map_stride = h.map_stride;
        if (!placed) {
            a.tables = TabRef{{base + h.off_tables, nullptr, nullptr}, 1};
        } else if (c.tabs[b].pieces.size() == 3) {  // by region: every third slot in each piece
            char* t0 = static_cast<char*>(c.tabs[b].base);
            const size_t P = c.tabs[b].piece;
            a.tables = TabRef{{t0, t0 + P, t0 + 2 * P}, 3};
        } else {
            a.tables = TabRef{{static_cast<char*>(c.tabs[b].base), nullptr, nullptr}, 1};
        }
        a.table_bytes = h.table_bytes;
        a.walks = reinterpret_cast<uint64_t*>(base + h.off_walk);
        a.rnds = reinterpret_cast<uint64_t*>(base + h.off_rnd);
        a.cnt = reinterpret_cast<uint32_t*>(base + h.off_cnt);
        a.fail = reinterpret_cast<uint32_t*>(base + h.off_fail);
        a.probes = out.probes || h.nwin > 1 ? reinterpret_cast<uint32_t*>(base + h.off_probes) : nullptr;
        a.keys = out.keys ? reinterpret_cast<uint32_t*>(base + h.off_key) : nullptr;
        a.digests = out.digests ? reinterpret_cast<unsigned long long*>(base + h.off_digest) : nullptr;
        a.item0 = 0;
        a.item1 = p.pnum;
        a.state = h.nwin > 1 ? reinterpret_cast<uint32_t*>(base + h.off_state) : nullptr;
    }
    if (st) {
        st->tier = lane ? CK_TIER_LANE : CK_TIER_HBM;
        st->chunk = chunk;
    }
    const uint32_t passes = (count + chunk - 1) / chunk;
    // Tables are reused across passes only in a workspace this context owns.
    const bool taggable = layout == CK_LAYOUT_DESC_PLAIN && !out.digests && !c.ws_static;
    const uint32_t tag_bits = taggable ? 62 - p.k * args[0].S.dbits : 0;
    const uint64_t plan = table_plan(c, p, h, nbuf, taggable);
    if (plan != c.table_plan) {
        forget_tables(c);
        c.table_plan = plan;
    }
    if (graph)
        return run_hbm_graph(c, p, first, count, layout, nbuf, chunk, h, bp, args, taggable, tag_bits, out, st);
    // The call's device span: from before the first build to the end of the
    // last chain (the chain-end event recorded last).
    CKH(hipEventRecord(c.ev[2], c.build_stream));
    hipEvent_t last_end = nullptr;
    auto span = [&]() -> int {
        if (st && last_end) {
            float ms = 0;
            CKH(hipEventSynchronize(last_end));
            CKH(hipEventElapsedTime(&ms, c.ev[2], last_end));
            st->span_ms += ms;
        }
        return CK_OK;
    };
    auto enqueue = [&](uint32_t i) -> int {
        const int b = nbuf == 2 ? int(i & 1) : 0;
        PassArgs& a = args[b];
        a.first = first + i * chunk;
        a.count = std::min(chunk, count - i * chunk);
        pass_tags(c, b, taggable, tag_bits, a);
        if (i >= 2) CKH(hipStreamWaitEvent(c.build_stream, c.ev_chain[b][1], 0));
        CKH(hipEventRecord(c.ev_build[b][0], c.build_stream));
        Dispatch bd, cd;
        bd.stream = c.build_stream;
        cd.stream = c.stream;
        launch_build(p.k, bp, h, a, bd);
        CKH(bd.err);
        CKH(hipGetLastError());
        CKH(hipEventRecord(c.ev_build[b][1], c.build_stream));
        CKH(hipStreamWaitEvent(c.stream, c.ev_build[b][1], 0));
        cd(&k_zero, dim3(1), dim3(64), 0, a.cnt, CNT_WORDS);
        CKH(hipEventRecord(c.ev_chain[b][0], c.stream));
        launch_chain(p.k, bgh, h, a, cd);
        CKH(cd.err);
        CKH(hipGetLastError());
        CKH(hipEventRecord(c.ev_chain[b][1], c.stream));
        last_end = c.ev_chain[b][1];
        if (a.digests) {
            switch (p.k) {
                case 2: launch_digest_k<2>(h, a, cd); break;
                case 3: launch_digest_k<3>(h, a, cd); break;
                default: launch_digest_k<4>(h, a, cd); break;
            }
            CKH(cd.err);
            CKH(hipGetLastError());
        }
        return CK_OK;
    };
    auto gather = [&](uint32_t i, bool timed) -> int {
        const int b = nbuf == 2 ? int(i & 1) : 0;
        const PassArgs& a = args[b];
        CKH(hipEventSynchronize(c.ev_chain[b][1]));
        if (a.digests) CKH(hipStreamSynchronize(c.stream));
        int r = take_counters(a.cnt, a.fail, a.count, out, st);
        if (r) return r;
        const size_t done = size_t(i) * chunk;
        if (out.probes) CKH(hipMemcpy(out.probes + done, a.probes, size_t(a.count) * 4, hipMemcpyDeviceToHost));
        if (out.tables &&
            (r = copy_rows(out.tables + done * p.buckets, a.tables.base[0], h.table_bytes, p.buckets, a.count, h.table16,
                           true)))
            return r;
        const bool map64 = h.desc || h.stream64;
        if (out.map && !map64 &&
            (r = copy_rows(out.map + done * p.pnum * p.k, a.map, h.map_bytes, size_t(p.pnum) * p.k, a.count, h.map16,
                           false)))
            return r;
        if (out.map && map64) {
            std::vector<uint64_t> d(size_t(p.pnum) * a.count);
            CKH(hipMemcpy2D(d.data(), size_t(p.pnum) * 8, a.map, h.map_bytes, size_t(p.pnum) * 8, a.count,
                            hipMemcpyDeviceToHost));
            const uint32_t bb = a.S.dbits;
            for (size_t i = 0; i < d.size(); ++i)
                for (uint32_t w = 0; w < p.k; ++w)
                    out.map[done * p.pnum * p.k + i * p.k + w] = static_cast<uint32_t>((d[i] >> (w * bb)) & ((1ull << bb) - 1));
        }
        if (out.digests)
            CKH(hipMemcpy(out.digests + done, a.digests, size_t(a.count) * 8, hipMemcpyDeviceToHost));
        if (out.keys) CKH(hipMemcpy(out.keys + done * 44, a.keys, size_t(a.count) * 44 * 4, hipMemcpyDeviceToHost));
        if (out.walks) CKH(hipMemcpy(out.walks + done, a.walks, size_t(a.count) * 8, hipMemcpyDeviceToHost));
        if (out.rnds) CKH(hipMemcpy(out.rnds + done, a.rnds, size_t(a.count) * 8, hipMemcpyDeviceToHost));
        if (st && timed) {
            float bms = 0, cms = 0;
            CKH(hipEventElapsedTime(&bms, c.ev_build[b][0], c.ev_build[b][1]));
            CKH(hipEventElapsedTime(&cms, c.ev_chain[b][0], c.ev_chain[b][1]));
            st->build_ms += bms;
            st->chain_ms += cms;
            st->launches += 2;
        }
        return CK_OK;
    };
    // Item windows: pass i runs as h.nwin (build, chain) pairs over one buffer.
    // The build of window j fills map half j & 1 on the build stream while the
    // chain of window j - 1 reads the other half; window j's build is enqueued
    // once the chain of window j - 2, the last reader of its half, has ended.
    // Only window 0 clears the tables and writes the walks.  Window parity e
    // uses the event pairs ev_build[e] and ev_chain[e].
    auto windowed = [&](uint32_t i) -> int {
        PassArgs& a = args[0];
        a.first = first + i * chunk;
        a.count = std::min(chunk, count - i * chunk);
        pass_tags(c, 0, false, 0, a);
        auto account = [&](uint32_t j) -> int {
            const int e = int(j & 1);
            CKH(hipEventSynchronize(c.ev_chain[e][1]));
            if (st) {
                float bms = 0, cms = 0;
                CKH(hipEventElapsedTime(&bms, c.ev_build[e][0], c.ev_build[e][1]));
                CKH(hipEventElapsedTime(&cms, c.ev_chain[e][0], c.ev_chain[e][1]));
                st->build_ms += bms;
                st->chain_ms += cms;
                st->launches += 2;
            }
            return CK_OK;
        };
        for (uint32_t j = 0; j < h.nwin; ++j) {
            const int e = int(j & 1);
            if (j >= 2 && (rc = account(j - 2))) return rc;
            PassArgs w = a;
            w.map = a.map + size_t(e) * h.map_bytes * chunk;
            w.item0 = j * h.window;
            w.item1 = std::min(p.pnum, w.item0 + h.window);
            w.clear = j == 0 ? 1u : 0u;
            CKH(hipEventRecord(c.ev_build[e][0], c.build_stream));
            Dispatch bd, cd;
            bd.stream = c.build_stream;
            cd.stream = c.stream;
            launch_build(p.k, bp, h, w, bd);
            CKH(bd.err);
            CKH(hipGetLastError());
            CKH(hipEventRecord(c.ev_build[e][1], c.build_stream));
            CKH(hipStreamWaitEvent(c.stream, c.ev_build[e][1], 0));
            if (j == 0) cd(&k_zero, dim3(1), dim3(64), 0, a.cnt, CNT_WORDS);
            CKH(hipEventRecord(c.ev_chain[e][0], c.stream));
            launch_chain(p.k, bgh, h, w, cd);
            CKH(cd.err);
            CKH(hipGetLastError());
            CKH(hipEventRecord(c.ev_chain[e][1], c.stream));
        }
        last_end = c.ev_chain[(h.nwin - 1) & 1][1];
        for (uint32_t j = h.nwin - std::min<uint32_t>(h.nwin, 2); j < h.nwin; ++j)
            if ((rc = account(j))) return rc;
        if (a.digests) {
            Dispatch cd;
            cd.stream = c.stream;
            switch (p.k) {
                case 2: launch_digest_k<2>(h, a, cd); break;
                case 3: launch_digest_k<3>(h, a, cd); break;
                default: launch_digest_k<4>(h, a, cd); break;
            }
            CKH(cd.err);
            CKH(hipGetLastError());
        }
        return gather(i, false);
    };
    if (h.nwin > 1) {
        for (uint32_t i = 0; i < passes; ++i)
            if ((rc = windowed(i))) return rc;
        return span();
    }
    if (nbuf == 1) {
        for (uint32_t i = 0; i < passes; ++i)
            if ((rc = enqueue(i)) || (rc = gather(i, true))) return rc;
        return span();
    }
    if ((rc = enqueue(0))) return rc;
    for (uint32_t i = 1; i < passes; ++i) {
        if ((rc = enqueue(i))) return rc;
        if ((rc = gather(i - 1, true))) return rc;
    }
    if ((rc = gather(passes - 1, true))) return rc;
    return span();
}

// LDS tier: one fused persistent launch per chunk.
static int run_lds(Ctx& c, const ck_params& p, const uint32_t rk[44], uint32_t first, uint32_t count, HostOut& out,
                   ck_stats* st) {
    if (c.prp == CK_PRP_DIRECT)
        return fail(CK_EINVAL, "the LDS tier builds a round cache: it cannot run with direct AES evaluation");
    const LdsPlan plan = plan_lds(p, c.lds_block, c.lds_cu, c.lds_variant == CK_LDS_ROUND);
    if (!plan.ok) return fail(CK_EINVAL, "shape does not fit the LDS tier (half_bits > 12 or table too large)");
    forget_tables(c);  // the workspace becomes this tier's scratch
    const Shape S = make_shape(p);
    RoundKeys bk;
    std::memcpy(bk.w, rk, sizeof bk.w);
    const uint32_t chunk = lds_chunk(c, count, out.tables != nullptr, p);
    const LdsScratch L = lds_scratch(p, chunk, out.tables != nullptr);
    int rc = ensure_ws(c, L.total);
    if (rc) return rc;
    char* base = static_cast<char*>(c.ws.base);
    auto* cnt = reinterpret_cast<uint32_t*>(base + L.off_cnt);
    auto* probes = reinterpret_cast<uint32_t*>(base + L.off_probes);
    auto* fo = reinterpret_cast<uint32_t*>(base + L.off_fail);
    auto* tables = reinterpret_cast<uint32_t*>(base + L.off_tables);
    auto* digests = reinterpret_cast<unsigned long long*>(base + L.off_digest);
    const bool bgh = p.mode == CK_MODE_BGH25;
    const uint32_t resident = plan.blocks_per_cu * c.cus;
    if (st) {
        st->tier = CK_TIER_LDS;
        st->chunk = chunk;
    }
    for (uint32_t done = 0; done < count;) {
        const uint32_t n = std::min(chunk, count - done);
        const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>(resident, (n + plan.waves - 1) / plan.waves));
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, c.stream, cnt, CNT_WORDS);
        CKH(hipEventRecord(c.ev[0], c.stream));
        uint32_t* pout = out.probes ? probes : nullptr;
        uint32_t* tout = out.tables ? tables : nullptr;
        unsigned long long* dout = out.digests ? digests : nullptr;
        switch (p.k) {
            case 2: launch_lds<2>(bgh, plan, blocks, S, bk, first + done, n, cnt, fo, pout, tout, dout, c.stream); break;
            case 3: launch_lds<3>(bgh, plan, blocks, S, bk, first + done, n, cnt, fo, pout, tout, dout, c.stream); break;
            default: launch_lds<4>(bgh, plan, blocks, S, bk, first + done, n, cnt, fo, pout, tout, dout, c.stream); break;
        }
        CKH(hipGetLastError());
        CKH(hipEventRecord(c.ev[1], c.stream));
        CKH(hipStreamSynchronize(c.stream));
        rc = take_counters(cnt, fo, n, out, st);
        if (rc) return rc;
        if (out.probes) CKH(hipMemcpy(out.probes + done, probes, size_t(n) * 4, hipMemcpyDeviceToHost));
        if (out.tables)
            CKH(hipMemcpy(out.tables + size_t(done) * p.buckets, tables, size_t(n) * p.buckets * 4,
                          hipMemcpyDeviceToHost));
        if (out.digests) CKH(hipMemcpy(out.digests + done, digests, size_t(n) * 8, hipMemcpyDeviceToHost));
        if (st) {
            float f = 0;
            CKH(hipEventElapsedTime(&f, c.ev[0], c.ev[1]));
            st->fused_ms += f;
            st->span_ms += f;
            st->launches += 1;
        }
        done += n;
    }
    return CK_OK;
}

// The plan of a launch of `count` trials under the context's settings
// (c.budget must be set); no device work.
static int plan_launch(const Ctx& c, const ck_params& p, uint32_t count, ck_plan& out) {
    std::memset(&out, 0, sizeof out);
    if (choose_tier(c, p) == CK_TIER_LDS) {
        if (c.prp == CK_PRP_DIRECT)
            return fail(CK_EINVAL, "the LDS tier builds a round cache: it cannot run with direct AES evaluation");
        const LdsPlan L = plan_lds(p, c.lds_block, c.lds_cu, c.lds_variant == CK_LDS_ROUND);
        if (!L.ok) return fail(CK_EINVAL, "shape does not fit the LDS tier (half_bits > 12 or table too large)");
        out.tier = CK_TIER_LDS;
        out.layout = CK_LAYOUT_AUTO;
        out.round_cache = 1;
        out.lds_variant = L.round ? CK_LDS_ROUND : CK_LDS_MAP;
        out.trials_per_pass = lds_chunk(c, std::max<uint32_t>(count, 1), false, p);
        out.buffers = 1;
        out.windows = 1;
        out.chain_waves = L.waves;
        out.chain_lds_bytes = static_cast<uint32_t>(L.lds);
        out.workspace_bytes = lds_scratch(p, out.trials_per_pass, false).total;
        return CK_OK;
    }
    HbmDecision d;
    const bool lane = choose_tier(c, p) == CK_TIER_LANE;
    const int rc = decide_hbm(c, p, std::max<uint32_t>(count, 1), false, c.graph_on, true, d, lane);
    if (rc) return rc;
    out.tier = lane ? CK_TIER_LANE : CK_TIER_HBM;
    out.layout = lane ? CK_LAYOUT_RESIDUAL : d.layout;
    out.round_cache = d.bp.cached ? 1 : 0;
    out.lds_variant = CK_LDS_AUTO;
    out.trials_per_pass = d.chunk;
    out.buffers = d.nbuf;
    out.windows = d.h.nwin;
    out.chain_waves = lane ? 1 : 4;  // k_chain_lane blocks are one wave, k_chain blocks 256 lanes
    out.chain_lds_bytes = lane ? static_cast<uint32_t>(d.h.lp.lds) : 0;
    out.build_lds_bytes = static_cast<uint32_t>(d.bp.lds);
    out.workspace_bytes = uint64_t(d.nbuf) * d.h.buffer;
    return CK_OK;
}

static int apply_placement(Ctx& c, const ck_placement* w) {
    if (!w || w->tier > CK_TIER_LANE || w->layout > CK_LAYOUT_RESIDUAL || w->prp > CK_PRP_DIRECT ||
        w->lds_variant > CK_LDS_ROUND)
        return fail(CK_EINVAL, "bad placement");
    c.tier = w->tier;
    c.layout = w->layout;
    c.prp = w->prp;
    c.lds_variant = w->lds_variant;
    c.budget_fixed = static_cast<size_t>(w->budget_bytes);
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
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->c.build_stream, hipStreamNonBlocking);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipEventCreate(&ctx->c.ev[i]);
    for (int b = 0; b < 2 && e == hipSuccess; ++b)
        for (int j = 0; j < 2 && e == hipSuccess; ++j) {
            e = hipEventCreate(&ctx->c.ev_build[b][j]);
            if (e == hipSuccess) e = hipEventCreate(&ctx->c.ev_chain[b][j]);
        }
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_te0), te0(), 256 * sizeof(uint32_t));
    if (e != hipSuccess) {
        ck_ctx_destroy(ctx);
        return fail(CK_EDEVICE, std::string("context setup: ") + hipGetErrorString(e));
    }
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, device) == hipSuccess && v > 0)
        ctx->c.lds_block = static_cast<size_t>(v);
    else if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
        ctx->c.lds_block = static_cast<size_t>(v);
    ctx->c.lds_cu = ctx->c.lds_block;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) == hipSuccess && v > 0)
        ctx->c.lds_cu = std::max(ctx->c.lds_block, static_cast<size_t>(v));
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0)
        ctx->c.cus = static_cast<uint32_t>(v);
    int rc = set_lds_limits(ctx->c.lds_block);
    if (rc == CK_OK) rc = static_scratch_of(device, &ctx->c.static_scratch);
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
    for (HbmGraph& g : ctx->c.graphs) {
        g.drop();
        if (g.aux) (void)hipFree(g.aux);
    }
    if (ctx->c.ws.base) {
        if (ctx->c.ws_graphed)
            retire_ws(ctx->c.device, ctx->c.ws);
        else
            (void)hipFree(ctx->c.ws.base);
    }
    for (TableRange& t : ctx->c.tabs) (void)free_tables(ctx->c, t);
    for (auto& ev : ctx->c.ev)
        if (ev) (void)hipEventDestroy(ev);
    for (int b = 0; b < 2; ++b)
        for (int j = 0; j < 2; ++j) {
            if (ctx->c.ev_build[b][j]) (void)hipEventDestroy(ctx->c.ev_build[b][j]);
            if (ctx->c.ev_chain[b][j]) (void)hipEventDestroy(ctx->c.ev_chain[b][j]);
        }
    if (ctx->c.build_stream) (void)hipStreamDestroy(ctx->c.build_stream);
    if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

int ck_ctx_set_tier(ck_ctx* ctx, uint32_t tier) {
    if (!ctx || tier > CK_TIER_LANE) return fail(CK_EINVAL, "bad tier");
    ctx->c.tier = tier;
    return CK_OK;
}

int ck_ctx_set_graph(ck_ctx* ctx, uint32_t on) {
    if (!ctx || on > 1) return fail(CK_EINVAL, "bad graph setting");
    ctx->c.graph_on = on != 0;
    return CK_OK;
}

int ck_ctx_set_chunk(ck_ctx* ctx, uint32_t trials) {
    if (!ctx) return fail(CK_EINVAL, "null ctx");
    ctx->c.chunk_cap = trials;
    return CK_OK;
}

int ck_ctx_set_place_search(ck_ctx* ctx, uint32_t candidates) {
    if (!ctx || candidates == 1 || candidates > 64) return fail(CK_EINVAL, "candidates must be 0 or 2..64");
    ctx->c.place_cands = candidates;
    return CK_OK;
}

int ck_ctx_place_report(ck_ctx* ctx, float* walk_ms, uint32_t cap, uint32_t* n, int32_t kept[2]) {
    if (!ctx || !n || !kept || (cap && !walk_ms)) return fail(CK_EINVAL, "null argument");
    const Ctx& c = ctx->c;
    *n = static_cast<uint32_t>(c.place_ms.size());
    for (uint32_t i = 0; i < std::min(cap, *n); ++i) walk_ms[i] = c.place_ms[i];
    kept[0] = c.placed_for ? c.place_kept[0] : -1;
    kept[1] = c.placed_for ? c.place_kept[1] : -1;
    return CK_OK;
}

int ck_ctx_place_regions(ck_ctx* ctx, int32_t* region, uint32_t cap, uint32_t* n, int32_t used[6]) {
    if (!ctx || !n || !used || (cap && !region)) return fail(CK_EINVAL, "null argument");
    const Ctx& c = ctx->c;
    const bool on = c.placed_for && !c.place_region.empty();
    *n = on ? static_cast<uint32_t>(c.place_region.size()) : 0;
    for (uint32_t i = 0; i < std::min(cap, *n); ++i) region[i] = c.place_region[i];
    for (int i = 0; i < 6; ++i) used[i] = on ? c.place_used[i] : -1;
    return CK_OK;
}

int ck_ctx_set_window(ck_ctx* ctx, uint32_t items) {
    if (!ctx) return fail(CK_EINVAL, "null ctx");
    ctx->c.window = items;
    return CK_OK;
}

int ck_aes_self_test(ck_ctx* ctx) {
    if (!ctx) return fail(CK_EINVAL, "null ctx");
    CKH(hipSetDevice(ctx->c.device));
    uint32_t* d = nullptr;
    CKH(hipMalloc(&d, 2 * 64 * 4 * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_aes_kat, dim3(1), dim3(64), tab_bytes(kRepl), ctx->c.stream, d, kRepl);
    hipLaunchKernelGGL(k_aes_kat, dim3(1), dim3(64), tab_bytes(0), ctx->c.stream, d + 256, 0u);
    std::vector<uint32_t> h(2 * 64 * 4);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, h.size() * 4, hipMemcpyDeviceToHost, ctx->c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->c.stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CK_EDEVICE, std::string("AES self test: ") + hipGetErrorString(e));
    const uint32_t want[4] = {0x69c4e0d8u, 0x6a7b0430u, 0xd8cdb780u, 0x70b4c55au};
    for (int lane = 0; lane < 128; ++lane)
        if (std::memcmp(h.data() + lane * 4, want, sizeof want))
            return fail(CK_ESELFTEST, "device AES-128 FIPS-197 known-answer test failed");
    return CK_OK;
}

static int launch_common(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t first, uint32_t count,
                         uint32_t* failed_offsets, uint32_t* n_failed, uint32_t* probes, uint32_t* tables,
                         uint64_t* digests, ck_stats* stats) {
    if (!ctx || !rk || !n_failed || (count && !failed_offsets)) return fail(CK_EINVAL, "null argument");
    int rc = check_params(p);
    if (rc) return rc;
    if (uint64_t(first) + count > (uint64_t(1) << 32)) return fail(CK_EINVAL, "trial range exceeds 32-bit offsets");
    CKH(hipSetDevice(ctx->c.device));
    if (stats) std::memset(stats, 0, sizeof *stats);
    std::vector<uint32_t> fails;
    HostOut out{&fails, probes, tables, nullptr, nullptr, nullptr, nullptr, digests};
    if (count) {
        const uint32_t tier = choose_tier(ctx->c, *p, tables != nullptr);
        rc = tier == CK_TIER_LDS ? run_lds(ctx->c, *p, rk, first, count, out, stats)
                                 : run_hbm(ctx->c, *p, rk, first, count, out, stats, tier == CK_TIER_LANE);
        if (rc) {
            // A failed call may have planned clears that never ran: the next
            // call starts from cleared tables.
            forget_tables(ctx->c);
            return rc;
        }
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
    return launch_common(ctx, p, rk, first, count, failed_offsets, n_failed, nullptr, nullptr, nullptr, stats);
}

int ck_launch_static(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t first, uint32_t count,
                     uint32_t* failed_offsets, uint32_t* n_failed, ck_stats* stats) {
    if (!ctx || !ctx->c.static_scratch) return fail(CK_EINVAL, "null ctx");
    Ctx& c = ctx->c;
    // The context's own workspace and settings are set aside for the call.
    const Workspace own = c.ws;
    const size_t own_budget = c.budget_fixed;
    const bool own_graph = c.graph_on;
    forget_tables(c);
    c.ws = Workspace{c.static_scratch, CK_STATIC_SCRATCH_BYTES};
    c.ws_static = true;
    c.budget_fixed = CK_STATIC_SCRATCH_BYTES;
    c.graph_on = false;
    const int rc = launch_common(ctx, p, rk, first, count, failed_offsets, n_failed, nullptr, nullptr, nullptr, stats);
    forget_tables(c);
    c.ws = own;
    c.ws_static = false;
    c.budget_fixed = own_budget;
    c.graph_on = own_graph;
    return rc;
}

int ck_launch_debug(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t first, uint32_t count,
                    uint32_t* failed_offsets, uint32_t* n_failed, uint32_t* probes, uint32_t* tables,
                    ck_stats* stats) {
    return launch_common(ctx, p, rk, first, count, failed_offsets, n_failed, probes, tables, nullptr, stats);
}

int ck_launch_digest(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t first, uint32_t count,
                     uint32_t* failed_offsets, uint32_t* n_failed, uint32_t* probes, uint64_t* digests,
                     ck_stats* stats) {
    if (!digests) return fail(CK_EINVAL, "null digests");
    return launch_common(ctx, p, rk, first, count, failed_offsets, n_failed, probes, nullptr, digests, stats);
}

// Autotuner.  Each applicable (tier, layout, trials per pass) candidate runs
// the first min(sample, 65536) trials once (workspace growth, placement
// search, warm caches), then the whole sample with wall-clock timing.  Every candidate must report the same
// failing offsets; the fastest becomes the context's setting.
int ck_ctx_tune(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t sample, ck_tune_entry* entries,
                uint32_t cap, uint32_t* n) {
    if (!ctx || !rk || !n || (cap && !entries)) return fail(CK_EINVAL, "null argument");
    if (sample == 0) return fail(CK_EINVAL, "tuning needs at least one trial");
    int rc = check_params(p);
    if (rc) return rc;
    Ctx& c = ctx->c;
    struct Cand {
        uint32_t tier, layout, chunk;
    };
    std::vector<Cand> cands;
    const uint32_t half = std::max<uint32_t>(64, std::min<uint32_t>(sample, kHbmChunk) / 2);
    for (uint32_t layout : {CK_LAYOUT_DESC_PLAIN, CK_LAYOUT_RESIDUAL, CK_LAYOUT_MAP, CK_LAYOUT_MAP_XCHG}) {
        if (layout == CK_LAYOUT_DESC_PLAIN && !desc_ok(*p)) continue;
        if (layout == CK_LAYOUT_RESIDUAL && !res_ok(*p)) continue;
        cands.push_back({CK_TIER_HBM, layout, 0});
        if (half < sample) cands.push_back({CK_TIER_HBM, layout, half});
    }
    if (c.prp != CK_PRP_DIRECT && plan_lds(*p, c.lds_block, c.lds_cu, c.lds_variant == CK_LDS_ROUND).ok)
        cands.push_back({CK_TIER_LDS, CK_LAYOUT_AUTO, 0});
    // The lane tier only where 16 or more trials fit a CU: with fewer (n = 2^14
    // fits one) a sample would run for hours.
    if (plan_lane(*p, c.lds_block, c.lds_cu).lanes_per_cu >= 16) cands.push_back({CK_TIER_LANE, CK_LAYOUT_AUTO, 0});
    const uint32_t saved_tier = c.tier, saved_layout = c.layout, saved_chunk = c.chunk_cap;
    std::vector<uint32_t> want, got(sample);
    bool have_want = false;
    double best_rate = -1;
    Cand best = {saved_tier, saved_layout, saved_chunk};
    *n = 0;
    for (const Cand& k : cands) {
        c.tier = k.tier;
        c.layout = k.layout;
        c.chunk_cap = k.chunk;
        uint32_t nf = 0;
        // Warm-up: a launch planned like the timed one (the same trials per
        // pass), so that workspace growth and the table placement happen here.
        rc = launch_common(ctx, p, rk, 0, std::min(sample, kHbmChunk), got.data(), &nf, nullptr, nullptr, nullptr,
                           nullptr);
        const auto t0 = std::chrono::steady_clock::now();
        if (rc == CK_OK) rc = launch_common(ctx, p, rk, 0, sample, got.data(), &nf, nullptr, nullptr, nullptr, nullptr);
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc != CK_OK) break;
        std::vector<uint32_t> f(got.begin(), got.begin() + nf);
        if (!have_want) {
            want = f;
            have_want = true;
        } else if (f != want) {
            rc = fail(CK_ERANGE, "tuning candidates disagree on the failing trials");
            break;
        }
        const double rate = sample / std::max(dt, 1e-9);
        if (*n < cap) entries[*n] = ck_tune_entry{k.tier, k.layout, k.chunk, 0, rate};
        ++*n;
        if (rate > best_rate) {
            best_rate = rate;
            best = k;
        }
    }
    c.tier = rc == CK_OK ? best.tier : saved_tier;
    c.layout = rc == CK_OK ? best.layout : saved_layout;
    c.chunk_cap = rc == CK_OK ? best.chunk : saved_chunk;
    return rc;
}

int ck_ctx_set_placement(ck_ctx* ctx, const ck_placement* want) {
    if (!ctx) return fail(CK_EINVAL, "null ctx");
    return apply_placement(ctx->c, want);
}

int ck_ctx_plan(ck_ctx* ctx, const ck_params* p, uint32_t count, ck_plan* out) {
    if (!ctx || !out) return fail(CK_EINVAL, "null argument");
    int rc = check_params(p);
    if (rc) return rc;
    CKH(hipSetDevice(ctx->c.device));
    ctx->c.budget = hbm_budget(ctx->c);
    return plan_launch(ctx->c, *p, count, *out);
}

int ck_plan_describe(const ck_params* p, const ck_placement* want, uint32_t count, ck_plan* out) {
    if (!out) return fail(CK_EINVAL, "null output");
    int rc = check_params(p);
    if (rc) return rc;
    Ctx c;  // settings only: no streams, no memory
    if ((rc = apply_placement(c, want))) return rc;
    c.lds_block = c.lds_cu = size_t(160) << 10;
    c.cus = 256;
    c.budget = c.budget_fixed ? c.budget_fixed : size_t(96) << 30;
    return plan_launch(c, *p, count, *out);
}

int ck_ctx_set_layout(ck_ctx* ctx, uint32_t layout) {
    if (!ctx || layout > CK_LAYOUT_RESIDUAL) return fail(CK_EINVAL, "bad layout");
    ctx->c.layout = layout;
    return CK_OK;
}

int ck_trial_map(ck_ctx* ctx, const ck_params* p, const uint32_t rk[44], uint32_t offset, uint32_t key_words[44],
                 uint64_t* walk, uint64_t* randomness, uint32_t* map) {
    if (!ctx || !rk || !key_words || !walk || !randomness || !map) return fail(CK_EINVAL, "null argument");
    int rc = check_params(p);
    if (rc) return rc;
    CKH(hipSetDevice(ctx->c.device));
    std::vector<uint32_t> fails;
    HostOut out{&fails, nullptr, nullptr, key_words, walk, randomness, map, nullptr};
    const uint32_t saved = ctx->c.chunk_cap;
    ctx->c.chunk_cap = 1;
    rc = run_hbm(ctx->c, *p, rk, offset, 1, out, nullptr);
    ctx->c.chunk_cap = saved;
    if (rc) forget_tables(ctx->c);
    return rc;
}

}  // extern "C"
