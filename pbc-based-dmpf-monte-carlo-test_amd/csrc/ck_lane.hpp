// ck_lane.hpp -- the lane tier's insertion chains: one trial per lane with the
// trial's whole table in LDS, the table layout of the reference's per-thread
// shared-memory bins (cuckoo_bench.cu:217-220, 876-880).
//
// Table.  Residual entries (the occupant's other k-1 bucket indices and its
// choice, ResCodec) of 16 or 32 bits.  The W lanes of a block interleave their
// tables: entry b of lane j is element b*W + j of the block's table region, so
// a probe is one LDS read and at most one LDS write at a lane-private address,
// and the lanes of a wave that probe the same bucket index hit distinct banks.
//
// Candidate stream.  New items take their candidates from the build's u64
// stream in HBM (one packed word per item, in insertion order).  A lane reads
// it through a ring of kRing words in LDS, refilled kChunk items at a time:
// every kChunk iterations of the chain loop (a wave-uniform point) the wave
// first copies the chunks loaded at the previous such point from registers
// into the rings, then issues the global loads of the next chunks.  A chunk's
// loads are thus in flight for kChunk probes before they are waited for, and
// the loop itself holds no global memory access, so no probe waits on HBM.
//
// Ring bounds.  Let b = fill - item (items in the ring from the one in hand)
// right after the copy of a refill point.  A lane consumes at most one item
// per iteration and reads the candidates of the item after the one in hand,
// so it needs b >= kChunk + 2 at every refill point.  It loads a chunk when
// b <= kChunk * 2 + 1: if it loads, the next point sees b - u + kChunk >= b
// (u <= kChunk items consumed), else b >= 2*kChunk + 2 - kChunk.  A ring of
// kRing >= 3*kChunk + 1 words holds the items in hand plus a chunk in flight.
// The chain starts with 2*kChunk items in the ring.
#pragma once
#include "ck_chain.hpp"

namespace ck {

constexpr uint32_t kLaneRing = 64;   // ring words per lane (a power of two >= 3*kLaneChunk + 1)
constexpr uint32_t kLaneChunk = 16;  // items per refill (128 bytes, eight 16-byte loads)

template <class E>
struct LdsRef;
template <>
struct LdsRef<uint16_t> {
    typedef __attribute__((address_space(3))) uint16_t* type;
};
template <>
struct LdsRef<uint32_t> {
    typedef __attribute__((address_space(3))) uint32_t* type;
};
typedef __attribute__((address_space(3))) uint64_t* lds_u64p;
typedef __attribute__((address_space(3))) uint32_t* lds_u32p;

// 32-bit LDS byte address of a pointer into the kernel's dynamic LDS.
__device__ __forceinline__ uint32_t lds_offset(const void* p) {
    return static_cast<uint32_t>(reinterpret_cast<uintptr_t>((const __attribute__((address_space(3))) char*)p));
}

// One lane's table: entry b at byte base + b * stride.
template <class E>
struct LaneTable {
    uint32_t base, stride;
    __device__ __forceinline__ E get(uint32_t b) const {
        return *reinterpret_cast<typename LdsRef<E>::type>(static_cast<uintptr_t>(base + b * stride));
    }
    __device__ __forceinline__ void put(uint32_t b, E v) const {
        *reinterpret_cast<typename LdsRef<E>::type>(static_cast<uintptr_t>(base + b * stride)) = v;
    }
};

// One lane's candidate stream: its row of the build's u64 stream in HBM, its
// ring in LDS (word i of the ring at byte ring + (i % kLaneRing) * ring_stride)
// and the chunk in flight.
template <int K>
struct LaneFeed {
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
    typedef const __attribute__((address_space(1))) u64x2* grow;
    grow row;              // the lane's stream row in global memory (16-byte aligned)
    uint32_t ring, ring_stride, bb;
    uint64_t bmask;
    uint32_t fill = 0;     // items [0, fill) have been copied into the ring
    uint32_t pend_at = 0;  // first item of the chunk in flight
    bool pending = false;
    u64x2 pend[kLaneChunk / 2];

    __device__ __forceinline__ void issue(uint32_t at) {
#pragma unroll
        for (uint32_t i = 0; i < kLaneChunk / 2; ++i) pend[i] = row[at / 2 + i];
        pend_at = at;
        pending = true;
    }
    __device__ __forceinline__ void land() {
#pragma unroll
        for (uint32_t i = 0; i < kLaneChunk; ++i) {
            const uint32_t w = (pend_at + i) & (kLaneRing - 1);
            *reinterpret_cast<lds_u64p>(static_cast<uintptr_t>(ring + w * ring_stride)) = pend[i / 2][i & 1];
        }
        fill = pend_at + kLaneChunk;
        pending = false;
    }
    __device__ __forceinline__ void cands(uint32_t item, Cands& c) const {
        const uint64_t d =
            *reinterpret_cast<lds_u64p>(static_cast<uintptr_t>(ring + (item & (kLaneRing - 1)) * ring_stride));
#pragma unroll
        for (int w = 0; w < K; ++w) c[w] = static_cast<uint32_t>((d >> (w * bb)) & bmask);
    }
    // The refill point: the copy of the chunk loaded at the previous point, then
    // the next chunk's loads for a lane that is still inserting and has at most
    // 2*kLaneChunk + 1 items in its ring.
    __device__ __forceinline__ void refill(bool inserting, uint32_t item, uint32_t end) {
        if (pending) land();
        if (inserting && fill < end && fill - item <= 2 * kLaneChunk + 1) issue(fill);
    }
};

// BGH25 with the lane's table in LDS: the decisions and probe count of
// chain_bgh_res (cuckoo::insert_core_k), one probe per loop iteration, the loop
// running while any lane of the wave is inserting so that the refill points
// are wave-uniform.
template <int K, class E>
__device__ __forceinline__ bool chain_bgh_lane(const Shape& S, const ResCodec<K, E>& R, LaneFeed<K>& f,
                                               const LaneTable<E>& t, bool mine, uint64_t walk, uint64_t rnd,
                                               uint32_t& probes) {
    constexpr E kEmpty = static_cast<E>(~E(0));
    const uint32_t end = S.pnum;
    uint32_t item = 0, coll = 0, pr = 0, stash_left = S.stash;
    bool active = mine, failed = false;
    Cands c = {}, nx = {};
    if (mine) {
        f.cands(0, c);
        if (end > 1) f.cands(1, nx);
    }
    for (uint32_t it = 0; __builtin_amdgcn_ballot_w64(active) != 0; ++it) {
        if ((it & (kLaneChunk - 1)) == 0) f.refill(active, item, end);
        if (!active) continue;
        if (rnd < K) rnd = splitmix(walk);
        const uint32_t which = take_choice<K>(rnd);
        const uint32_t b = pick<K>(c, which);
        const E old = t.get(b);
        ++pr;
        bool advance = old == kEmpty;
        if (!advance && ++coll > S.kick) {
            if (!stash_left) {
                failed = true;
                active = false;
                continue;
            }
            --stash_left;  // the item in hand goes to the stash; nothing is written
            advance = true;
        } else {
            t.put(b, R.enc(c, which));
        }
        if (advance) {
            ++item;
            c = nx;
            if (item + 1 < end) f.cands(item + 1, nx);
            coll = 0;
            active = item < end;
        } else {
            R.dec(old, b, c);
        }
    }
    probes = pr;
    return failed;
}

// DRRT18 with the lane's table in LDS: the decisions and probe count of
// chain_drrt_res as a state machine of one probe per iteration -- first the
// new item's candidates in order (`w`), then the kick walk from candidate 0.
template <int K, class E>
__device__ __forceinline__ bool chain_drrt_lane(const Shape& S, const ResCodec<K, E>& R, LaneFeed<K>& f,
                                                const LaneTable<E>& t, bool mine, uint32_t& probes) {
    constexpr E kEmpty = static_cast<E>(~E(0));
    const uint32_t end = S.pnum;
    uint32_t item = 0, pr = 0, stash_left = S.stash;
    uint32_t w = 0, choice = 0, target = 0, kicks = 0;
    bool trying = true, active = mine, failed = false;
    Cands c = {}, nx = {};
    if (mine) {
        f.cands(0, c);
        if (end > 1) f.cands(1, nx);
    }
    for (uint32_t it = 0; __builtin_amdgcn_ballot_w64(active) != 0; ++it) {
        if ((it & (kLaneChunk - 1)) == 0) f.refill(active, item, end);
        if (!active) continue;
        const uint32_t b = trying ? pick<K>(c, w) : target;
        const E e = t.get(b);
        ++pr;
        // 0: the item in hand stays in hand; 1: it is placed or stashed; 2: the trial fails.
        uint32_t outcome = 0;
        if (e == kEmpty) {
            t.put(b, R.enc(c, trying ? w : choice));
            outcome = 1;
        } else if (trying) {
            if (++w == uint32_t(K)) {
                trying = false;
                choice = 0;
                target = c.x;
                kicks = 0;
            }
        } else if (kicks >= S.kick) {
            outcome = stash_left ? 1 : 2;
        } else {
            t.put(b, R.enc(c, choice));
            choice = R.dec(e, b, c);
            bool found = false;
#pragma unroll
            for (int s = 0; s < K; ++s) {
                if (found) break;
                choice = (choice + 1 == uint32_t(K)) ? 0 : choice + 1;
                const uint32_t nb = pick<K>(c, choice);
                if (nb != b) {
                    target = nb;
                    found = true;
                }
            }
            if (found)
                ++kicks;
            else
                outcome = stash_left ? 1 : 2;
        }
        if (outcome == 2) {
            failed = true;
            active = false;
            continue;
        }
        if (outcome == 1) {
            if (e != kEmpty) --stash_left;  // stashed, not placed
            ++item;
            c = nx;
            if (item + 1 < end) f.cands(item + 1, nx);
            trying = true;
            w = 0;
            active = item < end;
        }
    }
    probes = pr;
    return failed;
}

}  // namespace ck
