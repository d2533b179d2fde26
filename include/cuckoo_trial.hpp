/*
 * cuckoo_trial.hpp -- the device-level trial API of the MI355X cuckoo Monte
 * Carlo library: one cuckoo-insertion trial over a caller's table ("Bins")
 * and bucket function ("Buckets"), callable from HIP kernels and from host
 * code.  The library's own chain kernels instantiate it (the bucket-map and
 * LDS-tier placements); the other HBM-tier layouts are specialisations of
 * the same decisions over table entries that carry candidate buckets.
 *
 * Replaces cuckoo_core.hpp:17-161 of the reference:
 *   insert_core<BGH>(pnum, maxkick, bins, bucket, walk, randomness) -> failed
 * with the same concepts:
 *   Bins     value_type; value_type& operator[](uint32_t bucket) const;
 *            static value_type empty(); static value_type pack(ordinal, choice);
 *            static uint32_t ordinal(value_type); static uint32_t choice(value_type)
 *   Buckets  uint32_t operator()(uint32_t ordinal, uint32_t which) const
 * and ready-made models of them: PackedBins (StridedBins = u32 ordinal |
 * choice << 16, SharedBins = u16 ordinal | choice << 14), PrecomputedBuckets
 * (a bucket map) and CachedBuckets (the Feistel PRP over a round cache).
 *
 * Extension (no reference counterpart; with K = 3 and stash = 0 it is the
 * reference exactly): insert_core_k<BGH, K>(pnum, maxkick, stash, bins,
 * bucket, walk, randomness, probes) with K candidate buckets per item
 * (which = r mod K, r /= K; drrt18 cycles (choice + 1) mod K), `stash` slots
 * for items whose walk ends at the kick bound, and the count of table reads.
 *
 * Per-trial inputs (walk, randomness, the PRP key and round cache, the affine
 * shift and step) come from ck_trial_inputs_of / ck_round_cache in
 * cuckoo_mc.h, or from the library's own build kernels.
 */
#ifndef CUCKOO_TRIAL_HPP
#define CUCKOO_TRIAL_HPP
#include <cstddef>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CUCKOO_TRIAL_FN __host__ __device__ __forceinline__
#else
#define CUCKOO_TRIAL_FN inline
#endif

namespace cuckoo {

// ---------------------------------------------------------------- Bins

// A table of packed entries ordinal | choice << ChoiceShift, element i at
// base[i * stride]; all-ones is empty.  A valid entry is never all-ones: the
// choice is at most 2 for K = 3 (at most K - 1 < 3 in general, and an
// ordinal below 2^ChoiceShift - 1 when K = 4).
template <class Value, unsigned ChoiceShift>
struct PackedBins {
    using value_type = Value;
    Value* base;
    size_t stride;
    CUCKOO_TRIAL_FN Value& operator[](uint32_t index) const { return base[size_t(index) * stride]; }
    CUCKOO_TRIAL_FN static Value empty() { return static_cast<Value>(~Value(0)); }
    CUCKOO_TRIAL_FN static Value pack(uint32_t ordinal, uint32_t choice) {
        return static_cast<Value>(ordinal | (choice << ChoiceShift));
    }
    CUCKOO_TRIAL_FN static uint32_t ordinal(Value e) { return uint32_t(e) & ((uint32_t(1) << ChoiceShift) - 1u); }
    CUCKOO_TRIAL_FN static uint32_t choice(Value e) { return (uint32_t(e) >> ChoiceShift) & 3u; }
};
using StridedBins = PackedBins<uint32_t, 16>;  // the reference's u32 table
using SharedBins = PackedBins<uint16_t, 14>;   // its u16 shared-memory table (pnum < 2^14)

// ---------------------------------------------------------------- Buckets

// bucket(o, w) read from a precomputed map: values[o * K + w].
template <class Value = uint16_t, int K = 3>
struct PrecomputedBuckets {
    const Value* values;
    CUCKOO_TRIAL_FN uint32_t operator()(uint32_t ordinal, uint32_t which) const {
        return uint32_t(values[size_t(ordinal) * K + which]);
    }
};

// bucket(o, w) through the Feistel PRP over a round cache
// (cuckoo_core.hpp:42-72): the element is o (sequential input) or
// (shift + o * step) mod 2^ell_i (affine); y = K * element + w runs 8 rounds
// L, R <- R, L ^ rounds[r << half_bits | R] over half_bits-bit halves,
// cycle-walked until y < domain; the bucket is y / width.  rounds holds the
// low half_bits bits of word 3 of AES_prp("CESP", r, half_bits, R)
// (ck_round_cache fills it), element i at rounds[i * stride].
template <class Value = uint16_t, int K = 3>
struct CachedBuckets {
    const Value* rounds;
    size_t stride;
    uint64_t domain, width, mask, range_mask, shift, step;
    uint32_t half_bits;
    bool sequential;
    CUCKOO_TRIAL_FN uint32_t operator()(uint32_t ordinal, uint32_t which) const {
        const uint64_t element = sequential ? uint64_t(ordinal) : (shift + uint64_t(ordinal) * step) & range_mask;
        uint64_t y = element * K + which;
        do {
            uint64_t left = y >> half_bits, right = y & mask;
            for (unsigned r = 0; r < 8; ++r) {
                const uint64_t next = left ^ uint64_t(rounds[((size_t(r) << half_bits) + right) * stride]);
                left = right;
                right = next & mask;
            }
            y = (left << half_bits) | right;
        } while (y >= domain);
        // The quotient is below the bucket count; 32-bit division when it can be.
        if (domain <= 0xffffffffull) return uint32_t(y) / uint32_t(width);
        return uint32_t(y / width);
    }
};

// ---------------------------------------------------------------- insertion

// SplitMix64 (Steele, Lea & Flood, 2014): the walk's random words.
CUCKOO_TRIAL_FN uint64_t splitmix64(uint64_t& state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// One trial: inserts ordinals 0 .. pnum-1 into `bins`; returns true when the
// trial fails.  probes receives the number of table reads.
//
// bgh25 (BGH = true): each step draws which = r mod K from the random word r
// (refilled with splitmix64(walk) when r < K, then r /= K) and probes
// bucket(cur, which).  An empty bucket takes the item in hand and the next
// item starts.  Otherwise the step is a collision: past `maxkick` collisions
// of one item the item in hand goes to the stash (the next item starts) or,
// with the stash full, the trial fails -- before evicting; within the bound
// the item in hand takes the bucket and the occupant becomes the item in hand.
//
// drrt18 (BGH = false): a new item tries buckets w = 0 .. K-1 in order and
// takes the first empty one.  Otherwise it evicts the occupant of its bucket
// 0; each displaced item moves to its next candidate after its stored choice,
// (choice + 1) mod K, whose bucket differs from the one it left, and evicts
// that bucket's occupant in turn.  The walk fails (or stashes the item in
// hand) when `maxkick` evictions have happened or no such candidate exists.
template <bool BGH, int K, class Bins, class Buckets>
CUCKOO_TRIAL_FN bool insert_core_k(uint32_t pnum, uint32_t maxkick, uint32_t stash, Bins bins, Buckets bucket,
                                   uint64_t walk, uint64_t randomness, uint32_t& probes) {
    using Value = typename Bins::value_type;
    const Value empty = Bins::empty();
    uint32_t reads = 0, stash_left = stash;
    bool failed = false;
    if constexpr (BGH) {
        uint32_t item = 0, cur = 0, collisions = 0;
        while (item < pnum) {
            if (randomness < uint64_t(K)) randomness = splitmix64(walk);
            const uint32_t which = static_cast<uint32_t>(randomness % K);
            randomness /= K;
            const uint32_t target = bucket(cur, which);
            const Value old = bins[target];
            ++reads;
            if (old == empty) {
                bins[target] = Bins::pack(cur, which);
                cur = ++item;
                collisions = 0;
                continue;
            }
            if (++collisions > maxkick) {
                if (stash_left == 0) {
                    failed = true;
                    break;
                }
                --stash_left;  // the item in hand goes to the stash; nothing is written
                cur = ++item;
                collisions = 0;
                continue;
            }
            bins[target] = Bins::pack(cur, which);
            cur = Bins::ordinal(old);
        }
    } else {
        for (uint32_t item = 0; item < pnum && !failed; ++item) {
            const uint32_t first = bucket(item, 0);
            bool placed = false;
            for (uint32_t w = 0; w < uint32_t(K) && !placed; ++w) {
                const uint32_t b = w == 0 ? first : bucket(item, w);
                ++reads;
                if (bins[b] == empty) {
                    bins[b] = Bins::pack(item, w);
                    placed = true;
                }
            }
            if (placed) continue;
            uint32_t cur = item, choice = 0, target = first, kicks = 0;
            for (;;) {
                const Value displaced = bins[target];
                ++reads;
                if (displaced == empty) {
                    bins[target] = Bins::pack(cur, choice);
                    break;
                }
                if (kicks >= maxkick) {
                    if (stash_left) {
                        --stash_left;
                        break;
                    }
                    failed = true;
                    break;
                }
                bins[target] = Bins::pack(cur, choice);
                cur = Bins::ordinal(displaced);
                choice = Bins::choice(displaced);
                const uint32_t previous = target;
                bool found = false;
                for (uint32_t s = 0; s < uint32_t(K) && !found; ++s) {
                    choice = (choice + 1 == uint32_t(K)) ? 0u : choice + 1;
                    const uint32_t next = bucket(cur, choice);
                    if (next != previous) {
                        target = next;
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
    }
    probes = reads;
    return failed;
}

// The reference's form: three candidates, no stash.
template <bool BGH, class Bins, class Buckets>
CUCKOO_TRIAL_FN bool insert_core(uint32_t pnum, uint32_t maxkick, Bins bins, Buckets bucket, uint64_t walk,
                                 uint64_t randomness) {
    uint32_t probes = 0;
    return insert_core_k<BGH, 3>(pnum, maxkick, 0u, bins, bucket, walk, randomness, probes);
}

}  // namespace cuckoo

#endif
