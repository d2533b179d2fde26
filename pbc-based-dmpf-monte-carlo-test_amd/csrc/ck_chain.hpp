// ck_chain.hpp -- the serial insertion chains, generic over where the table and
// the bucket map live (HBM or LDS) and over their widths.
//
// A table entry packs the item ordinal and its current choice:
//   u32: ordinal | choice << 30      u16: ordinal | choice << 14 (ordinal < 2^14)
// and all-ones is empty.  map[o*K + w] is bucket(o, w), precomputed by the build.
#pragma once
#include "ck_device.hpp"

namespace ck {

template <class E>
struct EntryCodec;
template <>
struct EntryCodec<uint32_t> {
    static constexpr uint32_t SHIFT = 30;
};
template <>
struct EntryCodec<uint16_t> {
    static constexpr uint32_t SHIFT = 14;
};

template <class E>
__device__ __forceinline__ E pack_entry(uint32_t ordinal, uint32_t choice) {
    return static_cast<E>(ordinal | (choice << EntryCodec<E>::SHIFT));
}
template <class E>
__device__ __forceinline__ uint32_t entry_ordinal(E e) {
    return uint32_t(e) & ((1u << EntryCodec<E>::SHIFT) - 1u);
}
template <class E>
__device__ __forceinline__ uint32_t entry_choice(E e) {
    return (uint32_t(e) >> EntryCodec<E>::SHIFT) & 3u;
}
template <class E>
__device__ __forceinline__ constexpr E empty_entry() {
    return static_cast<E>(~0u);
}

// BGH25 random-walk insertion (cuckoo_core.hpp:123-133) with the k/stash
// extension: each step draws which = r % K (refilling r from SplitMix64 when
// r < K); an empty target takes the item; otherwise the walk fails (or stashes
// the item in hand) once collisions exceed kick, else evicts the occupant.
template <int K, class M, class E>
__device__ __forceinline__ bool chain_bgh(const Shape& S, const M* __restrict__ m, E* __restrict__ t, uint64_t walk,
                                          uint64_t rnd, uint32_t& probes) {
    uint32_t item = 0, cur = 0, coll = 0, stash_left = S.stash, pr = 0;
    bool failed = false;
    while (item < S.pnum) {
        if (rnd < K) rnd = splitmix(walk);
        const uint32_t which = take_choice<K>(rnd);
        const uint32_t b = m[cur * K + which];
        const E old = t[b];
        ++pr;
        if (old == empty_entry<E>()) {
            t[b] = pack_entry<E>(cur, which);
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
        t[b] = pack_entry<E>(cur, which);
        cur = entry_ordinal<E>(old);
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
// extension: try candidates 0..K-1 in order; otherwise evict from candidate 0
// and move the displaced item to its next candidate (stored choice + 1, mod K)
// whose bucket differs from the one it left; fail (or stash) at kick evictions
// or when no such candidate exists.
template <int K, class M, class E>
__device__ __forceinline__ bool chain_drrt(const Shape& S, const M* __restrict__ m, E* __restrict__ t,
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
            if (t[c[w]] == empty_entry<E>()) {
                t[c[w]] = pack_entry<E>(item, uint32_t(w));
                placed = true;
            }
        }
        if (placed) continue;
        uint32_t cur = item, choice = 0, target = c[0], kicks = 0;
        for (;;) {
            const E disp = t[target];
            ++pr;
            if (disp == empty_entry<E>()) {
                t[target] = pack_entry<E>(cur, choice);
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
            t[target] = pack_entry<E>(cur, choice);
            cur = entry_ordinal<E>(disp);
            choice = entry_choice<E>(disp);
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

}  // namespace ck
