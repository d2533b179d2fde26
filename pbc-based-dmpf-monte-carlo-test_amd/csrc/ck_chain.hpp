// ck_chain.hpp -- the serial insertion chains, generic over where the table and
// the bucket map live (HBM or LDS) and over their widths.
//
// A table entry packs the item ordinal and its current choice:
//   u32: ordinal | choice << 30      u16: ordinal | choice << 14 (ordinal < 2^14)
// and all-ones is empty.  map[o*K + w] is bucket(o, w), precomputed by the build.
#pragma once
#include <type_traits>

#include "ck_device.hpp"
#include "cuckoo_trial.hpp"

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

// The bucket-map layout runs the trial API of include/cuckoo_trial.hpp
// directly: the slot's table as PackedBins of its entry width and its map rows
// as PrecomputedBuckets.  (The other layouts below make the same decisions
// over entries that carry the occupant's candidates.)
template <int K, class M, class E>
__device__ __forceinline__ bool chain_bgh(const Shape& S, const M* __restrict__ m, E* __restrict__ t, uint64_t walk,
                                          uint64_t rnd, uint32_t& probes) {
    return cuckoo::insert_core_k<true, K>(S.pnum, S.kick, S.stash, cuckoo::PackedBins<E, EntryCodec<E>::SHIFT>{t, 1},
                                          cuckoo::PrecomputedBuckets<M, K>{m}, walk, rnd, probes);
}
template <int K, class M, class E>
__device__ __forceinline__ bool chain_drrt(const Shape& S, const M* __restrict__ m, E* __restrict__ t,
                                           uint32_t& probes) {
    return cuckoo::insert_core_k<false, K>(S.pnum, S.kick, S.stash, cuckoo::PackedBins<E, EntryCodec<E>::SHIFT>{t, 1},
                                           cuckoo::PrecomputedBuckets<M, K>{m}, 0, 0, probes);
}

// An item's K <= 4 bucket indices, as one vector value.  A plain array would be
// a stack object: an index known only at run time (pick(), or a loop the
// optimiser has not unrolled yet) keeps it in memory, and the backend then
// places it in LDS (K*4 bytes per lane).  A chain kernel holding LDS cannot
// share a CU with a 160 KB build block, which slowed overlapped passes
// 2.5-fold.  A vector stays in registers.
typedef uint32_t Cands __attribute__((ext_vector_type(4)));

// c[i] for a run-time i < K, as selects on the vector's elements.
template <int K>
__device__ __forceinline__ uint32_t pick(const Cands& c, uint32_t i) {
    uint32_t v = c.x;
    if constexpr (K > 1) v = (i == 1u) ? c.y : v;
    if constexpr (K > 2) v = (i == 2u) ? c.z : v;
    if constexpr (K > 3) v = (i == 3u) ? c.w : v;
    return v;
}

// ---------------------------------------------------------------------------
// Descriptor layout (HBM tier, when k*bb + 2 <= 63 with bb = bit width of the
// largest bucket index): a table entry holds the occupant's k bucket indices,
// its current choice and the pass's tag,
//   desc | choice << (k*bb) | tag << (k*bb + 2),
// and an entry whose tag is not the pass's is empty.  A cleared table is
// all-ones, whose tag field (all-ones) no pass uses; a table reused without a
// clear holds only entries of earlier tags.  An eviction reads the occupant's
// candidates from the slot it just loaded instead of from the bucket map; new
// items take their descriptors from a per-trial stream in insertion order,
// prefetched one item ahead.
struct DescCodec {
    uint32_t bb;        // bits per bucket index
    uint32_t cs;        // choice shift = k*bb
    uint64_t bmask;     // (1 << bb) - 1
    uint64_t dmask;     // (1 << cs) - 1
    uint32_t ts = cs + 2;  // tag shift
    uint64_t tag = 0;      // this pass's tag (0 after a clear)
    __device__ __forceinline__ uint32_t bucket(uint64_t d, uint32_t w) const {
        return static_cast<uint32_t>((d >> (w * bb)) & bmask);
    }
    __device__ __forceinline__ bool empty(uint64_t e) const { return (e >> ts) != tag; }
    __device__ __forceinline__ uint64_t entry(uint64_t d, uint32_t w) const {
        return d | (uint64_t(w) << cs) | (tag << ts);
    }
};
constexpr uint64_t EMPTY64 = ~0ull;

template <int K>
__device__ __forceinline__ bool chain_bgh_desc(const Shape& S, const DescCodec& D, const uint64_t* __restrict__ desc,
                                               uint64_t* __restrict__ t, uint64_t walk, uint64_t rnd,
                                               uint32_t& probes) {
    uint32_t item = 0, coll = 0, stash_left = S.stash, pr = 0;
    uint64_t cur = desc[0];
    uint64_t next = S.pnum > 1 ? desc[1] : 0;
    bool failed = false;
    while (item < S.pnum) {
        if (rnd < K) rnd = splitmix(walk);
        const uint32_t which = take_choice<K>(rnd);
        const uint32_t b = D.bucket(cur, which);
        const uint64_t old = t[b];
        ++pr;
        bool advance = D.empty(old);
        if (!advance && ++coll > S.kick) {
            if (!stash_left) {
                failed = true;
                break;
            }
            --stash_left;  // the item in hand goes to the stash; nothing is written
            advance = true;
        } else {
            t[b] = D.entry(cur, which);
        }
        if (advance) {
            ++item;
            cur = next;
            if (item + 1 < S.pnum) next = desc[item + 1];
            coll = 0;
        } else {
            cur = old & D.dmask;
        }
    }
    probes = pr;
    return failed;
}

template <int K>
__device__ __forceinline__ bool chain_drrt_desc(const Shape& S, const DescCodec& D, const uint64_t* __restrict__ desc,
                                                uint64_t* __restrict__ t, uint32_t& probes) {
    uint32_t pr = 0, stash_left = S.stash;
    bool failed = false;
    uint64_t next = desc[0];
    for (uint32_t item = 0; item < S.pnum && !failed; ++item) {
        const uint64_t d = next;
        if (item + 1 < S.pnum) next = desc[item + 1];
        bool placed = false;
#pragma unroll
        for (int w = 0; w < K; ++w) {
            if (placed) break;
            ++pr;
            const uint32_t b = D.bucket(d, w);
            if (D.empty(t[b])) {
                t[b] = D.entry(d, w);
                placed = true;
            }
        }
        if (placed) continue;
        uint64_t cur = d;
        uint32_t choice = 0, target = D.bucket(d, 0), kicks = 0;
        for (;;) {
            const uint64_t disp = t[target];
            ++pr;
            if (D.empty(disp)) {
                t[target] = D.entry(cur, choice);
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
            t[target] = D.entry(cur, choice);
            cur = disp & D.dmask;
            choice = static_cast<uint32_t>(disp >> D.cs) & 3u;
            const uint32_t prev = target;
            bool found = false;
#pragma unroll
            for (int s = 0; s < K; ++s) {
                if (found) break;
                choice = (choice + 1 == K) ? 0 : choice + 1;
                const uint32_t nb = D.bucket(cur, choice);
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

// ---------------------------------------------------------------------------
// Residual layout: a table entry holds the occupant's other k-1 bucket indices
// (in candidate order, skipping its current one) and its choice,
//   others | choice << (k-1)*bb,
// and the occupant's current bucket is the slot the entry sits in.  An
// eviction therefore rebuilds the displaced item's whole candidate list from
// the entry it just read and the slot it read it from.  (k-1)*bb + 2 bits fit
// a u32 entry at n = 2^14, k = 3 (2*15 + 2 = 32, choice <= 2 keeps the entry
// below all-ones) and a u64 entry up to n = 2^20 and beyond, where the full
// descriptor (k*bb + 2 bits) no longer fits.  Between the two, 44-bit entries
// (Entries44) take 5.5 bytes -- n = 2^20, k = 3 needs 2*21 + 2 = 44 bits -- and
// 48-bit entries (Entries48) 6 bytes.  New
// items take their candidates from a stream in insertion order: one packed u64
// per item when k*bb <= 64, else the item's k map entries; it is prefetched
// one item ahead.

// A slot's table of 48-bit residual entries, 6 bytes apart from a 4-byte
// aligned base.  `t[b]` is a handle to entry b that reads and writes like an
// element of a u64 array.  A read loads the 8 bytes at the dword-aligned
// address at or 2 bytes below the entry (so the table carries 8 bytes of slack
// past the last entry) and maps the 48-bit all-ones (empty) to the u64
// all-ones.  A write through a handle that has been read stores the same 8
// bytes back with the entry replaced: the 2 neighbouring bytes are rewritten
// with the values just read, which is safe because one lane owns a table and
// writes it only through the handle of the probe in progress.  A write before
// any read is a 16-bit and a 32-bit store that leave the neighbours untouched.
struct Entries48 {
    char* base;
    struct Ref {
        char* a;
        mutable uint64_t raw = 0;  // the 8 bytes read, once read
        mutable bool read = false;
        __host__ __device__ __forceinline__ uint32_t mis() const {
            return static_cast<uint32_t>(reinterpret_cast<uintptr_t>(a)) & 3u;
        }
        __host__ __device__ __forceinline__ operator uint64_t() const {
            // pointer arithmetic, not an integer round trip, keeps the global address space
            const uint32_t m = mis();
            const uint32_t* w = reinterpret_cast<const uint32_t*>(a - m);
            raw = w[0] | (uint64_t(w[1]) << 32);
            read = true;
            const uint64_t v = (raw >> (m * 8)) & 0xFFFFFFFFFFFFull;
            return v == 0xFFFFFFFFFFFFull ? ~0ull : v;
        }
        __host__ __device__ __forceinline__ void operator=(uint64_t v) const {
            const uint32_t m = mis();
            if (read) {
                const uint64_t keep = ~(0xFFFFFFFFFFFFull << (m * 8));
                const uint64_t x = (raw & keep) | ((v & 0xFFFFFFFFFFFFull) << (m * 8));
                uint32_t* w = reinterpret_cast<uint32_t*>(a - m);
                w[0] = static_cast<uint32_t>(x);
                w[1] = static_cast<uint32_t>(x >> 32);
            } else if (m) {
                *reinterpret_cast<uint16_t*>(a) = static_cast<uint16_t>(v);
                *reinterpret_cast<uint32_t*>(a + 2) = static_cast<uint32_t>(v >> 16);
            } else {
                *reinterpret_cast<uint32_t*>(a) = static_cast<uint32_t>(v);
                *reinterpret_cast<uint16_t*>(a + 4) = static_cast<uint16_t>(v >> 32);
            }
        }
    };
    __host__ __device__ __forceinline__ Ref operator[](uint32_t b) const { return Ref{base + size_t(b) * 6}; }
};

// A slot's table of 44-bit residual entries packed end to end (5.5 bytes per
// entry): entry b is bits [44b, 44b + 44) of the table read as little-endian
// dwords from a 4-byte aligned base.  44b is a multiple of 4, so the entry
// starts at bit off = 44b mod 32, one of 0, 4, ..., 28, of dword 44b >> 5, and
// spans two dwords when off <= 20, three when off is 24 or 28 (2 entries in
// 8).  A read loads exactly the dwords the entry spans (two as one 8-byte
// load) and maps the 44-bit all-ones (empty) to the u64 all-ones.  A write
// stores the same dwords back with the entry's bits replaced and the
// neighbours' bits as they were read; as with Entries48 that is safe because
// one lane owns a table and nothing else writes it between the probe's read
// and its write.  A write before any read reads first.
struct Entries44 {
    static constexpr uint64_t kMask = (1ull << 44) - 1;
    char* base;
    struct Ref {
        uint32_t* w;
        uint32_t off;
        mutable uint64_t lo = 0;  // the entry's first two dwords, once read
        mutable uint32_t hi = 0;  // its third dword (off > 20), once read
        mutable bool read = false;
        __host__ __device__ __forceinline__ bool three() const { return off > 20; }
        __host__ __device__ __forceinline__ void load() const {
            lo = w[0] | (uint64_t(w[1]) << 32);
            hi = three() ? w[2] : 0u;
            read = true;
        }
        __host__ __device__ __forceinline__ operator uint64_t() const {
            load();
            const uint64_t top = three() ? uint64_t(hi) << (64 - off) : 0ull;
            const uint64_t v = ((lo >> off) | top) & kMask;
            return v == kMask ? ~0ull : v;
        }
        __host__ __device__ __forceinline__ void operator=(uint64_t v) const {
            if (!read) load();
            v &= kMask;
            const uint64_t x = (lo & ~(kMask << off)) | (v << off);
            w[0] = static_cast<uint32_t>(x);
            w[1] = static_cast<uint32_t>(x >> 32);
            if (three()) {
                const uint32_t hm = static_cast<uint32_t>(kMask >> (64 - off));  // the entry's bits in dword 2
                w[2] = (hi & ~hm) | static_cast<uint32_t>(v >> (64 - off));
            }
        }
    };
    __host__ __device__ __forceinline__ Ref operator[](uint32_t b) const {
        const uint64_t bit = uint64_t(b) * 44;
        return Ref{reinterpret_cast<uint32_t*>(base) + (bit >> 5), static_cast<uint32_t>(bit & 31)};
    }
    // Table bytes for m entries: whole dwords, plus 8 bytes of slack.
    static constexpr size_t bytes(size_t m) { return (m * 44 + 31) / 32 * 4 + 8; }
};

// Residual table entry types of the kernels: u32, u64, E44 (Entries44) or E48
// (Entries48), the last two decoded as u64.
struct E44 {};
struct E48 {};
template <class E>
struct ResWord {
    using type = E;
};
template <>
struct ResWord<E44> {
    using type = uint64_t;
};
template <>
struct ResWord<E48> {
    using type = uint64_t;
};
template <class E>
__device__ __forceinline__ auto res_table(char* p) {
    if constexpr (std::is_same<E, E48>::value)
        return Entries48{p};
    else if constexpr (std::is_same<E, E44>::value)
        return Entries44{p};
    else
        return reinterpret_cast<E*>(p);
}

template <int K, class E>
struct ResCodec {
    uint32_t bb;     // bits per bucket index
    uint32_t cs;     // choice shift = (K-1)*bb
    uint64_t bmask;  // (1 << bb) - 1
    __device__ __forceinline__ E enc(const Cands& c, uint32_t w) const {
        uint64_t e = uint64_t(w) << cs;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const uint32_t sh = (uint32_t(i) - (uint32_t(i) > w ? 1u : 0u)) * bb;
            e |= (uint32_t(i) == w) ? 0ull : (uint64_t(c[i]) << sh);
        }
        return static_cast<E>(e);
    }
    // Candidates of the occupant of bucket b; returns its choice.
    __device__ __forceinline__ uint32_t dec(E e, uint32_t b, Cands& c) const {
        const uint64_t v = e;
        const uint32_t w = static_cast<uint32_t>(v >> cs) & 3u;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const uint32_t sh = (uint32_t(i) - (uint32_t(i) > w ? 1u : 0u)) * bb;
            c[i] = (uint32_t(i) == w) ? b : static_cast<uint32_t>((v >> sh) & bmask);
        }
        return w;
    }
};

// Candidate stream of new items: M = u64 is one packed descriptor per item
// (bucket w at bits [w*bb, (w+1)*bb)), otherwise K map entries per item.
// The chains read the next item's Word one item ahead (raw) and unpack it only
// when that item comes into hand: unpacking at load time would wait for the
// load there and then, a memory round trip on every item advance.
template <int K, class M>
struct CandStream {
    using Word = std::conditional_t<sizeof(M) == 8, uint64_t, Cands>;
    const M* m;
    uint32_t bb;
    uint64_t bmask;
    __device__ __forceinline__ Word raw(uint32_t item) const {
        if constexpr (sizeof(M) == 8) {
            return m[item];
        } else {
            Cands c = {};
#pragma unroll
            for (int w = 0; w < K; ++w) c[w] = m[item * K + w];
            return c;
        }
    }
    __device__ __forceinline__ void unpack(const Word& d, Cands& c) const {
        if constexpr (sizeof(M) == 8) {
#pragma unroll
            for (int w = 0; w < K; ++w) c[w] = static_cast<uint32_t>((d >> (w * bb)) & bmask);
        } else {
            c = d;
        }
    }
    __device__ __forceinline__ void load(uint32_t item, Cands& c) const { unpack(raw(item), c); }
};

// Items [begin, end) that one residual-layout launch inserts; the candidate
// stream holds them from index 0.  A trial may run as several such windows in
// order: each ends at an item boundary, where nothing is in hand and the whole
// chain state is the walk, the random word, the stash count and the probe count.
struct ItemSpan {
    uint32_t begin, end;
};

// BGH25 over the residual layout (the decisions of cuckoo::insert_core_k).
// T is the table: an E* or Entries44 / Entries48 (E = u64).
template <int K, class E, class M, class T>
__device__ __forceinline__ bool chain_bgh_res(const Shape& S, const ResCodec<K, E>& R, const CandStream<K, M>& st,
                                              T t, ItemSpan span, uint64_t& walk, uint64_t& rnd,
                                              uint32_t& stash_left, uint32_t& probes) {
    constexpr E kEmpty = static_cast<E>(~E(0));
    uint32_t item = span.begin, coll = 0, pr = 0;
    Cands c = {};
    typename CandStream<K, M>::Word nx = {};
    st.load(0, c);
    if (item + 1 < span.end) nx = st.raw(1);
    bool failed = false;
    while (item < span.end) {
        if (rnd < K) rnd = splitmix(walk);
        const uint32_t which = take_choice<K>(rnd);
        const uint32_t b = pick<K>(c, which);
        auto&& slot = t[b];
        const E old = slot;
        ++pr;
        bool advance = old == kEmpty;
        if (!advance && ++coll > S.kick) {
            if (!stash_left) {
                failed = true;
                break;
            }
            --stash_left;  // the item in hand goes to the stash; nothing is written
            advance = true;
        } else {
            slot = R.enc(c, which);
        }
        if (advance) {
            ++item;
            st.unpack(nx, c);
            if (item + 1 < span.end) nx = st.raw(item + 1 - span.begin);
            coll = 0;
        } else {
            R.dec(old, b, c);
        }
    }
    probes = pr;
    return failed;
}

// DRRT18 over the residual layout (the decisions of cuckoo::insert_core_k).
template <int K, class E, class M, class T>
__device__ __forceinline__ bool chain_drrt_res(const Shape& S, const ResCodec<K, E>& R, const CandStream<K, M>& st,
                                               T t, ItemSpan span, uint32_t& stash_left, uint32_t& probes) {
    constexpr E kEmpty = static_cast<E>(~E(0));
    uint32_t pr = 0;
    bool failed = false;
    typename CandStream<K, M>::Word nx = st.raw(0);
    for (uint32_t item = span.begin; item < span.end && !failed; ++item) {
        Cands c = {};
        st.unpack(nx, c);
        if (item + 1 < span.end) nx = st.raw(item + 1 - span.begin);
        bool placed = false;
#pragma unroll
        for (int w = 0; w < K; ++w) {
            if (placed) break;
            ++pr;
            auto&& slot = t[c[w]];
            if (static_cast<E>(slot) == kEmpty) {
                slot = R.enc(c, uint32_t(w));
                placed = true;
            }
        }
        if (placed) continue;
        uint32_t choice = 0, target = c[0], kicks = 0;
        for (;;) {
            auto&& slot = t[target];
            const E disp = slot;
            ++pr;
            if (disp == kEmpty) {
                slot = R.enc(c, choice);
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
            slot = R.enc(c, choice);
            choice = R.dec(disp, target, c);
            const uint32_t prev = target;
            bool found = false;
#pragma unroll
            for (int s = 0; s < K; ++s) {
                if (found) break;
                choice = (choice + 1 == K) ? 0 : choice + 1;
                const uint32_t nb = pick<K>(c, choice);
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

// ---------------------------------------------------------------------------
// One-round-trip probes.  On this GPU family vector-memory loads and stores
// retire in order under one counter, so a probe written as load + store makes
// the next probe's load wait for the store's acknowledgement as well: two memory
// round trips per probe.  Here every probe is a single read-modify-write at the
// L2 -- an exchange that installs the item and returns the occupant, or, for a
// drrt18 first-placement attempt, a compare-and-swap against empty.  Where the
// reference would have left the slot untouched (the kick bound trips), the
// occupant is written back, so tables match the load/store form exactly.
template <class E>
__device__ __forceinline__ E xchg(E* p, E v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class E>
__device__ __forceinline__ E cas_empty(E* p, E v) {
    E expected = static_cast<E>(~E(0));
    __hip_atomic_compare_exchange_strong(p, &expected, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expected;  // the value found: all-ones when the slot was empty and now holds v
}

// Entry access used by the atomic chains: the map layout (ordinal | choice)
// with a bucket map, or the descriptor layout (candidates | choice).
template <int K, class M, class E>
struct MapView {
    const M* m;
    __device__ __forceinline__ uint32_t bucket_of(uint64_t cur, uint32_t w) const {
        return m[uint32_t(cur) * K + w];
    }
    __device__ __forceinline__ E pack(uint64_t cur, uint32_t choice) const {
        return pack_entry<E>(uint32_t(cur), choice);
    }
    __device__ __forceinline__ uint64_t who(E e) const { return entry_ordinal<E>(e); }
    __device__ __forceinline__ uint32_t choice(E e) const { return entry_choice<E>(e); }
    __device__ __forceinline__ uint64_t item(uint32_t o) const { return o; }
};
template <int K>
struct DescView {
    DescCodec D;
    const uint64_t* desc;
    __device__ __forceinline__ uint32_t bucket_of(uint64_t cur, uint32_t w) const { return D.bucket(cur, w); }
    __device__ __forceinline__ uint64_t pack(uint64_t cur, uint32_t choice) const {
        return cur | (uint64_t(choice) << D.cs);
    }
    __device__ __forceinline__ uint64_t who(uint64_t e) const { return e & D.dmask; }
    __device__ __forceinline__ uint32_t choice(uint64_t e) const { return uint32_t(e >> D.cs) & 3u; }
    __device__ __forceinline__ uint64_t item(uint32_t o) const { return desc[o]; }
};

// BGH25 with exchange probes (the decisions of cuckoo::insert_core_k).
template <int K, class E, class V>
__device__ __forceinline__ bool chain_bgh_xchg(const Shape& S, const V& v, E* __restrict__ t, uint64_t walk,
                                               uint64_t rnd, uint32_t& probes) {
    uint32_t item = 0, coll = 0, stash_left = S.stash, pr = 0;
    uint64_t cur = v.item(0);
    uint64_t next = S.pnum > 1 ? v.item(1) : 0;
    bool failed = false;
    while (item < S.pnum) {
        if (rnd < K) rnd = splitmix(walk);
        const uint32_t which = take_choice<K>(rnd);
        const uint32_t b = v.bucket_of(cur, which);
        const E old = xchg(t + b, static_cast<E>(v.pack(cur, which)));
        ++pr;
        bool advance = old == static_cast<E>(~E(0));
        if (!advance && ++coll > S.kick) {
            t[b] = old;  // the reference stops before evicting: restore the occupant
            if (!stash_left) {
                failed = true;
                break;
            }
            --stash_left;
            advance = true;
        }
        if (advance) {
            ++item;
            cur = next;
            if (item + 1 < S.pnum) next = v.item(item + 1);
            coll = 0;
        } else {
            cur = v.who(old);
        }
    }
    probes = pr;
    return failed;
}

// DRRT18 with compare-and-swap first placements and exchange evictions.
template <int K, class E, class V>
__device__ __forceinline__ bool chain_drrt_xchg(const Shape& S, const V& v, E* __restrict__ t, uint32_t& probes) {
    constexpr E kEmpty = static_cast<E>(~E(0));
    uint32_t pr = 0, stash_left = S.stash;
    bool failed = false;
    uint64_t next = v.item(0);
    for (uint32_t item = 0; item < S.pnum && !failed; ++item) {
        const uint64_t d = next;
        if (item + 1 < S.pnum) next = v.item(item + 1);
        bool placed = false;
#pragma unroll
        for (int w = 0; w < K; ++w) {
            if (placed) break;
            ++pr;
            if (cas_empty(t + v.bucket_of(d, w), static_cast<E>(v.pack(d, w))) == kEmpty) placed = true;
        }
        if (placed) continue;
        uint64_t cur = d;
        uint32_t choice = 0, target = v.bucket_of(d, 0), kicks = 0;
        for (;;) {
            const E disp = xchg(t + target, static_cast<E>(v.pack(cur, choice)));
            ++pr;
            if (disp == kEmpty) break;
            if (kicks >= S.kick) {
                t[target] = disp;  // bound reached before evicting: restore the occupant
                if (stash_left) {
                    --stash_left;
                    break;
                }
                failed = true;
                break;
            }
            cur = v.who(disp);
            choice = v.choice(disp);
            const uint32_t prev = target;
            bool found = false;
#pragma unroll
            for (int s = 0; s < K; ++s) {
                if (found) break;
                choice = (choice + 1 == K) ? 0 : choice + 1;
                const uint32_t nb = v.bucket_of(cur, choice);
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

// Order-independent digest of a final table: the wrapping sum over occupied
// buckets b of mix64(b * golden ^ H ^ choice << 61), where H is the wrapping
// sum over w < k of mix64(w << 32 | bucket(occupant, w)).  mix64 is the
// SplitMix64 output function.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// Hash of an item's candidate list, independent of how it is stored.
__host__ __device__ __forceinline__ uint64_t cand_hash_term(uint32_t w, uint32_t bucket) {
    return mix64((uint64_t(w) << 32) | bucket);
}
__host__ __device__ __forceinline__ uint64_t slot_digest(uint32_t b, uint64_t cand_hash, uint32_t choice) {
    return mix64(uint64_t(b) * 0x9E3779B97F4A7C15ull ^ cand_hash ^ (uint64_t(choice) << 61));
}

}  // namespace ck
