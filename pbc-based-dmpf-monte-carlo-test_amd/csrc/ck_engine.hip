// ck_engine.hip -- batched and served sweeps: many experiment shapes ("jobs")
// share device passes.
//
// A pass is a list of segments; a segment is up to 64 consecutive trials of
// one job, so one chain wave (and 64 build waves) always sees one shape and
// reads it with scalar loads.  Segments of different jobs -- different n,
// buckets, kick, stash, round-cache size, seed -- share a pass as long as they
// share the kernel instance: k, insertion policy, table layout and the
// cached/direct PRP form (the job's "group").  Each pass has its own arena for
// the maps and tables of its slots, its walk/randomness arrays and its failure
// list.  A ring of pass buffers lets the builds of later passes run on one
// stream while earlier passes chain on two others: consecutive passes chain on
// alternate streams, so the slowest trials of one pass (its tail) overlap the
// bulk of the next and the random-access path stays saturated.
//
// Per-job counters (failures, probes, trials) live in device memory: the
// chain waves add their ballot counts to their job's slot; the host reads them
// once every pass holding the job has finished.  A scheduler thread admits queued jobs as
// counter slots free up, plans passes into free buffers (continuous batching:
// jobs that arrive while others run join the next pass), and gathers finished
// passes in order.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ck_build.hpp"
#include "ck_chain.hpp"
#include "ck_device.hpp"
#include "ck_internal.hpp"
#include "ck_regions.hpp"

namespace ck {
namespace eng {

// One segment: live trials first .. first+count-1 of one batch of one job.
// Slot i of the segment keeps its bucket map at arena + map_off + i*map_bytes
// and its table at arena + table_off + i*table_bytes.
struct alignas(16) SegDev {
    Shape S;
    RoundKeys bk;
    uint64_t map_off, table_off;
    uint32_t map_bytes, table_bytes;
    uint32_t first, count;
    uint32_t job;  // device counter slot
    uint32_t pad[3];
};
static_assert(sizeof(SegDev) % 16 == 0, "SegDev must stay 16-byte aligned in arrays");

struct JobCounters {
    unsigned long long fails, probes, trials, pad;
};

// Pass counters (u32 words): [0] failures  [2..3] probe total (u64).
constexpr uint32_t PC_FAIL = 0, PC_PROBES = 2, PC_WORDS = 4;
constexpr uint32_t SEG = 64;

// ------------------------------------------------------------------ kernels

// build: one wave per trial slot (slot = segment*64 + i).
template <int K, bool CACHED, class M>
__global__ __launch_bounds__(1024) CK_BUILD_REGS void k_build_mix(const SegDev* __restrict__ segs, uint32_t slots,
                                                   const uint32_t* __restrict__ te0, uint32_t tab_shift,
                                                   uint32_t cache_shift, char* __restrict__ arena,
                                                   uint64_t* __restrict__ walks, uint64_t* __restrict__ rnds) {
    extern __shared__ uint32_t lds[];
    const LdsTab T = stage_tab_from(te0, lds, tab_shift);
    const uint32_t waves = blockDim.x >> 6;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint32_t slot = blockIdx.x * waves + wave;
    if (slot >= slots) return;  // wave-uniform; no block barrier follows
    const SegDev& G = segs[slot / SEG];
    const uint32_t i = slot % SEG;
    if (i >= G.count) return;
    const Shape S = G.S;
    TrialSeed ts;
    derive_uniform(T, G.bk, S, G.first + i, ts);  // batch key words read from the segment as used

    M* m = reinterpret_cast<M*>(arena + G.map_off + size_t(i) * G.map_bytes);
    if constexpr (CACHED) {
        uint16_t* cache = reinterpret_cast<uint16_t*>(lds + (256u << tab_shift)) + (size_t(wave) << cache_shift);
        build_cache(T, ts, S, cache, lane);
        for (uint32_t o = lane; o < S.pnum; o += 64) {
            uint32_t c[K];
            feistel_cached_k<K>(cache, S, element_of(S, ts.shift, ts.step, o), c);
            store_candidates<K>(m, S, o, c);
        }
    } else {
        for (uint32_t o = lane; o < S.pnum; o += 64) {
            const uint64_t el = element_of(S, ts.shift, ts.step, o);
            uint32_t c[K];
            for (int w = 0; w < K; ++w) c[w] = feistel_direct(T, ts.rk, S, el, K, w);
            store_candidates<K>(m, S, o, c);
        }
    }
    clear_table(arena + G.table_off + size_t(i) * G.table_bytes, G.table_bytes, lane);
    if (lane == 0) {
        walks[slot] = ts.walk;
        rnds[slot] = ts.rnd;
    }
}

// Table layouts of a job (its group keeps one per pass):
//   LAY_MAP  u32 bucket-map rows, u32 entries ordinal | choice << 30;
//   LAY_DESC u64 packed candidate stream, u64 entries candidates | choice;
//   LAY_RES  u64 packed candidate stream, u64 entries holding the occupant's
//            other k-1 candidates | choice (its current bucket is the slot):
//            one random read per probe where the descriptor is too wide;
//   LAY_RES32 the same entries in 32 bits where they fit: half the
//            descriptor's table, so a pass's tables take less memory and
//            address space.
constexpr uint32_t LAY_MAP = 0, LAY_DESC = 1, LAY_RES = 2, LAY_RES32 = 3;

// chain: one wave per segment, one lane per trial.
template <int K, bool BGH, uint32_t LAY>
__global__ __launch_bounds__(256) void k_chain_mix(const SegDev* __restrict__ segs, uint32_t nsegs,
                                                   char* __restrict__ arena, const uint64_t* __restrict__ walks,
                                                   const uint64_t* __restrict__ rnds, JobCounters* __restrict__ jc,
                                                   uint32_t* __restrict__ pc, uint32_t* __restrict__ fail_slots) {
    const uint32_t seg = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (seg >= nsegs) return;  // wave-uniform
    const uint32_t lane = threadIdx.x & 63;
    const SegDev& G = segs[seg];
    const Shape S = G.S;
    const uint32_t slot = seg * SEG + lane;
    bool failed = false;
    uint32_t pr = 0;
    if (lane < G.count) {
        if constexpr (LAY == LAY_DESC) {
            const uint64_t* m = reinterpret_cast<const uint64_t*>(arena + G.map_off + size_t(lane) * G.map_bytes);
            uint64_t* t = reinterpret_cast<uint64_t*>(arena + G.table_off + size_t(lane) * G.table_bytes);
            const DescCodec D{S.dbits, K * S.dbits, (1ull << S.dbits) - 1, (1ull << (K * S.dbits)) - 1};
            if constexpr (BGH)
                failed = chain_bgh_desc<K>(S, D, m, t, walks[slot], rnds[slot], pr);
            else
                failed = chain_drrt_desc<K>(S, D, m, t, pr);
        } else if constexpr (LAY == LAY_RES || LAY == LAY_RES32) {
            using E = std::conditional_t<LAY == LAY_RES32, uint32_t, uint64_t>;
            const uint64_t* m = reinterpret_cast<const uint64_t*>(arena + G.map_off + size_t(lane) * G.map_bytes);
            const auto t = res_table<E>(arena + G.table_off + size_t(lane) * G.table_bytes);
            const ResCodec<K, typename ResWord<E>::type> R{S.dbits, (K - 1) * S.dbits, (1ull << S.dbits) - 1};
            const CandStream<K, uint64_t> st{m, S.dbits, (1ull << S.dbits) - 1};
            uint32_t stash_left = S.stash;
            if constexpr (BGH) {
                uint64_t walk = walks[slot], rnd = rnds[slot];
                failed = chain_bgh_res<K>(S, R, st, t, ItemSpan{0, S.pnum}, walk, rnd, stash_left, pr);
            } else {
                failed = chain_drrt_res<K>(S, R, st, t, ItemSpan{0, S.pnum}, stash_left, pr);
            }
        } else {
            const uint32_t* m = reinterpret_cast<const uint32_t*>(arena + G.map_off + size_t(lane) * G.map_bytes);
            uint32_t* t = reinterpret_cast<uint32_t*>(arena + G.table_off + size_t(lane) * G.table_bytes);
            if constexpr (BGH)
                failed = chain_bgh<K>(S, m, t, walks[slot], rnds[slot], pr);
            else
                failed = chain_drrt<K>(S, m, t, pr);
        }
    }
    append_failures(failed, slot, pc + PC_FAIL, fail_slots);
    const uint64_t fm = __ballot(failed);
    add_wave_total(pr, reinterpret_cast<unsigned long long*>(pc + PC_PROBES));
    add_wave_total(pr, &jc[G.job].probes);
    if (lane == 0) {
        if (fm) atomicAdd(&jc[G.job].fails, static_cast<unsigned long long>(__popcll(fm)));
        atomicAdd(&jc[G.job].trials, static_cast<unsigned long long>(G.count));
    }
}

// ---------------------------------------------------------------- host side

using Clock = std::chrono::steady_clock;

constexpr uint32_t kRepl = 5;           // T0 copies = 2^5 (conflict-free lookups)
constexpr uint32_t kMaxCacheHalf = 12;  // larger round caches do not fit next to the T-table
constexpr uint32_t kBuildWaves = 16;
static size_t tab_bytes(uint32_t shift) { return size_t(256) << shift << 2; }
static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Integer-op weights of one trial's build (the cost model only orders passes;
// its scale is calibrated from measured pass times).
static double build_weight(const ck_params& p, bool cached) {
    const double aes = 336.0;
    const double evals = double(p.k) * p.pnum;
    return (cached ? double(8u << p.half_bits) * aes + evals * 48.0 : evals * 8.0 * 1.34 * aes);
}

struct Job {
    uint64_t id = 0;
    ck_job spec{};
    Shape S{};
    bool cached = false, bgh = true;
    uint32_t lay = LAY_MAP;
    uint32_t group = 0;
    size_t map_bytes = 0, table_bytes = 0;  // per slot
    uint64_t scheduled = 0, completed = 0;
    std::vector<uint64_t> fails;
    int32_t cslot = -1;
    double bw = 0;           // build weight per trial
    double ppt = 0;          // probes per trial (estimate, then measured)
    uint64_t last_batch = ~0ull;
    RoundKeys key{};
    bool admitted = false, done = false;
    ck_job_result res{};
};

struct SegMeta {
    Job* job;
    uint64_t trial0;
    uint32_t count;
};

struct PassBuf {
    char* arena = nullptr;
    SegDev* segs_d = nullptr;
    SegDev* segs_h = nullptr;  // pinned staging
    uint64_t* walks = nullptr;
    uint64_t* rnds = nullptr;
    uint32_t* fail_d = nullptr;
    uint32_t* pc_d = nullptr;
    int cs = 0;  // chain stream of the pass in flight
    hipEvent_t eb0 = nullptr, eb1 = nullptr, ec0 = nullptr, ec1 = nullptr;
    std::vector<SegMeta> meta;
    uint32_t nsegs = 0, slots = 0;
    double est_build = 0, est_chain = 0, weight = 0;
    bool keep = false;
};

// Spread arenas (ck_engine_config.spread): pieces of 1 GiB, at most 256 held
// while they are sorted into regions.
static constexpr size_t kArenaPiece = size_t(1) << 30;
static constexpr size_t kArenaMaxPieces = 256;

struct Engine {
    int device = 0;
    ck_engine_config cfg{};
    hipStream_t bstream = nullptr, cstream[2] = {nullptr, nullptr};
    uint64_t pass_seq = 0;
    std::vector<JobCounters> jc_h;  // host copy of the counter slots
    char* arena_base = nullptr;
    size_t arena_per = 0;
    // A spread arena (cfg.spread): kArenaPiece pieces mapped at arena_base,
    // from the memory regions in turn; arena_count[r] pieces of region r
    // (arena_count[3]: unclear).  Empty: one hipMalloc.
    std::vector<hipMemGenericAllocationHandle_t> arena_pieces;
    uint32_t arena_count[4] = {0, 0, 0, 0};
    uint32_t* te0_d = nullptr;
    JobCounters* jc_d = nullptr;
    size_t lds_block = 64 * 1024;
    std::vector<PassBuf> bufs;
    std::vector<int> free_bufs;
    std::deque<int> inflight;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> pending;
    std::vector<Job*> active;  // admitted, unfinished, in arrival order
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
    std::vector<int32_t> free_cslots;
    uint64_t next_id = 1;
    bool stop = false;
    int err = CK_OK;
    std::string err_msg;
    std::thread driver;
    Clock::time_point t0;
    ck_engine_stats st{};
    double ms_per_bw = 1.0e-10;    // build ms per weight unit (calibrated)
    double ms_per_probe = 4.7e-8;  // chain ms per probe (calibrated)

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

#define EH(expr)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return set_fail(E, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static int set_fail(Engine& E, const std::string& msg) {
    if (E.err == CK_OK) {
        E.err = CK_EDEVICE;
        E.err_msg = msg;
    }
    return CK_EDEVICE;
}

static uint32_t group_of(const Job& j) {
    return (j.spec.p.k << 8) | (j.bgh ? 0x80u : 0u) | (j.lay << 1) | (j.cached ? 1u : 0u);
}

// Slots of job j that fit in `slots_left` slots and `arena_left` bytes,
// in whole segments except for the job's last.
static uint64_t fit_trials(const Job& j, uint32_t slots_left, size_t arena_left) {
    const uint64_t remaining = j.spec.trials - j.scheduled;
    const size_t per = j.map_bytes + j.table_bytes;
    uint64_t segs = std::min<uint64_t>(slots_left / SEG, arena_left / (per * SEG + 512));
    return std::min<uint64_t>(remaining, segs * SEG);
}

static double est_chain_ms(const Engine& E, const Job& j, uint64_t trials) { return trials * j.ppt * E.ms_per_probe; }
static double est_build_ms(const Engine& E, const Job& j, uint64_t trials) { return trials * j.bw * E.ms_per_bw; }

// Chooses the job that leads the next pass.
static Job* pick_lead(Engine& E, const std::vector<Job*>& cand) {
    if (E.cfg.policy == CK_SCHED_FIFO) return cand.front();
    if (E.cfg.policy == CK_SCHED_SRPT) {
        Job* best = nullptr;
        double bc = 0;
        for (Job* j : cand) {
            const uint64_t rem = j->spec.trials - j->scheduled;
            const double c = est_build_ms(E, *j, rem) + est_chain_ms(E, *j, rem);
            if (!best || c < bc) {
                best = j;
                bc = c;
            }
        }
        return best;
    }
    // Balance: the chain stream has Q ms of committed work ahead of it.  The
    // next build runs beside that work, so prefer the most build-heavy pass
    // whose build still ends before the chains run dry; if none does, the pass
    // whose build is shortest (it feeds the chain stream soonest).
    double Q = 0;
    for (int b : E.inflight) Q += E.bufs[b].est_chain;
    Job* best = nullptr;
    double bb = -1;
    Job* quick = nullptr;
    double qb = 0;
    for (Job* j : cand) {
        const uint64_t n = fit_trials(*j, E.cfg.max_slots, E.arena_per);
        const double b = est_build_ms(E, *j, n), c = est_chain_ms(E, *j, n);
        const double heavy = b / std::max(c, 1e-9);
        if (b <= Q && (!best || heavy > bb)) {
            best = j;
            bb = heavy;
        }
        if (!quick || b < qb) {
            quick = j;
            qb = b;
        }
    }
    return best ? best : quick;
}

static int enqueue_pass(Engine& E, int bi);

// Plans the next pass into buffer bi; false when no admitted job has trials left.
static bool plan_pass(Engine& E, int bi, int& rc) {
    rc = CK_OK;
    std::vector<Job*> cand;
    for (Job* j : E.active) {
        if (j->scheduled >= j->spec.trials) continue;
        cand.push_back(j);
        if (E.cfg.window && cand.size() >= E.cfg.window) break;
    }
    if (cand.empty()) return false;
    Job* lead = pick_lead(E, cand);
    // Fill order: the lead, then jobs of the same group in the policy's order.
    std::vector<Job*> order{lead};
    std::vector<Job*> rest;
    for (Job* j : cand)
        if (j != lead && j->group == lead->group) rest.push_back(j);
    if (E.cfg.policy == CK_SCHED_SRPT)
        std::stable_sort(rest.begin(), rest.end(), [&](Job* a, Job* b) {
            return est_chain_ms(E, *a, a->spec.trials - a->scheduled) + est_build_ms(E, *a, a->spec.trials - a->scheduled) <
                   est_chain_ms(E, *b, b->spec.trials - b->scheduled) + est_build_ms(E, *b, b->spec.trials - b->scheduled);
        });
    order.insert(order.end(), rest.begin(), rest.end());

    PassBuf& B = E.bufs[bi];
    B.meta.clear();
    B.nsegs = 0;
    B.slots = 0;
    B.est_build = B.est_chain = B.weight = 0;
    B.keep = false;
    size_t arena_used = 0;
    uint32_t slots_left = E.cfg.max_slots;
    const double t = E.now_ms();
    for (Job* j : order) {
        const uint64_t n = fit_trials(*j, slots_left, E.arena_per - arena_used);
        if (n == 0) continue;
        if (j->res.start_ms < 0) j->res.start_ms = t;
        B.est_build += est_build_ms(E, *j, n);
        B.est_chain += est_chain_ms(E, *j, n);
        B.weight += j->bw * n;
        B.keep = B.keep || j->spec.keep_failures;
        const uint64_t mask = (j->spec.bt >= 64) ? ~0ull : ((uint64_t(1) << j->spec.bt) - 1);
        for (uint64_t done = 0; done < n; done += SEG) {
            const uint64_t t0 = j->scheduled + done;
            const uint32_t cnt = static_cast<uint32_t>(std::min<uint64_t>(SEG, n - done));
            const uint64_t batch = t0 >> j->spec.bt;
            if (batch != j->last_batch) {
                batch_key(j->spec.seed, batch, j->key.w);
                j->last_batch = batch;
            }
            SegDev& G = B.segs_h[B.nsegs];
            std::memset(&G, 0, sizeof G);
            G.S = j->S;
            G.bk = j->key;
            G.map_bytes = static_cast<uint32_t>(j->map_bytes);
            G.table_bytes = static_cast<uint32_t>(j->table_bytes);
            G.map_off = arena_used;
            arena_used = align256(arena_used + size_t(cnt) * j->map_bytes);
            G.table_off = arena_used;
            arena_used = align256(arena_used + size_t(cnt) * j->table_bytes);
            G.first = static_cast<uint32_t>(t0 & mask);
            G.count = cnt;
            G.job = static_cast<uint32_t>(j->cslot);
            B.meta.push_back(SegMeta{j, t0, cnt});
            ++B.nsegs;
        }
        j->scheduled += n;
        slots_left -= static_cast<uint32_t>((n + SEG - 1) / SEG * SEG);
        B.slots += static_cast<uint32_t>(n);
        if (slots_left < SEG) break;
    }
    if (B.nsegs == 0) return false;
    rc = enqueue_pass(E, bi);
    return rc == CK_OK;
}

template <int K>
static void launch_build(Engine& E, const PassBuf& B, bool cached, bool stream64, uint32_t max_half) {
    const uint32_t slots = B.nsegs * SEG;
    const size_t tab = tab_bytes(kRepl);
    if (cached) {
        const size_t cache = size_t(8) << max_half << 1;
        const uint32_t waves = static_cast<uint32_t>(std::min<size_t>(kBuildWaves, (E.lds_block - tab) / cache));
        const size_t lds = tab + waves * cache;
        const dim3 grid((slots + waves - 1) / waves), block(64 * waves);
        const uint32_t cs = max_half + 3;
        if (stream64)
            hipLaunchKernelGGL((k_build_mix<K, true, uint64_t>), grid, block, lds, E.bstream, B.segs_d, slots, E.te0_d,
                               kRepl, cs, B.arena, B.walks, B.rnds);
        else
            hipLaunchKernelGGL((k_build_mix<K, true, uint32_t>), grid, block, lds, E.bstream, B.segs_d, slots, E.te0_d,
                               kRepl, cs, B.arena, B.walks, B.rnds);
    } else {
        const uint32_t waves = 4;
        const dim3 grid((slots + waves - 1) / waves), block(64 * waves);
        if (stream64)
            hipLaunchKernelGGL((k_build_mix<K, false, uint64_t>), grid, block, tab, E.bstream, B.segs_d, slots, E.te0_d,
                               kRepl, 0u, B.arena, B.walks, B.rnds);
        else
            hipLaunchKernelGGL((k_build_mix<K, false, uint32_t>), grid, block, tab, E.bstream, B.segs_d, slots, E.te0_d,
                               kRepl, 0u, B.arena, B.walks, B.rnds);
    }
}

template <int K>
static void launch_chain(Engine& E, const PassBuf& B, bool bgh, uint32_t lay) {
    const dim3 grid((B.nsegs + 3) / 4), block(256);
#define CH(BG, L)                                                                                                \
    hipLaunchKernelGGL((k_chain_mix<K, BG, L>), grid, block, 0, E.cstream[B.cs], B.segs_d, B.nsegs, B.arena, B.walks, \
                       B.rnds, E.jc_d, B.pc_d, B.fail_d)
    if (bgh && lay == LAY_DESC)
        CH(true, LAY_DESC);
    else if (bgh && lay == LAY_RES)
        CH(true, LAY_RES);
    else if (bgh && lay == LAY_RES32)
        CH(true, LAY_RES32);
    else if (bgh)
        CH(true, LAY_MAP);
    else if (lay == LAY_DESC)
        CH(false, LAY_DESC);
    else if (lay == LAY_RES)
        CH(false, LAY_RES);
    else if (lay == LAY_RES32)
        CH(false, LAY_RES32);
    else
        CH(false, LAY_MAP);
#undef CH
}

static int enqueue_pass(Engine& E, int bi) {
    PassBuf& B = E.bufs[bi];
    const Job& lead = *B.meta.front().job;
    uint32_t max_half = 0;
    for (uint32_t s = 0; s < B.nsegs; ++s) max_half = std::max(max_half, B.segs_h[s].S.half);
    EH(hipMemcpyAsync(B.segs_d, B.segs_h, size_t(B.nsegs) * sizeof(SegDev), hipMemcpyHostToDevice, E.bstream));
    EH(hipEventRecord(B.eb0, E.bstream));
    switch (lead.spec.p.k) {
        case 2: launch_build<2>(E, B, lead.cached, lead.lay != LAY_MAP, max_half); break;
        case 3: launch_build<3>(E, B, lead.cached, lead.lay != LAY_MAP, max_half); break;
        default: launch_build<4>(E, B, lead.cached, lead.lay != LAY_MAP, max_half); break;
    }
    EH(hipGetLastError());
    EH(hipEventRecord(B.eb1, E.bstream));
    B.cs = static_cast<int>(E.pass_seq++ & 1);
    hipStream_t cst = E.cstream[B.cs];
    EH(hipStreamWaitEvent(cst, B.eb1, 0));
    EH(hipMemsetAsync(B.pc_d, 0, PC_WORDS * 4, cst));
    EH(hipEventRecord(B.ec0, cst));
    switch (lead.spec.p.k) {
        case 2: launch_chain<2>(E, B, lead.bgh, lead.lay); break;
        case 3: launch_chain<3>(E, B, lead.bgh, lead.lay); break;
        default: launch_chain<4>(E, B, lead.bgh, lead.lay); break;
    }
    EH(hipGetLastError());
    EH(hipEventRecord(B.ec1, cst));
    E.inflight.push_back(bi);
    return CK_OK;
}

static void finish_job(Engine& E, Job& j, const JobCounters& c) {
    j.res.trials = c.trials;
    j.res.failed = c.fails;
    j.res.probes = c.probes;
    j.res.end_ms = E.now_ms();
    if (c.trials != j.spec.trials) set_fail(E, "job counter holds a different trial count than was run");
    if (j.spec.keep_failures) {
        std::sort(j.fails.begin(), j.fails.end());
        if (j.fails.size() != c.fails) set_fail(E, "failure list and device failure counter disagree");
    }
    j.done = true;
    E.free_cslots.push_back(j.cslot);
    j.cslot = -1;
    E.st.jobs_done += 1;
    E.st.trials += c.trials;
    E.st.failed += c.fails;
    E.st.probes += c.probes;
}

static int gather_pass(Engine& E, int bi) {
    PassBuf& B = E.bufs[bi];
    EH(hipEventSynchronize(B.ec1));
    uint32_t pc[PC_WORDS];
    EH(hipMemcpy(pc, B.pc_d, sizeof pc, hipMemcpyDeviceToHost));
    const uint32_t nf = pc[PC_FAIL];
    if (nf > B.slots) return set_fail(E, "device failure count out of range");
    if (nf && B.keep) {
        std::vector<uint32_t> f(nf);
        EH(hipMemcpy(f.data(), B.fail_d, size_t(nf) * 4, hipMemcpyDeviceToHost));
        for (uint32_t s : f) {
            const uint32_t seg = s / SEG, lane = s % SEG;
            if (seg >= B.nsegs || lane >= B.meta[seg].count) return set_fail(E, "device failure slot out of range");
            Job* j = B.meta[seg].job;
            if (j->spec.keep_failures) j->fails.push_back(B.meta[seg].trial0 + lane);
        }
    }
    float bms = 0, cms = 0;
    EH(hipEventElapsedTime(&bms, B.eb0, B.eb1));
    EH(hipEventElapsedTime(&cms, B.ec0, B.ec1));
    const uint64_t probes = (uint64_t(pc[PC_PROBES + 1]) << 32) | pc[PC_PROBES];
    E.st.passes += 1;
    E.st.segments += B.nsegs;
    E.st.slots += B.slots;
    E.st.build_ms += bms;
    E.st.chain_ms += cms;
    if (B.weight > 0 && bms > 0) E.ms_per_bw = 0.7 * E.ms_per_bw + 0.3 * (bms / B.weight);
    if (probes > 0 && cms > 0) E.ms_per_probe = 0.7 * E.ms_per_probe + 0.3 * (cms / double(probes));
    for (const SegMeta& m : B.meta) m.job->completed += m.count;
    // Passes finish in order here (each earlier one was gathered first), so the
    // counters of a job whose last pass this is are final; the others are a
    // running estimate for the scheduler.
    int32_t lo = INT32_MAX, hi = -1;
    for (const SegMeta& m : B.meta) {
        lo = std::min(lo, m.job->cslot);
        hi = std::max(hi, m.job->cslot);
    }
    EH(hipMemcpy(E.jc_h.data() + lo, E.jc_d + lo, size_t(hi - lo + 1) * sizeof(JobCounters), hipMemcpyDeviceToHost));
    Job* prev = nullptr;
    for (const SegMeta& m : B.meta) {
        Job* j = m.job;
        if (j == prev) continue;
        prev = j;
        const JobCounters& c = E.jc_h[j->cslot];
        if (c.trials) j->ppt = double(c.probes) / double(c.trials);
        if (!j->done && j->completed == j->spec.trials) finish_job(E, *j, c);
    }
    E.active.erase(std::remove_if(E.active.begin(), E.active.end(), [](Job* j) { return j->done; }), E.active.end());
    E.cv_done.notify_all();
    return E.err;
}

static int admit(Engine& E) {
    while (!E.pending.empty() && !E.free_cslots.empty()) {
        Job* j = E.pending.front();
        E.pending.pop_front();
        j->cslot = E.free_cslots.back();
        E.free_cslots.pop_back();
        // On the build stream: every chain of the job waits for a build after it.
        EH(hipMemsetAsync(E.jc_d + j->cslot, 0, sizeof(JobCounters), E.bstream));
        j->admitted = true;
        E.active.push_back(j);
    }
    return CK_OK;
}

static void driver_main(Engine* ep) {
    Engine& E = *ep;
    (void)hipSetDevice(E.device);
    std::unique_lock<std::mutex> lk(E.mu);
    for (;;) {
        if (E.err) break;
        if (!E.stop) {
            if (admit(E)) break;
            int rc = CK_OK;
            while (!E.free_bufs.empty()) {
                const int bi = E.free_bufs.back();
                if (!plan_pass(E, bi, rc)) break;
                E.free_bufs.pop_back();
            }
            if (rc) break;
        }
        if (E.inflight.empty()) {
            if (E.stop) break;
            E.cv_work.wait(lk, [&] {
                if (E.stop || !E.pending.empty()) return true;
                for (Job* j : E.active)
                    if (j->scheduled < j->spec.trials) return true;
                return false;
            });
            continue;
        }
        const int bi = E.inflight.front();
        const hipError_t q = hipEventQuery(E.bufs[bi].ec1);
        if (q == hipSuccess) {
            E.inflight.pop_front();
            if (gather_pass(E, bi)) break;
            E.free_bufs.push_back(bi);
            continue;
        }
        if (q != hipErrorNotReady) {
            set_fail(E, std::string("pass event: ") + hipGetErrorString(q));
            break;
        }
        E.cv_work.wait_for(lk, std::chrono::microseconds(200));
    }
    // Drain what is still on the device so buffers can be freed safely.
    for (int bi : E.inflight) (void)hipEventSynchronize(E.bufs[bi].ec1);
    E.inflight.clear();
    if (E.err == CK_OK && E.stop) {
        E.err = CK_EINVAL;
        E.err_msg = "engine stopped";
    }
    E.cv_done.notify_all();
}

// A spread arena: whole kArenaPiece pieces (the arena rounded down to them),
// sorted into the device's memory regions by pair walks and mapped in the
// regions' turn, so that every pass buffer's maps and tables are spread over
// the regions; the pass buffers keep their offsets.  Leaves arena_base null
// (one hipMalloc follows) where fewer than two regions show.
static int spread_arena(Engine& E) {
    const size_t n = E.arena_per * E.cfg.buffers / kArenaPiece;
    size_t free_b = 0, total_b = 0;
    EH(hipMemGetInfo(&free_b, &total_b));
    const size_t max_pieces = std::min<size_t>(kArenaMaxPieces, free_b / 20 * 19 / kArenaPiece);
    if (max_pieces < n) return CK_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    unsigned long long* sink = nullptr;
    EH(hipMalloc(&sink, 256));
    EH(hipEventCreate(&e0));
    EH(hipEventCreate(&e1));
    RegionPool rp;
    std::string err;
    int rc = region_pool(E.device, E.cstream[0], e0, e1, sink, kArenaPiece, max_pieces, (n + 2) / 3, &rp, &err);
    if (rc == CK_OK && rp.two) {
        const std::vector<size_t> order = interleave(rp, n);
        void* va = nullptr;
        if (order.size() == n && (rc = map_pieces(E.device, rp, order, &va, &err)) == CK_OK) {
            E.arena_base = static_cast<char*>(va);
            for (size_t j : order) {
                E.arena_pieces.push_back(rp.h[j]);
                ++E.arena_count[rp.region[j] < 0 ? 3 : rp.region[j]];
            }
            E.arena_per = n * kArenaPiece / E.cfg.buffers & ~size_t(255);
        }
    }
    drop_pool(rp, E.arena_pieces);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    if (rc) return set_fail(E, "spread arena: " + err);
    return CK_OK;
}

static int engine_init(Engine& E) {
    EH(hipSetDevice(E.device));
    EH(hipStreamCreateWithFlags(&E.bstream, hipStreamNonBlocking));
    EH(hipStreamCreateWithFlags(&E.cstream[0], hipStreamNonBlocking));
    EH(hipStreamCreateWithFlags(&E.cstream[1], hipStreamNonBlocking));
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, E.device) == hipSuccess && v > 0)
        E.lds_block = size_t(v);
    const int lim = static_cast<int>(E.lds_block);
#define SETL(K, C, M) \
    EH(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_build_mix<K, C, M>), hipFuncAttributeMaxDynamicSharedMemorySize, lim))
    SETL(2, true, uint64_t); SETL(2, true, uint32_t); SETL(2, false, uint64_t); SETL(2, false, uint32_t);
    SETL(3, true, uint64_t); SETL(3, true, uint32_t); SETL(3, false, uint64_t); SETL(3, false, uint32_t);
    SETL(4, true, uint64_t); SETL(4, true, uint32_t); SETL(4, false, uint64_t); SETL(4, false, uint32_t);
#undef SETL
    EH(hipMalloc(&E.te0_d, 256 * sizeof(uint32_t)));
    EH(hipMemcpy(E.te0_d, te0(), 256 * sizeof(uint32_t), hipMemcpyHostToDevice));
    EH(hipMalloc(&E.jc_d, size_t(E.cfg.max_jobs) * sizeof(JobCounters)));
    EH(hipMemset(E.jc_d, 0, size_t(E.cfg.max_jobs) * sizeof(JobCounters)));
    E.jc_h.assign(E.cfg.max_jobs, JobCounters{});
    if (!E.cfg.arena_bytes) {
        // Default: up to 192 GiB, leaving a sixth of the free memory to others.
        size_t free_b = 0, total_b = 0;
        EH(hipMemGetInfo(&free_b, &total_b));
        E.cfg.arena_bytes = std::min<uint64_t>(uint64_t(192) << 30, free_b / 6 * 5);
    }
    E.arena_per = (E.cfg.arena_bytes / E.cfg.buffers) & ~size_t(255);
    if (E.cfg.spread && E.arena_per * E.cfg.buffers >= 3 * kArenaPiece) {
        if (int rc = spread_arena(E)) return rc;
    }
    if (!E.arena_base) EH(hipMalloc(&E.arena_base, E.arena_per * E.cfg.buffers));
    const uint32_t maxseg = E.cfg.max_slots / SEG;
    E.bufs.resize(E.cfg.buffers);
    for (uint32_t b = 0; b < E.cfg.buffers; ++b) {
        PassBuf& B = E.bufs[b];
        B.arena = E.arena_base + size_t(b) * E.arena_per;
        EH(hipMalloc(&B.segs_d, size_t(maxseg) * sizeof(SegDev)));
        EH(hipHostMalloc(&B.segs_h, size_t(maxseg) * sizeof(SegDev), hipHostMallocDefault));
        EH(hipMalloc(&B.walks, size_t(E.cfg.max_slots) * 8));
        EH(hipMalloc(&B.rnds, size_t(E.cfg.max_slots) * 8));
        EH(hipMalloc(&B.fail_d, size_t(E.cfg.max_slots) * 4));
        EH(hipMalloc(&B.pc_d, PC_WORDS * 4));
        EH(hipEventCreate(&B.eb0));
        EH(hipEventCreate(&B.eb1));
        EH(hipEventCreate(&B.ec0));
        EH(hipEventCreate(&B.ec1));
        E.free_bufs.push_back(static_cast<int>(b));
    }
    for (int32_t s = static_cast<int32_t>(E.cfg.max_jobs) - 1; s >= 0; --s) E.free_cslots.push_back(s);
    return CK_OK;
}

static void engine_free(Engine& E) {
    (void)hipSetDevice(E.device);
    for (PassBuf& B : E.bufs) {
        if (B.segs_d) (void)hipFree(B.segs_d);
        if (B.segs_h) (void)hipHostFree(B.segs_h);
        if (B.walks) (void)hipFree(B.walks);
        if (B.rnds) (void)hipFree(B.rnds);
        if (B.fail_d) (void)hipFree(B.fail_d);
        if (B.pc_d) (void)hipFree(B.pc_d);
        for (hipEvent_t ev : {B.eb0, B.eb1, B.ec0, B.ec1})
            if (ev) (void)hipEventDestroy(ev);
    }
    if (E.arena_base && !E.arena_pieces.empty()) {
        (void)hipDeviceSynchronize();
        unmap_pieces(E.arena_base, kArenaPiece, E.arena_pieces.size());
        for (hipMemGenericAllocationHandle_t h : E.arena_pieces) (void)hipMemRelease(h);
    } else if (E.arena_base) {
        (void)hipFree(E.arena_base);
    }
    if (E.jc_d) (void)hipFree(E.jc_d);
    if (E.te0_d) (void)hipFree(E.te0_d);
    if (E.bstream) (void)hipStreamDestroy(E.bstream);
    for (hipStream_t cs : E.cstream)
        if (cs) (void)hipStreamDestroy(cs);
}

static int engine_error(Engine& E) { return fail(E.err, E.err_msg); }

}  // namespace eng
}  // namespace ck

using namespace ck;
using namespace ck::eng;

struct ck_engine {
    Engine e;
};

extern "C" {

int ck_engine_create(int device, const ck_engine_config* cfg, ck_engine** out) {
    if (!out) return fail(CK_EINVAL, "null output");
    *out = nullptr;
    ck_engine_config c{};
    if (cfg) c = *cfg;
    if (!c.buffers) c.buffers = 4;
    if (!c.max_slots) c.max_slots = 65536;
    if (!c.max_jobs) c.max_jobs = 4096;
    if (c.buffers < 2 || c.buffers > 8) return fail(CK_EINVAL, "buffers must be 2..8");
    if (c.max_slots % SEG || c.max_slots > (1u << 22)) return fail(CK_EINVAL, "max_slots must be a multiple of 64, <= 2^22");
    if (c.policy > CK_SCHED_SRPT) return fail(CK_EINVAL, "bad scheduling policy");
    if (c.max_jobs > (1u << 20)) return fail(CK_EINVAL, "max_jobs too large");
    if (c.spread > 1) return fail(CK_EINVAL, "spread must be 0 or 1");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return fail(CK_EINVAL, "no such HIP device");
    auto* h = new ck_engine();
    Engine& E = h->e;
    E.device = device;
    E.cfg = c;
    E.t0 = Clock::now();
    const int rc = engine_init(E);
    if (rc) {
        const std::string msg = E.err_msg;
        engine_free(E);
        delete h;
        return fail(rc, msg);
    }
    E.driver = std::thread(driver_main, &E);
    *out = h;
    return CK_OK;
}

void ck_engine_destroy(ck_engine* h) {
    if (!h) return;
    Engine& E = h->e;
    {
        std::lock_guard<std::mutex> lk(E.mu);
        E.stop = true;
    }
    E.cv_work.notify_all();
    if (E.driver.joinable()) E.driver.join();
    engine_free(E);
    delete h;
}

int ck_engine_submit(ck_engine* h, const ck_job* job, uint64_t* job_id) {
    if (!h || !job || !job_id) return fail(CK_EINVAL, "null argument");
    int rc = check_params(&job->p);
    if (rc) return rc;
    if (job->bt < 6 || job->bt > 32) return fail(CK_EINVAL, "bt must be 6..32 (segments of 64 trials stay in one batch)");
    if (job->trials == 0) return fail(CK_EINVAL, "a job needs at least one trial");
    Engine& E = h->e;
    auto j = std::make_unique<Job>();
    j->spec = *job;
    j->S = make_shape(job->p);
    j->cached = job->p.half_bits <= kMaxCacheHalf &&
                tab_bytes(kRepl) + (size_t(8) << job->p.half_bits) * 2 <= E.lds_block;
    // Residual entries in 32 bits where they fit (half the descriptor's table;
    // faster at n = 2^14), then the descriptor (faster than 48-bit residual
    // entries: 361K vs 329K trials/s at n = 30000, k = 3), then residual
    // entries in 64 bits, then the bucket map.  A residual entry of `need` bits
    // must stay below all-ones: fewer bits than its width, or exactly its width
    // with a choice below 3.
    const uint32_t kb = job->p.k * j->S.dbits, need = kb - j->S.dbits + 2;
    const auto fits = [&](uint32_t width) { return need < width || (need == width && job->p.k <= 3); };
    if (kb <= 64 && fits(32))
        j->lay = LAY_RES32;
    else if (kb + 2 <= 63)
        j->lay = LAY_DESC;
    else if (kb <= 64 && fits(64))
        j->lay = LAY_RES;
    else
        j->lay = LAY_MAP;
    j->bgh = job->p.mode == CK_MODE_BGH25;
    j->group = group_of(*j);
    const size_t entry = j->lay == LAY_RES32 || j->lay == LAY_MAP ? 4 : 8;
    j->map_bytes = align256(size_t(job->p.pnum) * (j->lay != LAY_MAP ? 8 : size_t(job->p.k) * 4));
    j->table_bytes = align256(size_t(job->p.buckets) * entry);
    if ((j->map_bytes + j->table_bytes) * SEG + 1024 > E.arena_per)
        return fail(CK_ENOMEM, "64 trials of this shape do not fit one pass buffer");
    j->bw = build_weight(job->p, j->cached);
    j->ppt = double(job->p.pnum) * (j->bgh ? 2.5 : 1.8);
    j->res.start_ms = -1;
    std::lock_guard<std::mutex> lk(E.mu);
    if (E.err) return engine_error(E);
    j->id = E.next_id++;
    j->res.submit_ms = E.now_ms();
    *job_id = j->id;
    E.pending.push_back(j.get());
    E.jobs.emplace(j->id, std::move(j));
    E.st.jobs_submitted += 1;
    E.cv_work.notify_all();
    return CK_OK;
}

static int find_job(Engine& E, uint64_t id, Job*& j) {
    auto it = E.jobs.find(id);
    if (it == E.jobs.end()) return fail(CK_EINVAL, "unknown job id");
    j = it->second.get();
    return CK_OK;
}

int ck_engine_poll(ck_engine* h, uint64_t job_id, ck_job_result* out) {
    if (!h) return fail(CK_EINVAL, "null engine");
    Engine& E = h->e;
    std::lock_guard<std::mutex> lk(E.mu);
    Job* j;
    if (int rc = find_job(E, job_id, j)) return rc;
    if (j->done) {
        if (out) *out = j->res;
        return CK_OK;
    }
    if (E.err) return engine_error(E);
    return CK_PENDING;
}

int ck_engine_wait(ck_engine* h, uint64_t job_id, ck_job_result* out) {
    if (!h) return fail(CK_EINVAL, "null engine");
    Engine& E = h->e;
    std::unique_lock<std::mutex> lk(E.mu);
    Job* j;
    if (int rc = find_job(E, job_id, j)) return rc;
    E.cv_done.wait(lk, [&] { return j->done || E.err != CK_OK; });
    if (!j->done) return engine_error(E);
    if (out) *out = j->res;
    return CK_OK;
}

int ck_engine_failures(ck_engine* h, uint64_t job_id, uint64_t* out, uint64_t cap, uint64_t* n) {
    if (!h || !n || (cap && !out)) return fail(CK_EINVAL, "null argument");
    Engine& E = h->e;
    std::lock_guard<std::mutex> lk(E.mu);
    Job* j;
    if (int rc = find_job(E, job_id, j)) return rc;
    if (!j->done) return fail(CK_EINVAL, "job has not finished");
    if (!j->spec.keep_failures) return fail(CK_EINVAL, "job was submitted without keep_failures");
    *n = j->fails.size();
    std::copy_n(j->fails.begin(), std::min<uint64_t>(cap, j->fails.size()), out);
    return CK_OK;
}

int ck_engine_release(ck_engine* h, uint64_t job_id) {
    if (!h) return fail(CK_EINVAL, "null engine");
    Engine& E = h->e;
    std::lock_guard<std::mutex> lk(E.mu);
    Job* j;
    if (int rc = find_job(E, job_id, j)) return rc;
    if (!j->done) return fail(CK_EINVAL, "job has not finished");
    E.jobs.erase(job_id);
    return CK_OK;
}

int ck_engine_stats_get(ck_engine* h, ck_engine_stats* out) {
    if (!h || !out) return fail(CK_EINVAL, "null argument");
    Engine& E = h->e;
    std::lock_guard<std::mutex> lk(E.mu);
    *out = E.st;
    out->wall_ms = E.now_ms();
    return CK_OK;
}

int ck_engine_arena(ck_engine* h, uint32_t counts[4]) {
    if (!h || !counts) return fail(CK_EINVAL, "null argument");
    for (int i = 0; i < 4; ++i) counts[i] = h->e.arena_count[i];
    return CK_OK;
}

}  // extern "C"
