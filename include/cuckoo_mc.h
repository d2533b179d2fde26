/*
 * cuckoo_mc.h -- C ABI of the MI355X cuckoo-hashing Monte Carlo library
 * (pbc-based-dmpf-monte-carlo-test_amd/lib/libcuckoo_mc.so).
 *
 * The library replaces the reference's device half: the launch callable that
 * process_batch() drives (cuckoo_bench.cu:724-745, :979-1034), the AES/PRP
 * kernels (cuckoo_bench.cu:102-317, gpu_cached.cuh, gpu_shared.cuh) and their
 * known-answer self tests (cuckoo_bench.cu:322-354).  Plain pointers and sizes
 * only; no exceptions cross the ABI: every call returns CK_OK or a negative
 * code, and ck_last_error() describes the last failure on the calling thread.
 *
 * Threading: one ck_ctx per host thread (the reference binds one worker
 * thread per GPU, cuckoo_bench.cu:1102-1118).  Calls on one ctx are
 * synchronous unless named *_async.
 */
#ifndef CUCKOO_MC_H
#define CUCKOO_MC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CK_OK = 0,
    CK_EINVAL = -1,   /* invalid argument or shape */
    CK_EDEVICE = -2,  /* HIP runtime error */
    CK_ESELFTEST = -3,/* AES known-answer test failed */
    CK_ERANGE = -4,   /* device counter out of range (cuckoo_bench.cu:1027) */
    CK_ENOMEM = -5
};

enum { CK_MODE_BGH25 = 1, CK_MODE_DRRT18 = 2 }; /* journal encoding, cuckoo_bench.cu:607 */
#define CK_DERIVED 0xffffffffu

/* Experiment shape.  Mirrors Options+Params of the reference
 * (cuckoo_bench.cu:406-419, :499-526) plus the extension: k candidate buckets
 * (reference: 3), explicit buckets/kick (reference: derived from mode), stash
 * slots (reference: 0). */
typedef struct ck_params {
    uint32_t mode, k, pnum, buckets, kick, stash, ell_i, input_mode;
    uint32_t half_bits, ell_d;
    uint64_t bucket_size, domain, mask, range_mask;
} ck_params;

/* Table placement tier for the insertion chains. */
enum {
    CK_TIER_AUTO = 0, /* HBM (the fastest tier on every measured shape; ck_ctx_tune measures all) */
    CK_TIER_HBM = 1,  /* one lane per trial, table and bucket map in HBM */
    CK_TIER_LDS = 2,  /* trials per workgroup, table and bucket map in LDS */
    CK_TIER_LANE = 3  /* one lane per trial, its residual-entry table (<= 32-bit entries) in LDS, its new
                         items' candidates streamed from HBM: the reference's per-thread shared-memory
                         bins (cuckoo_bench.cu:217-220, 876-880) */
};

/* HBM-tier table layout and probe form. */
enum {
    CK_LAYOUT_AUTO = 0,       /* tables not requested: RESIDUAL when its entries fit 32 bits
                                 ((k-1)*bits+2 <= 32) and k*bits <= 64; else DESC_PLAIN when it fits;
                                 else RESIDUAL when it fits ((k-1)*bits+2 <= 63); otherwise MAP_XCHG
                                 below 2^14 trials per pass and MAP above */
    CK_LAYOUT_MAP = 1,        /* entry = ordinal | choice (u16 when it fits); load + store probes */
    CK_LAYOUT_DESC = 2,       /* entry = occupant's packed candidates | choice (k*bits+2 <= 63); exchange probes */
    CK_LAYOUT_MAP_XCHG = 3,   /* u32 entry = ordinal | choice; exchange probes */
    CK_LAYOUT_DESC_PLAIN = 4, /* descriptor entries; load + store probes */
    CK_LAYOUT_RESIDUAL = 5    /* entry = occupant's other k-1 candidates | choice (u32 when it fits, else
                                 44 bits packed end to end, else 48 bits in 6 bytes, else u64), its current
                                 bucket implied by the slot; load + store probes */
};

typedef struct ck_stats {
    uint64_t trials;      /* trials completed by the call */
    uint64_t failed;      /* failing trials */
    uint64_t probes;      /* table-slot reads of the insertion chains */
    uint32_t tier;        /* tier actually used */
    uint32_t chunk;       /* trials per device pass */
    double build_ms;      /* device time of the key/round-cache/map kernels */
    double chain_ms;      /* device time of the insertion kernels */
    double fused_ms;      /* device time of fused (LDS tier) kernels */
    uint32_t launches;    /* kernel launches */
    uint32_t graph;       /* 0: launched kernel by kernel; 1: replayed this context's graph; 2: built and ran a new one */
    double span_ms;       /* device time from the start of the call's first kernel to the end of its last
                             (overlapped builds and chains counted once) */
} ck_stats;

typedef struct ck_ctx ck_ctx;

/* parameters() (cuckoo_bench.cu:500-526) plus the extension.  Pass
 * CK_DERIVED as buckets/kick for the reference's mode-derived values. */
int ck_params_derive(uint32_t mode, uint32_t pnum, uint32_t ell_i, uint32_t k, uint32_t buckets,
                     uint32_t kick, uint32_t stash, uint32_t input_mode, ck_params* out);

/* batch_key(seed, batch_id) (cuckoo_bench.cu:393-404): 44 round-key words. */
void ck_batch_key(uint64_t seed, uint64_t batch_id, uint32_t round_keys[44]);

/* Host AES-128 on four big-endian words (host_aes128, cuckoo_bench.cu:360-392). */
void ck_host_aes(const uint32_t round_keys[44], const uint32_t in[4], uint32_t out[4]);

/* Inputs of one trial for the device-level trial API (include/cuckoo_trial.hpp),
 * derived on the host exactly as the library's build kernels derive them
 * (cuckoo_bench.cu:222-238): PRP key = schedule of AES_batch("KEY1", 0,
 * offset, "PRP1"); (a, b, c, d) = AES_batch("WALK", 0, offset, "RNG1"); the
 * walk starts at a<<32|b and randomness is its first SplitMix64 draw (the walk
 * advanced once); shift = (c<<32|d) mod 2^ell_i and step = (SplitMix64 of
 * shift ^ 0xd1b54a32d192ed03, | 1) mod 2^ell_i for affine input. */
typedef struct ck_trial_inputs {
    uint32_t prp_key[44];
    uint64_t walk, randomness, shift, step;
} ck_trial_inputs;
int ck_trial_inputs_of(const ck_params* p, const uint32_t batch_round_keys[44], uint32_t offset,
                       ck_trial_inputs* out);

/* The trial's round cache (gpu_shared.cuh:68-77, cpu_trials.hpp:139-142):
 * rounds[r << half_bits | right] = word 3 of AES_prp("CESP", r, half_bits,
 * right), masked to half_bits bits, for r < 8 and right < 2^half_bits.
 * rounds holds 8 << half_bits entries; CK_EINVAL when half_bits > 16. */
int ck_round_cache(const ck_params* p, const uint32_t prp_key[44], uint16_t* rounds);

/* One context per host thread and device (gpu_worker, cuckoo_bench.cu:813-973).
 * Uploads the AES tables to the device (upload_aes_tables, :64-75). */
int ck_ctx_create(int device, ck_ctx** out);
void ck_ctx_destroy(ck_ctx* ctx);

/* Forces a tier (CK_TIER_*) for later launches; CK_TIER_AUTO picks by shape. */
int ck_ctx_set_tier(ck_ctx* ctx, uint32_t tier);
/* Caps the trials per device pass (0 = library default). */
int ck_ctx_set_chunk(ck_ctx* ctx, uint32_t trials);
/* Residual-layout item windows: each HBM-tier pass inserts the items in
 * windows of `items` (0 = automatic: windows of pnum/64 when memory caps a pass
 * below its target, else none).  A window's build writes the candidates of its
 * items only, so a slot needs less memory and more trials fit in a pass.
 * Results do not depend on it.  Not applied in graph mode or to launches that
 * return the bucket map, walks or random words.  No reference counterpart (the
 * reference keeps no per-trial candidate stream). */
int ck_ctx_set_window(ck_ctx* ctx, uint32_t items);
/* Table placement for double-buffered HBM-tier passes of 16,384 or more
 * trials whose pass buffers are 256 MiB or larger (in graph mode too; not for
 * ck_launch_static and launches that return their tables); 0, the default,
 * turns it off.  The cuckoo tables of the two pass buffers are then held apart
 * from them.  MI355X memory falls into three regions of about 96 GB, and
 * random accesses spread over the regions run faster than accesses confined
 * to one (a random read-write walk: 20.5 G steps/s in one region, 27.3 G
 * spread over three), while an ordinary allocation of a few GB lies in one
 * region.  So the library allocates pieces of a third of a pass's tables (1
 * GiB at least; virtual-memory API), as many as its memory budget holds,
 * sorts them into regions by timing random walks over pairs of pieces, and
 * gives each pass buffer one piece from each region, every third trial slot's
 * table in each.  Where fewer than two regions show or the budget holds too
 * few pieces, it instead allocates up to `candidates` whole table
 * allocations, times the chains' access pattern over each and keeps the two
 * fastest.  Placement runs again when a plan needs tables of another size, so
 * it suits contexts that run many passes of one shape.  Results are
 * unaffected.  No reference counterpart (the reference allocates per worker
 * and never places). */
int ck_ctx_set_place_search(ck_ctx* ctx, uint32_t candidates);
/* The last placement: the walk's ms per candidate in the order tried (region
 * placement: candidate table ranges of one piece per region; search: whole
 * candidate allocations; up to cap of them, *n is their number) and the
 * candidates kept for pass buffers 0 and 1 (-1 when the context holds no
 * placed tables). */
int ck_ctx_place_report(ck_ctx* ctx, float* walk_ms, uint32_t cap, uint32_t* n, int32_t kept[2]);
/* The last region placement: each piece's region (0, 1, 2 in order of
 * discovery; up to cap of them, *n = their number, 0 when the tables were not
 * placed by region) and the pieces mapped as the tables of pass buffer 0
 * (used[0..2], in address order) and pass buffer 1 (used[3..5]). */
int ck_ctx_place_regions(ck_ctx* ctx, int32_t* region, uint32_t cap, uint32_t* n, int32_t used[6]);

/* Forces an HBM-tier table layout (CK_LAYOUT_*). */
int ck_ctx_set_layout(ck_ctx* ctx, uint32_t layout);
/* Graph mode (on = 1): HBM-tier ck_launch calls run as one hipGraph of all
 * their passes.  A graph is built on the first call for a shape, trial count
 * and placement, and replayed by later calls with only the batch key and
 * offsets rewritten -- the reference's per-micro-range launch of a fixed shape
 * (cuckoo_bench.cu:980-1031) with the host launch work of every pass taken out.
 * A context keeps its four most recently used graphs.
 * Results are those of the kernel-by-kernel form.  Launches that ask for
 * per-trial outputs (ck_launch_debug, ck_launch_digest) stay kernel by kernel.
 * With graph mode, stats->chain_ms is the device time of the whole replay and
 * stats->build_ms is 0.  On ROCm 7.2 the HIP runtime keeps an allocation that
 * an instantiated graph's kernel nodes pointed into after hipFree, until the
 * process ends; the library therefore keeps such workspaces itself and hands
 * them to the next context on the device, so a process holds one of them per
 * device, not one per graph-mode context.  Placed tables (virtual-memory
 * pieces) are not affected: they are unmapped and released as outside graph
 * mode. */
int ck_ctx_set_graph(ck_ctx* ctx, uint32_t on);

/* How bucket evaluations get their Feistel round function. */
enum {
    CK_PRP_AUTO = 0,   /* a round cache in LDS when half_bits <= 12, else AES evaluated directly */
    CK_PRP_CACHE = 1,  /* a round cache (CK_EINVAL at launch when half_bits > 12) */
    CK_PRP_DIRECT = 2  /* AES evaluated for every round (the reference's --prp-cache off) */
};
/* LDS-tier chain form (the reference's --shared-variant, cuckoo_bench.cu:527-563). */
enum {
    CK_LDS_AUTO = 0,  /* MAP */
    CK_LDS_MAP = 1,   /* the build writes the bucket map into LDS; a probe reads it */
    CK_LDS_ROUND = 2  /* no map: each probe evaluates its bucket over the LDS round cache */
};

/* Placement request of a context (the reference's --table-layout,
 * --shared-variant, --prp-cache and --scratch-mib, cuckoo_bench.cu:406-419). */
typedef struct ck_placement {
    uint32_t tier;          /* CK_TIER_* */
    uint32_t layout;        /* CK_LAYOUT_* (HBM tier) */
    uint32_t prp;           /* CK_PRP_* */
    uint32_t lds_variant;   /* CK_LDS_* */
    uint64_t budget_bytes;  /* device memory for a launch's workspace (0: 95 % of the free memory) */
} ck_placement;
int ck_ctx_set_placement(ck_ctx* ctx, const ck_placement* want);

/* What a launch of `count` trials would do. */
typedef struct ck_plan {
    uint32_t tier;             /* CK_TIER_HBM, CK_TIER_LDS or CK_TIER_LANE */
    uint32_t layout;           /* CK_LAYOUT_* chosen (HBM tier; RESIDUAL for the lane tier; AUTO for the LDS tier) */
    uint32_t round_cache;      /* 1: bucket evaluations read a round cache; 0: direct AES */
    uint32_t lds_variant;      /* CK_LDS_MAP or CK_LDS_ROUND (LDS tier) */
    uint32_t trials_per_pass;  /* HBM: chains per pass; LDS: trials per fused launch */
    uint32_t buffers;          /* HBM pass buffers */
    uint32_t windows;          /* HBM item windows per pass (1: none) */
    uint32_t chain_waves;      /* wave64s per chain (LDS: fused) block */
    uint32_t chain_lds_bytes;  /* LDS per chain (LDS: fused) block */
    uint32_t build_lds_bytes;  /* LDS per build block (HBM tier) */
    uint64_t workspace_bytes;  /* device workspace of the launch */
} ck_plan;
/* The context's plan, with its device's free memory (no kernel runs). */
int ck_ctx_plan(ck_ctx* ctx, const ck_params* p, uint32_t count, ck_plan* out);
/* The same without a device: a budget of 0 plans against 96 GiB and 160 KB of
 * LDS per block (MI355X). */
int ck_plan_describe(const ck_params* p, const ck_placement* want, uint32_t count, ck_plan* out);

/* One autotuner measurement. */
typedef struct ck_tune_entry {
    uint32_t tier, layout, chunk; /* CK_TIER_*, CK_LAYOUT_*, trials per pass (0: library default) */
    uint32_t pad;
    double trials_per_s;          /* wall-clock rate on the sample */
} ck_tune_entry;

/* Autotuner (per-shape tuning of the placement that the
 * reference exposes as --table-layout/--shared-variant/--block-threads,
 * cuckoo_bench.cu:819-901): runs trials [0, sample) of the batch under every
 * applicable tier / table layout / trials-per-pass candidate, checks that all
 * of them report the same failing trials (CK_ERANGE otherwise), keeps the
 * fastest as this context's setting and reports each measurement in
 * entries[0 .. min(cap, *n)-1]. */
int ck_ctx_tune(ck_ctx* ctx, const ck_params* p, const uint32_t round_keys[44], uint32_t sample,
                ck_tune_entry* entries, uint32_t cap, uint32_t* n);

/* FIPS-197 AES-128 known answer through the device AES path
 * (aes_self_test / cached_aes_self_test, cuckoo_bench.cu:332-354). */
int ck_aes_self_test(ck_ctx* ctx);

/* The launch callable of process_batch (cuckoo_bench.cu:980-1031): runs trials
 * [first, first+count) of the batch whose key schedule is round_keys.
 * failed_offsets (capacity >= count) receives the failing offsets in
 * ascending order, *n_failed their number.  stats may be NULL. */
int ck_launch(ck_ctx* ctx, const ck_params* p, const uint32_t round_keys[44], uint32_t first,
              uint32_t count, uint32_t* failed_offsets, uint32_t* n_failed, ck_stats* stats);

/* Allocation-free form of ck_launch.  It has no counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops or harnesses that forbid hipMalloc).  The launch's
 * workspace is a CK_STATIC_SCRATCH_BYTES device array that the library holds on
 * every GPU it runs on, from load until the process ends, instead of memory it
 * allocates: the call makes no hipMalloc, hipFree or other allocating call.
 * Results equal ck_launch's.  Passes are sized to the array (a few hundred
 * trials per pass at n = 2^14, so slower than ck_launch), graph mode is not
 * used and tables are cleared every pass.  CK_EINVAL, before any device work,
 * when one trial of the shape needs more workspace than the array holds.  All
 * calls on one GPU share its array and leave their contents in it: calls that
 * overlap on one GPU (other threads or contexts) race, so the caller must
 * serialise them. */
#define CK_STATIC_SCRATCH_BYTES (256ull << 20)
int ck_launch_static(ck_ctx* ctx, const ck_params* p, const uint32_t round_keys[44], uint32_t first,
                     uint32_t count, uint32_t* failed_offsets, uint32_t* n_failed, ck_stats* stats);

/* Test hook: like ck_launch, but also returns per-trial probe counts
 * (probes[i] for trial first+i) and, when tables is non-NULL and the shape
 * allows, the final table of every trial (tables[i*buckets + b], entry =
 * ordinal | choice<<30, empty = 0xffffffff). */
int ck_launch_debug(ck_ctx* ctx, const ck_params* p, const uint32_t round_keys[44], uint32_t first,
                    uint32_t count, uint32_t* failed_offsets, uint32_t* n_failed, uint32_t* probes,
                    uint32_t* tables, ck_stats* stats);

/* Test hook: like ck_launch, plus per-trial probe counts and an
 * order-independent digest of every trial's final table: the wrapping sum over
 * occupied buckets b of mix64(b*0x9E3779B97F4A7C15 ^ H ^ choice<<61), with
 * H = sum over w<k of mix64(w<<32 | bucket(occupant, w)) and mix64 the
 * SplitMix64 output function.  Works for every tier and layout. */
int ck_launch_digest(ck_ctx* ctx, const ck_params* p, const uint32_t round_keys[44], uint32_t first,
                     uint32_t count, uint32_t* failed_offsets, uint32_t* n_failed, uint32_t* probes,
                     uint64_t* digests, ck_stats* stats);

/* Test hook: the device key derivation and bucket map of one trial:
 * words[0..43] = PRP key schedule, walk/randomness of the first BGH draw,
 * map[o*k+w] = bucket(o, w). */
int ck_trial_map(ck_ctx* ctx, const ck_params* p, const uint32_t round_keys[44], uint32_t offset,
                 uint32_t key_words[44], uint64_t* walk, uint64_t* randomness, uint32_t* map);

const char* ck_last_error(void);

/* ------------------------------------------------------------------------
 * Batched and served sweeps.  The reference runs one experiment shape per
 * process (cuckoo_bench.cu:1062-1142, driven per point by tune_bench.py and
 * compare_cpu_gpu.py); the engine below runs many shapes ("jobs") through
 * shared device passes.  Each job is the same trial sequence as the launch
 * callable: trial t is offset t mod 2^bt of batch t >> bt, keyed by
 * batch_key(seed, batch) (cuckoo_bench.cu:393-404, :728-734), so a job's
 * failing trials are exactly those ck_launch reports for the same ranges.
 * ---------------------------------------------------------------------- */

enum { CK_PENDING = 1 }; /* ck_engine_poll: the job has not finished */

enum {
    CK_SCHED_BALANCE = 0, /* overlap build-heavy and chain-heavy passes */
    CK_SCHED_FIFO = 1,    /* oldest job first */
    CK_SCHED_SRPT = 2     /* least remaining estimated work first: the lowest latency, and on the
                             512-point mixed sweep also the highest throughput (passes hold jobs of
                             similar n, so their chains end together) */
};

typedef struct ck_job {
    ck_params p;            /* shape, as returned by ck_params_derive */
    uint64_t seed;          /* master seed of the batch keys */
    uint64_t trials;        /* trials 0 .. trials-1 */
    uint32_t bt;            /* batch bits, 6..32 */
    uint32_t keep_failures; /* nonzero: keep the failing trial indices (ck_engine_failures) */
} ck_job;

typedef struct ck_job_result {
    uint64_t trials;   /* trials run (device counter) */
    uint64_t failed;   /* failing trials (device counter) */
    uint64_t probes;   /* table-slot reads of the insertion chains (device counter) */
    double submit_ms;  /* engine clock at submission */
    double start_ms;   /* first pass holding the job enqueued */
    double end_ms;     /* last pass gathered; counters final */
} ck_job_result;

typedef struct ck_engine_config {
    uint64_t arena_bytes; /* device memory for maps and tables over all pass buffers
                             (0: 192 GiB or five sixths of the free memory, the smaller) */
    uint32_t buffers;     /* pass buffers in flight, 2..8 (0: 4); passes chain on two streams */
    uint32_t max_slots;   /* trial slots per pass, multiple of 64 (0: 65536) */
    uint32_t max_jobs;    /* jobs admitted at once: device counter slots (0: 4096) */
    uint32_t policy;      /* CK_SCHED_* */
    uint32_t window;      /* jobs considered per pass, in arrival order (0: all) */
    uint32_t spread;      /* 1: the arena is 1 GiB pieces taken from the device's memory regions in
                             turn (random read + write traffic spread over the regions runs faster
                             than traffic in one; see ck_ctx_set_place_search), or one allocation
                             where the regions cannot be told apart; 0: one allocation */
} ck_engine_config;

typedef struct ck_engine_stats {
    uint64_t passes, segments, slots;       /* passes run, 64-trial segments, live trial slots */
    uint64_t jobs_submitted, jobs_done;
    uint64_t trials, failed, probes;        /* totals over finished jobs */
    double build_ms, chain_ms;              /* device time summed over passes */
    double wall_ms;                         /* engine clock now */
} ck_engine_stats;

typedef struct ck_engine ck_engine;

/* An engine owns a device context, its pass buffers (ring of `buffers`),
 * device-resident per-job counters and a scheduler thread. */
int ck_engine_create(int device, const ck_engine_config* cfg, ck_engine** out);
/* Finishes the passes in flight, then stops; unfinished jobs are abandoned. */
void ck_engine_destroy(ck_engine* e);
/* Queues a job (thread-safe).  Jobs are admitted as device counter slots
 * free up and join the next pass planned.  CK_ENOMEM: 64 trials of the shape
 * do not fit one pass buffer. */
int ck_engine_submit(ck_engine* e, const ck_job* job, uint64_t* job_id);
/* CK_OK and the result when the job has finished, CK_PENDING otherwise. */
int ck_engine_poll(ck_engine* e, uint64_t job_id, ck_job_result* out);
/* Blocks until the job has finished. */
int ck_engine_wait(ck_engine* e, uint64_t job_id, ck_job_result* out);
/* Failing trial indices of a finished job kept with keep_failures, ascending:
 * out[0 .. min(cap, n)-1], *n = their number. */
int ck_engine_failures(ck_engine* e, uint64_t job_id, uint64_t* out, uint64_t cap, uint64_t* n);
/* Drops a finished job's host record. */
int ck_engine_release(ck_engine* e, uint64_t job_id);
int ck_engine_stats_get(ck_engine* e, ck_engine_stats* out);
/* The arena's 1 GiB pieces per memory region: counts[0..2] for the three
 * regions, counts[3] for pieces whose region was unclear; all zero when the
 * arena is one allocation.  No reference counterpart. */
int ck_engine_arena(ck_engine* e, uint32_t counts[4]);

#ifdef __cplusplus
}
#endif
#endif
