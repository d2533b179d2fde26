// cuckoo_bench.cpp -- the cuckoo_bench command line (the reference's
// cuckoo_bench.cu:1062-1142 interface: same flags, journals, summary.json and
// failures.csv) on the MI355X library.
//
// A run: lock the output directory, rebuild progress from the journals, then
// start one worker thread per context; workers claim batch ids from a shared
// cursor, skip committed ones, and run each batch through process_batch over
// ck_launch, committing it to their own journal.  When they stop (plan done,
// failure threshold, stop request or an error) the journals are scanned again
// and the reports are written from that scan.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <exception>
#include <iostream>
#include <memory>
#include <mutex>
#include <thread>

#include "ck_ledger.hpp"

using namespace ckrun;

namespace {

void check(int rc, const char* call) {
    if (rc != CK_OK) throw std::runtime_error(std::string(call) + ": " + ck_last_error());
}

// The library placement a spec asks for.  --table-layout shared is the LDS
// tier (its --shared-variant picks the map or the round-cache form);
// --table-layout global and --global-table the HBM tier; otherwise --tier.
ck_placement placement_of(const RunSpec& s, uint64_t budget_bytes) {
    ck_placement w{};
    if (s.table_layout == "shared")
        w.tier = CK_TIER_LDS;
    else if (s.table_layout == "global" || s.global_table)
        w.tier = CK_TIER_HBM;
    else
        w.tier = s.tier == "hbm"    ? CK_TIER_HBM
                 : s.tier == "lds"  ? CK_TIER_LDS
                 : s.tier == "lane" ? CK_TIER_LANE
                                    : CK_TIER_AUTO;
    w.layout = CK_LAYOUT_AUTO;
    w.prp = s.prp_cache == "on" ? CK_PRP_CACHE : s.prp_cache == "off" ? CK_PRP_DIRECT : CK_PRP_AUTO;
    w.lds_variant = s.shared_variant == "round" ? CK_LDS_ROUND : s.shared_variant == "map" ? CK_LDS_MAP : CK_LDS_AUTO;
    w.budget_bytes = budget_bytes;
    return w;
}

const char* layout_name(uint32_t layout) {
    switch (layout) {
        case CK_LAYOUT_MAP: return "map";
        case CK_LAYOUT_DESC: return "desc";
        case CK_LAYOUT_MAP_XCHG: return "map-xchg";
        case CK_LAYOUT_DESC_PLAIN: return "desc-plain";
        case CK_LAYOUT_RESIDUAL: return "residual";
        default: return "auto";
    }
}

// The planned placement in the reference's vocabulary where one exists
// (shared-map / shared-trial for workgroup-LDS tables, cuckoo_bench.cu:527-563),
// <cached|direct>-lane-shared for lane-private LDS tables and
// <cached|direct>-global-<table layout> for the HBM tier.
std::string placement_name(const ck_plan& pl) {
    if (pl.tier == CK_TIER_LDS) return pl.lds_variant == CK_LDS_ROUND ? "shared-trial" : "shared-map";
    if (pl.tier == CK_TIER_LANE) return std::string(pl.round_cache ? "cached" : "direct") + "-lane-shared";
    return std::string(pl.round_cache ? "cached" : "direct") + "-global-" + layout_name(pl.layout);
}

class Context {
   public:
    explicit Context(int device) { check(ck_ctx_create(device, &h_), "ck_ctx_create"); }
    ~Context() { ck_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    ck_ctx* get() const { return h_; }

   private:
    ck_ctx* h_ = nullptr;
};

// Everything the workers share.
struct Shared {
    RunSpec spec;
    ck_params shape{};
    std::vector<uint8_t> committed;  // batches already in the journals at start
    std::vector<uint64_t> budgets;   // device memory per worker, by device (0: library decides)
    RunTally tally;
    std::mutex log;
    std::vector<std::exception_ptr> errors;
};

void worker_main(Shared& sh, int index, int device) {
    const RunSpec& s = sh.spec;
    Context ctx(device);
    const ck_placement want = placement_of(s, sh.budgets[device]);
    check(ck_ctx_set_placement(ctx.get(), &want), "ck_ctx_set_placement");
    check(ck_ctx_set_graph(ctx.get(), s.graph ? 1 : 0), "ck_ctx_set_graph");
    // A worker runs many passes of one shape: the tables of its pass buffers
    // are spread over the memory regions (or, where pieces would be small, the
    // two fastest of --place-search candidates) where the plan qualifies.
    check(ck_ctx_set_place_search(ctx.get(), s.place_search), "ck_ctx_set_place_search");
    const uint32_t micro = static_cast<uint32_t>(std::min<uint64_t>(s.micro, uint64_t(1) << s.bt));
    std::string tuned;
    if (s.tune) {
        // Every placement on a sample of batch 0 (nothing is committed); the
        // fastest stays as this context's setting.
        const uint32_t sample = static_cast<uint32_t>(std::min<uint64_t>(batch_length(s, 0), 1u << 17));
        const BatchKey key = batch_key_of(s.seed, 0);
        ck_tune_entry entries[32];
        uint32_t n = 0;
        check(ck_ctx_tune(ctx.get(), &sh.shape, key.data(), sample, entries, 32, &n), "ck_ctx_tune");
        const ck_tune_entry* best = nullptr;
        for (uint32_t i = 0; i < std::min<uint32_t>(n, 32); ++i)
            if (!best || entries[i].trials_per_s > best->trials_per_s) best = &entries[i];
        if (best) {
            char buf[160];
            std::snprintf(buf, sizeof buf, " tuned=%s/%s/pass%u@%.0f/s",
                          best->tier == CK_TIER_LDS ? "lds" : best->tier == CK_TIER_LANE ? "lane" : "hbm",
                          layout_name(best->layout), best->chunk, best->trials_per_s);
            tuned = buf;
        }
    }
    ck_plan plan{};
    check(ck_ctx_plan(ctx.get(), &sh.shape, micro, &plan), "ck_ctx_plan");
    JournalWriter journal(s, sh.shape, index);
    {
        hipDeviceProp_t prop{};
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) prop.name[0] = '\0';
        std::lock_guard<std::mutex> lock(sh.log);
        // The fields the reference's harnesses read from this line
        // (compare_cpu_gpu.py:211-215): micro=, table=, warps=, shared_bytes=,
        // scratch_bytes=.
        std::cerr << "Worker " << index << " GPU " << device << " (" << prop.name << ") bins=" << sh.shape.buckets
                  << " table=" << placement_name(plan) << " micro=" << micro << " block=" << 64 * plan.chain_waves
                  << " warps=" << plan.chain_waves << " shared_bytes=" << plan.chain_lds_bytes
                  << " scratch_bytes=" << plan.workspace_bytes << " trials_per_pass=" << plan.trials_per_pass
                  << " build_lds_bytes=" << plan.build_lds_bytes << " k=" << sh.shape.k
                  << " stash=" << sh.shape.stash << " kick=" << sh.shape.kick << " half_bits=" << sh.shape.half_bits
                  << " CUs=" << prop.multiProcessorCount << tuned << (s.graph ? " graph=on" : "") << "\n";
    }
    std::vector<uint32_t> piece_failures(micro);
    const auto launch = [&](uint32_t count, uint32_t offset, const BatchKey& key, std::vector<uint32_t>& failing) {
        uint32_t nf = 0;
        check(ck_launch(ctx.get(), &sh.shape, key.data(), offset, count, piece_failures.data(), &nf, nullptr),
              "ck_launch");
        failing.insert(failing.end(), piece_failures.begin(), piece_failures.begin() + nf);
    };
    const auto commit = [&](uint64_t batch, uint32_t trials, const std::vector<uint32_t>& offsets) {
        journal.commit(batch, trials, offsets);
    };
    const uint64_t batches = batch_count(s);
    for (;;) {
        if (sh.tally.halted.load()) break;
        if (stop_requested(s)) {
            sh.tally.halted.store(true);
            break;
        }
        const uint64_t batch = sh.tally.cursor.fetch_add(1);
        if (batch >= batches) break;
        if (sh.committed[batch]) continue;
        if (!process_batch(s, batch, micro, sh.tally, launch, commit)) break;
    }
}

// One status line a second on stderr while workers run.
void show_progress(const Shared& sh, std::chrono::steady_clock::time_point start) {
    const uint64_t planned = planned_trials(sh.spec), at_start = sh.tally.trials.load();
    while (sh.tally.live_workers.load() > 0) {
        std::this_thread::sleep_for(std::chrono::seconds(1));
        const uint64_t done = sh.tally.trials.load();
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        char line[200];
        std::snprintf(line, sizeof line, "\r%6.2f%%  %llu/%llu trials  %llu failed  %.0f trials/s  %.1f s", 100.0 * done / planned,
                      static_cast<unsigned long long>(done), static_cast<unsigned long long>(planned),
                      static_cast<unsigned long long>(sh.tally.failures.load()),
                      (done - at_start) / (secs > 1e-6 ? secs : 1e-6), secs);
        std::cerr << line << std::flush;
    }
    std::cerr << '\n';
}

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    return n;
}

// Device memory per worker: --scratch-mib when given; else 95 % of each
// device's free memory shared evenly by the workers placed on it (0 = the
// library's own budget when a device holds one worker).
std::vector<uint64_t> worker_budgets(const RunSpec& s, int gpus) {
    std::vector<uint64_t> out(gpus, 0);
    for (int d = 0; d < gpus; ++d) {
        const int sharing = s.workers / gpus + (d < s.workers % gpus ? 1 : 0);
        if (s.scratch_given) {
            out[d] = uint64_t(s.scratch_mib) << 20;
        } else if (sharing > 1) {
            size_t free_b = 0, total_b = 0;
            if (hipSetDevice(d) == hipSuccess && hipMemGetInfo(&free_b, &total_b) == hipSuccess)
                out[d] = free_b / 20 * 19 / sharing;
        }
    }
    return out;
}

std::string status_word(bool error, const LedgerState& st, const RunSpec& s) {
    if (error) return "error";
    if (st.trials == planned_trials(s)) return "complete";
    if (stop_requested(s)) return "paused";
    if (st.failures > s.stop_after) return "threshold";
    return "ready";
}

int run(RunSpec s) {
    const ck_params shape = derive_shape(s);
    fs::create_directories(s.out);
    DirectoryLock lock(s.out);
    StopSignals signals;
    const auto start = std::chrono::steady_clock::now();
    LedgerState before = scan_ledger(s, shape, false);
    std::cout << "mode=" << s.mode << " ET=" << (s.trials ? 0 : s.et) << " BT=" << s.bt << " PNUM=" << s.pnum
              << " buckets=" << shape.buckets << " kick=" << shape.kick << " bucket_size=" << shape.bucket_size
              << " ell_d=" << shape.ell_d << " input_mode=" << s.input_mode << " k=" << shape.k
              << " stash=" << shape.stash << " planned=" << planned_trials(s) << "\n"
              << "Journals: " << before.trials << " trials in " << before.batches << " batches committed, "
              << before.failures << " failed\n";

    RunReport report;
    report.workers = s.workers;
    {
        ck_plan pl{};
        const ck_placement want = placement_of(s, 0);
        check(ck_plan_describe(&shape, &want, std::min<uint32_t>(s.micro, 1u << std::min(s.bt, 31u)), &pl),
              "ck_plan_describe");
        report.round_cache = pl.round_cache != 0;
        report.placement = placement_name(pl);
    }
    const bool nothing_to_do = before.trials == planned_trials(s) || before.failures > s.stop_after;
    if (s.status_only || nothing_to_do || stop_requested(s)) {
        const LedgerState st = scan_ledger(s, shape, true);
        report.status = status_word(false, st, s);
        write_reports(s, shape, st, report);
        std::cout << "Status: " << report.status << "; pending batches=" << (batch_count(s) - st.batches) << "\n";
        return 0;
    }

    const int gpus = device_count();
    if (gpus == 0) throw std::runtime_error("no HIP GPU is visible");
    if (s.workers == 0) s.workers = gpus;
    if (s.workers > gpus) std::cerr << "note: " << s.workers << " workers share " << gpus << " GPU(s)\n";
    report.workers = s.workers;
    // Known-answer test of the device AES on every GPU used, before any trial.
    for (int d = 0; d < std::min(gpus, s.workers); ++d) {
        Context ctx(d);
        check(ck_aes_self_test(ctx.get()), "device AES-128 known-answer test");
    }

    Shared sh;
    sh.spec = s;
    sh.shape = shape;
    sh.committed = std::move(before.committed);
    sh.budgets = worker_budgets(s, gpus);
    sh.tally.trials = before.trials;
    sh.tally.failures = before.failures;
    sh.tally.batches = before.batches;
    sh.errors.resize(s.workers);
    const int64_t work_start = steady_now_ns();
    std::vector<std::thread> threads;
    threads.reserve(s.workers);
    for (int w = 0; w < s.workers; ++w) {
        sh.tally.live_workers.fetch_add(1);
        try {
            threads.emplace_back([&sh, w, gpus] {
                try {
                    worker_main(sh, w, w % gpus);
                } catch (...) {
                    sh.errors[w] = std::current_exception();
                    sh.tally.aborted.store(true);
                    sh.tally.halted.store(true);
                }
                sh.tally.note_worker_exit();
                sh.tally.live_workers.fetch_sub(1);
            });
        } catch (...) {
            sh.tally.live_workers.fetch_sub(1);
            sh.errors[w] = std::current_exception();
            sh.tally.aborted.store(true);
            sh.tally.halted.store(true);
            break;
        }
    }
    show_progress(sh, start);
    for (auto& t : threads) t.join();

    // The journals, not the live tallies, are the result.
    const LedgerState st = scan_ledger(s, shape, true);
    std::exception_ptr first_error;
    for (const auto& e : sh.errors)
        if (e && !first_error) first_error = e;
    report.status = status_word(bool(first_error), st, s);
    if (report.status == "ready") report.status = "threshold";
    report.elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    report.work_s = std::max<int64_t>(0, sh.tally.last_exit_ns.load() - work_start) * 1e-9;
    write_reports(s, shape, st, report);
    std::cout << "RESULT status=" << report.status << " completed=" << st.trials << " failed=" << st.failures
              << " elapsed_s=" << report.elapsed_s << " work_s=" << report.work_s << "\n";
    if (report.status == "paused")
        std::cout << "Stopped on request: batches in progress were dropped and will rerun. Run the same command "
                     "again to resume (delete the PAUSE file first if one was used).\n";
    if (first_error) std::rethrow_exception(first_error);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(parse_command_line(argc, argv));
    } catch (const std::exception& e) {
        std::cerr << "FATAL: " << e.what() << "\n";
        return 1;
    }
}
