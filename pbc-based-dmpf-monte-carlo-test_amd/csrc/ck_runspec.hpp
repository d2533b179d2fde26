// ck_runspec.hpp -- what one cuckoo_bench run is: its specification (the
// command line of the reference's cuckoo_bench, cuckoo_bench.cu:406-480, plus
// the k / stash / explicit-size extension), the trial plan it implies, the
// shared tallies of a multi-worker run, stop requests (SIGINT, SIGTERM or a
// PAUSE file in the output directory), the output-directory lock, and the
// commit boundary `process_batch` (cuckoo_bench.cu:724-745) that every worker
// drives its device launches through.
//
// Host-only C++17: no device code; the tests drive process_batch with a fake
// launch.
#pragma once
#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

#include "cuckoo_mc.h"

namespace ckrun {

namespace fs = std::filesystem;

// One run's specification.  Defaults are the reference's (cuckoo_bench.cu:406-419).
struct RunSpec {
    // Experiment identity: these fields name the trial sequence and the shape.
    std::string mode = "drrt18";
    uint32_t pnum = 16;
    uint32_t ell_i = 32;
    uint64_t seed = 1;
    std::string input_mode = "sequential";
    uint32_t et = 30;      // planned = 2^et trials ...
    uint64_t trials = 0;   // ... unless this extension count is nonzero
    uint32_t bt = 20;      // batches of 2^bt trials
    uint32_t k = 3;        // extension: candidate buckets per item
    uint32_t buckets = CK_DERIVED, kick = CK_DERIVED;  // extension: explicit sizes
    uint32_t stash = 0;    // extension: stash slots
    // Execution: how the trials are run; results do not depend on these.
    std::string out = "cuckoo_results";
    int workers = 0;                   // 0: one per GPU
    uint32_t micro = 1u << 17;         // trials per device launch
    uint64_t stop_after = 1ull << 20;  // stop once more failures than this are committed
    uint32_t scratch_mib = 256;        // device memory per worker, when given
    bool scratch_given = false;        // otherwise the library sizes its own workspace
    uint32_t block_threads = 256, warps_per_block = 0;
    std::string prp_cache = "auto", table_layout = "auto", shared_variant = "auto";
    std::string tier = "auto";
    uint32_t place_search = 32;        // table placement on, candidates of its fallback search (0: off)
    bool global_table = false, status_only = false, tune = false, graph = false;
};

// Parses and validates a command line (argv[0] is skipped).  Throws
// std::runtime_error naming the offending flag or rule.  --help prints the
// usage and exits.
RunSpec parse_command_line(int argc, char** argv);
std::string usage_text();

// Unsigned decimal / 0x-hex / 0-octal, the whole string, no sign.
uint64_t parse_unsigned(const std::string& text, uint64_t max_value);

// ck_params of the spec (mode-derived buckets and kick unless given).
ck_params derive_shape(const RunSpec& s);

// Trial plan: trial t is offset t mod 2^bt of batch t >> bt.
uint64_t planned_trials(const RunSpec& s);
uint64_t batch_count(const RunSpec& s);
uint64_t batch_length(const RunSpec& s, uint64_t batch);

// Journal encoding of the mode and input mode (cuckoo_bench.cu:567-571, :607).
inline uint32_t mode_code(const RunSpec& s) { return s.mode == "bgh25" ? CK_MODE_BGH25 : CK_MODE_DRRT18; }
inline uint32_t input_code(const RunSpec& s) { return s.input_mode == "sequential" ? 0u : 1u; }

// What the run has committed so far, shared by all workers.  Only batches
// whose journal record is durable are counted (the commit happens first).
struct RunTally {
    std::atomic<uint64_t> trials{0}, failures{0}, batches{0};
    std::atomic<uint64_t> cursor{0};        // next batch id to claim
    std::atomic<bool> halted{false};        // no worker starts another batch
    std::atomic<bool> aborted{false};       // a worker failed: others drop their batch in progress
    std::atomic<int> live_workers{0};
    std::atomic<int64_t> last_exit_ns{0};   // when the last worker stopped (steady clock)

    void credit(uint64_t batch_trials, uint64_t batch_failures, uint64_t stop_after) {
        trials.fetch_add(batch_trials, std::memory_order_relaxed);
        const uint64_t total = failures.fetch_add(batch_failures) + batch_failures;
        batches.fetch_add(1);
        if (total > stop_after) halted.store(true);
    }
    void note_worker_exit();
};

int64_t steady_now_ns();

// Stop requests.  While a StopSignals object lives, SIGINT and SIGTERM set a
// flag instead of ending the process (the previous dispositions come back when
// it is destroyed).  A file named PAUSE in the output directory is the same
// request from outside.  Workers look between device launches, so an
// interrupted batch is dropped whole and rerun on resume.
class StopSignals {
   public:
    StopSignals();
    ~StopSignals();
    StopSignals(const StopSignals&) = delete;
    StopSignals& operator=(const StopSignals&) = delete;
    static bool raised();
    static void clear();
    static void raise_for_test();

   private:
    struct Saved;
    Saved* saved_;
};
bool stop_requested(const RunSpec& s);

// Exclusive advisory lock of the output directory (flock on <out>/.run.lock,
// the file and lock kind the reference uses, checkpoint.hpp:44-71, so the two
// programs exclude each other).  Released by the kernel when the process ends.
class DirectoryLock {
   public:
    explicit DirectoryLock(const fs::path& directory);
    ~DirectoryLock();
    DirectoryLock(const DirectoryLock&) = delete;
    DirectoryLock& operator=(const DirectoryLock&) = delete;

   private:
    int fd_ = -1;
};

using BatchKey = std::array<uint32_t, 44>;
BatchKey batch_key_of(uint64_t seed, uint64_t batch);

// The commit boundary (cuckoo_bench.cu:724-745).  Runs batch `batch` through
// launch(count, offset, key, failures) in pieces of at most `micro` trials;
// failures appends the failing offsets of its piece in any order.  Before each
// piece it checks for a stop request or another worker's abort: then it halts
// the run and returns false, committing nothing.  When every piece has run it
// sorts the failures, calls commit(batch, trials, failures) -- which must make
// the record durable -- and only then credits the tally; more than stop_after
// committed failures halt the run.  A throw from launch or commit propagates
// and leaves the tally untouched.
template <class Launch, class Commit>
bool process_batch(const RunSpec& spec, uint64_t batch, uint32_t micro, RunTally& tally, Launch&& launch,
                   Commit&& commit) {
    if (micro == 0) throw std::invalid_argument("process_batch: micro must be at least 1");
    const BatchKey key = batch_key_of(spec.seed, batch);
    const uint32_t length = static_cast<uint32_t>(batch_length(spec, batch));
    std::vector<uint32_t> failing;
    uint32_t offset = 0;
    while (offset < length) {
        if (tally.aborted.load() || stop_requested(spec)) {
            tally.halted.store(true);
            return false;
        }
        const uint32_t piece = std::min(micro, length - offset);
        launch(piece, offset, key, failing);
        offset += piece;
    }
    std::sort(failing.begin(), failing.end());
    commit(batch, length, failing);
    tally.credit(length, failing.size(), spec.stop_after);
    return true;
}

}  // namespace ckrun
