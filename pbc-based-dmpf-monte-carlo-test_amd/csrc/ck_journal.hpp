// ck_journal.hpp -- run controller above the C ABI: experiment identity, the
// append-only checksummed journal (v3 format of cuckoo_bench.cu:565-698, plus a
// v4 header for the k/stash/trial-count extension), resume, pause handling,
// the per-batch commit boundary process_batch (cuckoo_bench.cu:724-745) and the
// summary.json / failures.csv export (cuckoo_bench.cu:746-800).
//
// Host-only C++17; usable without a GPU (tests drive process_batch with a mock
// launch, as the reference's host tests do).
#pragma once
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <csignal>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "cuckoo_mc.h"

namespace ckh {

namespace fs = std::filesystem;

constexpr uint32_t JOURNAL_MAGIC = 0x43554b33u;  // "CUK3"
constexpr uint32_t RECORD_MAGIC = 0x454e4433u;   // "END3"

// Command-line options (cuckoo_bench.cu:406-419) plus the extension flags.
struct Options {
    uint32_t et = 30, bt = 20, pnum = 16, ell_i = 32, micro = 1u << 17;
    uint64_t seed = 1, stop_after = 1ULL << 20;
    std::string mode = "drrt18", out = "cuckoo_results", input_mode = "sequential";
    int workers = 0;
    uint32_t scratch_mib = 256;
    uint32_t block_threads = 256;
    uint32_t warps_per_block = 0;
    std::string prp_cache = "auto", table_layout = "auto", shared_variant = "auto";
    std::string backend = "hip";
    bool force_global = false, status_only = false;
    // Extension (no reference counterpart): k choices, explicit buckets/kick,
    // stash slots, a non-power-of-two trial count, the chain tier.
    uint32_t k = 3, buckets = CK_DERIVED, kick = CK_DERIVED, stash = 0;
    uint64_t trials = 0;  // 0: planned = 2^et
    std::string tier = "auto";
    bool tune = false;   // --tune: each worker times the placements on a sample first
    bool graph = false;  // --graph: each worker's launches replay a hipGraph per micro-range size
};

Options parse(int argc, char** argv);
uint64_t parse_u64(const std::string& s);
uint32_t parse_u32(const std::string& s);

ck_params parameters(const Options& o);  // throws std::runtime_error on an invalid shape
uint64_t planned(const Options& o);
uint64_t total_batches(const Options& o);
uint64_t batch_length(const Options& o, uint64_t id);

#pragma pack(push, 1)
struct FileHeaderV3 {
    uint32_t magic = JOURNAL_MAGIC, version = 3;
    uint64_t seed = 0, planned = 0;
    uint32_t et = 0, bt = 0, pnum = 0, ell_i = 0, mode = 0, buckets = 0, kick = 0, input_mode = 0;
};
struct HeaderExt {  // follows the v3 fields when version == 4
    uint32_t k = 3, stash = 0;
};
#pragma pack(pop)
static_assert(sizeof(FileHeaderV3) == 56, "v3 header is 56 bytes");
struct RecordHeader {
    uint64_t batch;
    uint32_t tried, failed;
};
struct RecordFooter {
    uint32_t check, magic;
};
static_assert(sizeof(RecordHeader) == 16 && sizeof(RecordFooter) == 8, "record layout");

struct Header {
    FileHeaderV3 base;
    HeaderExt ext;
    bool v4 = false;
    size_t size() const { return v4 ? sizeof(FileHeaderV3) + sizeof(HeaderExt) : sizeof(FileHeaderV3); }
};
Header create_header(const Options& o, const ck_params& p);
uint32_t checksum(uint64_t batch, uint32_t tried, const std::vector<uint32_t>& offsets);

struct Counters {
    std::atomic<uint64_t> completed{0}, failed{0}, batches{0};
    std::atomic<bool> stop{false}, abort_inflight{false};
    std::atomic<int> active{0};
    std::atomic<uint64_t> next_batch{0};
    std::atomic<int64_t> finished_ns{0};
};
int64_t steady_ns();
void mark_finished(Counters& s);

std::string journal_name(const Options& o, int worker);
std::vector<std::string> journal_paths(const Options& o);
void append_record(FILE* f, uint64_t id, uint32_t attempted, const std::vector<uint32_t>& offsets);
FILE* open_journal(const Options& o, const ck_params& p, int worker);
std::vector<uint8_t> resume(const Options& o, const ck_params& p, Counters& s);
void final_outputs(const Options& o, const ck_params& p, const Counters& s, double seconds,
                   const std::string& status = "complete", double work_seconds = 0);

// Pause: SIGINT/SIGTERM set a lock-free flag; a PAUSE file in the output dir
// has the same effect (checkpoint.hpp:19-40, cuckoo_bench.cu:717-719).
extern std::atomic<bool> pause_signal;
class PauseSignals {
    using Handler = void (*)(int);
    Handler old_int_, old_term_;

public:
    PauseSignals();
    ~PauseSignals();
    PauseSignals(const PauseSignals&) = delete;
};
bool pause_requested(const Options& o);

// Exclusive advisory lock on <out>/.run.lock (checkpoint.hpp:44-71).
class RunLock {
    int fd_ = -1;

public:
    explicit RunLock(const fs::path& directory);
    ~RunLock();
    RunLock(const RunLock&) = delete;
};
void durable_flush(FILE* f);

using BatchKey = std::array<uint32_t, 44>;
BatchKey batch_key(uint64_t seed, uint64_t id);

// The commit boundary: runs the whole batch through `launch` in micro-sized
// pieces, sorts the failures and commits them only when every piece finished
// (cuckoo_bench.cu:724-745).  A pause or another worker's abort between pieces
// discards the batch and returns false.
template <class Launch, class Commit>
bool process_batch(const Options& o, uint64_t id, uint32_t micro, Counters& s, Launch&& launch, Commit&& commit) {
    if (micro == 0) throw std::runtime_error("micro must be positive");
    const BatchKey key = batch_key(o.seed, id);
    const auto actual = static_cast<uint32_t>(batch_length(o, id));
    std::vector<uint32_t> failures;
    for (uint32_t offset = 0; offset < actual;) {
        if (s.abort_inflight.load() || pause_requested(o)) {
            s.stop.store(true);
            return false;
        }
        const auto n = static_cast<uint32_t>(std::min<uint64_t>(micro, actual - offset));
        launch(n, offset, key, failures);
        offset += n;
    }
    std::sort(failures.begin(), failures.end());
    commit(id, actual, failures);
    s.completed.fetch_add(actual, std::memory_order_relaxed);
    const auto total = s.failed.fetch_add(failures.size()) + failures.size();
    s.batches.fetch_add(1);
    if (total > o.stop_after) s.stop.store(true);
    return true;
}

}  // namespace ckh
