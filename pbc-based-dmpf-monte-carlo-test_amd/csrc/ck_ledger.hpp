// ck_ledger.hpp -- the run's durable record: one append-only journal file per
// worker (<out>/worker_<n>.bin), the scan that rebuilds a run's progress from
// them, and the summary.json / failures.csv export.
//
// The byte format is the reference's journal v3 (cuckoo_bench.cu:565-581), so
// either program resumes the other's journals:
//   file header, 56 bytes, little-endian:
//     u32 magic 0x43554b33 ("CUK3")  u32 version
//     u64 seed  u64 planned trials
//     u32 et, bt, pnum, ell_i, mode (1 bgh25, 2 drrt18), buckets, kick, input_mode (0 sequential, 1 affine)
//   version 4 (this build's extension: k != 3, stash != 0 or a non-power-of-two
//   plan, with et = 0) appends u32 k, u32 stash: 64 bytes.
//   record: u64 batch, u32 trials, u32 count, u32 offsets[count] (ascending,
//     each < trials), u32 checksum, u32 magic 0x454e4433 ("END3").
//   checksum: FNV-1a over the u32 words batch lo, batch hi, trials, count,
//     offsets... (offset basis 0x811c9dc5, prime 0x01000193).
// A record counts once its last byte is on disk; a shorter tail is a write
// that a crash interrupted and is cut off on the next scan.
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "ck_runspec.hpp"

namespace ckrun {

constexpr uint32_t kJournalMagic = 0x43554b33u;
constexpr uint32_t kRecordMagic = 0x454e4433u;
constexpr size_t kHeaderBytesV3 = 56, kHeaderBytesV4 = 64;
constexpr size_t kRecordHeadBytes = 16, kRecordTailBytes = 8;

// The experiment a journal belongs to: the decoded file header.
struct JournalIdentity {
    uint32_t version = 3;
    uint64_t seed = 0, planned = 0;
    uint32_t et = 0, bt = 0, pnum = 0, ell_i = 0, mode = 0, buckets = 0, kick = 0, input_mode = 0;
    uint32_t k = 3, stash = 0;  // version 4 only
    size_t bytes() const { return version == 4 ? kHeaderBytesV4 : kHeaderBytesV3; }
};
JournalIdentity identity_of(const RunSpec& s, const ck_params& p);
std::vector<uint8_t> encode_identity(const JournalIdentity& id);
// Decodes `size` bytes (at least kHeaderBytesV3; the v4 fields when present).
JournalIdentity decode_identity(const uint8_t* bytes, size_t size);
// Empty when equal, else the name of the first field that differs.
std::string identity_mismatch(const JournalIdentity& want, const JournalIdentity& got);

uint32_t record_checksum(uint64_t batch, uint32_t trials, const uint32_t* offsets, size_t count);
// One complete record's bytes.
std::vector<uint8_t> encode_record(uint64_t batch, uint32_t trials, const std::vector<uint32_t>& offsets);

fs::path journal_path(const RunSpec& s, int worker);
// worker_<digits>.bin files of a directory, by name.
std::vector<fs::path> journal_files(const fs::path& directory);

// What a directory's journals hold.
struct LedgerState {
    std::vector<uint8_t> committed;  // per batch of the plan: 1 once a complete record exists
    uint64_t batches = 0, trials = 0, failures = 0;
    std::vector<std::pair<uint64_t, uint32_t>> failing;  // (batch, offset), ascending, when kept
};

// Reads every journal of the run's directory.  Throws when a journal belongs
// to another experiment, when a complete record is corrupt or does not fit the
// plan, or when two records commit the same batch; files are left as they are
// then.  Otherwise, with `repair`, cuts each file back to its last complete
// record (a file shorter than a header becomes empty).
LedgerState scan_ledger(const RunSpec& s, const ck_params& p, bool keep_failures, bool repair = true);

// Appends records to one worker's journal: one write of the whole record,
// then fdatasync, before commit() returns.  Creates the file with its header
// when it is missing or empty (scan_ledger has emptied any torn header).
class JournalWriter {
   public:
    JournalWriter(const RunSpec& s, const ck_params& p, int worker);
    ~JournalWriter();
    JournalWriter(const JournalWriter&) = delete;
    JournalWriter& operator=(const JournalWriter&) = delete;
    // offsets ascending, unique and below trials (process_batch sorts them).
    void commit(uint64_t batch, uint32_t trials, const std::vector<uint32_t>& offsets);

   private:
    std::string path_;
    int fd_ = -1;
};

// How the run went and how it was placed, for summary.json.
struct RunReport {
    std::string status = "ready";  // complete | paused | threshold | error | ready
    std::string backend = "hip";
    int workers = 0;
    bool round_cache = false;                    // whether bucket evaluations used a round cache
    std::string placement = "auto";              // the planned table placement
    double elapsed_s = 0, work_s = 0;
};
// summary.json (the reference's keys in its order, cuckoo_bench.cu:753-776,
// then this build's) and failures.csv (batch_id,offset rows, ascending) from a
// scan that kept the failures.
void write_reports(const RunSpec& s, const ck_params& p, const LedgerState& state, const RunReport& r);

}  // namespace ckrun
