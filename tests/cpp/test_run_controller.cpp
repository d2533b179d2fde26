// test_run_controller.cpp -- host tests of the cuckoo_bench run controller
// (ck_runspec, ck_ledger): no GPU; a deterministic fake device stands in for
// ck_launch.
//
// Plan:
//   codec      header and record bytes at their offsets, FNV-1a known answers
//   crash      the journal cut at every byte length: the scan keeps exactly
//              the records that end inside it and trims the rest
//   corrupt    every byte of the record area changed in turn: the scan throws
//              or reports only records whose contents are intact
//   identity   journals of another experiment and duplicate commits refused,
//              files left as they are
//   commit     process_batch: pieces, sorting, stop / pause / abort / error
//              boundaries, durable-before-counted, failure threshold
//   resume     a plan with holes finished by three workers with other micro
//              sizes equals one straight run
//   lock       one run per output directory
//   parse      defaults, numbers, every validation rule
//   reports    summary.json key order and values, failures.csv rows
//
//   ck_run_tests <scratch directory>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <set>
#include <sstream>
#include <thread>

#include "ck_ledger.hpp"

using namespace ckrun;

namespace {

int g_checks = 0, g_failures = 0;
#define EXPECT(cond)                                                                     \
    do {                                                                                 \
        ++g_checks;                                                                      \
        if (!(cond)) {                                                                   \
            ++g_failures;                                                                \
            std::fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #cond);     \
        }                                                                                \
    } while (0)

template <class F>
std::string error_of(F&& f) {
    try {
        f();
    } catch (const std::exception& e) {
        return e.what()[0] ? e.what() : "(empty message)";
    }
    return {};
}
bool contains(const std::string& s, const std::string& part) { return s.find(part) != std::string::npos; }

fs::path g_root;

// 16 batches of 16 trials of a small bgh25 shape.
RunSpec small_plan(const std::string& name) {
    RunSpec s;
    s.mode = "bgh25";
    s.pnum = 100;
    s.ell_i = 12;
    s.seed = 77;
    s.et = 8;
    s.bt = 4;
    s.out = (g_root / name).string();
    fs::remove_all(s.out);
    fs::create_directories(s.out);
    return s;
}

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// The fake device's outcome of (batch, offset): about one trial in five fails.
bool fake_fails(uint64_t seed, uint64_t batch, uint32_t offset) {
    return mix64(seed ^ (batch << 32) ^ offset) % 5 == 0;
}
std::vector<uint32_t> expected_failures(const RunSpec& s, uint64_t batch) {
    std::vector<uint32_t> f;
    for (uint32_t o = 0; o < batch_length(s, batch); ++o)
        if (fake_fails(s.seed, batch, o)) f.push_back(o);
    return f;
}

// Stands in for ck_launch: checks the key, appends failures in descending
// order (process_batch must sort), records the pieces it was given.
struct FakeDevice {
    const RunSpec& s;
    uint64_t batch;
    std::vector<std::pair<uint32_t, uint32_t>> pieces;  // (offset, count)
    std::function<void(size_t)> on_piece;               // called with the piece index before it runs
    void operator()(uint32_t count, uint32_t offset, const BatchKey& key, std::vector<uint32_t>& out) {
        if (on_piece) on_piece(pieces.size());
        pieces.emplace_back(offset, count);
        if (key != batch_key_of(s.seed, batch)) throw std::runtime_error("wrong batch key");
        for (uint32_t o = offset + count; o-- > offset;)
            if (fake_fails(s.seed, batch, o)) out.push_back(o);
    }
};

ck_params shape_of(const RunSpec& s) { return derive_shape(s); }

bool run_one(const RunSpec& s, uint64_t batch, uint32_t micro, RunTally& tally, JournalWriter& w) {
    FakeDevice dev{s, batch, {}, {}};
    return process_batch(s, batch, micro, tally, std::ref(dev),
                         [&](uint64_t b, uint32_t n, const std::vector<uint32_t>& f) { w.commit(b, n, f); });
}

std::vector<uint8_t> read_bytes(const fs::path& p) {
    std::ifstream f(p, std::ios::binary);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}
void write_bytes(const fs::path& p, const uint8_t* data, size_t n) {
    std::ofstream f(p, std::ios::binary | std::ios::trunc);
    f.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(n));
}
uint32_t le32(const std::vector<uint8_t>& b, size_t at) {
    return uint32_t(b[at]) | uint32_t(b[at + 1]) << 8 | uint32_t(b[at + 2]) << 16 | uint32_t(b[at + 3]) << 24;
}
uint64_t le64(const std::vector<uint8_t>& b, size_t at) { return le32(b, at) | uint64_t(le32(b, at + 4)) << 32; }

// ------------------------------------------------------------------ codec

void test_codec() {
    JournalIdentity id;
    id.version = 3;
    id.seed = 0x0102030405060708ull;
    id.planned = 1ull << 33;
    id.et = 33;
    id.bt = 20;
    id.pnum = 16384;
    id.ell_i = 16;
    id.mode = 1;
    id.buckets = 24576;
    id.kick = 500;
    id.input_mode = 1;
    auto b = encode_identity(id);
    EXPECT(b.size() == 56);
    EXPECT(b[0] == 0x33 && b[1] == 0x4b && b[2] == 0x55 && b[3] == 0x43);  // "3KUC": 0x43554b33 little-endian
    EXPECT(le32(b, 4) == 3 && le64(b, 8) == id.seed && le64(b, 16) == id.planned);
    EXPECT(le32(b, 24) == 33 && le32(b, 28) == 20 && le32(b, 32) == 16384 && le32(b, 36) == 16);
    EXPECT(le32(b, 40) == 1 && le32(b, 44) == 24576 && le32(b, 48) == 500 && le32(b, 52) == 1);
    EXPECT(identity_mismatch(id, decode_identity(b.data(), b.size())).empty());
    id.version = 4;
    id.k = 4;
    id.stash = 2;
    b = encode_identity(id);
    EXPECT(b.size() == 64 && le32(b, 56) == 4 && le32(b, 60) == 2);
    EXPECT(identity_mismatch(id, decode_identity(b.data(), b.size())).empty());
    JournalIdentity other = id;
    other.stash = 3;
    EXPECT(identity_mismatch(id, other) == "stash");

    // FNV-1a word checksums, computed independently.
    EXPECT(record_checksum(0, 0, nullptr, 0) == 0x4b95f515u);
    const uint32_t offs[3] = {1, 2, 3};
    EXPECT(record_checksum(5, 16, offs, 3) == 0xa7972497u);
    EXPECT(record_checksum(0x0123456789abcdefull, 256, nullptr, 0) == 0xba783273u);
    const auto r = encode_record(5, 16, {1, 2, 3});
    EXPECT(r.size() == 16 + 12 + 8);
    EXPECT(le64(r, 0) == 5 && le32(r, 8) == 16 && le32(r, 12) == 3 && le32(r, 16) == 1 && le32(r, 24) == 3);
    EXPECT(le32(r, 28) == 0xa7972497u && le32(r, 32) == 0x454e4433u);  // "END3"

    // Plan arithmetic.
    RunSpec s;
    s.et = 8;
    s.bt = 4;
    EXPECT(planned_trials(s) == 256 && batch_count(s) == 16 && batch_length(s, 15) == 16 && batch_length(s, 16) == 0);
    s.trials = 1000;
    s.bt = 8;
    EXPECT(batch_count(s) == 4 && batch_length(s, 0) == 256 && batch_length(s, 3) == 232);
}

// ------------------------------------------------------------------ crash consistency

void test_crash_at_every_byte() {
    const RunSpec src = small_plan("crash_src");
    const ck_params p = shape_of(src);
    const std::vector<uint64_t> order = {3, 0, 7, 12};
    std::vector<uint64_t> ends;  // file length after each record
    {
        JournalWriter w(src, p, 0);
        RunTally t;
        for (uint64_t b : order) {
            EXPECT(run_one(src, b, 5, t, w));
            ends.push_back(fs::file_size(journal_path(src, 0)));
        }
    }
    const auto bytes = read_bytes(journal_path(src, 0));
    EXPECT(bytes.size() == ends.back());
    RunSpec s = small_plan("crash_cut");
    for (size_t cut = 0; cut <= bytes.size(); ++cut) {
        fs::remove_all(s.out);
        fs::create_directories(s.out);
        write_bytes(journal_path(s, 0), bytes.data(), cut);
        const LedgerState st = scan_ledger(s, p, true);
        size_t whole = 0;
        while (whole < ends.size() && ends[whole] <= cut) ++whole;
        std::set<uint64_t> want(order.begin(), order.begin() + whole), got;
        for (uint64_t b = 0; b < st.committed.size(); ++b)
            if (st.committed[b]) got.insert(b);
        EXPECT(got == want);
        EXPECT(st.trials == 16 * whole && st.batches == whole);
        uint64_t fails = 0;
        for (size_t i = 0; i < whole; ++i) fails += expected_failures(s, order[i]).size();
        EXPECT(st.failures == fails && st.failing.size() == fails);
        const uint64_t kept = cut < kHeaderBytesV3 ? 0 : whole ? ends[whole - 1] : kHeaderBytesV3;
        EXPECT(fs::file_size(journal_path(s, 0)) == kept);
        if (g_failures) {
            std::fprintf(stderr, "  (journal cut at %zu of %zu bytes)\n", cut, bytes.size());
            return;
        }
    }
    // A journal cut inside its header is restarted by the writer.
    fs::remove_all(s.out);
    fs::create_directories(s.out);
    write_bytes(journal_path(s, 0), bytes.data(), 3);
    scan_ledger(s, p, false);
    {
        JournalWriter w(s, p, 0);
        RunTally t;
        EXPECT(run_one(s, 9, 16, t, w));
    }
    EXPECT(scan_ledger(s, p, false).committed[9] == 1);
}

// ------------------------------------------------------------------ corruption

void test_corruption_never_accepted() {
    const RunSpec src = small_plan("corrupt_src");
    const ck_params p = shape_of(src);
    std::map<uint64_t, std::vector<uint32_t>> truth;
    {
        JournalWriter w(src, p, 0);
        RunTally t;
        for (uint64_t b : {2, 4, 6}) {
            EXPECT(run_one(src, b, 16, t, w));
            truth[b] = expected_failures(src, b);
        }
    }
    const auto bytes = read_bytes(journal_path(src, 0));
    RunSpec s = small_plan("corrupt");
    size_t threw = 0;
    for (size_t at = kHeaderBytesV3; at < bytes.size(); ++at) {
        auto bad = bytes;
        bad[at] ^= 0x5a;
        write_bytes(journal_path(s, 0), bad.data(), bad.size());
        LedgerState st;
        const std::string err = error_of([&] { st = scan_ledger(s, p, true, false); });
        if (!err.empty()) {
            ++threw;
            continue;
        }
        // Accepted: only records whose bytes are intact, each with its true failures.
        std::vector<std::pair<uint64_t, uint32_t>> want;
        for (uint64_t b = 0; b < st.committed.size(); ++b) {
            if (!st.committed[b]) continue;
            EXPECT(truth.count(b) == 1);
            for (uint32_t o : truth[b]) want.emplace_back(b, o);
        }
        EXPECT(st.failing == want);
        EXPECT(read_bytes(journal_path(s, 0)) == bad);  // no repair was asked for
    }
    EXPECT(threw > (bytes.size() - kHeaderBytesV3) / 2);
    // Corruptions that leave a plausible record -- another valid batch id, an
    // offset moved to another free in-range value -- only the checksum catches.
    {
        auto moved = bytes;
        moved[kHeaderBytesV3] = 9;  // record 0 (batch 2) now claims batch 9
        write_bytes(journal_path(s, 0), moved.data(), moved.size());
        EXPECT(contains(error_of([&] { scan_ledger(s, p, true, false); }), "checksum"));
        const auto& f = truth[2];
        EXPECT(!f.empty());
        uint32_t free_value = 0;
        while (std::find(f.begin(), f.end(), free_value) != f.end()) ++free_value;
        std::vector<uint32_t> g = f;
        g[0] = free_value;
        std::sort(g.begin(), g.end());
        moved = bytes;
        for (size_t i = 0; i < g.size(); ++i)
            for (int k = 0; k < 4; ++k)
                moved[kHeaderBytesV3 + 16 + 4 * i + k] = static_cast<uint8_t>(g[i] >> (8 * k));
        EXPECT(moved != bytes);
        write_bytes(journal_path(s, 0), moved.data(), moved.size());
        EXPECT(contains(error_of([&] { scan_ledger(s, p, true, false); }), "checksum"));
    }
    // A complete record with a bad checksum is an error, and the file stays.
    auto bad = bytes;
    bad.back() ^= 1;  // the last record's magic
    write_bytes(journal_path(s, 0), bad.data(), bad.size());
    EXPECT(contains(error_of([&] { scan_ledger(s, p, false); }), "checksum"));
    EXPECT(read_bytes(journal_path(s, 0)) == bad);
}

// ------------------------------------------------------------------ identity and duplicates

void test_identity_and_duplicates() {
    RunSpec s = small_plan("identity");
    const ck_params p = shape_of(s);
    {
        JournalWriter w(s, p, 0);
        RunTally t;
        EXPECT(run_one(s, 1, 16, t, w));
    }
    const auto before = read_bytes(journal_path(s, 0));
    RunSpec other = s;
    other.seed = 78;
    EXPECT(contains(error_of([&] { scan_ledger(other, p, false); }), "seed"));
    other = s;
    other.pnum = 101;
    EXPECT(contains(error_of([&] { scan_ledger(other, shape_of(other), false); }), "pnum"));
    other = s;
    other.stash = 1;  // a version-4 experiment
    EXPECT(contains(error_of([&] { scan_ledger(other, shape_of(other), false); }), "version"));
    EXPECT(read_bytes(journal_path(s, 0)) == before);

    // The same batch committed by two workers.
    {
        JournalWriter w(s, p, 5);
        RunTally t;
        EXPECT(run_one(s, 1, 4, t, w));
    }
    EXPECT(contains(error_of([&] { scan_ledger(s, p, false); }), "committed twice"));

    // Version-4 journals carry k and stash.
    RunSpec e = small_plan("identity_v4");
    e.k = 4;
    e.stash = 2;
    e.trials = 200;
    const ck_params pe = shape_of(e);
    {
        JournalWriter w(e, pe, 0);
        RunTally t;
        EXPECT(run_one(e, 12, 3, t, w));  // the short last batch: 200 - 12 * 16 = 8 trials
    }
    const LedgerState st = scan_ledger(e, pe, true);
    EXPECT(st.trials == 8 && st.committed[12]);
    EXPECT(fs::file_size(journal_path(e, 0)) == kHeaderBytesV4 + 16 + 4 * expected_failures(e, 12).size() + 8);
    RunSpec e2 = e;
    e2.k = 3;
    EXPECT(contains(error_of([&] { scan_ledger(e2, shape_of(e2), false); }), "k"));
}

// ------------------------------------------------------------------ the commit boundary

void test_process_batch() {
    RunSpec s = small_plan("commit");
    const ck_params p = shape_of(s);
    JournalWriter w(s, p, 0);
    const auto commit_to = [&](std::vector<std::vector<uint32_t>>& seen) {
        return [&seen, &w](uint64_t b, uint32_t n, const std::vector<uint32_t>& f) {
            seen.push_back(f);
            w.commit(b, n, f);
        };
    };
    // Pieces of every size cover the batch in order; the commit is sorted.
    uint64_t batch = 0;
    for (uint32_t micro : {1u, 3u, 5u, 15u, 16u, 17u, 1000u}) {
        RunTally t;
        FakeDevice dev{s, batch, {}, {}};
        std::vector<std::vector<uint32_t>> seen;
        EXPECT(process_batch(s, batch, micro, t, std::ref(dev), commit_to(seen)));
        uint32_t next = 0;
        for (const auto& [off, n] : dev.pieces) {
            EXPECT(off == next && n >= 1 && n <= micro);
            next += n;
        }
        EXPECT(next == 16 && dev.pieces.size() == (16 + micro - 1) / micro);
        EXPECT(seen.size() == 1 && seen[0] == expected_failures(s, batch));
        EXPECT(t.trials == 16 && t.batches == 1 && t.failures == seen[0].size());
        ++batch;
    }
    EXPECT(error_of([&] {
               RunTally t;
               FakeDevice dev{s, 0, {}, {}};
               process_batch(s, 0, 0, t, std::ref(dev), [](uint64_t, uint32_t, const std::vector<uint32_t>&) {});
           }) != "");

    // A stop request during a piece: that piece finishes, no further piece
    // runs, nothing is committed or counted.
    {
        StopSignals signals;
        RunTally t;
        FakeDevice dev{s, 9, {}, [](size_t i) {
                           if (i == 0) StopSignals::raise_for_test();
                       }};
        std::vector<std::vector<uint32_t>> seen;
        EXPECT(!process_batch(s, 9, 4, t, std::ref(dev), commit_to(seen)));
        EXPECT(dev.pieces.size() == 1 && seen.empty() && t.trials == 0 && t.batches == 0 && t.halted);
        StopSignals::clear();
        // ... and a request during the last piece still commits the finished batch.
        RunTally t2;
        FakeDevice dev2{s, 9, {}, [](size_t i) {
                            if (i == 3) StopSignals::raise_for_test();
                        }};
        EXPECT(process_batch(s, 9, 4, t2, std::ref(dev2), commit_to(seen)));
        EXPECT(seen.size() == 1 && t2.batches == 1);
        StopSignals::clear();
    }
    // A PAUSE file: no piece at all.  Another worker's abort: the same.
    {
        std::ofstream(fs::path(s.out) / "PAUSE").put('\n');
        RunTally t;
        FakeDevice dev{s, 10, {}, {}};
        std::vector<std::vector<uint32_t>> seen;
        EXPECT(!process_batch(s, 10, 4, t, std::ref(dev), commit_to(seen)));
        EXPECT(dev.pieces.empty() && seen.empty() && t.halted);
        fs::remove(fs::path(s.out) / "PAUSE");
        RunTally a;
        a.aborted = true;
        EXPECT(!process_batch(s, 10, 4, a, std::ref(dev), commit_to(seen)));
        EXPECT(dev.pieces.empty() && seen.empty());
    }
    // A device error propagates and commits nothing.
    {
        RunTally t;
        FakeDevice dev{s, 11, {}, [](size_t i) {
                           if (i == 1) throw std::runtime_error("device fault");
                       }};
        std::vector<std::vector<uint32_t>> seen;
        EXPECT(error_of([&] { process_batch(s, 11, 4, t, std::ref(dev), commit_to(seen)); }) == "device fault");
        EXPECT(seen.empty() && t.trials == 0);
    }
    // A crash after the durable write: the batch is in the journal though the
    // tally never counted it, so a resume does not run it again.
    {
        RunTally t;
        FakeDevice dev{s, 12, {}, {}};
        EXPECT(error_of([&] {
                   process_batch(s, 12, 16, t, std::ref(dev),
                                 [&](uint64_t b, uint32_t n, const std::vector<uint32_t>& f) {
                                     w.commit(b, n, f);
                                     throw std::runtime_error("crash after commit");
                                 });
               }) == "crash after commit");
        EXPECT(t.trials == 0 && scan_ledger(s, p, false).committed[12] == 1);
    }
    // More committed failures than --stop-after halt the run.
    {
        RunSpec low = s;
        low.stop_after = 1;
        RunTally t;
        uint64_t b = 13;
        while (expected_failures(low, b).size() < 2) ++b;
        FakeDevice dev{low, b, {}, {}};
        EXPECT(b < 16 && process_batch(low, b, 16, t, std::ref(dev),
                                        [](uint64_t, uint32_t, const std::vector<uint32_t>&) {}));
        EXPECT(t.halted);
    }
}

// ------------------------------------------------------------------ resume with other workers

void test_resume_with_holes() {
    const RunSpec straight = small_plan("straight");
    const ck_params p = shape_of(straight);
    {
        JournalWriter w(straight, p, 0);
        RunTally t;
        for (uint64_t b = 0; b < 16; ++b) EXPECT(run_one(straight, b, 16, t, w));
    }
    const LedgerState want = scan_ledger(straight, p, true);

    RunSpec s = small_plan("holes");
    {
        JournalWriter w(s, p, 3);
        RunTally t;
        for (uint64_t b : {13, 1, 9, 5}) EXPECT(run_one(s, b, 2, t, w));
    }
    LedgerState mid = scan_ledger(s, p, false);
    EXPECT(mid.batches == 4 && mid.trials == 64);
    RunTally tally;
    tally.trials = mid.trials;
    tally.batches = mid.batches;
    std::vector<std::thread> workers;
    std::vector<std::string> errors(3);
    const uint32_t micros[3] = {1, 7, 16};
    for (int i = 0; i < 3; ++i)
        workers.emplace_back([&, i] {
            try {
                JournalWriter w(s, p, i);
                for (;;) {
                    const uint64_t b = tally.cursor.fetch_add(1);
                    if (b >= 16) break;
                    if (mid.committed[b]) continue;
                    run_one(s, b, micros[i], tally, w);
                }
            } catch (const std::exception& e) {
                errors[i] = e.what();
            }
        });
    for (auto& t : workers) t.join();
    for (const auto& e : errors) EXPECT(e.empty());
    const LedgerState got = scan_ledger(s, p, true);
    EXPECT(got.batches == 16 && got.trials == 256 && tally.trials == 256);
    EXPECT(got.failing == want.failing && got.failures == want.failures);
    EXPECT(journal_files(s.out).size() == 4);
}

// ------------------------------------------------------------------ lock

void test_directory_lock() {
    const RunSpec s = small_plan("lock");
    {
        DirectoryLock first(s.out);
        EXPECT(contains(error_of([&] { DirectoryLock second(s.out); }), "in use"));
    }
    EXPECT(error_of([&] { DirectoryLock again(s.out); }).empty());
}

// ------------------------------------------------------------------ parse

RunSpec parse(std::vector<std::string> args) {
    args.insert(args.begin(), "cuckoo_bench");
    std::vector<char*> argv;
    for (auto& a : args) argv.push_back(a.data());
    return parse_command_line(static_cast<int>(argv.size()), argv.data());
}

void test_parse() {
    const RunSpec d = parse({});
    EXPECT(d.et == 30 && d.bt == 20 && d.pnum == 16 && d.ell_i == 32 && d.micro == (1u << 17));
    EXPECT(d.seed == 1 && d.stop_after == (1ull << 20) && d.mode == "drrt18" && d.out == "cuckoo_results");
    EXPECT(d.input_mode == "sequential" && d.workers == 0 && d.scratch_mib == 256 && !d.scratch_given);
    EXPECT(d.block_threads == 256 && d.warps_per_block == 0 && d.prp_cache == "auto");
    EXPECT(d.table_layout == "auto" && d.shared_variant == "auto" && d.k == 3 && d.stash == 0);
    EXPECT(d.place_search == 32 && parse({"--place-search", "0"}).place_search == 0);
    const RunSpec x = parse({"--pnum", "0x40", "--seed", "18446744073709551615", "--scratch-mib", "2048",
                             "--status", "--global-table", "--workers", "3"});
    EXPECT(x.pnum == 64 && x.seed == ~0ull && x.scratch_mib == 2048 && x.scratch_given && x.status_only);
    EXPECT(x.global_table && x.workers == 3);
    for (const auto& bad : std::vector<std::vector<std::string>>{
             {"--pnum", "-1"}, {"--pnum", "12x"}, {"--pnum", ""}, {"--micro", "4294967296"}, {"--workers", "257"},
             {"--nope", "1"}, {"--seed"}}) {
        EXPECT(!error_of([&] { parse(bad); }).empty());
    }
    const std::pair<std::vector<std::string>, std::string> rules[] = {
        {{"--et", "41"}, "--et"},
        {{"--bt", "0"}, "--bt is 1..30"},
        {{"--et", "8", "--bt", "9"}, "at most --et"},
        {{"--pnum", "2"}, "--pnum"},
        {{"--ell-i", "62"}, "--ell-i"},
        {{"--ell-i", "4", "--pnum", "17"}, "distinct"},
        {{"--k", "5"}, "--k"},
        {{"--micro", "0"}, "--micro"},
        {{"--scratch-mib", "0"}, "--scratch-mib"},
        {{"--block-threads", "100"}, "--block-threads"},
        {{"--warps-per-block", "33"}, "--warps-per-block"},
        {{"--mode", "x"}, "--mode"},
        {{"--input-mode", "x"}, "--input-mode"},
        {{"--prp-cache", "x"}, "--prp-cache"},
        {{"--table-layout", "x"}, "--table-layout"},
        {{"--shared-variant", "x"}, "--shared-variant"},
        {{"--tier", "x"}, "--tier"},
        {{"--place-search", "1"}, "--place-search"},
        {{"--place-search", "65"}, "--place-search"},
        {{"--table-layout", "shared", "--global-table"}, "exclude"},
        {{"--table-layout", "shared", "--prp-cache", "off"}, "round cache"},
        {{"--trials", "1099511627777"}, "--trials"},
    };
    for (const auto& [args, why] : rules) {
        const std::string err = error_of([&] { parse(args); });
        EXPECT(contains(err, "invalid configuration") && contains(err, why));
    }
    // The extension count relaxes the ET rules.
    const RunSpec t = parse({"--trials", "1000", "--et", "41", "--bt", "8"});
    EXPECT(t.trials == 1000 && planned_trials(t) == 1000);
}

// ------------------------------------------------------------------ reports

std::vector<std::string> json_keys(const std::string& text) {
    std::vector<std::string> keys;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        const auto a = line.find('"');
        if (a == std::string::npos) continue;
        const auto b = line.find('"', a + 1);
        if (line.compare(b + 1, 1, ":") == 0) keys.push_back(line.substr(a + 1, b - a - 1));
    }
    return keys;
}

void test_reports() {
    RunSpec s = small_plan("reports");
    const ck_params p = shape_of(s);
    {
        JournalWriter w(s, p, 1);
        RunTally t;
        for (uint64_t b : {7, 2, 11}) EXPECT(run_one(s, b, 3, t, w));
    }
    const LedgerState st = scan_ledger(s, p, true);
    RunReport r;
    r.status = "paused";
    r.workers = 2;
    r.round_cache = true;
    r.placement = "cached-global-desc-plain";
    r.elapsed_s = 1.5;
    r.work_s = 1.25;
    write_reports(s, p, st, r);
    std::ifstream js(fs::path(s.out) / "summary.json");
    const std::string text((std::istreambuf_iterator<char>(js)), std::istreambuf_iterator<char>());
    // The reference's summary keys, in its order, then this build's.
    const std::vector<std::string> want = {
        "status", "seed_scheme", "mode", "algorithm", "backend", "workers", "micro_requested", "scratch_mib",
        "prp_cache_requested", "prp_cache_enabled", "table_layout_requested", "table_layout_planned",
        "shared_variant_requested", "block_threads_requested", "warps_per_block_requested", "input_mode", "seed", "ET",
        "BT", "seed_decimal", "batch_size", "PNUM_T", "ELL_I", "BUCKETS_NUM", "MAX_KICK_TIMES", "BUCKET_SIZE",
        "PRP_DOMAIN", "planned", "completed", "failed", "completed_batches", "pending_batches", "failed_ratio",
        "elapsed_seconds", "work_elapsed_seconds", "threshold_exceeded", "K", "STASH", "journal_version"};
    EXPECT(json_keys(text) == want);
    EXPECT(contains(text, "\"status\": \"paused\"") && contains(text, "\"completed\": 48"));
    EXPECT(contains(text, "\"pending_batches\": 13") && contains(text, "\"prp_cache_enabled\": true"));
    EXPECT(contains(text, "\"seed_decimal\": \"77\"") && contains(text, "\"journal_version\": 3"));
    std::ifstream csv(fs::path(s.out) / "failures.csv");
    std::string line;
    std::getline(csv, line);
    EXPECT(line == "batch_id,offset");
    std::vector<std::pair<uint64_t, uint32_t>> rows;
    while (std::getline(csv, line)) {
        unsigned long long b = 0;
        unsigned o = 0;
        EXPECT(std::sscanf(line.c_str(), "%llu,%u", &b, &o) == 2);
        rows.emplace_back(b, o);
    }
    std::vector<std::pair<uint64_t, uint32_t>> want_rows;
    for (uint64_t b : {2, 7, 11})
        for (uint32_t o : expected_failures(s, b)) want_rows.emplace_back(b, o);
    EXPECT(rows == want_rows);
}

}  // namespace

int main(int argc, char** argv) {
    g_root = fs::absolute(argc > 1 ? argv[1] : "run_controller_test_data");
    fs::create_directories(g_root);
    const std::pair<const char*, void (*)()> tests[] = {
        {"codec", test_codec},
        {"crash", test_crash_at_every_byte},
        {"corrupt", test_corruption_never_accepted},
        {"identity", test_identity_and_duplicates},
        {"commit", test_process_batch},
        {"resume", test_resume_with_holes},
        {"lock", test_directory_lock},
        {"parse", test_parse},
        {"reports", test_reports},
    };
    for (const auto& [name, fn] : tests) {
        const int before = g_failures;
        const std::string err = error_of(fn);
        if (!err.empty()) {
            ++g_failures;
            std::fprintf(stderr, "%s: unexpected exception: %s\n", name, err.c_str());
        }
        std::printf("%s %s\n", g_failures == before ? "PASS" : "FAIL", name);
    }
    std::printf("%d checks, %d failed\n", g_checks, g_failures);
    return g_failures ? 1 : 0;
}
