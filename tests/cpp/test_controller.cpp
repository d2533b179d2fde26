// Host tests of the run controller (process_batch, journal, resume, export,
// pause, lock), driven by a mock launch so no GPU is needed.  The mock fails
// offset j of batch id when (id*8+j) % 7 == 0, and checks that the batch key
// it receives does not depend on scheduling.
#include <array>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <set>
#include <thread>

#include "ck_journal.hpp"

using namespace ckh;

static int g_checks = 0;
#define REQUIRE(cond, what)                                                   \
    do {                                                                      \
        ++g_checks;                                                           \
        if (!(cond)) throw std::runtime_error(std::string("FAILED: ") + what); \
    } while (0)

template <class F>
static void must_throw(F f, const char* what) {
    bool threw = false;
    try {
        f();
    } catch (const std::exception&) {
        threw = true;
    }
    REQUIRE(threw, what);
}

using Journal = std::unique_ptr<FILE, decltype(&fclose)>;
static Journal journal(const Options& o, const ck_params& p, int worker = 0) {
    return Journal(open_journal(o, p, worker), fclose);
}

static void mock_launch(uint64_t id, uint32_t n, uint32_t offset, const BatchKey& key, std::vector<uint32_t>& failures) {
    REQUIRE(key == batch_key(123, id), "batch key changed with scheduling");
    for (uint32_t j = offset; j < offset + n; ++j)
        if ((id * 8 + j) % 7 == 0) failures.push_back(j);
}
static bool run_batch(const Options& o, uint64_t id, uint32_t micro, Counters& s, FILE* f) {
    return process_batch(
        o, id, micro, s, [&](uint32_t n, uint32_t off, const BatchKey& key, std::vector<uint32_t>& fl) { mock_launch(id, n, off, key, fl); },
        [&](uint64_t b, uint32_t c, const std::vector<uint32_t>& fl) { append_record(f, b, c, fl); });
}
static Options make_options(const fs::path& root, const char* name) {
    Options o;
    o.et = 6;
    o.bt = 3;
    o.seed = 123;
    o.pnum = 16;
    o.ell_i = 8;
    o.out = (root / name).string();
    fs::create_directories(o.out);
    return o;
}
static std::set<std::string> csv_rows(const Options& o) {
    std::ifstream file(fs::path(o.out) / "failures.csv");
    std::string line;
    std::getline(file, line);
    std::set<std::string> rows;
    while (std::getline(file, line)) rows.insert(line);
    return rows;
}

static void check_keys() {
    const uint32_t want[2][4] = {{0x992dc0b2, 0xecc207f0, 0x5f9a7017, 0x1fb84a2f},
                                 {0x58f85b99, 0xfc3cbd02, 0x022bf19c, 0x10dad53d}};
    const uint64_t ids[2] = {0, 7};
    for (int i = 0; i < 2; ++i) {
        const BatchKey k = batch_key(123, ids[i]);
        const uint32_t zero[4] = {0, 0, 0, 0};
        uint32_t out[4];
        ck_host_aes(k.data(), zero, out);
        REQUIRE(std::memcmp(out, want[i], sizeof out) == 0, "batch-key golden vector");
    }
}

static void check_pause(const fs::path& root) {
    auto o = make_options(root, "pause");
    const auto p = parameters(o);
    auto f = journal(o, p);
    Counters s;
    PauseSignals signals;
    unsigned launches = 0;
    bool ok = process_batch(
        o, 0, 2, s,
        [&](uint32_t n, uint32_t off, const BatchKey& key, std::vector<uint32_t>& fl) {
            mock_launch(0, n, off, key, fl);
            ++launches;
            std::raise(SIGINT);
        },
        [&](uint64_t b, uint32_t c, const std::vector<uint32_t>& fl) { append_record(f.get(), b, c, fl); });
    REQUIRE(!ok && launches == 1 && s.completed == 0 && s.failed == 0, "partial batch was counted");
    f.reset();
    Counters r;
    auto done = resume(o, p, r);
    REQUIRE(!done[0] && r.completed == 0, "partial batch was persisted");
    pause_signal.store(false);
    s.stop.store(false);
    f = journal(o, p);
    REQUIRE(run_batch(o, 0, 3, s, f.get()), "unfinished batch did not restart");
    REQUIRE(s.completed == 8 && s.failed == 2, "restart counts");
    f.reset();
    Counters saved;
    done = resume(o, p, saved);
    REQUIRE(done[0] && saved.completed == 8, "completed batch not recovered");

    auto full = make_options(root, "pause_after_last_micro");
    const auto fp = parameters(full);
    auto ff = journal(full, fp);
    Counters fs_;
    ok = process_batch(
        full, 0, 8, fs_,
        [&](uint32_t n, uint32_t off, const BatchKey& key, std::vector<uint32_t>& fl) {
            mock_launch(0, n, off, key, fl);
            pause_signal.store(true);
        },
        [&](uint64_t b, uint32_t c, const std::vector<uint32_t>& fl) { append_record(ff.get(), b, c, fl); });
    REQUIRE(ok && fs_.completed == 8, "a fully finished batch was discarded on pause");
    pause_signal.store(false);

    auto fc = make_options(root, "pause_file");
    const auto fcp = parameters(fc);
    std::ofstream(fs::path(fc.out) / "PAUSE").close();
    Counters fcs;
    auto fcf = journal(fc, fcp);
    REQUIRE(!run_batch(fc, 0, 2, fcs, fcf.get()) && fcs.completed == 0, "PAUSE file ignored");
    fs::remove(fs::path(fc.out) / "PAUSE");

    auto ab = make_options(root, "worker_error");
    const auto abp = parameters(ab);
    Counters ac;
    {
        auto af = journal(ab, abp);
        unsigned calls = 0;
        must_throw(
            [&] {
                process_batch(
                    ab, 0, 2, ac,
                    [&](uint32_t n, uint32_t off, const BatchKey& key, std::vector<uint32_t>& fl) {
                        if (++calls == 2) throw std::runtime_error("mock device error");
                        mock_launch(0, n, off, key, fl);
                    },
                    [&](uint64_t b, uint32_t c, const std::vector<uint32_t>& fl) { append_record(af.get(), b, c, fl); });
            },
            "device error was ignored");
        ac.abort_inflight.store(true);
        REQUIRE(!run_batch(ab, 1, 2, ac, af.get()), "another worker's abort was ignored");
    }
    Counters ar;
    resume(ab, abp, ar);
    REQUIRE(ar.completed == 0, "failed batch was committed");
}

static void check_journals(const fs::path& root) {
    auto o = make_options(root, "holes");
    const auto p = parameters(o);
    Counters s;
    {
        auto f = journal(o, p, 3);
        run_batch(o, 6, 2, s, f.get());
        run_batch(o, 2, 2, s, f.get());
    }
    const auto path = journal_name(o, 3);
    const auto safe = fs::file_size(path);
    {
        std::ofstream f(path, std::ios::binary | std::ios::app);
        RecordHeader h{1, 8, 1};
        f.write(reinterpret_cast<char*>(&h), sizeof h);
        f.write("x", 1);
    }
    Counters restored;
    auto done = resume(o, p, restored);
    REQUIRE(done[6] && done[2] && !done[1] && restored.completed == 16, "out-of-order restore");
    REQUIRE(fs::file_size(path) == safe, "torn tail was not removed");
    std::vector<std::thread> workers;
    std::array<std::exception_ptr, 3> errors{};
    for (int w = 0; w < 3; ++w)
        workers.emplace_back([&, w] {
            try {
                auto f = journal(o, p, w);
                for (;;) {
                    const auto id = restored.next_batch.fetch_add(1);
                    if (id >= done.size()) break;
                    if (done[id]) continue;
                    run_batch(o, id, 5, restored, f.get());
                }
            } catch (...) {
                errors[w] = std::current_exception();
            }
        });
    for (auto& w : workers) w.join();
    for (auto& e : errors)
        if (e) std::rethrow_exception(e);
    Counters recovered;
    done = resume(o, p, recovered);
    REQUIRE(recovered.completed == 64 && recovered.batches == 8 && recovered.failed == 10,
            "resume double-counted or omitted work");
    final_outputs(o, p, recovered, 0, "complete");
    auto one = make_options(root, "one_shot");
    const auto onep = parameters(one);
    Counters ones;
    {
        auto f = journal(one, onep);
        for (uint64_t id = 0; id < 8; ++id) run_batch(one, id, 1, ones, f.get());
    }
    final_outputs(one, onep, ones, 0, "complete");
    REQUIRE(csv_rows(o) == csv_rows(one) && csv_rows(o).size() == 10, "resumed run differs from uninterrupted run");

    auto torn = make_options(root, "torn_file_header");
    const auto tp = parameters(torn);
    {
        std::ofstream f(journal_name(torn, 0), std::ios::binary);
        f.write("CUK", 3);
    }
    Counters tc;
    resume(torn, tp, tc);
    {
        auto f = journal(torn, tp);
        REQUIRE(run_batch(torn, 0, 2, tc, f.get()), "header repair");
    }
    Counters tr;
    resume(torn, tp, tr);
    REQUIRE(tr.completed == 8, "repaired header not usable");

    auto crash = make_options(root, "commit_before_counter_crash");
    const auto cp = parameters(crash);
    Counters cs;
    {
        auto f = journal(crash, cp);
        must_throw(
            [&] {
                process_batch(
                    crash, 0, 4, cs,
                    [&](uint32_t n, uint32_t off, const BatchKey& key, std::vector<uint32_t>& fl) { mock_launch(0, n, off, key, fl); },
                    [&](uint64_t b, uint32_t c, const std::vector<uint32_t>& fl) {
                        append_record(f.get(), b, c, fl);
                        throw std::runtime_error("crash");
                    });
            },
            "mock crash did not throw");
    }
    REQUIRE(cs.completed == 0, "crash published counters");
    Counters cr;
    resume(crash, cp, cr);
    REQUIRE(cr.completed == 8, "durable commit lost before the counter update");

    auto corrupt = make_options(root, "corrupt");
    const auto xp = parameters(corrupt);
    Counters xc;
    {
        auto f = journal(corrupt, xp);
        run_batch(corrupt, 0, 8, xc, f.get());
    }
    const auto xfile = journal_name(corrupt, 0);
    const auto xsize = fs::file_size(xfile);
    {
        std::fstream f(xfile, std::ios::binary | std::ios::in | std::ios::out);
        f.seekp(-1, std::ios::end);
        f.put('\0');
    }
    Counters xr;
    must_throw([&] { resume(corrupt, xp, xr); }, "corrupt complete record accepted");
    REQUIRE(fs::file_size(xfile) == xsize, "corrupt record silently truncated");

    auto dup = make_options(root, "duplicate");
    const auto dp = parameters(dup);
    Counters dc;
    {
        auto f = journal(dup, dp, 0);
        run_batch(dup, 0, 8, dc, f.get());
    }
    {
        auto f = journal(dup, dp, 1);
        run_batch(dup, 0, 8, dc, f.get());
    }
    Counters dr;
    must_throw([&] { resume(dup, dp, dr); }, "duplicate batch accepted");
    auto different = o;
    different.seed = 124;
    Counters mm;
    must_throw([&] { resume(different, p, mm); }, "mismatched seed accepted");
}

static void check_extension_journal(const fs::path& root) {
    // Non-power-of-two plan: 100 trials in batches of 8, the last one short.
    auto o = make_options(root, "trials_mode");
    o.trials = 100;
    const auto p = parameters(o);
    REQUIRE(total_batches(o) == 13 && batch_length(o, 12) == 4 && batch_length(o, 11) == 8, "short last batch");
    const auto h = create_header(o, p);
    REQUIRE(h.v4 && h.size() == 64, "extension uses the v4 header");
    Counters s;
    {
        auto f = journal(o, p);
        for (uint64_t id = 0; id < 13; ++id) run_batch(o, id, 3, s, f.get());
    }
    REQUIRE(fs::file_size(journal_name(o, 0)) >= 64, "v4 header written");
    Counters r;
    resume(o, p, r);
    REQUIRE(r.completed == 100 && r.batches == 13, "v4 resume");
    // The same directory under a different k is a different experiment.
    auto k2 = o;
    k2.k = 2;
    Counters m;
    must_throw([&] { resume(k2, parameters(k2), m); }, "k mismatch accepted");
    // A reference-shaped experiment keeps the 56-byte v3 header.
    auto v3 = make_options(root, "v3_shape");
    REQUIRE(!create_header(v3, parameters(v3)).v4 && create_header(v3, parameters(v3)).size() == 56, "v3 header");
}

static void check_locks_and_parameters(const fs::path& root) {
    auto o = make_options(root, "locks");
    {
        RunLock lock(o.out);
        must_throw([&] { RunLock another(o.out); }, "concurrent directory lock accepted");
    }
    { RunLock lock(o.out); }
    must_throw([] { parse_u64("-1"); }, "negative unsigned argument accepted");
    must_throw([] { parse_u32("4294967296"); }, "overflowing argument accepted");
    {
        char name[] = "bench", flag[] = "--scratch-mib", value[] = "2048";
        char* argv[] = {name, flag, value};
        REQUIRE(parse(3, argv).scratch_mib == 2048, "budget argument ignored");
        char zero[] = "0";
        argv[2] = zero;
        must_throw([&] { parse(3, argv); }, "zero budget accepted");
    }
    {
        char name[] = "bench", f1[] = "--et", v1[] = "41";
        char* argv[] = {name, f1, v1};
        must_throw([&] { parse(3, argv); }, "ET 41 accepted");
    }
    auto p = parameters(o);
    REQUIRE(p.buckets == 23, "DRRT18 bucket derivation changed");
    o.mode = "bgh25";
    REQUIRE(parameters(o).buckets == 32, "BGH25 bucket derivation changed");
}

int main(int argc, char** argv) {
    try {
        const fs::path base = fs::absolute(argc > 1 ? argv[1] : "controller_test_data");
        const fs::path root = base / ("run_" + std::to_string(steady_ns()));
        fs::create_directories(root);
        check_keys();
        check_pause(root);
        check_journals(root);
        check_extension_journal(root);
        check_locks_and_parameters(root);
        std::cout << "PASS controller tests (" << g_checks << " checks)\n";
        return 0;
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 1;
    }
}
