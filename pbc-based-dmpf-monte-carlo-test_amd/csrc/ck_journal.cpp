// ck_journal.cpp -- run controller: options, experiment identity, journal,
// resume, pause, run lock and the summary/CSV export.
#include "ck_journal.hpp"

#include <fcntl.h>
#include <sys/file.h>
#include <unistd.h>

#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>

namespace ckh {

std::atomic<bool> pause_signal{false};
static_assert(std::atomic<bool>::is_always_lock_free, "signal handler needs a lock-free flag");
static void request_pause(int) { pause_signal.store(true, std::memory_order_relaxed); }

PauseSignals::PauseSignals() {
    pause_signal.store(false);
    old_int_ = std::signal(SIGINT, request_pause);
    if (old_int_ == SIG_ERR) throw std::runtime_error("Cannot install SIGINT handler");
    old_term_ = std::signal(SIGTERM, request_pause);
    if (old_term_ == SIG_ERR) {
        std::signal(SIGINT, old_int_);
        throw std::runtime_error("Cannot install SIGTERM handler");
    }
}
PauseSignals::~PauseSignals() {
    std::signal(SIGINT, old_int_);
    std::signal(SIGTERM, old_term_);
}

bool pause_requested(const Options& o) {
    return pause_signal.load(std::memory_order_relaxed) || fs::exists(fs::path(o.out) / "PAUSE");
}

RunLock::RunLock(const fs::path& directory) {
    const auto path = (directory / ".run.lock").string();
    fd_ = ::open(path.c_str(), O_CREAT | O_RDWR, 0600);
    if (fd_ < 0) throw std::runtime_error("Cannot open run lock: " + path);
    if (::flock(fd_, LOCK_EX | LOCK_NB) != 0) {
        ::close(fd_);
        fd_ = -1;
        throw std::runtime_error("Another process is using the output directory: " + path);
    }
}
RunLock::~RunLock() {
    if (fd_ < 0) return;
    ::flock(fd_, LOCK_UN);
    ::close(fd_);
}

void durable_flush(FILE* f) {
    if (std::fflush(f) != 0) throw std::runtime_error("Journal flush failed");
    if (::fsync(::fileno(f)) != 0) throw std::runtime_error("Journal disk sync failed");
}

BatchKey batch_key(uint64_t seed, uint64_t id) {
    BatchKey k{};
    ck_batch_key(seed, id, k.data());
    return k;
}

// ------------------------------------------------------------------ options

uint64_t parse_u64(const std::string& s) {
    if (s.empty() || s[0] == '-') throw std::runtime_error("Invalid unsigned number: " + s);
    size_t at = 0;
    unsigned long long v = 0;
    try {
        v = std::stoull(s, &at, 0);
    } catch (const std::exception&) {
        throw std::runtime_error("Invalid number: " + s);
    }
    if (at != s.size()) throw std::runtime_error("Invalid number: " + s);
    return v;
}
uint32_t parse_u32(const std::string& s) {
    const auto v = parse_u64(s);
    if (v > std::numeric_limits<uint32_t>::max()) throw std::runtime_error("Number exceeds uint32: " + s);
    return static_cast<uint32_t>(v);
}

static const char* kUsage =
    "Usage: cuckoo_bench --mode bgh25|drrt18 --et 40 --bt 20 --pnum 16 --seed 123 --workers 1 --micro 131072\n"
    "       --out results [--input-mode sequential|affine] [--ell-i 32] [--scratch-mib 256]\n"
    "       [--prp-cache auto|on|off] [--block-threads 256] [--table-layout auto|global|shared]\n"
    "       [--shared-variant auto|round|map] [--warps-per-block N] [--stop-after 1048576]\n"
    "       [--global-table] [--status]\n"
    "  extension: [--k 2..4] [--buckets M] [--kick N] [--stash S] [--trials N] [--tier auto|hbm|lds] [--tune] [--graph]\n";

// Flags and validation of cuckoo_bench.cu:431-480; layout flags are accepted
// for compatibility (the tier is chosen by --tier / the library).
Options parse(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--help" || a == "-h") {
            std::cout << kUsage;
            std::exit(0);
        }
        if (a == "--global-table") {
            o.force_global = true;
            continue;
        }
        if (a == "--status") {
            o.status_only = true;
            continue;
        }
        if (a == "--tune") {
            o.tune = true;
            continue;
        }
        if (a == "--graph") {
            o.graph = true;
            continue;
        }
        if (i + 1 >= argc) throw std::runtime_error("Missing value after " + a);
        const std::string v = argv[++i];
        if (a == "--et") o.et = parse_u32(v);
        else if (a == "--bt") o.bt = parse_u32(v);
        else if (a == "--pnum") o.pnum = parse_u32(v);
        else if (a == "--ell-i") o.ell_i = parse_u32(v);
        else if (a == "--micro") o.micro = parse_u32(v);
        else if (a == "--scratch-mib") o.scratch_mib = parse_u32(v);
        else if (a == "--prp-cache") o.prp_cache = v;
        else if (a == "--block-threads") o.block_threads = parse_u32(v);
        else if (a == "--table-layout") o.table_layout = v;
        else if (a == "--shared-variant") o.shared_variant = v;
        else if (a == "--warps-per-block") o.warps_per_block = parse_u32(v);
        else if (a == "--seed") o.seed = parse_u64(v);
        else if (a == "--workers") {
            const auto n = parse_u64(v);
            if (n > 256) throw std::runtime_error("workers exceeds 256");
            o.workers = static_cast<int>(n);
        } else if (a == "--stop-after") o.stop_after = parse_u64(v);
        else if (a == "--mode") o.mode = v;
        else if (a == "--input-mode") o.input_mode = v;
        else if (a == "--out") o.out = v;
        else if (a == "--k") o.k = parse_u32(v);
        else if (a == "--buckets") o.buckets = parse_u32(v);
        else if (a == "--kick") o.kick = parse_u32(v);
        else if (a == "--stash") o.stash = parse_u32(v);
        else if (a == "--trials") o.trials = parse_u64(v);
        else if (a == "--tier") o.tier = v;
        else throw std::runtime_error("Unknown option " + a);
    }
    const bool count_ok = o.trials ? (o.trials <= (1ULL << 40) && o.bt >= 1 && o.bt <= 30)
                                   : (o.et >= 1 && o.et <= 40 && o.bt >= 1 && o.bt <= 30 && o.bt <= o.et);
    if (!count_ok || o.pnum < 3 || o.micro == 0 || o.scratch_mib == 0 || o.ell_i > 61 ||
        (o.pnum > 65534 && o.pnum > (1u << 30)) ||
        (o.block_threads != 64 && o.block_threads != 128 && o.block_threads != 256) ||
        (o.prp_cache != "auto" && o.prp_cache != "on" && o.prp_cache != "off") ||
        (o.table_layout != "auto" && o.table_layout != "global" && o.table_layout != "shared") ||
        (o.shared_variant != "auto" && o.shared_variant != "round" && o.shared_variant != "map") ||
        o.warps_per_block > 32 || (o.ell_i < 32 && (1ULL << o.ell_i) < o.pnum) ||
        (o.mode != "bgh25" && o.mode != "drrt18") ||
        (o.input_mode != "sequential" && o.input_mode != "affine") || o.k < 2 || o.k > 4 ||
        (o.tier != "auto" && o.tier != "hbm" && o.tier != "lds"))
        throw std::runtime_error(
            "Invalid config (ET 1..40, BT 1..min(ET,30), PNUM >= 3, ell-i <=61, 2^ell-i >=PNUM, "
            "input mode sequential|affine, k 2..4)");
    return o;
}

ck_params parameters(const Options& o) {
    ck_params p;
    const int rc = ck_params_derive(o.mode == "bgh25" ? CK_MODE_BGH25 : CK_MODE_DRRT18, o.pnum, o.ell_i, o.k,
                                    o.buckets, o.kick, o.stash, o.input_mode == "sequential" ? 0 : 1, &p);
    if (rc != CK_OK) throw std::runtime_error(std::string("Invalid experiment shape: ") + ck_last_error());
    return p;
}

uint64_t planned(const Options& o) { return o.trials ? o.trials : (1ULL << o.et); }
uint64_t total_batches(const Options& o) {
    const uint64_t b = 1ULL << o.bt;
    return (planned(o) + b - 1) / b;
}
uint64_t batch_length(const Options& o, uint64_t id) {
    const uint64_t b = 1ULL << o.bt, start = id * b, total = planned(o);
    return start >= total ? 0 : std::min<uint64_t>(b, total - start);
}

// --------------------------------------------------------------- journal

// v3 layout (cuckoo_bench.cu:567-571) when the experiment is representable by
// the reference; v4 appends {k, stash} and records a non-power-of-two plan.
Header create_header(const Options& o, const ck_params& p) {
    Header h;
    h.v4 = o.k != 3 || o.stash != 0 || o.trials != 0;
    h.base.version = h.v4 ? 4 : 3;
    h.base.seed = o.seed;
    h.base.planned = planned(o);
    h.base.et = o.trials ? 0 : o.et;
    h.base.bt = o.bt;
    h.base.pnum = o.pnum;
    h.base.ell_i = o.ell_i;
    h.base.mode = o.mode == "bgh25" ? 1 : 2;
    h.base.buckets = p.buckets;
    h.base.kick = p.kick;
    h.base.input_mode = o.input_mode == "sequential" ? 0 : 1;
    h.ext.k = o.k;
    h.ext.stash = o.stash;
    return h;
}

// Word-wise FNV-1a over (batch lo, batch hi, tried, count, offsets...)
// (cuckoo_bench.cu:574-581).
uint32_t checksum(uint64_t batch, uint32_t tried, const std::vector<uint32_t>& offsets) {
    uint32_t h = 0x811c9dc5u;
    auto mix = [&](uint32_t x) { h = (h ^ x) * 0x01000193u; };
    mix(static_cast<uint32_t>(batch));
    mix(static_cast<uint32_t>(batch >> 32));
    mix(tried);
    mix(static_cast<uint32_t>(offsets.size()));
    for (uint32_t x : offsets) mix(x);
    return h;
}

static bool sorted_unique_below(const std::vector<uint32_t>& v, uint32_t limit) {
    for (size_t i = 1; i < v.size(); ++i)
        if (v[i - 1] >= v[i]) return false;
    return v.empty() || v.back() < limit;
}

void append_record(FILE* f, uint64_t id, uint32_t attempted, const std::vector<uint32_t>& offsets) {
    if (offsets.size() > attempted) throw std::runtime_error("Invalid failure count");
    if (!sorted_unique_below(offsets, attempted)) throw std::runtime_error("Invalid failure offsets");
    const RecordHeader h{id, attempted, static_cast<uint32_t>(offsets.size())};
    const RecordFooter tail{checksum(id, attempted, offsets), RECORD_MAGIC};
    if (fwrite(&h, sizeof h, 1, f) != 1 ||
        (!offsets.empty() && fwrite(offsets.data(), sizeof(uint32_t), offsets.size(), f) != offsets.size()) ||
        fwrite(&tail, sizeof tail, 1, f) != 1)
        throw std::runtime_error("Journal write failed");
    durable_flush(f);
}

std::string journal_name(const Options& o, int worker) {
    return (fs::path(o.out) / ("worker_" + std::to_string(worker) + ".bin")).string();
}

std::vector<std::string> journal_paths(const Options& o) {
    std::vector<std::string> paths;
    for (const auto& e : fs::directory_iterator(o.out)) {
        if (!e.is_regular_file()) continue;
        const std::string name = e.path().filename().string();
        if (name.rfind("worker_", 0) != 0 || name.size() < 12 || name.substr(name.size() - 4) != ".bin") continue;
        const std::string digits = name.substr(7, name.size() - 11);
        if (digits.empty() || !std::all_of(digits.begin(), digits.end(), [](char c) { return c >= '0' && c <= '9'; }))
            continue;
        paths.push_back(e.path().string());
    }
    std::sort(paths.begin(), paths.end());
    return paths;
}

static void validate_header(const Header& got, const Header& want) {
    const auto& a = got.base;
    const auto& b = want.base;
    if (a.magic != b.magic || a.version != b.version || a.seed != b.seed || a.planned != b.planned ||
        a.et != b.et || a.bt != b.bt || a.pnum != b.pnum || a.ell_i != b.ell_i || a.mode != b.mode ||
        a.buckets != b.buckets || a.kick != b.kick || a.input_mode != b.input_mode ||
        (want.v4 && (got.ext.k != want.ext.k || got.ext.stash != want.ext.stash)))
        throw std::runtime_error("Existing journal has different experiment parameters");
}

static bool read_header(std::ifstream& in, const Header& want, Header& got) {
    got = Header{};
    got.v4 = want.v4;
    in.read(reinterpret_cast<char*>(&got.base), sizeof got.base);
    if (want.v4) in.read(reinterpret_cast<char*>(&got.ext), sizeof got.ext);
    return bool(in);
}

// Validates every journal, truncates torn tails, rejects corrupt or duplicate
// complete records, and returns the completed-batch bitmap
// (cuckoo_bench.cu:650-698).
std::vector<uint8_t> resume(const Options& o, const ck_params& p, Counters& s) {
    const uint64_t batches = total_batches(o);
    if (batches > (1ULL << 28)) throw std::runtime_error("Too many batches for the completion bitmap; increase --bt");
    std::vector<uint8_t> done(static_cast<size_t>(batches), 0);
    const Header want = create_header(o, p);
    const auto paths = journal_paths(o);
    for (const auto& path : paths) {
        if (fs::file_size(path) < want.size()) continue;
        std::ifstream in(path, std::ios::binary);
        Header got;
        if (!read_header(in, want, got)) throw std::runtime_error("Cannot read journal header: " + path);
        validate_header(got, want);
    }
    for (const auto& path : paths) {
        const uint64_t file_size = fs::file_size(path);
        if (file_size < want.size()) {
            fs::resize_file(path, 0);
            continue;
        }
        std::ifstream in(path, std::ios::binary);
        in.seekg(static_cast<std::streamoff>(want.size()));
        uint64_t safe = want.size();
        while (safe < file_size) {
            if (file_size - safe < sizeof(RecordHeader)) break;  // torn record header
            RecordHeader r{};
            in.read(reinterpret_cast<char*>(&r), sizeof r);
            if (!in) throw std::runtime_error("Journal read failed: " + path);
            if (r.batch >= batches || r.tried != batch_length(o, r.batch) || r.failed > r.tried)
                throw std::runtime_error("Invalid journal record header: " + path);
            const uint64_t record_bytes = sizeof r + uint64_t(r.failed) * 4 + sizeof(RecordFooter);
            if (file_size - safe < record_bytes) break;  // uncommitted partial tail
            std::vector<uint32_t> offsets(r.failed);
            if (r.failed) in.read(reinterpret_cast<char*>(offsets.data()), std::streamsize(r.failed) * 4);
            RecordFooter tail{};
            in.read(reinterpret_cast<char*>(&tail), sizeof tail);
            if (!in) throw std::runtime_error("Journal read failed: " + path);
            if (tail.magic != RECORD_MAGIC || tail.check != checksum(r.batch, r.tried, offsets))
                throw std::runtime_error("Corrupt complete journal record (file preserved): " + path);
            if (!sorted_unique_below(offsets, r.tried)) throw std::runtime_error("Invalid journal offsets: " + path);
            if (done[r.batch]) throw std::runtime_error("Duplicate committed batch: " + std::to_string(r.batch));
            done[r.batch] = 1;
            s.batches.fetch_add(1);
            s.completed.fetch_add(r.tried);
            s.failed.fetch_add(r.failed);
            safe += record_bytes;
        }
        in.close();
        if (file_size != safe) fs::resize_file(path, safe);
    }
    return done;
}

FILE* open_journal(const Options& o, const ck_params& p, int worker) {
    const auto path = journal_name(o, worker);
    const bool initialized = fs::exists(path) && fs::file_size(path) != 0;
    std::unique_ptr<FILE, decltype(&fclose)> f(fopen(path.c_str(), initialized ? "ab" : "wb"), fclose);
    if (!f) throw std::runtime_error("Unable to open: " + path);
    if (!initialized) {
        const Header h = create_header(o, p);
        bool ok = fwrite(&h.base, sizeof h.base, 1, f.get()) == 1;
        if (ok && h.v4) ok = fwrite(&h.ext, sizeof h.ext, 1, f.get()) == 1;
        if (!ok) throw std::runtime_error("Journal header write failed");
        durable_flush(f.get());
    }
    return f.release();
}

int64_t steady_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}
void mark_finished(Counters& s) {
    const auto now = steady_ns();
    auto prev = s.finished_ns.load();
    while (prev < now && !s.finished_ns.compare_exchange_weak(prev, now)) {
    }
}

static const char* tier_name(const Options& o) {
    if (o.tier == "hbm") return "hbm-lane";
    if (o.tier == "lds") return "lds-workgroup";
    return "auto";
}

// summary.json (same keys as cuckoo_bench.cu:753-776, extension keys appended)
// and failures.csv re-read from the journals.
void final_outputs(const Options& o, const ck_params& p, const Counters& s, double seconds, const std::string& status,
                   double work_seconds) {
    const auto completed = s.completed.load(), fails = s.failed.load(), batches = s.batches.load();
    const uint64_t plan = planned(o);
    {
        std::ofstream out;
        out.exceptions(std::ios::failbit | std::ios::badbit);
        out.open((fs::path(o.out) / "summary.json").string(), std::ios::trunc);
        out << "{\n  \"status\": \"" << status << "\",\n  \"seed_scheme\": \"aes-batch-v1\",\n"
            << "  \"mode\": \"" << o.mode << "\",\n  \"algorithm\": \"prp-cuckoo-aes-feistel-port\",\n"
            << "  \"backend\": \"" << o.backend << "\",\n  \"workers\": " << o.workers << ",\n"
            << "  \"micro_requested\": " << o.micro << ",\n  \"scratch_mib\": " << o.scratch_mib << ",\n"
            << "  \"prp_cache_requested\": \"" << o.prp_cache << "\",\n"
            << "  \"prp_cache_enabled\": " << (p.half_bits <= 12 ? "true" : "false") << ",\n"
            << "  \"table_layout_requested\": \"" << o.table_layout << "\",\n"
            << "  \"table_layout_planned\": \"" << tier_name(o) << "\",\n"
            << "  \"shared_variant_requested\": \"" << o.shared_variant << "\",\n"
            << "  \"block_threads_requested\": " << o.block_threads << ",\n"
            << "  \"warps_per_block_requested\": " << o.warps_per_block << ",\n"
            << "  \"input_mode\": \"" << o.input_mode << "\",\n"
            << "  \"seed\": " << o.seed << ",\n  \"ET\": " << (o.trials ? 0 : o.et) << ",\n  \"BT\": " << o.bt
            << ",\n"
            << "  \"seed_decimal\": \"" << o.seed << "\",\n  \"batch_size\": " << (1ULL << o.bt) << ",\n"
            << "  \"PNUM_T\": " << o.pnum << ",\n  \"ELL_I\": " << o.ell_i << ",\n"
            << "  \"BUCKETS_NUM\": " << p.buckets << ",\n  \"MAX_KICK_TIMES\": " << p.kick << ",\n"
            << "  \"BUCKET_SIZE\": " << p.bucket_size << ",\n  \"PRP_DOMAIN\": " << p.domain << ",\n"
            << "  \"planned\": " << plan << ",\n  \"completed\": " << completed << ",\n  \"failed\": " << fails
            << ",\n  \"completed_batches\": " << batches
            << ",\n  \"pending_batches\": " << (total_batches(o) - batches) << ",\n  \"failed_ratio\": "
            << std::setprecision(17) << (completed ? double(fails) / double(completed) : 0.0)
            << ",\n  \"elapsed_seconds\": " << seconds << ",\n  \"work_elapsed_seconds\": " << work_seconds
            << ",\n  \"threshold_exceeded\": " << (fails > o.stop_after ? "true" : "false") << ",\n"
            << "  \"K\": " << o.k << ",\n  \"STASH\": " << o.stash << ",\n  \"journal_version\": "
            << (create_header(o, p).v4 ? 4 : 3) << "\n}\n";
        out.flush();
        out.close();
    }
    const Header want = create_header(o, p);
    std::ofstream csv;
    csv.exceptions(std::ios::failbit | std::ios::badbit);
    csv.open((fs::path(o.out) / "failures.csv").string(), std::ios::trunc);
    csv << "batch_id,offset\n";
    uint64_t count = 0;
    for (const auto& path : journal_paths(o)) {
        std::ifstream in(path, std::ios::binary);
        if (!in) continue;
        in.seekg(static_cast<std::streamoff>(want.size()), std::ios::beg);
        for (;;) {
            RecordHeader r{};
            in.read(reinterpret_cast<char*>(&r), sizeof r);
            if (!in) break;
            for (uint32_t i = 0; i < r.failed; ++i) {
                uint32_t off = 0;
                in.read(reinterpret_cast<char*>(&off), sizeof off);
                csv << r.batch << ',' << off << '\n';
                ++count;
            }
            in.seekg(sizeof(RecordFooter), std::ios::cur);
            if (!in) throw std::runtime_error("Corrupt journal in final export");
        }
    }
    if (count != fails) throw std::runtime_error("Exported failure count mismatch");
    csv.flush();
    csv.close();
}

}  // namespace ckh
