// ck_runspec.cpp -- command-line parsing and validation, the trial plan,
// stop requests and the output-directory lock of a cuckoo_bench run.
#include "ck_runspec.hpp"

#include <fcntl.h>
#include <signal.h>
#include <sys/file.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>

namespace ckrun {

// ------------------------------------------------------------------ numbers

uint64_t parse_unsigned(const std::string& text, uint64_t max_value) {
    // strtoull accepts leading blanks and a sign; a count on a command line may
    // have neither.
    if (text.empty() || !(text[0] >= '0' && text[0] <= '9'))
        throw std::runtime_error("not an unsigned number: '" + text + "'");
    errno = 0;
    char* end = nullptr;
    const unsigned long long v = std::strtoull(text.c_str(), &end, 0);
    if (errno == ERANGE || v > max_value)
        throw std::runtime_error("number out of range: '" + text + "' (at most " + std::to_string(max_value) + ")");
    if (end == nullptr || *end != '\0') throw std::runtime_error("not an unsigned number: '" + text + "'");
    return v;
}

// ------------------------------------------------------------------ flags

namespace {

// One command-line option: a switch (no value) or a value stored into one
// RunSpec field.  `also` runs after the store (flags with side conditions).
struct Option {
    const char* flag;
    bool RunSpec::*toggle;
    uint32_t RunSpec::*u32;
    uint64_t RunSpec::*u64;
    std::string RunSpec::*text;
    void (*also)(RunSpec&, const std::string&);
};
constexpr Option switch_opt(const char* f, bool RunSpec::*m) { return {f, m, nullptr, nullptr, nullptr, nullptr}; }
constexpr Option u32_opt(const char* f, uint32_t RunSpec::*m, void (*also)(RunSpec&, const std::string&) = nullptr) {
    return {f, nullptr, m, nullptr, nullptr, also};
}
constexpr Option u64_opt(const char* f, uint64_t RunSpec::*m) { return {f, nullptr, nullptr, m, nullptr, nullptr}; }
constexpr Option text_opt(const char* f, std::string RunSpec::*m) { return {f, nullptr, nullptr, nullptr, m, nullptr}; }
constexpr Option custom_opt(const char* f, void (*set)(RunSpec&, const std::string&)) {
    return {f, nullptr, nullptr, nullptr, nullptr, set};
}

void set_workers(RunSpec& s, const std::string& v) { s.workers = static_cast<int>(parse_unsigned(v, 256)); }
void mark_scratch(RunSpec& s, const std::string&) { s.scratch_given = true; }

// The reference's flags (cuckoo_bench.cu:431-466) followed by the extension's.
const Option kOptions[] = {
    u32_opt("--et", &RunSpec::et),
    u32_opt("--bt", &RunSpec::bt),
    u32_opt("--pnum", &RunSpec::pnum),
    u32_opt("--ell-i", &RunSpec::ell_i),
    u32_opt("--micro", &RunSpec::micro),
    u32_opt("--scratch-mib", &RunSpec::scratch_mib, mark_scratch),
    text_opt("--prp-cache", &RunSpec::prp_cache),
    u32_opt("--block-threads", &RunSpec::block_threads),
    text_opt("--table-layout", &RunSpec::table_layout),
    text_opt("--shared-variant", &RunSpec::shared_variant),
    u32_opt("--warps-per-block", &RunSpec::warps_per_block),
    u64_opt("--seed", &RunSpec::seed),
    custom_opt("--workers", set_workers),
    u64_opt("--stop-after", &RunSpec::stop_after),
    text_opt("--mode", &RunSpec::mode),
    text_opt("--input-mode", &RunSpec::input_mode),
    text_opt("--out", &RunSpec::out),
    switch_opt("--global-table", &RunSpec::global_table),
    switch_opt("--status", &RunSpec::status_only),
    // extension
    u32_opt("--k", &RunSpec::k),
    u32_opt("--buckets", &RunSpec::buckets),
    u32_opt("--kick", &RunSpec::kick),
    u32_opt("--stash", &RunSpec::stash),
    u64_opt("--trials", &RunSpec::trials),
    text_opt("--tier", &RunSpec::tier),
    u32_opt("--place-search", &RunSpec::place_search),
    switch_opt("--tune", &RunSpec::tune),
    switch_opt("--graph", &RunSpec::graph),
};

bool one_of(const std::string& v, std::initializer_list<const char*> allowed) {
    for (const char* a : allowed)
        if (v == a) return true;
    return false;
}

// Validation rules, checked in order after parsing; the first that a spec
// breaks is reported.
struct Rule {
    bool (*broken)(const RunSpec&);
    const char* requirement;
};
const Rule kRules[] = {
    {[](const RunSpec& s) { return !one_of(s.mode, {"bgh25", "drrt18"}); }, "--mode is bgh25 or drrt18"},
    {[](const RunSpec& s) { return !one_of(s.input_mode, {"sequential", "affine"}); },
     "--input-mode is sequential or affine"},
    {[](const RunSpec& s) { return s.trials == 0 && (s.et < 1 || s.et > 40); }, "--et is 1..40"},
    {[](const RunSpec& s) { return s.bt < 1 || s.bt > 30; }, "--bt is 1..30"},
    {[](const RunSpec& s) { return s.trials == 0 && s.bt > s.et; }, "--bt is at most --et"},
    {[](const RunSpec& s) { return s.trials > (uint64_t(1) << 40); }, "--trials is at most 2^40"},
    {[](const RunSpec& s) { return s.pnum < 3 || s.pnum > (1u << 30); }, "--pnum is 3..2^30"},
    {[](const RunSpec& s) { return s.ell_i > 61; }, "--ell-i is at most 61"},
    {[](const RunSpec& s) { return (uint64_t(1) << s.ell_i) < s.pnum; },
     "2^ell-i is at least --pnum (the items must be distinct)"},
    {[](const RunSpec& s) { return s.k < 2 || s.k > 4; }, "--k is 2..4"},
    {[](const RunSpec& s) { return s.micro == 0; }, "--micro is at least 1"},
    {[](const RunSpec& s) { return s.scratch_mib == 0; }, "--scratch-mib is at least 1"},
    {[](const RunSpec& s) { return !one_of(std::to_string(s.block_threads), {"64", "128", "256"}); },
     "--block-threads is 64, 128 or 256"},
    {[](const RunSpec& s) { return s.warps_per_block > 32; }, "--warps-per-block is at most 32"},
    {[](const RunSpec& s) { return !one_of(s.prp_cache, {"auto", "on", "off"}); }, "--prp-cache is auto, on or off"},
    {[](const RunSpec& s) { return !one_of(s.table_layout, {"auto", "global", "shared"}); },
     "--table-layout is auto, global or shared"},
    {[](const RunSpec& s) { return !one_of(s.shared_variant, {"auto", "round", "map"}); },
     "--shared-variant is auto, round or map"},
    {[](const RunSpec& s) { return !one_of(s.tier, {"auto", "hbm", "lds", "lane"}); },
     "--tier is auto, hbm, lds or lane"},
    {[](const RunSpec& s) { return s.place_search == 1 || s.place_search > 64; },
     "--place-search is 0 (none) or 2..64 candidate table regions"},
    {[](const RunSpec& s) { return s.table_layout == "shared" && s.global_table; },
     "--table-layout shared and --global-table exclude each other"},
    {[](const RunSpec& s) { return s.table_layout == "shared" && s.prp_cache == "off"; },
     "--table-layout shared builds a round cache: it cannot run with --prp-cache off"},
};

}  // namespace

std::string usage_text() {
    return "usage: cuckoo_bench --mode bgh25|drrt18 [--et 30 | --trials N] [--bt 20] [--pnum 16] [--ell-i 32]\n"
           "                    [--seed 1] [--input-mode sequential|affine] [--workers 0] [--micro 131072]\n"
           "                    [--out cuckoo_results] [--stop-after 1048576] [--status]\n"
           "  placement:        [--scratch-mib MiB] [--prp-cache auto|on|off] [--table-layout auto|global|shared]\n"
           "                    [--shared-variant auto|round|map] [--global-table] [--tier auto|hbm|lds|lane]\n"
           "                    [--block-threads 64|128|256] [--warps-per-block N] [--tune] [--graph]\n"
           "                    [--place-search 32]\n"
           "  extension shape:  [--k 2..4] [--buckets M] [--kick N] [--stash S]\n";
}

RunSpec parse_command_line(int argc, char** argv) {
    RunSpec s;
    for (int i = 1; i < argc; ++i) {
        const std::string flag = argv[i];
        if (flag == "--help" || flag == "-h") {
            std::cout << usage_text();
            std::exit(0);
        }
        const Option* opt = nullptr;
        for (const Option& o : kOptions)
            if (flag == o.flag) opt = &o;
        if (!opt) throw std::runtime_error("unknown option " + flag);
        if (opt->toggle) {
            s.*(opt->toggle) = true;
            continue;
        }
        if (i + 1 >= argc) throw std::runtime_error(flag + " needs a value");
        const std::string value = argv[++i];
        try {
            if (opt->u32) s.*(opt->u32) = static_cast<uint32_t>(parse_unsigned(value, 0xffffffffull));
            if (opt->u64) s.*(opt->u64) = parse_unsigned(value, std::numeric_limits<uint64_t>::max());
            if (opt->text) s.*(opt->text) = value;
            if (opt->also) opt->also(s, value);
        } catch (const std::runtime_error& e) {
            throw std::runtime_error(flag + ": " + e.what());
        }
    }
    for (const Rule& r : kRules)
        if (r.broken(s)) throw std::runtime_error(std::string("invalid configuration: ") + r.requirement);
    return s;
}

// ------------------------------------------------------------------ shape and plan

ck_params derive_shape(const RunSpec& s) {
    ck_params p{};
    if (ck_params_derive(mode_code(s), s.pnum, s.ell_i, s.k, s.buckets, s.kick, s.stash, input_code(s), &p) != CK_OK)
        throw std::runtime_error(std::string("experiment shape rejected: ") + ck_last_error());
    return p;
}

uint64_t planned_trials(const RunSpec& s) { return s.trials ? s.trials : uint64_t(1) << s.et; }

uint64_t batch_count(const RunSpec& s) {
    const uint64_t size = uint64_t(1) << s.bt;
    return (planned_trials(s) + size - 1) >> s.bt;
}

uint64_t batch_length(const RunSpec& s, uint64_t batch) {
    const uint64_t size = uint64_t(1) << s.bt, planned = planned_trials(s);
    if (batch >= batch_count(s)) return 0;
    const uint64_t start = batch << s.bt;
    return planned - start < size ? planned - start : size;
}

BatchKey batch_key_of(uint64_t seed, uint64_t batch) {
    BatchKey k{};
    ck_batch_key(seed, batch, k.data());
    return k;
}

int64_t steady_now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

void RunTally::note_worker_exit() {
    const int64_t now = steady_now_ns();
    int64_t seen = last_exit_ns.load();
    while (seen < now && !last_exit_ns.compare_exchange_weak(seen, now)) {
    }
}

// ------------------------------------------------------------------ stop requests

namespace {
std::atomic<bool> g_stop_flag{false};
static_assert(std::atomic<bool>::is_always_lock_free, "the signal handler stores to a lock-free flag only");
extern "C" void ck_on_stop_signal(int) { g_stop_flag.store(true, std::memory_order_relaxed); }
}  // namespace

struct StopSignals::Saved {
    struct sigaction interrupt {};
    struct sigaction terminate {};
};

StopSignals::StopSignals() : saved_(new Saved) {
    g_stop_flag.store(false);
    struct sigaction act {};
    act.sa_handler = ck_on_stop_signal;
    sigemptyset(&act.sa_mask);
    act.sa_flags = SA_RESTART;
    if (sigaction(SIGINT, &act, &saved_->interrupt) != 0) {
        delete saved_;
        throw std::runtime_error("cannot take over SIGINT");
    }
    if (sigaction(SIGTERM, &act, &saved_->terminate) != 0) {
        sigaction(SIGINT, &saved_->interrupt, nullptr);
        delete saved_;
        throw std::runtime_error("cannot take over SIGTERM");
    }
}

StopSignals::~StopSignals() {
    sigaction(SIGINT, &saved_->interrupt, nullptr);
    sigaction(SIGTERM, &saved_->terminate, nullptr);
    delete saved_;
}

bool StopSignals::raised() { return g_stop_flag.load(std::memory_order_relaxed); }
void StopSignals::clear() { g_stop_flag.store(false); }
void StopSignals::raise_for_test() { g_stop_flag.store(true); }

bool stop_requested(const RunSpec& s) {
    if (StopSignals::raised()) return true;
    std::error_code ec;
    return fs::exists(fs::path(s.out) / "PAUSE", ec);
}

// ------------------------------------------------------------------ directory lock

DirectoryLock::DirectoryLock(const fs::path& directory) {
    const std::string path = (directory / ".run.lock").string();
    fd_ = ::open(path.c_str(), O_CREAT | O_RDWR | O_CLOEXEC, 0600);
    if (fd_ < 0) throw std::runtime_error("cannot open the run lock " + path + ": " + std::strerror(errno));
    if (::flock(fd_, LOCK_EX | LOCK_NB) != 0) {
        const int err = errno;
        ::close(fd_);
        fd_ = -1;
        if (err == EWOULDBLOCK)
            throw std::runtime_error("output directory " + directory.string() + " is in use by another run (" + path +
                                     " is locked)");
        throw std::runtime_error("cannot lock " + path + ": " + std::strerror(err));
    }
}

DirectoryLock::~DirectoryLock() {
    if (fd_ >= 0) ::close(fd_);  // closing the descriptor releases the flock
}

}  // namespace ckrun
