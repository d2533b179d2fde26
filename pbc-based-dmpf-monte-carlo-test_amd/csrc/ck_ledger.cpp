// ck_ledger.cpp -- journal encoding, the recovery scan, the journal writer and
// the result export.
#include "ck_ledger.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace ckrun {

// ------------------------------------------------------------------ byte codec

namespace {

void put_u32(uint8_t* p, uint32_t v) {
    for (int i = 0; i < 4; ++i) p[i] = static_cast<uint8_t>(v >> (8 * i));
}
void put_u64(uint8_t* p, uint64_t v) {
    for (int i = 0; i < 8; ++i) p[i] = static_cast<uint8_t>(v >> (8 * i));
}
uint32_t get_u32(const uint8_t* p) {
    uint32_t v = 0;
    for (int i = 3; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}
uint64_t get_u64(const uint8_t* p) { return uint64_t(get_u32(p)) | (uint64_t(get_u32(p + 4)) << 32); }

// Field offsets of the file header.
enum : size_t {
    kOffMagic = 0, kOffVersion = 4, kOffSeed = 8, kOffPlanned = 16, kOffEt = 24, kOffBt = 28, kOffPnum = 32,
    kOffEll = 36, kOffMode = 40, kOffBuckets = 44, kOffKick = 48, kOffInput = 52, kOffK = 56, kOffStash = 60
};

std::string os_error(const std::string& what, const std::string& path) {
    return what + " " + path + ": " + std::strerror(errno);
}

void write_fully(int fd, const uint8_t* data, size_t size, const std::string& path) {
    while (size) {
        const ssize_t n = ::write(fd, data, size);
        if (n < 0 && errno == EINTR) continue;
        if (n <= 0) throw std::runtime_error(os_error("cannot append to", path));
        data += n;
        size -= static_cast<size_t>(n);
    }
}

void sync_data(int fd, const std::string& path) {
    if (::fdatasync(fd) != 0) throw std::runtime_error(os_error("cannot flush", path));
}

// A journal file mapped read-only for the scan.
class MappedFile {
   public:
    explicit MappedFile(const fs::path& path) {
        const std::string name = path.string();
        fd_ = ::open(name.c_str(), O_RDONLY | O_CLOEXEC);
        if (fd_ < 0) throw std::runtime_error(os_error("cannot open", name));
        struct stat st {};
        if (::fstat(fd_, &st) != 0) throw std::runtime_error(os_error("cannot stat", name));
        size_ = static_cast<uint64_t>(st.st_size);
        if (size_) {
            void* m = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
            if (m == MAP_FAILED) throw std::runtime_error(os_error("cannot map", name));
            data_ = static_cast<const uint8_t*>(m);
        }
    }
    ~MappedFile() {
        if (data_) ::munmap(const_cast<uint8_t*>(data_), size_);
        if (fd_ >= 0) ::close(fd_);
    }
    MappedFile(const MappedFile&) = delete;
    MappedFile& operator=(const MappedFile&) = delete;
    const uint8_t* data() const { return data_; }
    uint64_t size() const { return size_; }

   private:
    int fd_ = -1;
    const uint8_t* data_ = nullptr;
    uint64_t size_ = 0;
};

}  // namespace

JournalIdentity identity_of(const RunSpec& s, const ck_params& p) {
    JournalIdentity id;
    id.version = (s.k != 3 || s.stash != 0 || s.trials != 0) ? 4 : 3;
    id.seed = s.seed;
    id.planned = planned_trials(s);
    id.et = s.trials ? 0 : s.et;
    id.bt = s.bt;
    id.pnum = s.pnum;
    id.ell_i = s.ell_i;
    id.mode = mode_code(s);
    id.buckets = p.buckets;
    id.kick = p.kick;
    id.input_mode = input_code(s);
    id.k = s.k;
    id.stash = s.stash;
    return id;
}

std::vector<uint8_t> encode_identity(const JournalIdentity& id) {
    std::vector<uint8_t> b(id.bytes(), 0);
    put_u32(&b[kOffMagic], kJournalMagic);
    put_u32(&b[kOffVersion], id.version);
    put_u64(&b[kOffSeed], id.seed);
    put_u64(&b[kOffPlanned], id.planned);
    put_u32(&b[kOffEt], id.et);
    put_u32(&b[kOffBt], id.bt);
    put_u32(&b[kOffPnum], id.pnum);
    put_u32(&b[kOffEll], id.ell_i);
    put_u32(&b[kOffMode], id.mode);
    put_u32(&b[kOffBuckets], id.buckets);
    put_u32(&b[kOffKick], id.kick);
    put_u32(&b[kOffInput], id.input_mode);
    if (id.version == 4) {
        put_u32(&b[kOffK], id.k);
        put_u32(&b[kOffStash], id.stash);
    }
    return b;
}

JournalIdentity decode_identity(const uint8_t* b, size_t size) {
    JournalIdentity id;
    id.version = get_u32(b + kOffVersion);
    id.seed = get_u64(b + kOffSeed);
    id.planned = get_u64(b + kOffPlanned);
    id.et = get_u32(b + kOffEt);
    id.bt = get_u32(b + kOffBt);
    id.pnum = get_u32(b + kOffPnum);
    id.ell_i = get_u32(b + kOffEll);
    id.mode = get_u32(b + kOffMode);
    id.buckets = get_u32(b + kOffBuckets);
    id.kick = get_u32(b + kOffKick);
    id.input_mode = get_u32(b + kOffInput);
    if (id.version == 4 && size >= kHeaderBytesV4) {
        id.k = get_u32(b + kOffK);
        id.stash = get_u32(b + kOffStash);
    }
    return id;
}

std::string identity_mismatch(const JournalIdentity& want, const JournalIdentity& got) {
    const struct {
        const char* name;
        uint64_t a, b;
    } fields[] = {
        {"version", want.version, got.version}, {"seed", want.seed, got.seed},
        {"planned", want.planned, got.planned}, {"et", want.et, got.et},
        {"bt", want.bt, got.bt},                {"pnum", want.pnum, got.pnum},
        {"ell_i", want.ell_i, got.ell_i},       {"mode", want.mode, got.mode},
        {"buckets", want.buckets, got.buckets}, {"kick", want.kick, got.kick},
        {"input_mode", want.input_mode, got.input_mode},
        {"k", want.version == 4 ? want.k : 0, want.version == 4 ? got.k : 0},
        {"stash", want.version == 4 ? want.stash : 0, want.version == 4 ? got.stash : 0},
    };
    for (const auto& f : fields)
        if (f.a != f.b) return f.name;
    return {};
}

uint32_t record_checksum(uint64_t batch, uint32_t trials, const uint32_t* offsets, size_t count) {
    uint32_t h = 0x811c9dc5u;
    const auto fold = [&h](uint32_t word) { h = (h ^ word) * 0x01000193u; };
    fold(static_cast<uint32_t>(batch));
    fold(static_cast<uint32_t>(batch >> 32));
    fold(trials);
    fold(static_cast<uint32_t>(count));
    for (size_t i = 0; i < count; ++i) fold(offsets[i]);
    return h;
}

std::vector<uint8_t> encode_record(uint64_t batch, uint32_t trials, const std::vector<uint32_t>& offsets) {
    std::vector<uint8_t> r(kRecordHeadBytes + 4 * offsets.size() + kRecordTailBytes);
    put_u64(&r[0], batch);
    put_u32(&r[8], trials);
    put_u32(&r[12], static_cast<uint32_t>(offsets.size()));
    uint8_t* at = &r[kRecordHeadBytes];
    for (uint32_t o : offsets) {
        put_u32(at, o);
        at += 4;
    }
    put_u32(at, record_checksum(batch, trials, offsets.data(), offsets.size()));
    put_u32(at + 4, kRecordMagic);
    return r;
}

fs::path journal_path(const RunSpec& s, int worker) {
    return fs::path(s.out) / ("worker_" + std::to_string(worker) + ".bin");
}

std::vector<fs::path> journal_files(const fs::path& directory) {
    std::vector<fs::path> out;
    std::error_code ec;
    for (const auto& entry : fs::directory_iterator(directory, ec)) {
        if (!entry.is_regular_file()) continue;
        const std::string name = entry.path().filename().string();
        const std::string head = "worker_", tail = ".bin";
        if (name.size() <= head.size() + tail.size() || name.compare(0, head.size(), head) != 0 ||
            name.compare(name.size() - tail.size(), tail.size(), tail) != 0)
            continue;
        const std::string digits = name.substr(head.size(), name.size() - head.size() - tail.size());
        if (std::all_of(digits.begin(), digits.end(), [](char c) { return c >= '0' && c <= '9'; }))
            out.push_back(entry.path());
    }
    if (ec) throw std::runtime_error("cannot list " + directory.string() + ": " + ec.message());
    std::sort(out.begin(), out.end());
    return out;
}

// ------------------------------------------------------------------ scan

LedgerState scan_ledger(const RunSpec& s, const ck_params& p, bool keep_failures, bool repair) {
    const uint64_t batches = batch_count(s);
    if (batches > (uint64_t(1) << 28)) throw std::runtime_error("the plan has more than 2^28 batches: raise --bt");
    const JournalIdentity want = identity_of(s, p);
    const size_t header = want.bytes();
    LedgerState st;
    st.committed.assign(static_cast<size_t>(batches), 0);
    const auto files = journal_files(s.out);

    // Every whole header first: a journal of another experiment stops the scan
    // before anything is read or cut.
    for (const auto& path : files) {
        if (fs::file_size(path) < header) continue;
        MappedFile f(path);
        if (f.size() < header || get_u32(f.data() + kOffMagic) != kJournalMagic)
            throw std::runtime_error("journal " + path.string() + " is not a cuckoo journal");
        const std::string field = identity_mismatch(want, decode_identity(f.data(), f.size()));
        if (!field.empty())
            throw std::runtime_error("journal " + path.string() + " records a different experiment (" + field +
                                     " differs)");
    }

    // Then the records, without changing any file.
    std::vector<std::pair<fs::path, uint64_t>> cuts;  // (file, length to keep)
    for (const auto& path : files) {
        MappedFile f(path);
        const uint64_t size = f.size();
        if (size < header) {
            if (size) cuts.emplace_back(path, 0);
            continue;
        }
        const uint8_t* base = f.data();
        uint64_t at = header;
        while (size - at >= kRecordHeadBytes) {
            const uint64_t batch = get_u64(base + at);
            const uint32_t trials = get_u32(base + at + 8), count = get_u32(base + at + 12);
            if (batch >= batches || trials != batch_length(s, batch) || count > trials)
                throw std::runtime_error("journal " + path.string() + ": the record at byte " + std::to_string(at) +
                                         " is not a batch of this plan");
            const uint64_t length = kRecordHeadBytes + 4 * uint64_t(count) + kRecordTailBytes;
            if (size - at < length) break;  // a write cut short: uncommitted
            std::vector<uint32_t> offsets(count);
            for (uint32_t i = 0; i < count; ++i) offsets[i] = get_u32(base + at + kRecordHeadBytes + 4 * i);
            const uint8_t* tail = base + at + kRecordHeadBytes + 4 * uint64_t(count);
            if (get_u32(tail + 4) != kRecordMagic ||
                get_u32(tail) != record_checksum(batch, trials, offsets.data(), count))
                throw std::runtime_error("journal " + path.string() + ": the complete record at byte " +
                                         std::to_string(at) + " fails its checksum (the file is left as it is)");
            for (uint32_t i = 0; i < count; ++i)
                if (offsets[i] >= trials || (i && offsets[i] <= offsets[i - 1]))
                    throw std::runtime_error("journal " + path.string() + ": the record at byte " +
                                             std::to_string(at) + " lists offsets out of order or range");
            if (st.committed[batch])
                throw std::runtime_error("batch " + std::to_string(batch) + " is committed twice (again in " +
                                         path.string() + ")");
            st.committed[batch] = 1;
            ++st.batches;
            st.trials += trials;
            st.failures += count;
            if (keep_failures)
                for (uint32_t o : offsets) st.failing.emplace_back(batch, o);
            at += length;
        }
        if (at != size) cuts.emplace_back(path, at);
    }

    if (repair)
        for (const auto& [path, keep] : cuts) fs::resize_file(path, keep);
    if (keep_failures) std::sort(st.failing.begin(), st.failing.end());
    return st;
}

// ------------------------------------------------------------------ writer

JournalWriter::JournalWriter(const RunSpec& s, const ck_params& p, int worker) : path_(journal_path(s, worker)) {
    fd_ = ::open(path_.c_str(), O_WRONLY | O_CREAT | O_APPEND | O_CLOEXEC, 0644);
    if (fd_ < 0) throw std::runtime_error(os_error("cannot open", path_));
    struct stat st {};
    if (::fstat(fd_, &st) != 0) {
        ::close(fd_);
        throw std::runtime_error(os_error("cannot stat", path_));
    }
    if (st.st_size == 0) {
        const auto header = encode_identity(identity_of(s, p));
        try {
            write_fully(fd_, header.data(), header.size(), path_);
            sync_data(fd_, path_);
        } catch (...) {
            ::close(fd_);
            throw;
        }
    }
}

JournalWriter::~JournalWriter() {
    if (fd_ >= 0) ::close(fd_);
}

void JournalWriter::commit(uint64_t batch, uint32_t trials, const std::vector<uint32_t>& offsets) {
    if (offsets.size() > trials) throw std::logic_error("more failing offsets than trials");
    for (size_t i = 0; i < offsets.size(); ++i)
        if (offsets[i] >= trials || (i && offsets[i] <= offsets[i - 1]))
            throw std::logic_error("failing offsets must be ascending, unique and below the batch length");
    const auto record = encode_record(batch, trials, offsets);
    write_fully(fd_, record.data(), record.size(), path_);
    sync_data(fd_, path_);
}

// ------------------------------------------------------------------ export

namespace {

// A JSON object written field by field, in insertion order.
class JsonFields {
   public:
    void text(const char* key, const std::string& v) { add(key, "\"" + v + "\""); }
    void number(const char* key, uint64_t v) { add(key, std::to_string(v)); }
    void real(const char* key, double v) {
        char buf[40];
        std::snprintf(buf, sizeof buf, "%.17g", v);
        add(key, buf);
    }
    void boolean(const char* key, bool v) { add(key, v ? "true" : "false"); }
    std::string str() const {
        std::string s = "{\n";
        for (size_t i = 0; i < rows_.size(); ++i)
            s += "  \"" + rows_[i].first + "\": " + rows_[i].second + (i + 1 < rows_.size() ? ",\n" : "\n");
        return s + "}\n";
    }

   private:
    void add(const char* key, std::string v) { rows_.emplace_back(key, std::move(v)); }
    std::vector<std::pair<std::string, std::string>> rows_;
};

void replace_file(const fs::path& path, const std::string& content) {
    const fs::path tmp = path.string() + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        f << content;
        f.flush();
        if (!f) throw std::runtime_error("cannot write " + tmp.string());
    }
    fs::rename(tmp, path);
}

}  // namespace

void write_reports(const RunSpec& s, const ck_params& p, const LedgerState& st, const RunReport& r) {
    const uint64_t planned = planned_trials(s);
    JsonFields j;
    j.text("status", r.status);
    j.text("seed_scheme", "aes-batch-v1");
    j.text("mode", s.mode);
    j.text("algorithm", "prp-cuckoo-aes-feistel-port");
    j.text("backend", r.backend);
    j.number("workers", static_cast<uint64_t>(r.workers));
    j.number("micro_requested", s.micro);
    j.number("scratch_mib", s.scratch_given ? s.scratch_mib : 0);
    j.text("prp_cache_requested", s.prp_cache);
    j.boolean("prp_cache_enabled", r.round_cache);
    j.text("table_layout_requested", s.table_layout);
    j.text("table_layout_planned", r.placement);
    j.text("shared_variant_requested", s.shared_variant);
    j.number("block_threads_requested", s.block_threads);
    j.number("warps_per_block_requested", s.warps_per_block);
    j.text("input_mode", s.input_mode);
    j.number("seed", s.seed);
    j.number("ET", s.trials ? 0 : s.et);
    j.number("BT", s.bt);
    j.text("seed_decimal", std::to_string(s.seed));
    j.number("batch_size", uint64_t(1) << s.bt);
    j.number("PNUM_T", s.pnum);
    j.number("ELL_I", s.ell_i);
    j.number("BUCKETS_NUM", p.buckets);
    j.number("MAX_KICK_TIMES", p.kick);
    j.number("BUCKET_SIZE", p.bucket_size);
    j.number("PRP_DOMAIN", p.domain);
    j.number("planned", planned);
    j.number("completed", st.trials);
    j.number("failed", st.failures);
    j.number("completed_batches", st.batches);
    j.number("pending_batches", batch_count(s) - st.batches);
    j.real("failed_ratio", st.trials ? double(st.failures) / double(st.trials) : 0.0);
    j.real("elapsed_seconds", r.elapsed_s);
    j.real("work_elapsed_seconds", r.work_s);
    j.boolean("threshold_exceeded", st.failures > s.stop_after);
    j.number("K", s.k);
    j.number("STASH", s.stash);
    j.number("journal_version", identity_of(s, p).version);
    replace_file(fs::path(s.out) / "summary.json", j.str());

    if (st.failing.size() != st.failures)
        throw std::logic_error("write_reports needs a scan that kept the failing trials");
    std::string csv = "batch_id,offset\n";
    csv.reserve(csv.size() + st.failing.size() * 16);
    char row[48];
    for (const auto& [batch, offset] : st.failing) {
        std::snprintf(row, sizeof row, "%" PRIu64 ",%" PRIu32 "\n", batch, offset);
        csv += row;
    }
    replace_file(fs::path(s.out) / "failures.csv", csv);
}

}  // namespace ckrun
