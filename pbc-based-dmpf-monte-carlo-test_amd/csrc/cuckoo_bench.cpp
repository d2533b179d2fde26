// cuckoo_bench.cpp -- the cuckoo_bench command line (cuckoo_bench.cu:1062-1142)
// on the MI355X library: resume from the journals, one worker thread per
// context, batches claimed through a shared atomic cursor, each batch run by
// process_batch over ck_launch and committed durably, then summary.json,
// failures.csv and the RESULT line.
#include <hip/hip_runtime_api.h>

#include <iomanip>
#include <iostream>
#include <memory>
#include <mutex>
#include <thread>

#include "ck_journal.hpp"

using namespace ckh;

namespace {

void check(int rc, const char* what) {
    if (rc != CK_OK) throw std::runtime_error(std::string(what) + ": " + ck_last_error());
}

struct Ctx {
    ck_ctx* h = nullptr;
    explicit Ctx(int device) { check(ck_ctx_create(device, &h), "ck_ctx_create"); }
    ~Ctx() { ck_ctx_destroy(h); }
    Ctx(const Ctx&) = delete;
};

uint32_t tier_code(const std::string& t) { return t == "hbm" ? CK_TIER_HBM : t == "lds" ? CK_TIER_LDS : CK_TIER_AUTO; }

void worker(int id, int device, const Options& o, const ck_params& p, const std::vector<uint8_t>& done, Counters& s,
            std::mutex& print_mutex) {
    Ctx ctx(device);
    check(ck_ctx_set_tier(ctx.h, tier_code(o.tier)), "ck_ctx_set_tier");
    check(ck_ctx_set_graph(ctx.h, o.graph ? 1 : 0), "ck_ctx_set_graph");
    const uint64_t batches = total_batches(o);
    const uint32_t micro = static_cast<uint32_t>(std::min<uint64_t>(o.micro, 1ULL << o.bt));
    std::string tuned;
    if (o.tune) {
        // A sample of batch 0 under every placement; results are not committed.
        const uint32_t sample = static_cast<uint32_t>(std::min<uint64_t>(batch_length(o, 0), 1u << 17));
        const BatchKey key = batch_key(o.seed, 0);
        ck_tune_entry ents[32];
        uint32_t n = 0;
        check(ck_ctx_tune(ctx.h, &p, key.data(), sample, ents, 32, &n), "ck_ctx_tune");
        const ck_tune_entry* best = nullptr;
        for (uint32_t i = 0; i < std::min<uint32_t>(n, 32); ++i)
            if (!best || ents[i].trials_per_s > best->trials_per_s) best = &ents[i];
        if (best)
            tuned = " tuned=tier" + std::to_string(best->tier) + "/layout" + std::to_string(best->layout) + "/chunk" +
                    std::to_string(best->chunk) + "@" + std::to_string(static_cast<uint64_t>(best->trials_per_s)) +
                    "/s";
    }
    std::unique_ptr<FILE, decltype(&fclose)> journal(open_journal(o, p, id), fclose);
    std::vector<uint32_t> chunk(micro);
    {
        std::lock_guard<std::mutex> lock(print_mutex);
        hipDeviceProp_t prop{};
        (void)hipGetDeviceProperties(&prop, device);
        std::cerr << "Worker " << id << " GPU " << device << " (" << prop.name << ") bins=" << p.buckets
                  << " table=" << (o.tier == "auto" ? "auto" : o.tier) << " micro=" << micro << " k=" << p.k
                  << " stash=" << p.stash << " kick=" << p.kick << " half_bits=" << p.half_bits
                  << " CUs=" << prop.multiProcessorCount << " lds_per_block=" << prop.sharedMemPerBlock << tuned
                  << (o.graph ? " graph=on" : "") << "\n";
    }
    while (!s.stop.load(std::memory_order_relaxed)) {
        if (pause_requested(o)) {
            s.stop.store(true);
            break;
        }
        const uint64_t bid = s.next_batch.fetch_add(1);
        if (bid >= batches) break;
        if (done[bid]) continue;
        const bool committed = process_batch(
            o, bid, micro, s,
            [&](uint32_t n, uint32_t offset, const BatchKey& key, std::vector<uint32_t>& failures) {
                uint32_t nf = 0;
                check(ck_launch(ctx.h, &p, key.data(), offset, n, chunk.data(), &nf, nullptr), "ck_launch");
                failures.insert(failures.end(), chunk.begin(), chunk.begin() + nf);
            },
            [&](uint64_t batch, uint32_t actual, const std::vector<uint32_t>& failures) {
                append_record(journal.get(), batch, actual, failures);
            });
        if (!committed) break;
    }
}

void print_progress(const Options& o, const Counters& s, std::chrono::steady_clock::time_point start) {
    const uint64_t initial = s.completed.load(), total = planned(o);
    while (s.active.load() > 0) {
        std::this_thread::sleep_for(std::chrono::seconds(1));
        const auto done = s.completed.load(), fail = s.failed.load();
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        const double rate = double(done - initial) / std::max(elapsed, 1e-6);
        const unsigned filled = static_cast<unsigned>(std::min<uint64_t>(30, done * 30 / total));
        std::cerr << '\r' << '[' << std::string(filled, '#') << std::string(30 - filled, '-') << "] " << std::fixed
                  << std::setprecision(2) << 100.0 * double(done) / double(total) << "% | " << done << "/" << total
                  << " | Fail " << fail << " | " << std::setprecision(0) << rate << " avg exp/s | elapsed "
                  << std::setprecision(1) << elapsed << "s" << std::flush;
    }
    std::cerr << '\n';
}

}  // namespace

int main(int argc, char** argv) {
    try {
        auto o = parse(argc, argv);
        const auto p = parameters(o);
        fs::create_directories(o.out);
        RunLock run_lock(o.out);
        PauseSignals signals;
        Counters s;
        auto done = resume(o, p, s);
        const uint64_t plan = planned(o);
        const auto start = std::chrono::steady_clock::now();
        std::cout << "mode=" << o.mode << " ET=" << o.et << " BT=" << o.bt << " PNUM=" << o.pnum
                  << " buckets=" << p.buckets << " kick=" << p.kick << " bucket_size=" << p.bucket_size
                  << " ell_d=" << p.ell_d << " input_mode=" << o.input_mode << " k=" << p.k << " stash=" << p.stash
                  << " planned=" << plan << "\n"
                  << "Resume: " << s.completed << " completed / " << s.failed << " failed\n";
        if (o.status_only || s.completed.load() == plan || pause_requested(o) || s.failed.load() > o.stop_after) {
            const std::string status = s.completed.load() == plan ? "complete"
                                       : pause_requested(o)       ? "paused"
                                       : s.failed.load() > o.stop_after ? "threshold"
                                                                         : "ready";
            final_outputs(o, p, s, 0, status);
            std::cout << "Status: " << status << "; pending batches=" << (done.size() - s.batches.load()) << "\n";
            return 0;
        }
        int gpus = 0;
        if (hipGetDeviceCount(&gpus) != hipSuccess || gpus == 0) throw std::runtime_error("No HIP GPUs");
        if (o.workers == 0) o.workers = gpus;
        if (o.workers < 1 || o.workers > 256) throw std::runtime_error("workers must be in 1..256");
        if (o.workers > gpus) std::cerr << "WARNING: Multiple workers share GPUs; usually use one per GPU.\n";
        for (int g = 0; g < std::min(gpus, o.workers); ++g) {
            Ctx ctx(g);
            check(ck_aes_self_test(ctx.h), "AES-128 FIPS known-answer test");
        }
        std::mutex print_mutex;
        std::vector<std::thread> threads;
        std::vector<std::exception_ptr> errors(o.workers);
        std::exception_ptr launch_error;
        const auto work_start = steady_ns();
        if (s.failed.load() <= o.stop_after) {
            for (int g = 0; g < o.workers; ++g) {
                s.active.fetch_add(1);
                try {
                    threads.emplace_back([&, g]() {
                        try {
                            worker(g, g % gpus, o, p, done, s, print_mutex);
                        } catch (...) {
                            errors[g] = std::current_exception();
                            s.abort_inflight.store(true);
                            s.stop.store(true);
                        }
                        mark_finished(s);
                        s.active.fetch_sub(1);
                    });
                } catch (...) {
                    s.active.fetch_sub(1);
                    launch_error = std::current_exception();
                    s.abort_inflight.store(true);
                    s.stop.store(true);
                    break;
                }
            }
            print_progress(o, s, start);
            for (auto& t : threads) t.join();
        }
        Counters recovered;
        resume(o, p, recovered);
        const bool error =
            bool(launch_error) || std::any_of(errors.begin(), errors.end(), [](const auto& e) { return bool(e); });
        const std::string status = error                                 ? "error"
                                   : recovered.completed.load() == plan ? "complete"
                                   : pause_requested(o)                  ? "paused"
                                                                         : "threshold";
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        const double work_seconds = std::max<int64_t>(0, s.finished_ns.load() - work_start) * 1e-9;
        final_outputs(o, p, recovered, elapsed, status, work_seconds);
        std::cout << "RESULT status=" << status << " completed=" << recovered.completed
                  << " failed=" << recovered.failed << " elapsed_s=" << elapsed << " work_s=" << work_seconds << "\n";
        if (pause_requested(o))
            std::cout << "Paused: unfinished batches were not counted. Restart the same command to resume. "
                         "If using the PAUSE file, remove it first.\n";
        if (launch_error) std::rethrow_exception(launch_error);
        for (auto& e : errors)
            if (e) std::rethrow_exception(e);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "FATAL: " << e.what() << "\n";
        return 1;
    }
}
