// ref_shim.cpp -- extern "C" entry points into the reference's own host code.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) against
// the reference sources where they lie (-I <reference dir>); the result goes to
// oracle/_ref/libref_shim.so and is used only by tests/ to pin the oracle
// against the real reference functions.  Nothing of the reference is copied:
// this file includes the reference's host half (cuckoo_bench.cu compiled with
// CUCKOO_HOST_TEST, as the reference's own cpu_tests.cpp does) and cpu_trials.hpp.
#define CUCKOO_HOST_TEST
#include "cuckoo_bench.cu"
#include "cpu_trials.hpp"

namespace {
Options make_options(int mode, int input_mode, uint32_t pnum, uint32_t ell_i, uint64_t seed) {
    Options o;
    o.mode = mode == 1 ? "bgh25" : "drrt18";
    o.input_mode = input_mode == 0 ? "sequential" : "affine";
    o.pnum = pnum;
    o.ell_i = ell_i;
    o.seed = seed;
    return o;
}
}  // namespace

extern "C" {

// parameters() (cuckoo_bench.cu:500-526).
int ref_parameters(int mode, uint32_t pnum, uint32_t ell_i, uint32_t* buckets, uint32_t* kick,
                   uint64_t* bucket_size, uint64_t* domain, uint32_t* half_bits, uint32_t* ell_d) {
    try {
        const Params p = parameters(make_options(mode, 0, pnum, ell_i, 0));
        *buckets = p.buckets;
        *kick = p.kick;
        *bucket_size = p.bucket_size;
        *domain = p.domain;
        *half_bits = p.half_bits;
        *ell_d = p.ell_d;
        return 0;
    } catch (...) {
        return -1;
    }
}

float ref_original_log2(uint64_t n) { return original_log2_u64(n); }

// batch_key (cuckoo_bench.cu:393-404).
void ref_batch_key(uint64_t seed, uint64_t id, uint32_t w[44]) {
    const AesKey k = batch_key(seed, id);
    std::memcpy(w, k.w, sizeof(k.w));
}

void ref_expand_key(uint64_t lo, uint64_t hi, uint32_t w[44]) {
    const AesKey k = expand_key(lo, hi);
    std::memcpy(w, k.w, sizeof(k.w));
}

// host_aes128 (cuckoo_bench.cu:360-392).
void ref_host_aes(const uint32_t w[44], const uint32_t in[4], uint32_t out[4]) {
    AesKey k;
    std::memcpy(k.w, w, sizeof(k.w));
    host_aes128(k, in[0], in[1], in[2], in[3], out);
}

// CpuAes::words (cpu_trials.hpp:19-43).
void ref_cpu_aes(const uint32_t w[44], const uint32_t in[4], uint32_t out[4]) {
    AesKey k;
    std::memcpy(k.w, w, sizeof(k.w));
    CpuAes(k).words(in[0], in[1], in[2], in[3], out);
}

// cpu_bucket (cpu_trials.hpp:56-68) for the PRP key of trial (seed, batch, offset).
uint32_t ref_bucket(int mode, int input_mode, uint32_t pnum, uint32_t ell_i, uint64_t seed, uint64_t batch,
                    uint32_t offset, uint32_t ordinal, uint32_t which) {
    const Options o = make_options(mode, input_mode, pnum, ell_i, seed);
    const Params p = parameters(o);
    const CpuAes bk(batch_key(seed, batch));
    uint32_t words[4];
    bk.words(0x4b455931u, 0, offset, 0x50525031u, words);
    const CpuAes aes(cpu_key_from_words(words));
    bk.words(0x57414c4bu, 0, offset, 0x524e4731u, words);
    const uint64_t range_mask = (uint64_t{1} << o.ell_i) - 1;
    const uint64_t shift = ((uint64_t(words[2]) << 32) | words[3]) & range_mask;
    uint64_t input_state = shift ^ 0xd1b54a32d192ed03ULL;
    const uint64_t step = (cpu_splitmix(input_state) | 1ULL) & range_mask;
    const uint64_t element = input_mode == 0 ? uint64_t(ordinal) : (shift + uint64_t(ordinal) * step) & range_mask;
    return cpu_bucket(element, which, aes, p);
}

// cpu_trial / cpu_trial_cached (cpu_trials.hpp:69-149): returns 1 on failure;
// bins_out (p.buckets entries, may be null) receives the reference table,
// entry = ordinal | choice<<16, empty = 0xffffffff.
int ref_trial(int mode, int input_mode, uint32_t pnum, uint32_t ell_i, uint64_t seed, uint64_t batch,
              uint32_t offset, int cached, uint32_t* bins_out) {
    const Options o = make_options(mode, input_mode, pnum, ell_i, seed);
    const Params p = parameters(o);
    const CpuAes bk(batch_key(seed, batch));
    std::vector<uint32_t> bins(p.buckets);
    bool failed;
    if (cached) {
        std::vector<uint16_t> rounds(size_t{8} << p.half_bits);
        failed = cpu_trial_cached(o, p, bk, offset, bins, rounds);
    } else {
        failed = cpu_trial(o, p, bk, offset, bins);
    }
    if (bins_out) std::memcpy(bins_out, bins.data(), bins.size() * sizeof(uint32_t));
    return failed ? 1 : 0;
}

}  // extern "C"
