// test_trial_api.cpp -- runs the trial API of include/cuckoo_trial.hpp on the
// host (the same templates the library's kernels instantiate) over
// caller-side Bins and Buckets, for trials listed one per line of <rows>:
//   mode input_mode pnum ell_i seed batch offset
// Per row it writes to <out>, for each combination of a table and a bucket
// function, a block of u32 words: combination id, failed, probes, buckets,
// then the final table as the reference's u32 entries (ordinal | choice << 16,
// empty all-ones).  tests/test_trial_api.py compares them with the
// reference's own outcomes and table digests.
//   combination 0: StridedBins (u32) + CachedBuckets (round cache)
//   combination 1: StridedBins (u32) + PrecomputedBuckets (bucket map)
//   combination 2: SharedBins (u16, pnum < 2^14) + PrecomputedBuckets
//   combination 3: SharedBins (u16, pnum < 2^14) + CachedBuckets
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "cuckoo_mc.h"
#include "cuckoo_trial.hpp"

namespace {

void put(std::vector<uint32_t>& out, uint32_t combo, bool failed, uint32_t probes, const std::vector<uint32_t>& t) {
    out.push_back(combo);
    out.push_back(failed ? 1u : 0u);
    out.push_back(probes);
    out.push_back(static_cast<uint32_t>(t.size()));
    out.insert(out.end(), t.begin(), t.end());
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: ck_trial_api_tests <rows> <out>\n");
        return 2;
    }
    std::ifstream rows(argv[1]);
    std::vector<uint32_t> out;
    std::string line;
    while (std::getline(rows, line)) {
        std::istringstream in(line);
        unsigned mode, input_mode, pnum, ell_i;
        unsigned long long seed, batch;
        unsigned offset;
        if (!(in >> mode >> input_mode >> pnum >> ell_i >> seed >> batch >> offset)) continue;
        ck_params p;
        if (ck_params_derive(mode, pnum, ell_i, 3, CK_DERIVED, CK_DERIVED, 0, input_mode, &p) != CK_OK) {
            std::fprintf(stderr, "shape: %s\n", ck_last_error());
            return 1;
        }
        uint32_t bk[44];
        ck_batch_key(seed, batch, bk);
        ck_trial_inputs ti;
        if (ck_trial_inputs_of(&p, bk, offset, &ti) != CK_OK) return 1;
        std::vector<uint16_t> rounds(size_t(8) << p.half_bits);
        if (ck_round_cache(&p, ti.prp_key, rounds.data()) != CK_OK) return 1;
        const cuckoo::CachedBuckets<uint16_t, 3> cached{rounds.data(), 1, p.domain, p.bucket_size, p.mask,
                                                        p.range_mask, ti.shift, ti.step, p.half_bits,
                                                        p.input_mode == 0};
        // The bucket map, evaluated through the cached form.
        std::vector<uint32_t> map32(size_t(pnum) * 3);
        for (uint32_t o = 0; o < pnum; ++o)
            for (uint32_t w = 0; w < 3; ++w) map32[o * 3 + w] = cached(o, w);
        const cuckoo::PrecomputedBuckets<uint32_t, 3> mapped{map32.data()};
        const bool bgh = mode == CK_MODE_BGH25;
        for (uint32_t combo = 0; combo < 4; ++combo) {
            const bool small = combo >= 2;
            if (small && pnum >= (1u << 14)) continue;
            uint32_t probes = 0;
            bool failed;
            std::vector<uint32_t> t32(p.buckets, ~0u);
            std::vector<uint16_t> t16(p.buckets, 0xffffu);
            const cuckoo::StridedBins wide{t32.data(), 1};
            const cuckoo::SharedBins narrow{t16.data(), 1};
            if (combo == 0)
                failed = bgh ? cuckoo::insert_core_k<true, 3>(pnum, p.kick, 0, wide, cached, ti.walk, ti.randomness, probes)
                             : cuckoo::insert_core_k<false, 3>(pnum, p.kick, 0, wide, cached, 0, 0, probes);
            else if (combo == 1)
                failed = bgh ? cuckoo::insert_core_k<true, 3>(pnum, p.kick, 0, wide, mapped, ti.walk, ti.randomness, probes)
                             : cuckoo::insert_core_k<false, 3>(pnum, p.kick, 0, wide, mapped, 0, 0, probes);
            else if (combo == 2)
                failed = bgh ? cuckoo::insert_core_k<true, 3>(pnum, p.kick, 0, narrow, mapped, ti.walk, ti.randomness, probes)
                             : cuckoo::insert_core_k<false, 3>(pnum, p.kick, 0, narrow, mapped, 0, 0, probes);
            else
                failed = bgh ? cuckoo::insert_core_k<true, 3>(pnum, p.kick, 0, narrow, cached, ti.walk, ti.randomness, probes)
                             : cuckoo::insert_core_k<false, 3>(pnum, p.kick, 0, narrow, cached, 0, 0, probes);
            if (small)
                for (size_t b = 0; b < t16.size(); ++b)
                    t32[b] = t16[b] == 0xffffu ? ~0u
                                               : (cuckoo::SharedBins::ordinal(t16[b]) |
                                                  (cuckoo::SharedBins::choice(t16[b]) << 16));
            // The reference form with its own argument list must agree.
            if (combo == 0) {
                std::vector<uint32_t> again(p.buckets, ~0u);
                const bool f2 = bgh ? cuckoo::insert_core<true>(pnum, p.kick, cuckoo::StridedBins{again.data(), 1},
                                                                 cached, ti.walk, ti.randomness)
                                    : cuckoo::insert_core<false>(pnum, p.kick, cuckoo::StridedBins{again.data(), 1},
                                                                  cached, 0, 0);
                if (f2 != failed || again != t32) {
                    std::fprintf(stderr, "insert_core and insert_core_k<.., 3> disagree\n");
                    return 1;
                }
            }
            put(out, combo, failed, probes, t32);
        }
    }
    std::ofstream f(argv[2], std::ios::binary | std::ios::trunc);
    f.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size() * 4));
    return f ? 0 : 1;
}
