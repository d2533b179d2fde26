// Host test of the memory-region helpers in ck_regions.hpp that need no
// device: split_modes (the two modes of a set of pair-walk times) and
// interleave (the order of a spread range's pieces).  The times follow what
// pair walks of 2 GiB pieces measured on MI355X: about 2.64 ms for two pieces
// in one region, 2.12 ms for two in different regions, a few in between for
// the reference's neighbours (2.31, 2.49) and for pieces straddling a
// boundary, and one faster outlier (1.89).  Built with hipcc for the HIP
// types in the header; it makes no HIP runtime calls.
//
//   test_regions            prints "<n> checks, <f> failed"
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "ck_regions.hpp"

namespace {

int g_checks = 0, g_failed = 0;

void check(bool ok, const char* what) {
    ++g_checks;
    if (!ok) {
        ++g_failed;
        std::printf("FAIL %s\n", what);
    }
}

bool near(float a, float b, float tol) { return std::fabs(a - b) <= tol; }

std::vector<size_t> all_but_first(size_t n) {
    std::vector<size_t> v;
    for (size_t j = 1; j < n; ++j) v.push_back(j);
    return v;
}

void test_split_modes() {
    std::mt19937 rng(7);
    std::normal_distribution<float> noise(0.f, 0.005f);
    // 36 pieces in the reference's region, 90 elsewhere, t[0] unused (0).
    std::vector<float> t(1, 0.f);
    for (int i = 0; i < 36; ++i) t.push_back(2.64f + noise(rng));
    for (int i = 0; i < 90; ++i) t.push_back(2.12f + noise(rng));
    float f = 0, s = 0;
    check(ck::split_modes(t, all_but_first(t.size()), 1.12f, &f, &s), "clean bimodal splits");
    check(near(f, 2.12f, 0.02f) && near(s, 2.64f, 0.02f), "clean bimodal modes");
    // Neighbours of the reference, a straddling piece and a fast outlier move neither mode.
    std::vector<float> u = t;
    u[1] = 2.31f;
    u[2] = 2.49f;
    u[3] = 1.89f;
    u.push_back(2.40f);
    check(ck::split_modes(u, all_but_first(u.size()), 1.12f, &f, &s), "outliers: still splits");
    check(near(f, 2.12f, 0.02f) && near(s, 2.64f, 0.02f), "outliers: modes unmoved");
    // The reference's region small (3 of 40): still two modes.
    std::vector<float> w(1, 0.f);
    for (int i = 0; i < 3; ++i) w.push_back(2.64f);
    for (int i = 0; i < 37; ++i) w.push_back(2.12f);
    check(ck::split_modes(w, all_but_first(w.size()), 1.12f, &f, &s) && near(s, 2.64f, 1e-3f), "small slow side");
    // One piece alone on a side is not a mode.
    std::vector<float> one(1, 0.f);
    one.push_back(2.64f);
    for (int i = 0; i < 20; ++i) one.push_back(2.12f);
    check(!ck::split_modes(one, all_but_first(one.size()), 1.12f, &f, &s), "a single slow piece is no mode");
    // One region only: no split.
    std::vector<float> uni(1, 0.f);
    for (int i = 0; i < 50; ++i) uni.push_back(2.64f + noise(rng));
    check(!ck::split_modes(uni, all_but_first(uni.size()), 1.12f, &f, &s), "unimodal: no split");
    // Two groups closer than the ratio: no split.
    std::vector<float> close(1, 0.f);
    for (int i = 0; i < 20; ++i) close.push_back(2.12f);
    for (int i = 0; i < 20; ++i) close.push_back(2.30f);
    check(!ck::split_modes(close, all_but_first(close.size()), 1.12f, &f, &s), "modes closer than the ratio");
    // Too few times, and zero (unmeasured) times ignored.
    std::vector<float> few = {0.f, 2.64f, 2.12f, 2.12f};
    check(!ck::split_modes(few, all_but_first(few.size()), 1.12f, &f, &s), "fewer than four times");
    std::vector<float> zeros = {0.f, 2.64f, 2.64f, 0.f, 0.f, 2.12f, 2.12f};
    check(ck::split_modes(zeros, all_but_first(zeros.size()), 1.12f, &f, &s) && near(f, 2.12f, 1e-3f),
          "zero times are skipped");
    // Only the listed indices count.
    std::vector<size_t> idx;
    for (size_t j = 37; j < t.size(); ++j) idx.push_back(j);
    check(!ck::split_modes(t, idx, 1.12f, &f, &s), "an index subset of one mode: no split");
}

ck::RegionPool pool_of(const std::vector<int32_t>& region) {
    ck::RegionPool p;
    p.region = region;
    p.ready = region.size();
    return p;
}

void test_interleave() {
    const ck::RegionPool p = pool_of({0, 0, 0, 1, 1, 2, -1, 2, 1, 0});
    check(ck::interleave(p, 7) == std::vector<size_t>({0, 3, 5, 1, 4, 7, 2}), "regions in turn");
    check(ck::interleave(p, 10) == std::vector<size_t>({0, 3, 5, 1, 4, 7, 2, 8, 9, 6}),
          "exhausted regions skipped, unclear last");
    check(ck::interleave(p, 20).size() == 10, "no more pieces than the pool holds");
    check(ck::interleave(p, 0).empty(), "an empty range");
    const ck::RegionPool two = pool_of({0, 0, 1, 1, 0, 1});
    check(ck::interleave(two, 6) == std::vector<size_t>({0, 2, 1, 3, 4, 5}), "two regions alternate");
    // Every region's share of a balanced interleave differs by at most one.
    std::vector<int32_t> big;
    for (int i = 0; i < 90; ++i) big.push_back(0);
    for (int i = 0; i < 96; ++i) big.push_back(1);
    for (int i = 0; i < 70; ++i) big.push_back(2);
    const ck::RegionPool b = pool_of(big);
    const std::vector<size_t> order = ck::interleave(b, 192);
    int cnt[3] = {0, 0, 0};
    for (size_t j : order) ++cnt[big[j]];
    check(cnt[0] == 64 && cnt[1] == 64 && cnt[2] == 64, "192 pieces: 64 per region");
    bool prefix_ok = true;  // every prefix of three is one piece per region
    for (size_t i = 0; i + 3 <= order.size(); i += 3)
        prefix_ok &= big[order[i]] == 0 && big[order[i + 1]] == 1 && big[order[i + 2]] == 2;
    check(prefix_ok, "192 pieces: A, B, C repeating");
}

}  // namespace

int main() {
    test_split_modes();
    test_interleave();
    std::printf("%d checks, %d failed\n", g_checks, g_failed);
    return g_failed ? 1 : 0;
}
