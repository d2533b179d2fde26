// Host test of the packed residual-entry tables (Entries44, Entries48 in
// ck_chain.hpp): every read and every write is checked against a bit-level
// model of the table, in which entry b of width W holds bits [W*b, W*b + W) of
// the table read as a little-endian bit string.  After each write the whole
// table image, slack bytes included, must equal the model's; a write must not
// change one bit outside its entry.  Built with hipcc (the handles are
// host + device code); it makes no HIP runtime calls.
//
//   test_res_entries            prints "<n> checks, <f> failed"
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "ck_chain.hpp"

namespace {

int g_checks = 0, g_failed = 0;

void check(bool ok, const char* what, unsigned width, size_t m, size_t b) {
    ++g_checks;
    if (!ok) {
        ++g_failed;
        if (g_failed <= 20) std::printf("FAIL %s: width %u, %zu entries, entry %zu\n", what, width, m, b);
    }
}

// The model: a byte image, written bit by bit.
struct Model {
    unsigned width;
    std::vector<uint8_t> img;
    uint64_t get(size_t b) const {
        uint64_t v = 0;
        for (unsigned j = 0; j < width; ++j) {
            const size_t bit = b * width + j;
            v |= uint64_t((img[bit >> 3] >> (bit & 7)) & 1u) << j;
        }
        return v;
    }
    void set(size_t b, uint64_t v) {
        for (unsigned j = 0; j < width; ++j) {
            const size_t bit = b * width + j;
            const uint8_t mask = uint8_t(1u << (bit & 7));
            img[bit >> 3] = ((v >> j) & 1u) ? (img[bit >> 3] | mask) : (img[bit >> 3] & ~mask);
        }
    }
};

size_t table_bytes(unsigned width, size_t m) {
    return width == 44 ? ck::Entries44::bytes(m) : m * 6 + 8;
}

template <class T>
void run(unsigned width, size_t m, uint64_t seed, int ops) {
    const uint64_t full = (1ull << width) - 1;
    const size_t bytes = table_bytes(width, m);
    check(bytes * 8 >= m * width + 32, "table holds its entries and a dword of slack", width, m, 0);
    // The table as the device sees it: 4-byte aligned (a 256-byte aligned slot).
    std::vector<uint32_t> store((bytes + 3) / 4 + 1);
    char* base = reinterpret_cast<char*>(store.data());
    std::memset(base, 0xFF, bytes);  // the clear
    Model model{width, std::vector<uint8_t>(bytes, 0xFF)};
    const T t{base};
    for (size_t b = 0; b < m; ++b) check(uint64_t(t[uint32_t(b)]) == ~0ull, "cleared entry reads empty", width, m, b);

    std::mt19937_64 rng(seed);
    for (int i = 0; i < ops; ++i) {
        const size_t b = rng() % m;
        uint64_t v = rng() & full;
        if (v == full) v ^= 1;  // all-ones is the empty entry, never stored
        auto&& slot = t[uint32_t(b)];
        const int kind = int(rng() % 4);
        if (kind) {  // the chains' form: read, then write through the same handle
            const uint64_t want = model.get(b);
            const uint64_t got = slot;
            check(got == (want == full ? ~0ull : want), "read", width, m, b);
            if (kind == 3) v = full;  // an entry emptied again (writes the empty pattern)
            slot = v;
        } else {  // a write before any read
            slot = v;
        }
        model.set(b, v & full);
        check(std::memcmp(base, model.img.data(), bytes) == 0, "table image after a write", width, m, b);
    }
    for (size_t b = 0; b < m; ++b) {
        const uint64_t want = model.get(b);
        check(uint64_t(t[uint32_t(b)]) == (want == full ? ~0ull : want), "final read", width, m, b);
    }
}

}  // namespace

int main() {
    // Every entry offset within a dword recurs with period 8 (44-bit) or 2
    // (48-bit); small tables put entries at the very end of the image.
    for (size_t m = 1; m <= 40; ++m) {
        run<ck::Entries44>(44, m, 1000 + m, 400);
        run<ck::Entries48>(48, m, 2000 + m, 400);
    }
    run<ck::Entries44>(44, 24576, 7, 20000);
    run<ck::Entries48>(48, 24576, 8, 20000);
    run<ck::Entries44>(44, 1363149 / 64, 9, 20000);
    std::printf("%d checks, %d failed\n", g_checks, g_failed);
    return g_failed ? 1 : 0;
}
