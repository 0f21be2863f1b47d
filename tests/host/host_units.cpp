// Host-only unit checks of the C ABI's bookkeeping, built with AddressSanitizer
// and UndefinedBehaviorSanitizer (tests/host/Makefile) and run by
// tests/test_host_units.py on the CPU:
//  - colorspace_key ignores the 4 padding bytes of Vship_Colorspace_t and
//    separates descriptors that differ in any named field;
//  - HandleTable ids: reuse after remove, stale ids fail, a reference held by
//    a call outlives remove();
//  - the front-end's tile division: (n * magic) >> shift equals n / d for
//    every divisor up to 16384 (tiles of images up to 2^20 columns) on the
//    numerators where a wrong reciprocal fails first (multiples of d and
//    their neighbours, up to 2^31 - 1) and on random ones.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "../../vship_amd/csrc/colorspace_key.hpp"
#include "../../vship_amd/csrc/handle_table.hpp"
#include "../../vship_amd/csrc/tile_div.hpp"

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                 \
        }                                                                 \
    } while (0)

static Vship_Colorspace_t filled(unsigned char pad)
{
    Vship_Colorspace_t cs;
    std::memset(&cs, pad, sizeof(cs));  // padding (and everything) = pad
    cs.width = 1920;
    cs.height = 1080;
    cs.target_width = -1;
    cs.target_height = -1;
    cs.sample = Vship_SampleUINT8;
    cs.range = Vship_RangeLimited;
    cs.subsampling.subw = 1;
    cs.subsampling.subh = 1;
    cs.chromaLocation = Vship_ChromaLoc_Left;
    cs.colorFamily = Vship_ColorYUV;
    cs.YUVMatrix = Vship_MATRIX_BT709;
    cs.transferFunction = Vship_TRC_BT709;
    cs.primaries = Vship_PRIMARIES_BT709;
    cs.crop.top = cs.crop.bottom = cs.crop.left = cs.crop.right = 0;
    return cs;
}

static void test_colorspace_key()
{
    static_assert(sizeof(Vship_Colorspace_t) == 88, "reference layout: 84 bytes of members, 4 of padding");
    const Vship_Colorspace_t a = filled(0x00), b = filled(0xa5);
    CHECK(std::memcmp(&a, &b, sizeof(a)) != 0);  // the raw bytes differ (padding)
    CHECK(vship::colorspace_key(a) == vship::colorspace_key(b));
    // every named field separates keys
    for (int f = 0; f < 17; f++) {
        Vship_Colorspace_t c = filled(0);
        switch (f) {
        case 0: c.width++; break;
        case 1: c.height++; break;
        case 2: c.target_width = 640; break;
        case 3: c.target_height = 360; break;
        case 4: c.sample = Vship_SampleUINT10; break;
        case 5: c.range = Vship_RangeFull; break;
        case 6: c.subsampling.subw = 0; break;
        case 7: c.subsampling.subh = 0; break;
        case 8: c.chromaLocation = Vship_ChromaLoc_Center; break;
        case 9: c.colorFamily = Vship_ColorRGB; break;
        case 10: c.YUVMatrix = Vship_MATRIX_BT2020_NCL; break;
        case 11: c.transferFunction = Vship_TRC_PQ; break;
        case 12: c.primaries = Vship_PRIMARIES_BT2020; break;
        case 13: c.crop.top = 2; break;
        case 14: c.crop.bottom = 2; break;
        case 15: c.crop.left = 2; break;
        case 16: c.crop.right = 2; break;
        }
        CHECK(vship::colorspace_key(c) != vship::colorspace_key(a));
    }
}

struct Counted {
    static int alive;
    int value;
    explicit Counted(int v) : value(v) { alive++; }
    ~Counted() { alive--; }
};
int Counted::alive = 0;

static void test_handle_table()
{
    vship::HandleTable<Counted> t;
    const unsigned a = t.insert(std::make_shared<Counted>(1), "a");
    const unsigned b = t.insert(std::make_shared<Counted>(2), "b");
    CHECK(a == 0 && b == 1);
    std::shared_ptr<Counted> p;
    CHECK(t.lookup(b, &p) && p->value == 2);
    std::string e;
    CHECK(t.get_error(a, &e) && e == "a");
    CHECK(!t.lookup(7, &p));
    // a reference held by a running call keeps the object alive past remove()
    std::shared_ptr<Counted> held;
    CHECK(t.lookup(a, &held));
    std::shared_ptr<Counted> out;
    CHECK(t.remove(a, &out));
    out.reset();
    CHECK(Counted::alive == 2 && held->value == 1);
    held.reset();
    CHECK(Counted::alive == 1);
    // a stale id fails everywhere, then its slot is reused
    CHECK(!t.lookup(a, &p) && !t.remove(a, &out) && !t.set_error(a, "x") && !t.get_error(a, &e));
    const unsigned c = t.insert(nullptr, "failed init");
    CHECK(c == a);
    CHECK(t.lookup(c, &p) && !p);  // live id of a failed init: no object
    CHECK(t.remove(b, &out));
    out.reset();
    p.reset();
    CHECK(Counted::alive == 0);
}

static void test_tile_div()
{
    uint64_t rng = 0x9e3779b97f4a7c15ull;
    auto next = [&] {
        rng ^= rng << 13;
        rng ^= rng >> 7;
        rng ^= rng << 17;
        return rng;
    };
    const uint32_t top = 0x7fffffffu;
    for (int d = 1; d <= 16384; d++) {
        uint32_t magic;
        int32_t shift;
        vship::tile_div_magic(d, &magic, &shift);
        auto ok = [&](uint32_t n) { return vship::tile_div_apply(n, magic, shift) == n / (uint32_t)d; };
        CHECK(ok(0) && ok(1) && ok(top) && ok(top - 1));
        // the largest multiples of d below 2^31 and the small ones, +-1
        const uint32_t kmax = top / (uint32_t)d;
        for (uint32_t k : {1u, 2u, 3u, kmax - 2, kmax - 1, kmax}) {
            if (k == 0 || k > kmax) continue;
            const uint32_t m = k * (uint32_t)d;
            CHECK(ok(m) && ok(m - 1) && (m == top || ok(m + 1)));
        }
        for (int i = 0; i < 64; i++) CHECK(ok((uint32_t)(next() & top)));
    }
}

int main()
{
    test_colorspace_key();
    test_handle_table();
    test_tile_div();
    std::puts("host units: ok");
    return 0;
}
