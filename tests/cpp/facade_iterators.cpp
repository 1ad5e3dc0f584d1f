// Host-side checks of the C++ facade (include/amethyst.hpp): bit_iterator and
// slab_iterator/sentinel semantics, and that every templated wrapper compiles.
// Built and run by tests/test_facade.py with g++ (no GPU calls are made).
#include <cstdio>
#include <cstring>
#include <vector>

#include "amethyst.hpp"

using namespace amethyst;

static int fails = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
            fails++;                                               \
        }                                                          \
    } while (0)

// instantiate the wrappers (compile-only: not called)
template int binary_op_slabs<11, 12, AM_OP_ADD>(const uint64_t*, const uint64_t*, uint64_t*, size_t, void*);
template int binary_op_nogaps<11, 12, AM_OP_ADD>(const uint64_t*, const uint64_t*, uint64_t*, size_t, void*);
template int binary_op_baseline<11, AM_OP_MUL>(const uint64_t*, const uint64_t*, uint64_t*, size_t, void*);
template int filter_slabs<11>(const uint64_t*, uint64_t*, size_t, am_predicate, void*);
template int filter_nogaps<17>(const uint64_t*, uint64_t*, size_t, am_predicate, void*);
template int filter_baseline<32>(const uint64_t*, uint64_t*, size_t, am_predicate, void*);
template int compressstore_slabs<11>(const uint8_t*, const uint64_t*, uint64_t*, size_t, uint64_t*, const Workspace&, void*);
template int hop_slabs<4>(const uint64_t*, size_t, uint64_t*, const Workspace&, void*);
template int hop_nogaps<11>(const uint64_t*, size_t, uint64_t*, const Workspace&, void*);
template int groupby_sm_array<10>(const uint64_t*, const uint64_t*, size_t, uint64_t*, uint64_t*, uint64_t*,
                                  const Workspace&, am_groupby_mode, void*);

int main(int argc, char** argv)
{
    // bit_iterator over bytes given on the command line (hex), printed as bits
    std::vector<uint8_t> bytes;
    for (int i = 1; i < argc; i++) bytes.push_back((uint8_t)std::strtoul(argv[i], nullptr, 16));
    bit_iterator it{bytes.data()};
    for (size_t i = 0; i < bytes.size() * 8; i++) std::printf("%d", (int)it[i]);
    std::printf("\n");
    bit_iterator w{bytes.data()};
    for (size_t i = 0; i < bytes.size() * 8; i++) {
        CHECK(*w == it[i]);
        w++;
    }
    CHECK((it + 9).idx == 9 && ((it + 9) - 9).idx == 0);

    // slab_iterator / sentinel: masked reads and writes over tight arrays
    std::vector<uint16_t> t12 = {0xFFFF, 0x1234, 0x0ABC};
    slab_iterator<12> s{t12.data()};
    CHECK((uint16_t)s[0] == 0x0FFF);
    CHECK((uint16_t)*(s + 1) == 0x0234);
    s[2] = (uint16_t)0xF123;
    CHECK(t12[2] == 0x0123);
    CHECK(s[1] < s[0]);
    CHECK(((s + 2) - s) == 2);
    std::vector<uint8_t> t3 = {0xFF, 0x05};
    slab_iterator<3> s3{t3.data()};
    CHECK((uint8_t)s3[0] == 7 && (uint8_t)s3[1] == 5);
    static_assert(std::is_same<tight_type<9>::type, uint16_t>::value, "tight 9");
    static_assert(std::is_same<tight_type<17>::type, uint32_t>::value, "tight 17");
    static_assert(std::is_same<tight_type<33>::type, uint64_t>::value, "tight 33");
    static_assert(add_outbits(11) == 12 && mul_outbits(40) == 64 && gen_bitmask(11) == 2047, "widths");
    std::printf(fails ? "FAILED\n" : "OK\n");
    return fails ? 1 : 0;
}
