// Every wrapper of the C++ facade (include/amethyst.hpp) called once on the
// GPU and checked against the CPU oracle (oracle/liboracle.so, linked as the
// checker): binary ops in both operator forms (library enum and the
// reference's operator functions), filters, compressstore + count, the
// reference's compressstore_baseline_cub, sums and the group-by.
// Built on the host by tests/cpp/Makefile (g++ against the HIP runtime API);
// run by tests/test_facade.py::test_facade_wrappers_on_gpu. Prints one
// "ok <wrapper>" line per check and "ALL OK" at the end; exit status 1 on
// any mismatch.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "amethyst.hpp"

extern "C" {
uint64_t or_packed_words(int fmt, int bits, uint64_t n);
void or_pack(int fmt, int bits, const uint64_t* vals, uint64_t n, uint64_t* out);
void or_filter(int fmt, int bits, const uint64_t* in, uint64_t n, int pop, uint64_t c, uint64_t* mask);
uint64_t or_compressstore(int ifmt, int ibits, const uint64_t* in, uint64_t n, const uint64_t* mask, int ofmt,
                          int obits, uint64_t* out, uint64_t out_cap, uint64_t* overflow);
void or_binary(int op, int fmt, int bits, const uint64_t* a, const uint64_t* b, uint64_t n, int ofmt, int obits,
               uint64_t* out);
uint64_t or_sum(int fmt, int bits, const uint64_t* in, uint64_t n);
uint64_t or_groupby(int kfmt, int kbits, const uint64_t* keys, int vfmt, int vbits, const uint64_t* vals, uint64_t n,
                    int use_pred, int pop, uint64_t pc, uint64_t* sums, uint64_t* counts);
}

using namespace amethyst;

static int fails = 0;
static void report(const char* what, bool ok)
{
    std::printf("%s %s\n", ok ? "ok" : "FAIL", what);
    std::fflush(stdout);
    if (!ok) fails++;
}

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64()
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Dev {
    uint64_t* p = nullptr;
    size_t words = 0;
    explicit Dev(size_t w) : words(w ? w : 1)
    {
        if (hipMalloc(&p, words * 8) != hipSuccess) p = nullptr;
        else (void)hipMemset(p, 0, words * 8);
    }
    Dev(const std::vector<uint64_t>& h) : Dev(h.size()) { (void)hipMemcpy(p, h.data(), h.size() * 8, hipMemcpyHostToDevice); }
    ~Dev() { (void)hipFree(p); }
    std::vector<uint64_t> get(size_t w) const
    {
        std::vector<uint64_t> h(w);
        (void)hipMemcpy(h.data(), p, w * 8, hipMemcpyDeviceToHost);
        return h;
    }
};

static std::vector<uint64_t> values(size_t n, int bits)
{
    std::vector<uint64_t> v(n);
    for (auto& x : v) x = next_u64() & gen_bitmask(bits);
    return v;
}
static std::vector<uint64_t> packed(int fmt, int bits, const std::vector<uint64_t>& v)
{
    std::vector<uint64_t> w(or_packed_words(fmt, bits, v.size()) + 1, 0);
    or_pack(fmt, bits, v.data(), v.size(), w.data());
    w.resize(or_packed_words(fmt, bits, v.size()));
    return w;
}

int main()
{
    if (hipSetDevice(0) != hipSuccess) {
        std::printf("no GPU\n");
        return 2;
    }
    const size_t n = 100003;
    const std::vector<uint64_t> a11 = values(n, 11), b11 = values(n, 11);
    const auto sa = packed(AM_FMT_SLAB, 11, a11), sb = packed(AM_FMT_SLAB, 11, b11);
    const auto ga = packed(AM_FMT_BITSTREAM, 11, a11), gb = packed(AM_FMT_BITSTREAM, 11, b11);
    const auto ba = packed(AM_FMT_BASELINE, 11, a11), bb = packed(AM_FMT_BASELINE, 11, b11);
    Dev dsa(sa), dsb(sb), dga(ga), dgb(gb), dba(ba), dbb(bb);

    // ---- binary ops: slabs (enum and operator-function forms), nogaps, baseline
    {
        const size_t w = or_packed_words(AM_FMT_SLAB, 12, n);
        std::vector<uint64_t> exp(w);
        or_binary(AM_OP_ADD, AM_FMT_SLAB, 11, sa.data(), sb.data(), n, AM_FMT_SLAB, 12, exp.data());
        Dev o1(w), o2(w);
        report("binary_op_slabs<11,12,AM_OP_ADD>", binary_op_slabs<11, 12, AM_OP_ADD>(dsa.p, dsb.p, o1.p, n) == AM_OK &&
                                                       o1.get(w) == exp);
        report("binary_op_slabs<11,12,op_add<12>,1024>",
               binary_op_slabs<11, 12, op_add<12>, 1024>(dsa.p, dsb.p, o2.p, n, 512) == AM_OK && o2.get(w) == exp);
    }
    {
        const size_t w = or_packed_words(AM_FMT_BITSTREAM, 11, n);
        std::vector<uint64_t> exp(w);
        or_binary(AM_OP_XOR, AM_FMT_BITSTREAM, 11, ga.data(), gb.data(), n, AM_FMT_BITSTREAM, 11, exp.data());
        Dev o1(w), o2(w);
        report("binary_op_nogaps<11,11,AM_OP_XOR>", binary_op_nogaps<11, 11, AM_OP_XOR>(dga.p, dgb.p, o1.p, n) == AM_OK &&
                                                        o1.get(w) == exp);
        report("binary_op_nogaps<11,11,op_xor<11>,256>",
               binary_op_nogaps<11, 11, op_xor<11>, 256>(dga.p, dgb.p, o2.p, n) == AM_OK && o2.get(w) == exp);
    }
    {
        std::vector<uint64_t> exp(n);
        or_binary(AM_OP_MUL, AM_FMT_BASELINE, 11, ba.data(), bb.data(), n, AM_FMT_BASELINE, 22, exp.data());
        Dev o1(n), o2(n);
        report("binary_op_baseline<11,AM_OP_MUL,22>",
               binary_op_baseline<11, AM_OP_MUL, 22>(dba.p, dbb.p, o1.p, n) == AM_OK && o1.get(n) == exp);
        report("binary_op_baseline<11,op_mul<22>>", binary_op_baseline<11, op_mul<22>>(dba.p, dbb.p, o2.p, n) == AM_OK &&
                                                        o2.get(n) == exp);
        // the reference's configured operator: op_mul at op_mul_outbits, into a baseline column
        std::vector<uint64_t> exp2(n);
        or_binary(AM_OP_MOD, AM_FMT_BASELINE, 11, ba.data(), bb.data(), n, AM_FMT_BASELINE, 11, exp2.data());
        Dev o3(n);
        report("binary_op_baseline<11,op_mod<11>>", binary_op_baseline<11, op_mod<11>>(dba.p, dbb.p, o3.p, n) == AM_OK &&
                                                        o3.get(n) == exp2);
    }

    // ---- filters (the reference predicate v >= 2^(bits-1), and v < 700)
    const size_t mw = (n + 63) / 64;
    std::vector<uint64_t> m_ref(mw), m_lt(mw);
    or_filter(AM_FMT_SLAB, 11, sa.data(), n, AM_PRED_GE, 1024, m_ref.data());
    or_filter(AM_FMT_SLAB, 11, sa.data(), n, AM_PRED_LT, 700, m_lt.data());
    {
        Dev m1(mw), m2(mw), m3(mw);
        report("filter_slabs<11>", filter_slabs<11>(dsa.p, m1.p, n) == AM_OK && m1.get(mw) == m_ref);
        report("filter_nogaps<11>", filter_nogaps<11>(dga.p, m2.p, n, am_predicate{AM_PRED_LT, 0, 700}) == AM_OK &&
                                        m2.get(mw) == m_lt);
        report("filter_baseline<11>", filter_baseline<11>(dba.p, m3.p, n) == AM_OK && m3.get(mw) == m_ref);
    }

    // ---- compressstore by mask (slab -> slab) and its count
    Workspace ws;
    ws.bytes = am_workspace_bytes(n);
    Dev dws((ws.bytes + 7) / 8);
    ws.ptr = dws.p;
    Dev dstats(4);
    {
        std::vector<uint64_t> exp(or_packed_words(AM_FMT_SLAB, 11, n) + 1), ovf(1);
        const uint64_t k = or_compressstore(AM_FMT_SLAB, 11, sa.data(), n, m_lt.data(), AM_FMT_SLAB, 11, exp.data(), n,
                                            ovf.data());
        Dev mask(m_lt), out(or_packed_words(AM_FMT_SLAB, 11, n));
        const int st = compressstore_slabs<11>((const uint8_t*)mask.p, dsa.p, out.p, n, dstats.p, ws);
        const size_t kk = compressstore_slabs_count(dstats.p);
        const size_t w = or_packed_words(AM_FMT_SLAB, 11, k);
        auto got = out.get(w);
        report("compressstore_slabs<11> + compressstore_slabs_count",
               st == AM_OK && kk == k && std::equal(got.begin(), got.end(), exp.begin()));
    }
    // ---- the reference's compressstore_baseline_cub (whole uint64 elements)
    {
        std::vector<uint64_t> words(n);
        for (auto& x : words) x = next_u64();
        std::vector<uint64_t> exp(n), ovf(1);
        const uint64_t k = or_compressstore(AM_FMT_BASELINE, 64, words.data(), n, m_lt.data(), AM_FMT_BASELINE, 64,
                                            exp.data(), n, ovf.data());
        Dev in(words), mask(m_lt), out(n);
        const size_t kk = compressstore_baseline_cub<11>((const uint8_t*)mask.p, in.p, out.p, n, dstats.p);
        auto got = out.get(k);
        report("compressstore_baseline_cub<11>", kk == k && std::equal(got.begin(), got.end(), exp.begin()));
    }

    // ---- sums
    {
        Dev s1(1), s2(1);
        const uint64_t e1 = or_sum(AM_FMT_SLAB, 11, sa.data(), n), e2 = or_sum(AM_FMT_BITSTREAM, 11, ga.data(), n);
        const int st1 = hop_slabs<11>(dsa.p, n, s1.p, ws), st2 = hop_nogaps<11>(dga.p, n, s2.p, ws);
        report("hop_slabs<11>", st1 == AM_OK && s1.get(1)[0] == e1);
        report("hop_nogaps<11>", st2 == AM_OK && s2.get(1)[0] == e2);
    }

    // ---- group-by over 10-bit keys, the reference's SUM mod 2^bits (AM_GB_REF)
    {
        const std::vector<uint64_t> keys = values(n, 10), vals = values(n, 10);
        const auto sk = packed(AM_FMT_SLAB, 10, keys), sv = packed(AM_FMT_SLAB, 10, vals);
        std::vector<uint64_t> sums(1024), counts(1024);
        or_groupby(AM_FMT_SLAB, 10, sk.data(), AM_FMT_SLAB, 10, sv.data(), n, 0, 0, 0, sums.data(), counts.data());
        Dev dk(sk), dv(sv), ok(1024), os(1024), res(4);
        const int st = groupby_sm_array<10>(dk.p, dv.p, n, ok.p, os.p, res.p, ws);
        const auto r = res.get(4);
        bool good = st == AM_OK && r[1] == 10 && r[3] == n;
        std::vector<uint64_t> ek, es;
        for (uint32_t g = 0; g < 1024; g++)
            if (counts[g]) {
                ek.push_back(g);
                es.push_back(sums[g] & 1023);
            }
        good = good && r[0] == ek.size();
        good = good && ok.get(or_packed_words(AM_FMT_SLAB, 10, ek.size())) == packed(AM_FMT_SLAB, 10, ek);
        good = good && os.get(or_packed_words(AM_FMT_SLAB, 10, es.size())) == packed(AM_FMT_SLAB, 10, es);
        report("groupby_sm_array<10>", good);
    }

    // every check above must have run on valid device buffers
    report("device allocations", dsa.p && dsb.p && dga.p && dgb.p && dba.p && dbb.p && dws.p && dstats.p);
    if (fails) {
        std::printf("%d FAILED\n", fails);
        return 1;
    }
    std::printf("ALL OK\n");
    return 0;
}
