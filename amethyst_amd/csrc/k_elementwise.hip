// k_elementwise.hip -- format conversion and packed binary operators.
//
// Reference: kernel_binary_op_baseline / _nogaps / _slabs
// (src/kernels/binary_op_baseline.cuh:4-13, binary_op_nogaps.cuh:5-89,
// binary_op_slabs.cuh:7-67). The reference stages input slabs in shared
// memory and has each thread assemble one output word; its operator is a
// compile-time function pointer masked to the out-width (src/config.cuh:5-42).
//
// Here one lane also owns one output word (so every store is a coalesced,
// whole-word write and dead/padding bits are written as zero). Columns whose
// input and output words hold the same elements go word-per-lane
// (k_binary_aligned); unaligned slab pairs and bitstream inputs are staged in
// LDS per block, and a lane reads one 64-bit window of fields per input
// column and output word, every field then a shift by a wave-uniform amount
// (k_binary_stream); pairs with a container side read two words per lane
// from global memory (k_binary_words). Conversion and the remaining cases
// decode each element through a per-block TileReader (k_elementwise). Any
// input format x any output format; the operator is a template argument so
// the switch folds away.
#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

__device__ inline uint64_t udiv(uint64_t l, uint64_t r) { return r ? l / r : ~0ull; }
__device__ inline uint64_t umod(uint64_t l, uint64_t r) { return r ? l % r : l; }

template <int OP>
__device__ inline uint64_t binop(uint64_t l, uint64_t r)
{
    if (OP == AM_OP_ADD) return l + r;
    if (OP == AM_OP_MUL) return l * r;
    if (OP == AM_OP_MOD) return umod(l, r);
    if (OP == AM_OP_DIV) return udiv(l, r);
    if (OP == AM_OP_MATH) return udiv(l * l + r * r + 4 * l, l * l + r * r + 5 * l);
    if (OP == AM_OP_BOOL_SIMPLE) return (l || r) ? 1 : 0;
    if (OP == AM_OP_BOOL_COMPLEX) return ((l && !r) || (!l && r)) ? 1 : 0;
    if (OP == AM_OP_AND) return l & r;
    if (OP == AM_OP_OR) return l | r;
    if (OP == AM_OP_XOR) return l ^ r;
    return l - r; // AM_OP_SUB
}

constexpr int OP_COPY = -1; // conversion: out = in & mask(out.bits)
constexpr uint32_t EW_UNITS = 1024; // output units (elements or words) per block

template <int AK, int OK, int OP>
__global__ __launch_bounds__(256) void k_elementwise(Col a, Col b, Col out)
{
    const uint64_t n = out.n;
    const uint32_t bo = out.bits;
    const uint64_t om = out.mask;
    const uint64_t u0 = (uint64_t)blockIdx.x * EW_UNITS;
    uint64_t t0; // first element touched by this block
    if (OK == K_STRIDED) t0 = u0;
    else if (OK == K_SLAB) t0 = u0 * out.eps;
    else t0 = (u0 * 64) / bo;
    if (t0 >= n) return;
    TileReader<AK> ra(a, t0);
    TileReader<AK> rb(OP == OP_COPY ? a : b, t0);
    auto value = [&](uint32_t i) -> uint64_t {
        if (OP == OP_COPY) return ra.get(i) & om;
        return binop<OP>(ra.get(i), rb.get(i)) & om;
    };
    for (uint32_t r = 0; r < EW_UNITS / 256; r++) {
        const uint64_t u = u0 + r * 256 + threadIdx.x;
        if (OK == K_STRIDED) {
            if (u >= n) break;
            uint64_t v = value((uint32_t)(u - t0));
            switch (out.cbytes) {
                case 1: st_nt((uint8_t*)out.p + u, (uint8_t)v); break;
                case 2: st_nt((uint16_t*)out.p + u, (uint16_t)v); break;
                case 4: st_nt((uint32_t*)out.p + u, (uint32_t)v); break;
                default: st_nt(out.p + u, v);
            }
        }
        else if (OK == K_SLAB) {
            const uint64_t e0 = u * out.eps;
            if (e0 >= n) break;
            const uint32_t cnt = (uint32_t)(n - e0 < out.eps ? n - e0 : out.eps);
            uint64_t word = 0;
            for (uint32_t s = 0; s < cnt; s++) word |= value((uint32_t)(e0 - t0) + s) << (s * bo);
            st_nt(out.p + u, word);
        }
        else {
            const uint64_t b0 = u * 64;
            if (b0 >= n * bo) break;
            uint64_t e0 = b0 / bo, e1 = (b0 + 63) / bo + 1;
            if (e1 > n) e1 = n;
            uint64_t word = 0;
            for (uint64_t e = e0; e < e1; e++) {
                int64_t pos = (int64_t)(e * bo) - (int64_t)b0;
                uint64_t v = value((uint32_t)(e - t0));
                word |= pos >= 0 ? (v << pos) : (v >> (-pos));
            }
            st_nt(out.p + u, word);
        }
    }
}

// Word-per-lane binary op for slab and container columns (a container
// column is a slab of fw = 8*cbytes-bit fields). Lane q builds output word q
// from its eps_o elements. ALIGNED: both inputs hold the same elements in
// word q (eps_i == eps_o), so the lane reads one word of each and applies the
// operator field by field at wave-uniform shifts. Otherwise the lane's first
// input word and field come from a per-block base and a 32-bit reciprocal
// multiply, and it reads two words of each input (eps_o <= eps_i: an output
// word's elements span at most two input words), joined into one window of
// fields.
struct BinWords {
    Col a, b, out;
    uint32_t fi, eps_i, fo, eps_o;
    uint64_t words_i, words_o;
    uint64_t fmask; // aligned bitwise ops at equal width: every field's bits
};

// Aligned columns (input and output words hold the same elements): a thread
// takes BA_PER words, loads them all before computing any (more bytes in
// flight per wave than one word per thread). AND/OR/XOR at equal width work
// on whole words under the fields' mask (dead bits stay zero).
constexpr uint32_t BA_PER_DEFAULT = 4; // words per thread (2, 4 or 8: AM_KNOB_BINARY_WORDS)

template <int OP, uint32_t BA_PER>
__global__ __launch_bounds__(256) void k_binary_aligned(BinWords g)
{
    constexpr bool BITWISE = OP == AM_OP_AND || OP == AM_OP_OR || OP == AM_OP_XOR;
    const uint64_t q0 = (uint64_t)blockIdx.x * (256 * BA_PER) + threadIdx.x;
    const uint64_t n = g.out.n;
    const uint32_t fi = g.fi, fo = g.fo, eo = g.eps_o;
    const uint64_t im = g.a.mask, om = g.out.mask;
    uint64_t x[BA_PER], y[BA_PER];
#pragma unroll
    for (uint32_t k = 0; k < BA_PER; k++) {
        const uint64_t q = q0 + 256 * k;
        x[k] = q < g.words_o ? __builtin_nontemporal_load(g.a.p + q) : 0;
        y[k] = q < g.words_o ? __builtin_nontemporal_load(g.b.p + q) : 0;
    }
#pragma unroll
    for (uint32_t k = 0; k < BA_PER; k++) {
        const uint64_t q = q0 + 256 * k;
        if (q >= g.words_o) break;
        const uint64_t e0 = q * eo;
        const uint32_t cnt = (uint32_t)(n - e0 < eo ? n - e0 : eo);
        uint64_t word = 0;
        if (BITWISE && g.fmask) {
            const uint64_t m = cnt == eo ? g.fmask : g.fmask & lowmask(cnt * fi);
            word = binop<OP>(x[k] & m, y[k] & m) & m;
        }
        else
            for (uint32_t j = 0, si = 0, so = 0; j < eo; j++, si += fi, so += fo) {
                const uint64_t r = binop<OP>((x[k] >> si) & im, (y[k] >> si) & im) & om;
                word |= (j < cnt ? r : 0) << so;
            }
        st_nt(g.out.p + q, word);
    }
}

// Unaligned columns: a thread builds BA_PER output words (one per 256-word
// sub-block of its block), loading the input words of all of them before
// combining any.
template <int OP, uint32_t BA_PER>
__global__ __launch_bounds__(256) void k_binary_words(BinWords g)
{
    const uint64_t n = g.out.n;
    const uint32_t fi = g.fi, fo = g.fo, ei = g.eps_i, eo = g.eps_o;
    const uint64_t im = g.a.mask, om = g.out.mask;
    uint64_t x0[BA_PER], y0[BA_PER], x1[BA_PER], y1[BA_PER];
    uint32_t s0[BA_PER];
#pragma unroll
    for (uint32_t k = 0; k < BA_PER; k++) {
        // sub-block base: element u0*eo lies in input word qb at field rb
        const uint64_t u0 = ((uint64_t)blockIdx.x * BA_PER + k) * 256;
        const uint64_t q = u0 + threadIdx.x;
        x0[k] = y0[k] = x1[k] = y1[k] = 0;
        s0[k] = 0;
        if (q >= g.words_o) continue;
        const uint64_t t0 = u0 * eo;
        const uint64_t qb = t0 / ei;
        const uint32_t rb = (uint32_t)(t0 - qb * ei);
        const uint32_t rel = threadIdx.x * eo + rb; // < 2^24
        const uint32_t dq = ei == 1 ? rel : __umulhi(rel, g.a.eps_magic);
        s0[k] = rel - dq * ei;
        const uint64_t qi = qb + dq;
        x0[k] = g.a.p[qi];
        y0[k] = g.b.p[qi];
        if (s0[k] + eo > ei && qi + 1 < g.words_i) {
            x1[k] = g.a.p[qi + 1];
            y1[k] = g.b.p[qi + 1];
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < BA_PER; k++) {
        const uint64_t q = ((uint64_t)blockIdx.x * BA_PER + k) * 256 + threadIdx.x;
        if (q >= g.words_o) break;
        const uint64_t e0 = q * eo;
        const uint32_t cnt = (uint32_t)(n - e0 < eo ? n - e0 : eo);
        // the word's eo fields side by side (eo * fi <= 64 bits): the fields
        // of x0 from s0 on, then those of x1 (dead bits of x0 cleared)
        const uint32_t s = s0[k], lo = (ei - s) * fi; // s < ei; lo < 64 when x1 is needed
        const bool two = s + eo > ei;
        const uint64_t keep = two ? lowmask(lo) : ~0ull;
        const uint64_t X = ((x0[k] >> (s * fi)) & keep) | (two ? x1[k] << lo : 0);
        const uint64_t Y = ((y0[k] >> (s * fi)) & keep) | (two ? y1[k] << lo : 0);
        uint64_t word = 0;
        for (uint32_t j = 0, si = 0, so = 0; j < eo; j++, si += fi, so += fo)
            word |= (binop<OP>((X >> si) & im, (Y >> si) & im) & om) << so;
        // fields past the column's end (its last word) are zero
        st_nt(g.out.p + q, cnt == eo ? word : word & lowmask(cnt * fo));
    }
}

template <int OP>
static bool launch_words(const Col& a, const Col& b, const Col& o, hipStream_t s, uint32_t fmt_a, uint32_t fmt_o)
{
    auto field = [](const Col& c) { return c.kind == K_SLAB ? c.bits : 8 * c.cbytes; };
    if (a.kind == K_BITSTREAM || o.kind == K_BITSTREAM) return false;
    BinWords g{};
    g.a = a;
    g.b = b;
    g.out = o;
    g.fi = field(a);
    g.eps_i = 64 / g.fi;
    g.fo = field(o);
    g.eps_o = 64 / g.fo;
    if (g.eps_o > g.eps_i) return false; // narrower output: the general kernel
    // slab -> slab with fewer fields per output word: the staged kernel
    // (measured faster: w=4 -> 5 0.101 vs 0.144 ms, w=10 -> 11 0.21 vs 0.26)
    if (g.eps_o != g.eps_i && a.kind == K_SLAB && o.kind == K_SLAB) return false;
    g.a.eps = g.eps_i;
    g.a.eps_magic = g.eps_i > 1 ? (uint32_t)((1ull << 32) / g.eps_i + 1) : 0;
    g.words_i = (a.n + g.eps_i - 1) / g.eps_i;
    g.words_o = (o.n + g.eps_o - 1) / g.eps_o;
    const uint32_t tp = am_tuned(AM_KNOB_BINARY_WORDS, a.bits, fmt_a, fmt_o);
    const uint32_t per = tp >= 8 ? 8 : (tp >= 4 || tp == 0 ? BA_PER_DEFAULT : 2);
    auto grid = [&](uint32_t pw) { return dim3((uint32_t)((g.words_o + 256 * pw - 1) / (256 * pw))); };
    if (g.eps_o == g.eps_i) {
        if (g.fi == g.fo && a.mask == o.mask) // equal field layout
            for (uint32_t j = 0; j < g.eps_i; j++) g.fmask |= a.mask << (j * g.fi);
        switch (per) {
            case 2: hipLaunchKernelGGL((k_binary_aligned<OP, 2>), grid(2), dim3(256), 0, s, g); break;
            case 8: hipLaunchKernelGGL((k_binary_aligned<OP, 8>), grid(8), dim3(256), 0, s, g); break;
            default: hipLaunchKernelGGL((k_binary_aligned<OP, 4>), grid(4), dim3(256), 0, s, g); break;
        }
        return true;
    }
    switch (per) {
        case 2: hipLaunchKernelGGL((k_binary_words<OP, 2>), grid(2), dim3(256), 0, s, g); break;
        case 8: hipLaunchKernelGGL((k_binary_words<OP, 8>), grid(8), dim3(256), 0, s, g); break;
        default: hipLaunchKernelGGL((k_binary_words<OP, 4>), grid(4), dim3(256), 0, s, g); break;
    }
    return true;
}

// Word-per-lane binary op for bitstream inputs (in-width wi <= 32) into a
// bitstream or slab output no narrower than the input (so a block's output
// words need at most as many input words). The block copies its contiguous
// input range of both columns into LDS with coalesced loads; each lane then
// builds whole output words, taking each element's wi bits with one funnel
// shift (v_alignbit) of two LDS dwords. Indices are block-relative 32-bit;
// divisions by the out-width are a reciprocal multiply (exact below 2^24).
// ISLAB: slab inputs with dead bits into a slab output of fewer fields per
// word, staged the same way; a lane takes its output word's fields from two
// staged words of each input (the window below) -- reading the words from
// global memory per lane (k_binary_words) measured slower, its loads being
// lane-strided.
constexpr uint32_t BS_WORDS = 1024;                // output words per block (4 per lane)
constexpr uint32_t BS_IN = BS_WORDS + 4;           // input words per column staged per block
struct BinStream {
    Col a, b, out;
    uint32_t wi, wo, eps_o, magic_o; // magic_o: floor(2^32/wo)+1 (64-bit multiply)
    uint32_t eps_i, magic_i;         // ISLAB: fields per input word, floor(2^32/eps_i)+1
    uint32_t win_e, max_e;           // bitstream output: floor(64/wi), floor(63/wo)+2
    uint64_t words_i, words_o;
};

template <int OP, bool OSLAB, bool ISLAB>
__global__ __launch_bounds__(256) void k_binary_stream(BinStream g)
{
    static_assert(OSLAB || !ISLAB, "slab inputs go to slab outputs");
    __shared__ uint64_t s_a[BS_IN + 1], s_b[BS_IN + 1];
    const uint32_t tid = threadIdx.x;
    const uint64_t n = g.out.n;
    const uint32_t wi = g.wi, wo = g.wo;
    const uint64_t q0 = (uint64_t)blockIdx.x * BS_WORDS;
    const uint64_t qend = q0 + BS_WORDS < g.words_o ? q0 + BS_WORDS : g.words_o;
    // the block's first element, and one past the last element its words hold
    const uint64_t t0 = OSLAB ? q0 * g.eps_o : (q0 * 64) / wo;
    uint64_t tend = OSLAB ? qend * g.eps_o : (qend * 64 - 1) / wo + 1;
    if (tend > n) tend = n;
    const uint64_t ib0 = ISLAB ? t0 / g.eps_i : (t0 * wi) >> 6;
    uint64_t ib1 = ISLAB ? (tend + g.eps_i - 1) / g.eps_i : (tend * wi + 63) >> 6;
    if (ib1 > g.words_i) ib1 = g.words_i;
    const uint32_t nin = (uint32_t)(ib1 - ib0); // <= BS_IN (wo >= wi; ISLAB: eps_o < eps_i)
    {
        // every load of the thread first (each LDS store would otherwise wait
        // for its own load: one memory round trip per step)
        constexpr uint32_t PER = (BS_IN + 1 + 255) / 256;
        uint64_t va[PER], vb[PER];
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            const uint32_t i = tid + 256 * k;
            va[k] = i < nin ? g.a.p[ib0 + i] : 0;
            vb[k] = i < nin ? g.b.p[ib0 + i] : 0;
        }
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            const uint32_t i = tid + 256 * k;
            if (i <= nin) {
                s_a[i] = va[k];
                s_b[i] = vb[k];
            }
        }
    }
    __syncthreads();
    const uint32_t* sa = (const uint32_t*)s_a;
    const uint32_t* sb = (const uint32_t*)s_b;
    const uint32_t rb = (uint32_t)(t0 * wi - ib0 * 64);      // input bit of element t0 in the staged run
    const uint32_t base = OSLAB ? 0 : (uint32_t)(q0 * 64 - t0 * wo); // output bit of word q0 past element t0's start
    const uint32_t ne = (uint32_t)(tend - t0);                // elements the block may touch
    const uint32_t im = (uint32_t)g.a.mask;
    const uint64_t om = g.out.mask;
    if (ISLAB) {
        // word qr holds elements qr*eps_o .. +eps_o, starting in staged word dq
        // at field s: the fields of that word from s on, then those of the
        // next (eps_o * wi <= 64 bits, dead bits of the first word cleared)
        const uint32_t eo = g.eps_o, ei = g.eps_i;
        const uint32_t r0 = (uint32_t)(t0 - ib0 * ei); // field of element t0 in staged word 0
        for (uint32_t k = 0; k < BS_WORDS / 256; k++) {
            const uint32_t qr = k * 256 + tid;
            if (q0 + qr >= qend) break;
            const uint32_t e = qr * eo;
            const uint32_t cnt = ne - e < eo ? ne - e : eo;
            const uint32_t rel = r0 + e; // < 2^17
            const uint32_t dq = __umulhi(rel, g.magic_i), s = rel - dq * ei; // dq + 1 <= nin
            const uint32_t lo = (ei - s) * wi;   // < 64 when the next word is needed (s >= 1)
            const bool two = s + eo > ei;
            const uint64_t keep = two ? lowmask(lo) : ~0ull;
            const uint64_t xa = ((s_a[dq] >> (s * wi)) & keep) | (two ? s_a[dq + 1] << lo : 0);
            const uint64_t xb = ((s_b[dq] >> (s * wi)) & keep) | (two ? s_b[dq + 1] << lo : 0);
            uint64_t word = 0;
            for (uint32_t f = 0, si = 0, so = 0; f < eo; f++, si += wi, so += wo) { // si <= 64 - wi
                const uint32_t va = (uint32_t)(xa >> si) & im, vb = (uint32_t)(xb >> si) & im;
                word |= (binop<OP>((uint64_t)va, (uint64_t)vb) & om) << so;
            }
            st_nt(g.out.p + q0 + qr, cnt == eo ? word : word & lowmask(cnt * wo));
        }
        return;
    }
    if (OSLAB) {
        // word qr holds elements qr*eps_o .. +eps_o, at most 64 input bits: one
        // three-dword window per column gives all of them, each field then a
        // shift by a wave-uniform amount (no LDS read per element)
        const uint32_t eo = g.eps_o;
        for (uint32_t k = 0; k < BS_WORDS / 256; k++) {
            const uint32_t qr = k * 256 + tid;
            if (q0 + qr >= qend) break;
            const uint32_t e = qr * eo;
            const uint32_t cnt = ne - e < eo ? ne - e : eo; // e < ne: the word exists
            const uint32_t ib = rb + e * wi, j = ib >> 5, sh = ib & 31; // j + 2 <= 2 * nin + 1
            const uint64_t xa = ((uint64_t)__builtin_amdgcn_alignbit(sa[j + 2], sa[j + 1], sh) << 32) |
                                __builtin_amdgcn_alignbit(sa[j + 1], sa[j], sh);
            const uint64_t xb = ((uint64_t)__builtin_amdgcn_alignbit(sb[j + 2], sb[j + 1], sh) << 32) |
                                __builtin_amdgcn_alignbit(sb[j + 1], sb[j], sh);
            uint64_t word = 0;
            for (uint32_t f = 0, si = 0, so = 0; f < eo; f++, si += wi, so += wo) { // si <= 64 - wi
                const uint32_t va = (uint32_t)(xa >> si) & im, vb = (uint32_t)(xb >> si) & im;
                word |= (binop<OP>((uint64_t)va, (uint64_t)vb) & om) << so;
            }
            // fields past the column's end (its last word) are zero
            st_nt(g.out.p + q0 + qr, cnt == eo ? word : word & lowmask(cnt * wo));
        }
        return;
    }
    // bitstream output: word qr's elements [e, e1) -- the first may start
    // before the word -- number at most max_e; their input bits are read as
    // two 64-bit windows, the first holding win_e = floor(64/wi) whole
    // elements from e on, the second the (at most two) after them, so each
    // field is a shift by a wave-uniform amount and only the output position
    // differs per lane
    auto div_o = [&](uint32_t x) { return (uint32_t)(((uint64_t)x * g.magic_o) >> 32); };
    auto window = [&](const uint32_t* sp, uint32_t ib) {
        const uint32_t j = ib >> 5 < 2 * BS_IN - 1 ? ib >> 5 : 2 * BS_IN - 1, sh = ib & 31;
        return ((uint64_t)__builtin_amdgcn_alignbit(sp[j + 2], sp[j + 1], sh) << 32) |
               __builtin_amdgcn_alignbit(sp[j + 1], sp[j], sh);
    };
    const uint32_t m = g.win_e, me = g.max_e;
    for (uint32_t k = 0; k < BS_WORDS / 256; k++) {
        const uint32_t qr = k * 256 + tid;
        if (q0 + qr >= qend) break;
        const uint32_t ob = qr * 64 + base; // the word's first output bit past element t0's start
        const uint32_t e = div_o(ob);
        uint32_t e1 = div_o(ob + 63) + 1;
        if (e1 > ne) e1 = ne;
        const uint32_t cnt = e1 - e;                          // 1 .. max_e
        const uint32_t p0 = ob - e * wo;                      // element e starts p0 bits before the word
        const uint32_t ib = rb + e * wi;
        const uint64_t xa = window(sa, ib), xb = window(sb, ib);
        auto field = [&](uint64_t x, uint64_t y, uint32_t si) {
            const uint32_t va = (uint32_t)(x >> si) & im, vb = (uint32_t)(y >> si) & im;
            return binop<OP>((uint64_t)va, (uint64_t)vb) & om;
        };
        uint64_t word = field(xa, xb, 0) >> p0;
        uint32_t f = 1, p = wo - p0; // output bit of element e + f
        for (uint32_t si = wi; f < m && f < me; f++, si += wi, p += wo) {
            const uint64_t r = field(xa, xb, si);
            word |= (f < cnt ? r : 0) << p; // p < 64 for f < cnt
        }
        if (f < me) { // uniform: the elements past the first window
            const uint64_t ya = window(sa, ib + m * wi), yb = window(sb, ib + m * wi);
            for (uint32_t si = 0; f < me; f++, si += wi, p += wo) {
                const uint64_t r = field(ya, yb, si);
                word |= (f < cnt ? r : 0) << p;
            }
        }
        st_nt(g.out.p + q0 + qr, word);
    }
}

template <int OP>
static bool launch_stream(const Col& a, const Col& b, const Col& o, hipStream_t s)
{
    // (a slab whose width divides 64 has the stream layout too)
    const bool stream_in = a.kind == K_BITSTREAM || (a.kind == K_SLAB && 64 % a.bits == 0);
    // other slabs: into a slab of fewer fields per word
    const bool slab_in = !stream_in && a.kind == K_SLAB && o.kind == K_SLAB && 64 / o.bits < 64 / a.bits;
    if (!(stream_in || slab_in) || a.bits > 32 || o.bits < a.bits) return false;
    if (o.kind != K_BITSTREAM && o.kind != K_SLAB) return false;
    BinStream g{};
    g.a = a;
    g.b = b;
    g.out = o;
    g.wi = a.bits;
    g.wo = o.bits;
    g.eps_o = o.kind == K_SLAB ? 64 / o.bits : 0;
    g.magic_o = (uint32_t)0; // unused for slab outputs
    g.words_i = a.words;
    g.words_o = o.words;
    const uint64_t mo = (1ull << 32) / o.bits + 1;
    if (mo >> 32) return false; // out-width 1 (only possible with in-width 1)
    g.magic_o = (uint32_t)mo;
    g.win_e = 64 / a.bits;
    g.max_e = 63 / o.bits + 2;
    g.eps_i = 64 / a.bits;
    g.magic_i = (uint32_t)((1ull << 32) / g.eps_i + 1); // eps_i >= 2 when used
    const dim3 grid((uint32_t)((g.words_o + BS_WORDS - 1) / BS_WORDS));
    if (slab_in) hipLaunchKernelGGL((k_binary_stream<OP, true, true>), grid, dim3(256), 0, s, g);
    else if (o.kind == K_SLAB) hipLaunchKernelGGL((k_binary_stream<OP, true, false>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((k_binary_stream<OP, false, false>), grid, dim3(256), 0, s, g);
    return true;
}

template <int OP>
static void launch_ew(const Col& a, const Col& b, const Col& o, hipStream_t s, uint32_t fmt_a, uint32_t fmt_o)
{
    if constexpr (OP != OP_COPY) {
        if (launch_words<OP>(a, b, o, s, fmt_a, fmt_o)) return;
        if (launch_stream<OP>(a, b, o, s)) return;
    }
    uint64_t units = o.kind == K_STRIDED ? o.n : (o.kind == K_SLAB ? (o.n + o.eps - 1) / o.eps : (o.n * o.bits + 63) / 64);
    dim3 grid((uint32_t)((units + EW_UNITS - 1) / EW_UNITS));
#define EW_CASE(AK, OK) hipLaunchKernelGGL((k_elementwise<AK, OK, OP>), grid, dim3(256), 0, s, a, b, o)
#define EW_OUT(AK)                                                  \
    switch (o.kind) {                                               \
        case K_STRIDED: EW_CASE(AK, K_STRIDED); break;              \
        case K_SLAB: EW_CASE(AK, K_SLAB); break;                    \
        default: EW_CASE(AK, K_BITSTREAM); break;                   \
    }
    switch (a.kind) {
        case K_STRIDED: EW_OUT(K_STRIDED); break;
        case K_SLAB: EW_OUT(K_SLAB); break;
        default: EW_OUT(K_BITSTREAM); break;
    }
#undef EW_OUT
#undef EW_CASE
}

} // namespace am

using namespace am;

extern "C" int am_convert(const am_column* in, am_column* out, void* stream)
{
    Col a, o;
    int st = make_col(in, &a);
    if (st) return st;
    if ((st = make_col(out, &o))) return st;
    if (a.n != o.n) return AM_E_INVALID;
    if (o.n == 0) return AM_OK;
    launch_ew<OP_COPY>(a, a, o, (hipStream_t)stream, in->format, out->format);
    return check_launch();
}

extern "C" int am_binary(uint32_t op, const am_column* a, const am_column* b, am_column* out, void* stream)
{
    Col ca, cb, o;
    int st = make_col(a, &ca);
    if (st) return st;
    if ((st = make_col(b, &cb))) return st;
    if ((st = make_col(out, &o))) return st;
    if (ca.n != cb.n || ca.n != o.n || a->bits != b->bits || a->format != b->format) return AM_E_INVALID;
    if (o.n == 0) return AM_OK;
    hipStream_t s = (hipStream_t)stream;
    switch (op) {
        case AM_OP_ADD: launch_ew<AM_OP_ADD>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_MUL: launch_ew<AM_OP_MUL>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_MOD: launch_ew<AM_OP_MOD>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_DIV: launch_ew<AM_OP_DIV>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_MATH: launch_ew<AM_OP_MATH>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_BOOL_SIMPLE: launch_ew<AM_OP_BOOL_SIMPLE>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_BOOL_COMPLEX: launch_ew<AM_OP_BOOL_COMPLEX>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_AND: launch_ew<AM_OP_AND>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_OR: launch_ew<AM_OP_OR>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_XOR: launch_ew<AM_OP_XOR>(ca, cb, o, s, a->format, out->format); break;
        case AM_OP_SUB: launch_ew<AM_OP_SUB>(ca, cb, o, s, a->format, out->format); break;
        default: return AM_E_INVALID;
    }
    return check_launch();
}
