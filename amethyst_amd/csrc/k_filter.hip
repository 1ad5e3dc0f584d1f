// k_filter.hip -- predicate filter to an LSB-first bitmask, and the u64 sum.
//
// Filter reference: kernel_filter_baseline / _nogaps / _slabs
// (src/kernels/filter_baseline.cuh:9-36, filter_nogaps.cuh:4-65,
// filter_slabs.cuh:4-57). There each thread decodes 32 consecutive elements
// serially and writes one u32 mask word; the baseline variant needs a 32x32
// shared-memory bit transpose.
// Here slab and container inputs go word-per-lane (k_filter_words: all
// fields of a word selected with 4 VALU each, mask words combined across
// lanes by DPP or in LDS); bitstream inputs go element-per-lane (k_filter: lane l
// of a wave decodes element base + 64j + l, so one wave64 ballot -- a single
// v_cmp into an SGPR pair -- *is* mask word j; no transpose).
//
// Sum reference: kernel_hop_slabs / kernel_hop_nogaps (src/kernels/hop.cuh:7-76),
// one global atomicAdd per thread. Here each lane decodes whole packed words
// (slab: eps fields with v_bfe; bitstream: the elements starting in its
// word, straddlers completed from the next word), blocks write partial sums
// to the workspace and a one-block kernel adds them: no contended atomics,
// deterministic (u64 addition wraps and commutes).
#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t FT_J = 16;                 // mask words per wave and tile
constexpr uint32_t FT_TILE = 4 * FT_J * 64;   // elements per 256-thread tile

// Persistent: block b filters tiles b, b+G, ... and loads tile t+G while it
// evaluates tile t, so every wave keeps a tile's worth of loads in flight.
template <int IK, bool WIDE>
__global__ __launch_bounds__(256) void k_filter(Col in, Pred pr, uint64_t* mask_out, uint32_t T)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t n = in.n, words = (n + 63) >> 6;
    uint32_t t = blockIdx.x;
    if (t >= T) return;
    Fetch<IK, WIDE> f[FT_J];
    auto chunk = [&](uint32_t tt) { return (uint64_t)tt * FT_TILE + wave * (FT_J * 64); };
    auto fetch_tile = [&](uint32_t tt, Fetch<IK, WIDE>* ff) {
        const uint64_t wb = (uint64_t)tt * FT_TILE + wave * (FT_J * 64);
        if (wb < n) {
            const uint32_t rem = (uint32_t)(n - wb < FT_J * 64 ? n - wb : FT_J * 64);
            TileReader<IK> rd(in, wb);
#pragma unroll
            for (uint32_t j = 0; j < FT_J; j++) {
                const uint32_t i = j * 64 + lane;
                ff[j].load(rd, i < rem ? i : rem - 1);
            }
        }
    };
    fetch_tile(t, f);
    for (; t < T; t += gridDim.x) {
        Fetch<IK, WIDE> nf[FT_J];
        if (t + gridDim.x < T) fetch_tile(t + gridDim.x, nf);
        const uint64_t wbase = chunk(t);
        const uint32_t wrem = wbase < n ? (uint32_t)(n - wbase < FT_J * 64 ? n - wbase : FT_J * 64) : 0;
        const TileReader<IK> rd(in, wbase < n ? wbase : 0);
        uint64_t mine = 0;
#pragma unroll
        for (uint32_t j = 0; j < FT_J; j++) {
            const uint32_t i = j * 64 + lane < wrem ? j * 64 + lane : (wrem ? wrem - 1 : 0);
            const uint64_t v = f[j].value(rd, i);
            const bool fl = (j * 64 + lane < wrem) && (WIDE ? pr(v) : pr.sel32((uint32_t)v));
            const uint64_t b = ballot(fl);
            if (lane == j) mine = b;
        }
        const uint64_t w = (wbase >> 6) + lane;
        if (lane < FT_J && w < words) st_nt(mask_out + w, mine);
#pragma unroll
        for (uint32_t j = 0; j < FT_J; j++) f[j] = nf[j];
    }
}


// Bitstream inputs of <= 32 bits: a block owns 4096 elements; it copies their
// input words into LDS with coalesced loads (one pad word in front), then
// lane l of a wave takes element 64j + l as the 32 stream bits ending at the
// element's last bit -- one funnel shift (v_alignbit) of two LDS dwords puts
// the field at the top with the previous element's bits beneath, which a
// range test scaled by 2^(32-bits) ignores -- and the wave's ballot of the
// test is mask word j. Chunk j+1's element sits 64*bits bits later, a whole
// number of dwords, so the shift is the lane's for every chunk and only the
// LDS address moves; the range is cyclic (one test for every operator, NE
// included). One tile per block.
constexpr uint32_t FS_J = 16;                   // mask words per wave
constexpr uint32_t FS_TILE = 4 * FS_J * 64;     // elements per block
constexpr uint32_t FS_IN = FS_TILE * 32 / 64 + 3; // staged input words (bits <= 32)

struct FilterStreamArgs {
    Col in;
    uint32_t clo, crg; // selected <=> (y - clo) mod 2^32 <= crg, y the field top-aligned
    uint32_t none;     // the predicate selects nothing
    uint64_t* mask_out;
};

template <bool LO0>
__global__ __launch_bounds__(256) void k_filter_stream(FilterStreamArgs a)
{
    __shared__ uint64_t s_w[FS_IN + 1]; // s_w[0]: pad below the first element
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const Col& in = a.in;
    const uint64_t n = in.n, t0 = (uint64_t)blockIdx.x * FS_TILE;
    const uint32_t w = in.bits;
    const uint64_t ib0 = (t0 * w) >> 6;
    const uint64_t tend = t0 + FS_TILE < n ? t0 + FS_TILE : n;
    uint64_t ib1 = (tend * w + 63) >> 6;
    if (ib1 > in.words) ib1 = in.words;
    const uint32_t nin = (uint32_t)(ib1 - ib0);
    {
        // every load of the thread first (each LDS store would otherwise wait
        // for its own load: one memory round trip per step)
        constexpr uint32_t PER = (FS_IN + 255) / 256;
        uint64_t v[PER];
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            const uint32_t i = tid + 256 * k;
            v[k] = i < nin ? gmem(in.p)[ib0 + i] : 0;
        }
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            const uint32_t i = tid + 256 * k;
            if (i < FS_IN) s_w[1 + i] = v[k];
        }
        if (tid == 0) s_w[0] = 0;
    }
    __syncthreads();
    const uint32_t* s32 = (const uint32_t*)s_w;
    const uint32_t rb = (uint32_t)(t0 * w - ib0 * 64); // stream bit of element t0 in the staged run
    const uint32_t ne = (uint32_t)(tend - t0);
    // end bit (exclusive) of the lane's element of chunk 0, relative to the run
    const uint32_t e0 = rb + (wave * (FS_J * 64) + lane + 1) * w;
    const uint32_t sh = e0 & 31;
    const uint32_t* const p = s32 + (e0 >> 5) + 1; // the window's low dword (run dword 0 is s32[2])
    const uint32_t CLO = a.clo, CRG = a.crg;
    const bool last = blockIdx.x == gridDim.x - 1;
    uint32_t mlo = 0, mhi = 0; // lane j: the wave's mask word j
    // every chunk's window first (one LDS round trip for the wave's 16)
    uint32_t y[FS_J];
    static_for(std::make_integer_sequence<uint32_t, FS_J>{}, [&](auto jc) {
        constexpr uint32_t j = decltype(jc)::value;
        const uint32_t* q = p + 2 * w * j;
        y[j] = __builtin_amdgcn_alignbit(q[1], q[0], sh);
    });
    static_for(std::make_integer_sequence<uint32_t, FS_J>{}, [&](auto jc) {
        constexpr uint32_t j = decltype(jc)::value;
        uint64_t b = ballot(LO0 ? (y[j] <= CRG) : (y[j] - CLO <= CRG));
        if (last) { // elements past n (only in the last block)
            const uint32_t e = wave * (FS_J * 64) + j * 64;
            b &= e >= ne ? 0 : (ne - e >= 64 ? ~0ull : lowmask(ne - e));
        }
        if (a.none) b = 0;
        writelane_imm64<j>(mlo, mhi, b);
    });
    const uint64_t mw = (t0 >> 6) + wave * FS_J + lane;
    if (lane < FS_J && mw < ((n + 63) >> 6)) st_nt(a.mask_out + mw, ((uint64_t)mhi << 32) | mlo);
}

// Word-per-lane filter for slab and container inputs: a block owns a tile
// of 256*R input words; each lane issues its R word loads at once and forms
// each word's selection bits with select_fields (4 VALU per field), or, for
// fields of 2, 4 or 8 bits under a range from 0 or up to the top value (<,
// <=, >, >=, and == / != of the top value), with select_fields_pow2 (two
// 64-bit adds for all fields, then a log-step gather of the bits). Mask
// words: one field per word -> the chunk's ballot; a power-of-two number of
// fields per word -> the lanes feeding one mask word are an aligned segment
// of 64/eps lanes, combined with a DPP OR-scan and stored by its last lane;
// other widths -> each lane puts its word's bits in LDS and lane m of the
// wave gathers the wave's mask words m, m+64, ... from the (about 64/eps)
// words each spans (wave-local, so no block barrier).
// One tile per block; no block waits on another.
struct FilterWordsArgs {
    Col in;
    Pred pr;
    TopPred tp;
    uint32_t fw, eps, R, pow2;
    uint64_t in_words;
    uint64_t* mask_out;
    uint32_t swar, swinv;     // fields of 2, 4 or 8 bits, range from 0 or to the top: select_fields_pow2
    uint64_t swE, swK, swG;   // its slot masks
};

constexpr uint32_t FW_RMAX = 8;
constexpr uint32_t FW_SELW = 4 * 64 * FW_RMAX; // LDS selection words (u32) per tile

template <bool WIDE, bool GATHER>
__global__ __launch_bounds__(256) void k_filter_words(FilterWordsArgs a)
{
    // non-power-of-two eps: selection bits per word (u32); otherwise the
    // wave's mask words (u64, at most 32*R of them), stored together at the end
    __shared__ uint64_t s_lds[FW_SELW / 2];
    uint32_t* const s_sel = (uint32_t*)s_lds;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t t = blockIdx.x;
    const uint32_t fw = a.fw, eps = a.eps, R = a.R, bw = a.in.bits;
    constexpr bool lds = GATHER; // non-power-of-two eps (never WIDE)
    const uint64_t n = a.in.n, nw = a.in_words;
    const uint64_t w0 = t * (uint64_t)(256 * R);
    const uint64_t nmw = (n + 63) >> 6;
    const uint64_t mw0 = (w0 * eps) >> 6;
    uint32_t* const sg = s_sel + wave * 64 * R; // this wave's words
    uint64_t* const sm = (uint64_t*)sg;         // (power-of-two eps) its mask words
    uint64_t x[FW_RMAX];
#pragma unroll
    for (uint32_t r = 0; r < FW_RMAX; r++) {
        x[r] = 0;
        const uint64_t wi = w0 + (uint64_t)(wave * R + r) * 64 + lane;
        if (r < R && wi < nw) x[r] = a.in.p[wi];
    }
    // (all R loads in flight before any word is used: without this the
    // compiler sinks each load to its chunk and the loads go out one by one)
    asm volatile("" ::: "memory");
    const bool last = t == gridDim.x - 1;
    const uint32_t S = 64 / eps; // lanes per mask word (power-of-two eps)
#pragma unroll
    for (uint32_t r = 0; r < FW_RMAX; r++) {
        if (r >= R) continue;
        const uint32_t ci = wave * R + r;
        uint32_t gg = WIDE ? (uint32_t)a.pr(x[r] & a.in.mask)
                           : (a.swar ? select_fields_pow2(x[r], fw, a.swE, a.swK, a.swG, a.swinv)
                                     : select_fields(x[r], eps, fw, bw, a.tp));
        if (last) {
            const uint64_t e0 = (w0 + ci * 64 + lane) * eps;
            const uint32_t nv = e0 >= n ? 0 : (n - e0 < eps ? (uint32_t)(n - e0) : eps);
            gg &= (uint32_t)lowmask(nv);
        }
        if (eps == 1) { // the chunk's mask word is the ballot
            const uint64_t m = ballot(gg & 1);
            if (lane == 0) sm[r] = m;
        }
        else if (!lds) {
            const uint64_t m = segment_or((uint64_t)gg << __umul24(lane & (S - 1), eps), S);
            if ((lane & (S - 1)) == S - 1) sm[r * eps + lane / S] = m;
        }
        else
            sg[r * 64 + lane] = gg;
    }
    if (!lds) {
        // the wave's R*eps mask words are consecutive: one store of them all
        // (per-chunk stores were 8-16 byte pieces of a line at eps = 1, 2)
        const uint32_t nwm = R * eps;
        for (uint32_t mw = lane; mw < nwm; mw += 64) {
            const uint64_t mi = mw0 + (uint64_t)wave * nwm + mw;
            if (mi < nmw) st_nt(a.mask_out + mi, sm[mw]);
        }
        return;
    }
    const uint32_t nwm = R * eps; // the wave's mask words (<= 8*21)
    for (uint32_t mw = lane; mw < nwm; mw += 64) {
        const uint32_t wa = (mw * 64) / eps;    // the word holding the mask word's first element
        const uint32_t sk = mw * 64 - wa * eps; // its elements before that one
        uint64_t m = sg[wa] >> sk;
        uint32_t w = wa + 1, sh = eps - sk;
        // two words per step (independent loads); both start inside the
        // mask word, so both exist
        for (; sh + eps < 64; w += 2, sh += 2 * eps) {
            const uint32_t g0 = sg[w], g1 = sg[w + 1];
            m |= ((uint64_t)g0 << sh) | ((uint64_t)g1 << (sh + eps));
        }
        if (sh < 64) m |= (uint64_t)sg[w] << sh;
        const uint64_t mi = mw0 + (uint64_t)wave * nwm + mw;
        if (mi < nmw) st_nt(a.mask_out + mi, m); // bits past n are zero (cleared above)
    }
}

// ---- sum: lane-per-word decode, per-block partials, then one final block.
constexpr uint32_t SUM_UNROLL = 4;

// sum of the elements of slab word q (elements q*eps .. q*eps+eps-1, < n)
__device__ inline uint64_t slab_word_sum(uint64_t w, uint64_t q, const Col& c)
{
    uint64_t first = q * c.eps;
    uint32_t cnt = c.eps;
    if (first + cnt > c.n) cnt = (uint32_t)(c.n - first);
    uint64_t s = 0;
    if (c.bits == 32) {
        // v_bfe encodes the width in 5 bits: a 32-bit field is the half itself
        s = (uint32_t)w;
        if (cnt > 1) s += (uint32_t)(w >> 32);
    }
    else if (c.bits < 32) {
        const uint32_t lo = (uint32_t)w, hi = (uint32_t)(w >> 32);
        for (uint32_t k = 0; k < cnt; k++) {
            const uint32_t off = k * c.bits;
            uint32_t v;
            if (off + c.bits <= 32) v = __builtin_amdgcn_ubfe(lo, off, c.bits);
            else if (off >= 32) v = __builtin_amdgcn_ubfe(hi, off - 32, c.bits);
            else v = (uint32_t)((w >> off) & c.mask);
            s += v;
        }
    }
    else {
        for (uint32_t k = 0; k < cnt; k++) s += (w >> (k * c.bits)) & c.mask;
    }
    return s;
}

// SWAR sum for slab and container columns with at most 32-bit fields: the
// even-numbered fields of a word (x & E) and the odd-numbered ones
// ((x >> fw) & E') land in 2*fw-bit slots, each with fw bits of headroom
// above its value, and are added as whole 64-bit words; the slots are
// summed out every K words, before any can overflow. With an odd number of
// fields the top one (whose headroom is only the dead bits) is added
// separately. About 10 VALU per word whatever the number of fields.
struct SumSwar {
    const uint64_t* p;
    uint64_t words, n_last; // words; valid fields of the last word
    uint64_t E, O;          // value bits of the even fields / of the odd fields shifted down by fw
    uint32_t fw, eps, top_sh, has_top;
    uint32_t bits, K;       // value bits; words per lane between slot flushes
};

__device__ inline uint64_t swar_slots(uint64_t acc, uint32_t fw, uint32_t nslots)
{
    const uint64_t sm = lowmask(2 * fw);
    uint64_t s = 0;
    for (uint32_t j = 0; j < nslots; j++) s += (acc >> (2 * j * fw)) & sm;
    return s;
}

__global__ __launch_bounds__(256) void k_sum_swar(SumSwar g, uint64_t* partial)
{
    __shared__ uint64_t s_part[4];
    const uint32_t nslots = (g.eps + 1) / 2;
    const uint64_t vmask = lowmask(g.bits);
    uint64_t total = 0, ae = 0, ao = 0;
    uint32_t cnt = 0;
    const uint64_t gstride = (uint64_t)gridDim.x * 256 * SUM_UNROLL;
    for (uint64_t q0 = (uint64_t)blockIdx.x * 256 * SUM_UNROLL + threadIdx.x; q0 < g.words; q0 += gstride) {
        uint64_t wv[SUM_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < SUM_UNROLL; u++) {
            const uint64_t q = q0 + u * 256;
            uint64_t x = q < g.words ? __builtin_nontemporal_load(g.p + q) : 0;
            if (q == g.words - 1) x &= lowmask((uint32_t)(g.n_last * g.fw)); // fields past n
            wv[u] = x;
        }
#pragma unroll
        for (uint32_t u = 0; u < SUM_UNROLL; u++) {
            const uint64_t x = wv[u];
            ae += x & g.E;
            ao += (x >> g.fw) & g.O;
            if (g.has_top) total += (x >> g.top_sh) & vmask;
        }
        cnt += SUM_UNROLL;
        if (cnt >= g.K) {
            total += swar_slots(ae, g.fw, nslots) + swar_slots(ao, g.fw, nslots);
            ae = ao = 0;
            cnt = 0;
        }
    }
    total += swar_slots(ae, g.fw, nslots) + swar_slots(ao, g.fw, nslots);
    total = wave_sum_u64(total);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = total;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
}

// Bitstream sum, SWAR form (2 <= w <= 32, w not dividing 64): a lane takes
// word q and the fields that START in it, at offsets off, off+w, ... The
// 128-bit pair (word q+1 : word q) shifted right by `off` holds the first
// F = floor(64/w) of them at bits 0, w, 2w, ... -- a slab word of F fields,
// summed by the same even/odd slot accumulation as k_sum_swar -- and at most
// one more field starts below bit 64, taken alone. The last word (fields
// past n) takes the per-field path. About 5 lane operations per element
// instead of about 10. Block steps away from the end run branch-free, with
// the next word's low half taken from the next lane (DPP) rather than
// loaded a second time.
__global__ __launch_bounds__(256) void k_sum_bits_swar(SumSwar g, uint64_t nbits, uint64_t* partial)
{
    __shared__ uint64_t s_part[4];
    const uint32_t w = g.fw, F = g.eps, nslots = (F + 1) / 2;
    const uint64_t vmask = lowmask(w);
    const uint32_t vm = (uint32_t)vmask;
    const uint64_t wmagic = (1ull << 32) / w + 1; // x / w exact below 2^24
    uint64_t total = 0, ae = 0, ao = 0;
    uint32_t cnt = 0;
    const uint64_t gstride = (uint64_t)gridDim.x * 256 * SUM_UNROLL;
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t qb = (uint64_t)blockIdx.x * 256 * SUM_UNROLL; qb < g.words; qb += gstride) {
        const uint32_t r0 = (uint32_t)((qb * 64) % w); // bits of an element before word qb's bit 0, mod w
        if (qb + 256 * SUM_UNROLL < g.words) {
            // every word of the step and the one after each exist: the next
            // word's low half comes from the next lane (DPP wave shift), lane
            // 63 loading it; no per-word branches
            uint64_t wv[SUM_UNROLL];
            uint32_t xv[SUM_UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < SUM_UNROLL; u++) {
                const uint64_t q = qb + u * 256 + threadIdx.x;
                wv[u] = __builtin_nontemporal_load(g.p + q);
                xv[u] = lane == 63 ? (uint32_t)g.p[q + 1] : 0;
            }
#pragma unroll
            for (uint32_t u = 0; u < SUM_UNROLL; u++) {
                const uint32_t d = u * 256 + threadIdx.x;
                const uint32_t xo = r0 + d * 64; // < 2^17
                const uint32_t rx = xo - (uint32_t)(((uint64_t)xo * wmagic) >> 32) * w;
                const uint32_t off = rx ? w - rx : 0; // first element start in the word (< 32)
                const uint32_t l0 = (uint32_t)wv[u], h0 = (uint32_t)(wv[u] >> 32);
                const uint32_t nx = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)l0, 0x130, 0xf, 0xf, false); // lane+1
                const uint32_t l1 = lane == 63 ? xv[u] : nx;
                const uint64_t x = ((uint64_t)__builtin_amdgcn_alignbit(l1, h0, off) << 32) |
                                   __builtin_amdgcn_alignbit(h0, l0, off);
                ae += x & g.E;
                ao += (x >> w) & g.O;
                if (g.has_top) total += (x >> g.top_sh) & vmask;
                const uint32_t pe = off + F * w; // >= 33 (F*w > 64 - w)
                total += pe < 64 ? (__builtin_amdgcn_alignbit(l1, h0, pe - 32) & vm) : 0;
            }
            cnt += SUM_UNROLL;
            if (cnt >= g.K) {
                total += swar_slots(ae, w, nslots) + swar_slots(ao, w, nslots);
                ae = ao = 0;
                cnt = 0;
            }
            continue;
        }
        uint64_t wv[SUM_UNROLL], nv[SUM_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < SUM_UNROLL; u++) {
            const uint64_t q = qb + u * 256 + threadIdx.x;
            wv[u] = q < g.words ? g.p[q] : 0;
            nv[u] = q + 1 < g.words ? g.p[q + 1] : 0;
        }
#pragma unroll
        for (uint32_t u = 0; u < SUM_UNROLL; u++) {
            const uint32_t d = u * 256 + threadIdx.x;
            const uint64_t q = qb + d;
            if (q >= g.words) continue;
            const uint32_t xo = r0 + d * 64; // < 2^17
            const uint32_t rx = xo - (uint32_t)(((uint64_t)xo * wmagic) >> 32) * w;
            const uint32_t off = rx ? w - rx : 0; // first element start in the word
            const uint32_t l0 = (uint32_t)wv[u], h0 = (uint32_t)(wv[u] >> 32), l1 = (uint32_t)nv[u];
            if (q + 1 < g.words) {
                const uint64_t x = off ? (wv[u] >> off) | (nv[u] << (64 - off)) : wv[u];
                ae += x & g.E;
                ao += (x >> w) & g.O;
                if (g.has_top) total += (x >> g.top_sh) & vmask;
                const uint32_t pe = off + F * w; // the one more field, if it starts in the word (pe >= 33)
                if (pe < 64) total += __builtin_amdgcn_alignbit(l1, h0, pe - 32) & vm;
            }
            else { // the last word: only fields starting below n*w
                const uint64_t left = nbits - q * 64;
                const uint32_t lim = left < 64 ? (uint32_t)left : 64;
                for (uint32_t o = off; o < lim; o += w) {
                    const uint32_t hi = o < 32 ? h0 : l1, lo = o < 32 ? l0 : h0;
                    total += __builtin_amdgcn_alignbit(hi, lo, o & 31) & vm;
                }
            }
        }
        cnt += SUM_UNROLL;
        if (cnt >= g.K) {
            total += swar_slots(ae, w, nslots) + swar_slots(ao, w, nslots);
            ae = ao = 0;
            cnt = 0;
        }
    }
    total += swar_slots(ae, w, nslots) + swar_slots(ao, w, nslots);
    total = wave_sum_u64(total);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = total;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
}

template <int IK, bool WIDE>
__global__ __launch_bounds__(256) void k_sum_partial(Col in, uint64_t* partial)
{
    __shared__ uint64_t s_part[4];
    uint64_t acc = 0;
    const uint64_t gstride = (uint64_t)gridDim.x * 256 * SUM_UNROLL;
    if (IK == K_SLAB) {
        const uint64_t words = (in.n + in.eps - 1) / in.eps;
        for (uint64_t q0 = (uint64_t)blockIdx.x * 256 * SUM_UNROLL + threadIdx.x; q0 < words; q0 += gstride) {
            uint64_t wv[SUM_UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < SUM_UNROLL; u++) {
                const uint64_t q = q0 + u * 256;
                wv[u] = q < words ? in.p[q] : 0;
            }
#pragma unroll
            for (uint32_t u = 0; u < SUM_UNROLL; u++) {
                const uint64_t q = q0 + u * 256;
                if (q < words) acc += slab_word_sum(wv[u], q, in);
            }
        }
    }
    else if (IK == K_BITSTREAM && !WIDE) {
        // lane owns the elements that START in its word; a straddler takes
        // its top bits from the next word. Offsets are 32-bit: per block and
        // step, (64*q0) mod bits once (scalar), then each lane's residue by
        // a reciprocal multiply; fields (<= 32 bits) by one funnel shift.
        const uint32_t w = in.bits;
        const uint64_t words = (in.n * w + 63) >> 6;
        const uint64_t nbits = in.n * w;
        const uint64_t wmagic = (1ull << 32) / w + 1; // exact below 2^24
        const uint32_t vm = (uint32_t)in.mask;
        for (uint64_t qb = (uint64_t)blockIdx.x * 256 * SUM_UNROLL; qb < words; qb += gstride) {
            const uint32_t r0 = (uint32_t)((qb * 64) % w); // bits of an element before word qb's bit 0, mod w
            uint64_t wv[SUM_UNROLL], nv[SUM_UNROLL];
#pragma unroll
            for (uint32_t u = 0; u < SUM_UNROLL; u++) {
                const uint64_t q = qb + u * 256 + threadIdx.x;
                wv[u] = q < words ? in.p[q] : 0;
                nv[u] = q + 1 < words ? in.p[q + 1] : 0;
            }
#pragma unroll
            for (uint32_t u = 0; u < SUM_UNROLL; u++) {
                const uint32_t d = u * 256 + threadIdx.x;
                const uint64_t q = qb + d;
                if (q >= words) continue;
                const uint32_t x = r0 + d * 64; // < 2^17
                const uint32_t rx = x - (uint32_t)(((uint64_t)x * wmagic) >> 32) * w;
                uint32_t off = rx ? w - rx : 0; // first element start in the word
                const uint64_t left = nbits - q * 64;
                const uint32_t lim = left < 64 ? (uint32_t)left : 64; // starts before n*bits
                const uint32_t l0 = (uint32_t)wv[u], h0 = (uint32_t)(wv[u] >> 32), l1 = (uint32_t)nv[u];
                for (; off < lim; off += w) {
                    const uint32_t hi = off < 32 ? h0 : l1, lo = off < 32 ? l0 : h0;
                    acc += __builtin_amdgcn_alignbit(hi, lo, off & 31) & vm;
                }
            }
        }
    }
    else {
        // strided containers (and wide bitstreams): element per lane
        for (uint64_t t0 = (uint64_t)blockIdx.x * FT_TILE; t0 < in.n; t0 += (uint64_t)gridDim.x * FT_TILE) {
            TileReader<IK> rd(in, t0);
            const uint64_t rem = in.n - t0;
#pragma unroll 4
            for (uint32_t r = 0; r < FT_TILE / 256; r++) {
                const uint32_t i = r * 256 + threadIdx.x;
                if (i < rem) acc += rd.get(i);
            }
        }
    }
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
}

// 8-byte containers (the baseline format, 64-bit strided columns) at a
// 16-byte aligned base: each lane sums 16-byte pieces (two elements), eight
// in flight, with non-temporal loads; an odd last element is added by
// block 0's first thread
constexpr uint32_t SUM_U64_UNROLL = 8;
__global__ __launch_bounds__(256) void k_sum_u64(const uint64_t* p, uint64_t n, uint64_t vmask, uint64_t* partial)
{
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
    __shared__ uint64_t s_part[4];
    const u64x2* q = (const u64x2*)p;
    const uint64_t pairs = n / 2;
    const uint64_t gstride = (uint64_t)gridDim.x * 256 * SUM_U64_UNROLL;
    uint64_t acc = 0;
    for (uint64_t i0 = (uint64_t)blockIdx.x * 256 * SUM_U64_UNROLL + threadIdx.x; i0 < pairs; i0 += gstride) {
        u64x2 w[SUM_U64_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < SUM_U64_UNROLL; u++) {
            const uint64_t i = i0 + u * 256;
            w[u] = i < pairs ? __builtin_nontemporal_load(q + i) : u64x2{0, 0};
        }
#pragma unroll
        for (uint32_t u = 0; u < SUM_U64_UNROLL; u++) acc += (w[u].x & vmask) + (w[u].y & vmask);
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) acc += p[n - 1] & vmask;
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
}

__global__ __launch_bounds__(256) void k_sum_final(const uint64_t* partial, uint32_t nparts, uint64_t* out)
{
    __shared__ uint64_t s_part[4];
    // eight loads per thread in flight per step (a load-add loop waited one
    // round trip per partial: 5.6 us for 4096 of them)
    constexpr uint32_t U = 8;
    uint64_t acc = 0;
    for (uint32_t i0 = 0; i0 < nparts; i0 += 256 * U) {
        uint64_t v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t i = i0 + u * 256 + threadIdx.x;
            v[u] = i < nparts ? partial[i] : 0;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) acc += v[u];
    }
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
}

} // namespace am

using namespace am;

extern "C" int am_filter(const am_column* in, am_predicate pred, uint64_t* mask_out, void* stream)
{
    Col c;
    int st = make_col(in, &c);
    if (st) return st;
    if (pred.op > AM_PRED_NE) return AM_E_INVALID;
    if (!mask_out && c.n) return AM_E_INVALID;
    if (c.n == 0) return AM_OK;
    const uint64_t T = (c.n + FT_TILE - 1) / FT_TILE;
    if (T >= (1ull << 31)) return AM_E_UNSUPPORTED_CONFIG;
    const Pred pr = make_pred(pred, c.mask);
    hipStream_t s = (hipStream_t)stream;
    if (c.kind != K_BITSTREAM && !(c.kind == K_SLAB && c.bits == 1)) {
        FilterWordsArgs a{};
        a.in = c;
        a.pr = pr;
        a.fw = c.kind == K_SLAB ? c.bits : 8 * c.cbytes;
        a.eps = 64 / a.fw;
        a.pow2 = (a.eps & (a.eps - 1)) == 0;
        a.in_words = (c.n + a.eps - 1) / a.eps;
        // enough blocks to keep every CU busy to the end (~8192), at most
        // FW_RMAX words per lane
        uint64_t r = a.in_words / (256ull * 8192);
        if (r < 2) r = 2;
        a.R = (uint32_t)(r > FW_RMAX ? FW_RMAX : r);
        if (const uint32_t tr = am_tuned(AM_KNOB_FILTER_WORDS, c.bits, in->format, 0)) a.R = tr > FW_RMAX ? FW_RMAX : tr;
        if (c.bits <= 32) a.tp = make_top_pred(pr, c.bits, a.eps);
        SwarSel sw;
        if (swar_select(pr, a.tp, c.bits, a.fw, a.eps, &sw)) {
            a.swar = 1;
            a.swinv = sw.inv;
            a.swE = sw.E;
            a.swK = sw.K;
            a.swG = sw.G;
        }
        a.mask_out = mask_out;
        const uint64_t TW = 256ull * a.R;
        const uint64_t TWB = (a.in_words + TW - 1) / TW;
        if (TWB >= (1ull << 31)) return AM_E_UNSUPPORTED_CONFIG;
        if (c.bits > 32) hipLaunchKernelGGL((k_filter_words<true, false>), dim3((uint32_t)TWB), dim3(256), 0, s, a);
        else if (a.pow2) hipLaunchKernelGGL((k_filter_words<false, false>), dim3((uint32_t)TWB), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_filter_words<false, true>), dim3((uint32_t)TWB), dim3(256), 0, s, a);
        return check_launch();
    }
    if (c.kind == K_SLAB) c.kind = K_BITSTREAM; // 1-bit slabs are dense streams
    if (c.kind == K_BITSTREAM && c.bits <= 32) {
        const uint64_t TS = (c.n + FS_TILE - 1) / FS_TILE;
        if (TS >= (1ull << 31)) return AM_E_UNSUPPORTED_CONFIG;
        FilterStreamArgs fa{};
        fa.in = c;
        fa.none = cyclic_top_range(pr, c.bits, &fa.clo, &fa.crg) ? 0 : 1;
        fa.mask_out = mask_out;
        if (fa.clo == 0) hipLaunchKernelGGL(k_filter_stream<true>, dim3((uint32_t)TS), dim3(256), 0, s, fa);
        else hipLaunchKernelGGL(k_filter_stream<false>, dim3((uint32_t)TS), dim3(256), 0, s, fa);
        return check_launch();
    }
#define FT_GO(IK, W)                                                                                        \
    {                                                                                                       \
        static uint32_t capb = 0;                                                                           \
        if (!capb) capb = resident_blocks(k_filter<IK, W>, 256, 0);                                         \
        const uint32_t g = (uint32_t)(T < capb ? T : capb);                                                 \
        hipLaunchKernelGGL((k_filter<IK, W>), dim3(g), dim3(256), 0, s, c, pr, mask_out, (uint32_t)T); \
    }
    const bool wide = c.bits > 32;
    switch (c.kind) {
        case K_STRIDED: if (wide) FT_GO(K_STRIDED, true) else FT_GO(K_STRIDED, false) break;
        case K_SLAB: if (wide) FT_GO(K_SLAB, true) else FT_GO(K_SLAB, false) break;
        default: if (wide) FT_GO(K_BITSTREAM, true) else FT_GO(K_BITSTREAM, false) break;
    }
#undef FT_GO
    return check_launch();
}

extern "C" int am_sum(const am_column* in, uint64_t* sum_dev, void* ws, uint64_t ws_bytes, void* stream)
{
    Col c;
    int st = make_col(in, &c);
    if (st) return st;
    if (!sum_dev) return AM_E_INVALID;
    if (!ws || ws_bytes < WS_FIXED) return AM_E_WORKSPACE;
    uint64_t units = c.kind == K_SLAB ? (c.n + c.eps - 1) / c.eps
                   : (c.kind == K_BITSTREAM && c.bits <= 32 ? (c.n * c.bits + 63) / 64 : (c.n + 3) / 4);
    // at most 2048 blocks of partial sums, or the tuned count (64..8192)
    uint32_t bcap = am_tuned(AM_KNOB_SUM_BLOCKS, c.bits, in->format, 0);
    bcap = bcap ? (bcap < 64 ? 64 : (bcap > 8192 ? 8192 : bcap)) : 2048;
    uint64_t blocks = (units + 256 * SUM_UNROLL - 1) / (256 * SUM_UNROLL);
    if (blocks < 1) blocks = 1;
    if (blocks > bcap) blocks = bcap;
    uint64_t* partial = ws_small(ws);
    hipStream_t s = (hipStream_t)stream;
    const bool wide = c.bits > 32;
    dim3 g((uint32_t)blocks), b(256);
    const uint32_t fw = c.kind == K_SLAB ? c.bits : 8 * c.cbytes;
    if (c.kind != K_BITSTREAM && fw >= 2 && fw <= 32 && c.n) { // 2-bit slots need fw >= 2
        SumSwar sw{};
        sw.p = c.p;
        sw.fw = fw;
        sw.eps = 64 / fw;
        sw.bits = c.bits;
        sw.words = (c.n + sw.eps - 1) / sw.eps;
        sw.n_last = c.n - (sw.words - 1) * sw.eps;
        const uint64_t vm = lowmask(c.bits);
        const uint32_t ne = sw.eps & 1 ? sw.eps - 1 : sw.eps; // fields in even/odd pairs
        for (uint32_t k = 0; k < ne; k += 2) {
            sw.E |= vm << (k * fw);
            sw.O |= vm << (k * fw); // field k+1 shifted down by fw
        }
        sw.has_top = sw.eps & 1;
        sw.top_sh = (sw.eps - 1) * fw;
        // a slot (2*fw bits) holds K values < 2^bits while K*(2^bits - 1) < 2^(2*fw)
        const uint32_t head = 2 * fw - c.bits;
        sw.K = head >= 16 ? (1u << 16) : (1u << head);
        if (sw.K < SUM_UNROLL) sw.K = SUM_UNROLL;
        uint64_t sb = (sw.words + 256 * SUM_UNROLL - 1) / (256 * SUM_UNROLL);
        if (sb < 1) sb = 1;
        if (sb > bcap) sb = bcap;
        hipLaunchKernelGGL(k_sum_swar, dim3((uint32_t)sb), b, 0, s, sw, partial);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partial, (uint32_t)sb, sum_dev);
        return check_launch();
    }
    if (c.kind == K_BITSTREAM && c.bits >= 2 && c.bits <= 32 && c.n) {
        // the SWAR form over fields at a stride of `bits` (F = floor(64/bits) per word)
        SumSwar sw{};
        sw.p = c.p;
        sw.fw = c.bits;
        sw.eps = 64 / c.bits;
        sw.bits = c.bits;
        sw.words = (c.n * c.bits + 63) / 64;
        const uint64_t vm = lowmask(c.bits);
        const uint32_t ne = sw.eps & 1 ? sw.eps - 1 : sw.eps;
        for (uint32_t k = 0; k < ne; k += 2) {
            sw.E |= vm << (k * c.bits);
            sw.O |= vm << (k * c.bits);
        }
        sw.has_top = sw.eps & 1;
        sw.top_sh = (sw.eps - 1) * c.bits;
        // a 2w-bit slot holds K values < 2^w while K*(2^w - 1) < 2^(2w)
        sw.K = c.bits >= 16 ? (1u << 16) : (1u << c.bits);
        if (sw.K < SUM_UNROLL) sw.K = SUM_UNROLL;
        hipLaunchKernelGGL(k_sum_bits_swar, g, b, 0, s, sw, c.n * c.bits, partial);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partial, (uint32_t)blocks, sum_dev);
        return check_launch();
    }
    if (c.kind == K_STRIDED && c.cbytes == 8 && ((uintptr_t)c.p & 15) == 0 && c.n) {
        uint64_t ub = (c.n / 2 + 256 * SUM_U64_UNROLL - 1) / (256 * SUM_U64_UNROLL);
        if (ub < 1) ub = 1;
        if (ub > bcap) ub = bcap;
        hipLaunchKernelGGL(k_sum_u64, dim3((uint32_t)ub), b, 0, s, (const uint64_t*)c.p, c.n, c.mask, partial);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partial, (uint32_t)ub, sum_dev);
        return check_launch();
    }
    switch (c.kind) {
        case K_STRIDED: hipLaunchKernelGGL((k_sum_partial<K_STRIDED, false>), g, b, 0, s, c, partial); break;
        case K_SLAB: hipLaunchKernelGGL((k_sum_partial<K_SLAB, false>), g, b, 0, s, c, partial); break;
        default:
            if (wide) hipLaunchKernelGGL((k_sum_partial<K_BITSTREAM, true>), g, b, 0, s, c, partial);
            else hipLaunchKernelGGL((k_sum_partial<K_BITSTREAM, false>), g, b, 0, s, c, partial);
            break;
    }
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partial, (uint32_t)blocks, sum_dev);
    return check_launch();
}
