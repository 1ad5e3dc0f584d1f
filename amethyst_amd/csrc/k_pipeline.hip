// k_pipeline.hip -- the per-column operator chain
//     mask = filter(a, pred); s = compressstore(a, mask) + compressstore(b, mask);
//     total = sum(s)
// for a whole batch of columns of uneven length and width in ONE
// launch (plus a one-block epoch bump), capturable in a hipGraph.
//
// The reference has no fused chain: "op chaining" there means one operator's
// output format feeds the next launch (src/kernels/filter_slabs.cuh:47,55 ->
// src/kernels/compressstore_slabs.cuh:89-109 -> src/kernels/hop.cuh:61-76;
// src/op_chaining_demo.cu only runs the benches in sequence). The plain plan
// (amethyst_amd/planner.py: run_plain) reproduces that chain one operator
// launch at a time.
//
// Fused: compress(a)+compress(b) selects the same rows, so s = compress(a+b).
// Each 4096-element tile of every column loads a and b once, evaluates the
// predicate, forms v = (a+b) & mask(w+1) for the selected rows, and compacts
// v into s with the same decoupled look-back as k_compress -- here the scan
// element also carries the running sum, so the column's last tile knows its
// total and count without any atomics. Tiles of all columns are numbered
// globally (column by column) and walked in order by resident blocks, so the
// look-back always makes progress. s may be slab or bitstream at any width
// per column (a planner's choice); values that do not fit are counted so the
// caller can fall back to width w+1.
#include <algorithm>
#include <vector>

#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t PL_WAVES = 4;
// kernel variants: J elements per lane per tile -> tiles of 4*J*64 elements
__host__ __device__ constexpr uint32_t pl_tile(uint32_t J) { return PL_WAVES * J * 64; }

struct PipeJob {
    Col a, b;         // inputs: same n, bits, kind
    Col s;            // output: n = capacity; kind SLAB or BITSTREAM
    Pred pr;
    uint64_t* stats;  // device: {K, overflow, sum, 0}
    uint32_t tile0;   // first global tile
    uint32_t T;       // tiles of this column (>= 1)
    uint32_t vbits;   // width of a+b (w+1, capped at 64)
    uint32_t chk_ovf; // s narrower than vbits
    uint32_t gbase;   // first look-back group of this column
    uint32_t pad;
};

struct PipeArgs {
    const PipeJob* jobs;
    uint32_t njobs;
    uint32_t total_tiles;
    WsHeader* hdr;
    uint64_t* lb_tiles;  // tile descriptors, indexed by global tile
    uint64_t* lb_groups; // group descriptors, per column from its gbase
};

__device__ inline uint32_t pl_find(const PipeJob* jobs, uint32_t n, uint32_t tile)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int IK, uint32_t PL_J>
__global__ __launch_bounds__(256) void k_pipeline(PipeArgs a)
{
    constexpr uint32_t PL_TILE = pl_tile(PL_J);
    constexpr uint32_t F = LB_WIN | LB_OVF | LB_SUM;
    __shared__ uint32_t s_wcnt[PL_WAVES], s_wovf[PL_WAVES];
    __shared__ uint64_t s_wsum[PL_WAVES];
    __shared__ uint64_t s_P, s_carry;
    __shared__ uint32_t s_epoch, s_abort;
    __shared__ uint32_t s_vals[pl_tile(PL_J)]; // a+b of <= 31-bit inputs fits 32 bits

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t g = blockIdx.x; // global tile (one per block, in order)
    const PipeJob& J = a.jobs[pl_find(a.jobs, a.njobs, g)];
    const uint32_t t = g - J.tile0;
    if (tid == 0) {
        s_epoch = ld32(&a.hdr->epoch);
        s_abort = ld32(&a.hdr->abort);
    }
    __syncthreads();
    if (s_abort) {
        if (t == J.T - 1 && tid == 0 && J.stats) J.stats[0] = STATS_INVALID;
        return;
    }
    const uint64_t n = J.a.n;
    const uint32_t bo = J.s.bits;
    const uint64_t om = J.s.mask, vm = lowmask(J.vbits);
    const uint64_t wbase = (uint64_t)t * PL_TILE + wave * (PL_J * 64);
    const uint32_t wrem = wbase < n ? (uint32_t)(n - wbase < PL_J * 64 ? n - wbase : PL_J * 64) : 0;
    const TileReader<IK> ra(J.a, wbase < n ? wbase : 0), rb(J.b, wbase < n ? wbase : 0);
    auto idx = [&](uint32_t j) { return j * 64 + lane < wrem ? j * 64 + lane : (wrem ? wrem - 1 : 0); };

    // ---- 1. select, add, count, sum
    Fetch<IK, false> fa[PL_J], fb[PL_J];
    if (wrem) {
#pragma unroll
        for (uint32_t j = 0; j < PL_J; j++) {
            fa[j].load(ra, idx(j));
            fb[j].load(rb, idx(j));
        }
    }
    uint64_t mine = 0, vsum = 0;
    uint32_t cnt = 0, ovf = 0;
    uint32_t val[PL_J];
#pragma unroll
    for (uint32_t j = 0; j < PL_J; j++) {
        const uint64_t x = wrem ? fa[j].value(ra, idx(j)) : 0, y = wrem ? fb[j].value(rb, idx(j)) : 0;
        const bool fl = (j * 64 + lane < wrem) && J.pr.sel32((uint32_t)x);
        val[j] = (uint32_t)((x + y) & vm);
        const uint64_t bl = ballot(fl);
        if (lane == j) mine = bl;
        cnt += (uint32_t)__popcll(bl);
        vsum += fl ? val[j] : 0;
        if (J.chk_ovf) ovf += (uint32_t)__popcll(ballot(fl && ((uint64_t)val[j] & ~om) != 0));
    }
    vsum = wave_sum_u64(vsum);
    if (lane == 0) {
        s_wcnt[wave] = cnt;
        s_wovf[wave] = ovf;
        s_wsum[wave] = vsum;
    }
    __syncthreads();
    uint32_t wpre = 0, ctile = 0, otile = 0;
    uint64_t stile = 0;
#pragma unroll
    for (uint32_t w = 0; w < PL_WAVES; w++) {
        wpre += w < wave ? s_wcnt[w] : 0;
        ctile += s_wcnt[w];
        otile += s_wovf[w];
        stile += s_wsum[w];
    }
    {
        uint32_t run = wpre;
#pragma unroll
        for (uint32_t j = 0; j < PL_J; j++) {
            const uint64_t bl = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(mine >> 32), j) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((uint32_t)mine, j);
            if ((bl >> lane) & 1) s_vals[run + mbcnt(bl)] = (uint32_t)(val[j] & om);
            run += (uint32_t)__popcll(bl);
        }
    }
    __syncthreads();

    // ---- 2. look-back within the column (wave 0)
    if (wave == 0) {
        uint64_t own_w = 0;
        if (bo < 64) {
            const uint32_t k = (63 + bo - 1) / bo;
            const uint32_t take = ctile < k ? ctile : k;
            if (lane < take) {
                const uint32_t jx = ctile - take + lane;
                const int64_t shift = (int64_t)jx * bo - ((int64_t)ctile * bo - 63);
                const uint64_t v = s_vals[jx];
                own_w = shift >= 0 ? (v << shift) : (v >> (-shift));
            }
            own_w = wave_or_u64(own_w) & WIN_BITS;
        }
        Lookback<F> L{};
        L.epoch = s_epoch;
        L.bits = bo;
        L.hdr = a.hdr;
        L.tiles = a.lb_tiles + (uint64_t)J.tile0 * Lookback<F>::TS;
        L.groups = a.lb_groups + (uint64_t)J.gbase * Lookback<F>::GS;
        const Agg own{ctile, own_w, otile, stile};
        const Agg ex = L.run(t, own, lane);
        if (lane == 0) {
            s_P = ex.c;
            s_carry = ex.w;
            if (t == J.T - 1 && J.stats) {
                J.stats[0] = L.aborted() ? STATS_INVALID : ex.c + ctile;
                J.stats[1] = ovf_add(ex.o, otile);
                J.stats[2] = ex.s + stile;
                J.stats[3] = 0;
            }
        }
    }
    lds_barrier(); // (wave 0's descriptor and stats stores need not drain)
    const uint64_t P = s_P, carry = s_carry;
    const bool last = (t == J.T - 1);
    const uint64_t ocap = J.s.n;

    // ---- 3. write the words this tile completes
    if (J.s.kind == K_SLAB) write_packed_tile<K_SLAB>(J.s, ocap, s_vals, P, carry, ctile, last, tid, 256);
    else write_packed_tile<K_BITSTREAM>(J.s, ocap, s_vals, P, carry, ctile, last, tid, 256);
}

__global__ void k_epoch_bump(WsHeader* hdr)
{
    if (threadIdx.x == 0) st32(&hdr->epoch, ld32(&hdr->epoch) + 1);
}

} // namespace am

using namespace am;

// Workspace: header + small region, job table, then the look-back area:
// tile descriptors for every tile of the batch and (tiles/64 + 1) group
// descriptors per column.
using PlLookback = Lookback<LB_WIN | LB_OVF | LB_SUM>;

static uint64_t pl_lb_bytes(uint64_t tiles, uint64_t groups)
{
    return tiles * PlLookback::TS * 8 + groups * PlLookback::GS * 8;
}

extern "C" uint64_t am_pipeline_workspace_bytes(const uint64_t* n, uint32_t count)
{
    const uint32_t PL_TILE = pl_tile(8); // the smallest variant bounds the tile count
    uint64_t tiles = 1, groups = 1;
    for (uint32_t i = 0; i < count; i++) {
        const uint64_t t = n[i] ? (n[i] + PL_TILE - 1) / PL_TILE : 1;
        tiles += t;
        groups += t / 64 + 1;
    }
    return 256 + WS_SMALL_BYTES + ((uint64_t)count * sizeof(PipeJob) + 255) / 256 * 256 + pl_lb_bytes(tiles, groups);
}

// Host-side validation + job table upload (synchronous); after this,
// am_pipeline_run launches kernels only and can be captured in a graph.
extern "C" int am_pipeline_prepare(const am_pipeline_col* cols, uint32_t count, uint32_t variant, void* ws,
                                   uint64_t ws_bytes, uint32_t* total_tiles)
{
    if (!cols || count == 0 || (variant != 8 && variant != 16)) return AM_E_INVALID;
    const uint32_t PL_TILE = pl_tile(variant);
    if (!ws) return AM_E_WORKSPACE;
    std::vector<uint64_t> ns(count);
    for (uint32_t i = 0; i < count; i++) ns[i] = cols[i].a.n;
    if (ws_bytes < am_pipeline_workspace_bytes(ns.data(), count)) return AM_E_WORKSPACE;
    std::vector<PipeJob> jobs(count);
    uint64_t tiles = 0, groups = 0;
    int kind0 = -1; // input format of the batch
    for (uint32_t i = 0; i < count; i++) {
        const am_pipeline_col& c = cols[i];
        PipeJob& J = jobs[i];
        int st = make_col(&c.a, &J.a);
        if (st) return st;
        if ((st = make_col(&c.b, &J.b))) return st;
        if ((st = make_col(&c.s, &J.s))) return st;
        if (J.a.n != J.b.n || c.a.bits != c.b.bits || c.a.format != c.b.format) return AM_E_INVALID;
        if (c.a.bits > 31) return AM_E_UNSUPPORTED_CONFIG;
        if (J.s.kind == K_STRIDED) return AM_E_UNSUPPORTED_CONFIG;
        if (kind0 < 0) kind0 = (int)c.a.format;
        if ((int)c.a.format != kind0) return AM_E_INVALID; // one input format per batch (the launch's decoder)
        if (c.pred.op > AM_PRED_NE) return AM_E_INVALID;
        J.pr = make_pred(c.pred, J.a.mask);
        J.stats = c.stats;
        J.vbits = c.a.bits + 1;
        J.chk_ovf = c.s.bits < J.vbits ? 1 : 0;
        J.T = J.a.n ? (uint32_t)((J.a.n + PL_TILE - 1) / PL_TILE) : 1;
        J.tile0 = (uint32_t)tiles;
        J.gbase = (uint32_t)groups;
        tiles += J.T;
        groups += J.T / 64 + 1;
        if (tiles >= (1ull << 31)) return AM_E_UNSUPPORTED_CONFIG;
    }
    uint8_t* jt = (uint8_t*)ws + 256 + WS_SMALL_BYTES;
    if (hipMemcpy(jt, jobs.data(), count * sizeof(PipeJob), hipMemcpyHostToDevice) != hipSuccess) return AM_E_HIP;
    if (total_tiles) *total_tiles = (uint32_t)tiles;
    // the group area starts right after the tile area of this batch
    uint64_t meta[2] = {tiles, groups};
    if (hipMemcpy((uint8_t*)ws + 32, meta, sizeof(meta), hipMemcpyHostToDevice) != hipSuccess) return AM_E_HIP;
    return AM_OK;
}

extern "C" int am_pipeline_run(uint32_t count, uint32_t total_tiles, uint32_t in_format, uint32_t variant, void* ws,
                               uint64_t ws_bytes, void* stream)
{
    if (!ws || count == 0 || total_tiles == 0 || (variant != 8 && variant != 16)) return AM_E_INVALID;
    PipeArgs a{};
    a.jobs = (const PipeJob*)((uint8_t*)ws + 256 + WS_SMALL_BYTES);
    a.njobs = count;
    a.total_tiles = total_tiles;
    a.hdr = (WsHeader*)ws;
    const uint64_t off = 256 + WS_SMALL_BYTES + ((uint64_t)count * sizeof(PipeJob) + 255) / 256 * 256;
    if (ws_bytes < off + pl_lb_bytes(total_tiles + 1, 1)) return AM_E_WORKSPACE;
    // Every launch of a prepared batch rewrites all of its tile descriptors,
    // and group aggregates are validated by their status word's epoch, so
    // nothing here needs slice clearing.
    a.lb_tiles = (uint64_t*)((uint8_t*)ws + off);
    a.lb_groups = a.lb_tiles + (uint64_t)total_tiles * PlLookback::TS;
    hipStream_t s = (hipStream_t)stream;
#define PL_GO(IK, JJ) hipLaunchKernelGGL((k_pipeline<IK, JJ>), dim3(total_tiles), dim3(256), 0, s, a);
#define PL_V(IK) { if (variant == 8) { PL_GO(IK, 8) } else { PL_GO(IK, 16) } }
    switch (in_format) {
        case AM_FMT_SLAB: PL_V(K_SLAB); break;
        case AM_FMT_BITSTREAM: PL_V(K_BITSTREAM); break;
        default: PL_V(K_STRIDED); break;
    }
#undef PL_V
#undef PL_GO
    hipLaunchKernelGGL(k_epoch_bump, dim3(1), dim3(64), 0, s, (WsHeader*)ws);
    return check_launch();
}
