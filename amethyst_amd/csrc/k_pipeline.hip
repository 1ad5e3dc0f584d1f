// k_pipeline.hip -- the per-column operator chain
//     mask = filter(a, pred); s = compressstore(a, mask) + compressstore(b, mask);
//     total = sum(s)
// for a whole batch of columns of uneven length, width and tile size in ONE
// launch, capturable in a hipGraph.
//
// The reference has no fused chain: "op chaining" there means one operator's
// output format feeds the next launch (src/kernels/filter_slabs.cuh:47,55 ->
// src/kernels/compressstore_slabs.cuh:89-109 -> src/kernels/hop.cuh:61-76;
// src/op_chaining_demo.cu only runs the benches in sequence). The plain plan
// (amethyst_amd/planner.py: run_plain) reproduces that chain one operator
// launch at a time.
//
// Fused: compress(a)+compress(b) selects the same rows, so s = compress(a+b).
// One block per tile, tiles numbered column by column; a column's tile size
// (J elements per lane: 2048 or 4096 elements) is part of its plan. A block
//   1. has each lane fetch its J elements of a and of b (element base +
//      64j + lane; all loads in flight before any is used), evaluates the
//      predicate on a and forms v = (a+b) & mask(w+1); the wave64 ballot of
//      the predicate is the selection of 64 consecutive rows and v_mbcnt
//      gives the ranks;
//   2. compacts the selected v into LDS and looks back within its column
//      with the decoupled look-back of the compaction
//      kernel, whose scan element here also carries the running sum -- so the
//      column's last tile knows its count, overflow and total without atomics;
//   3. writes the output words it completes (slab or bitstream at the
//      column's planned width).
// The last tile of the last column to finish advances the launch epoch
// (stale descriptors are recognised by their epoch; nothing is cleared).
// Each launch checks the workspace's recorded tile count and size against
// its arguments and does nothing but mark every column's statistics invalid
// when they differ.
#include <algorithm>
#include <vector>

#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t PL_WAVES = 4;
constexpr uint32_t PL_JMAX = 16;
__host__ __device__ constexpr uint32_t pl_tile(uint32_t J) { return PL_WAVES * J * 64; }
constexpr uint32_t PL_TMAX = pl_tile(PL_JMAX);
constexpr uint32_t PL_MAGIC = 0x41504C31u; // "APL1": the workspace holds a prepared batch

struct PipeJob {
    Col a, b;         // inputs: same n, bits, format
    Col s;            // output: n = capacity; kind SLAB or BITSTREAM
    Pred pr;
    uint64_t* stats;  // device: {K, overflow, sum, 0}
    uint32_t tile0;   // first global tile
    uint32_t T;       // tiles of this column (>= 1)
    uint32_t J;       // elements per lane per tile (8 or 16)
    uint32_t vbits;   // width of a+b (w+1)
    uint32_t chk_ovf; // s narrower than vbits
    uint32_t gbase;   // first look-back group of this column
    uint32_t pad0, pad1;
};

// Recorded by prepare in the workspace header's spare words.
struct PipeMeta {
    uint32_t magic, count, tiles, done; // done: columns finished in this launch
    uint64_t need;                      // workspace bytes the batch needs
    uint64_t off_tiles, off_groups;     // byte offsets of the look-back areas
};
static_assert(sizeof(PipeMeta) <= sizeof(WsHeader) - 8, "meta fits the header");
__host__ __device__ inline PipeMeta* pl_meta(void* ws) { return (PipeMeta*)((uint8_t*)ws + 8); }

struct PipeArgs {
    uint8_t* ws;
    const PipeJob* jobs;
    WsHeader* hdr;
    PipeMeta* meta;
    uint32_t total_tiles;
    uint64_t ws_bytes;
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

// One tile of one column: decode, select, add, count and sum; compact the
// selected values into LDS; look back within the column; write the output
// words the tile completes. IK is the column's input kind (block-uniform).
template <int IK, bool WIDEJ>
__device__ inline void pl_tile_run(const PipeArgs& a, const PipeJob& J, uint32_t t, uint32_t epoch, uint32_t* s_vals,
                                   uint32_t* s_wcnt, uint32_t* s_wovf, uint64_t* s_wsum, uint64_t* s_P,
                                   uint64_t* s_carry)
{
    constexpr uint32_t F = LB_WIN | LB_OVF | LB_SUM;
    constexpr uint32_t JJ = WIDEJ ? 16 : 8;
    constexpr uint32_t TILE = pl_tile(JJ);
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t n = J.a.n;
    const uint32_t bo = J.s.bits;
    const uint64_t om = J.s.mask, vm = lowmask(J.vbits);
    const uint64_t wbase = (uint64_t)t * TILE + wave * (JJ * 64);
    const uint32_t wrem = wbase < n ? (uint32_t)(n - wbase < JJ * 64 ? n - wbase : JJ * 64) : 0;
    const TileReader<IK> ra(J.a, wbase < n ? wbase : 0), rb(J.b, wbase < n ? wbase : 0);
    auto idx = [&](uint32_t j) { return j * 64 + lane < wrem ? j * 64 + lane : (wrem ? wrem - 1 : 0); };

    // ---- 1. select, add, count, sum (every load of the lane in flight first)
    Fetch<IK, false> fa[JJ], fb[JJ];
    if (wrem) {
#pragma unroll
        for (uint32_t j = 0; j < JJ; j++) {
            fa[j].load(ra, idx(j));
            fb[j].load(rb, idx(j));
        }
    }
    uint64_t mine = 0, vsum = 0;
    uint32_t cnt = 0, ovf = 0;
    uint32_t val[JJ];
#pragma unroll
    for (uint32_t j = 0; j < JJ; j++) {
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
        for (uint32_t j = 0; j < JJ; j++) {
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
        L.epoch = epoch;
        L.bits = bo;
        L.hdr = a.hdr;
        L.tiles = (uint64_t*)(a.ws + a.meta->off_tiles) + (uint64_t)J.tile0 * Lookback<F>::TS;
        L.groups = (uint64_t*)(a.ws + a.meta->off_groups) + (uint64_t)J.gbase * Lookback<F>::GS;
        const Agg own{ctile, own_w, otile, stile};
        const Agg ex = L.run(t, own, lane);
        if (lane == 0) {
            *s_P = ex.c;
            *s_carry = ex.w;
            if (t == J.T - 1) {
                if (J.stats) {
                    J.stats[0] = L.aborted() ? STATS_INVALID : ex.c + ctile;
                    J.stats[1] = ovf_add(ex.o, otile);
                    J.stats[2] = ex.s + stile;
                    J.stats[3] = 0;
                }
                // every tile of this column has published (so has read the
                // epoch); the last column to get here advances it
                if (add32(&a.meta->done, 1u) == a.meta->count - 1) {
                    st32(&a.meta->done, 0u);
                    st32(&a.hdr->epoch, epoch + 1);
                }
            }
        }
    }
    lds_barrier(); // (wave 0's descriptor and stats stores need not drain)
    const uint64_t P = *s_P, carry = *s_carry;
    const bool last = (t == J.T - 1);
    const uint64_t ocap = J.s.n;

    // ---- 3. write the words this tile completes
    if (J.s.kind == K_SLAB) write_packed_tile<K_SLAB>(J.s, ocap, s_vals, P, carry, ctile, last, tid, 256);
    else write_packed_tile<K_BITSTREAM>(J.s, ocap, s_vals, P, carry, ctile, last, tid, 256);
}

__global__ __launch_bounds__(256) void k_pipeline(PipeArgs a)
{
    __shared__ uint32_t s_vals[PL_TMAX]; // a+b of <= 31-bit inputs fits 32 bits
    __shared__ uint32_t s_wcnt[PL_WAVES], s_wovf[PL_WAVES];
    __shared__ uint64_t s_wsum[PL_WAVES];
    __shared__ uint64_t s_P, s_carry;
    __shared__ uint32_t s_epoch, s_abort;

    const uint32_t tid = threadIdx.x;
    const uint32_t g = blockIdx.x; // global tile (one per block, in order)
    // ---- the launch must match the prepared batch
    {
        const PipeMeta m = *a.meta;
        if (m.magic != PL_MAGIC) return; // never prepared: nothing to run or report
        if (m.tiles != a.total_tiles || m.need > a.ws_bytes) {
            if (g == 0)
                for (uint32_t c = tid; c < m.count; c += 256)
                    if (a.jobs[c].stats) a.jobs[c].stats[0] = STATS_INVALID;
            return;
        }
    }
    const PipeJob& J = a.jobs[pl_find(a.jobs, a.meta->count, g)];
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
#define PL_RUN(IK)                                                                                             \
    {                                                                                                          \
        if (J.J == 16) pl_tile_run<IK, true>(a, J, t, s_epoch, s_vals, s_wcnt, s_wovf, s_wsum, &s_P, &s_carry);  \
        else pl_tile_run<IK, false>(a, J, t, s_epoch, s_vals, s_wcnt, s_wovf, s_wsum, &s_P, &s_carry);          \
    }
    switch (J.a.kind) { // (one kind per column; in practice one per batch)
        case K_SLAB: PL_RUN(K_SLAB); break;
        case K_BITSTREAM: PL_RUN(K_BITSTREAM); break;
        default: PL_RUN(K_STRIDED); break;
    }
#undef PL_RUN
}

} // namespace am

using namespace am;

// Workspace: header (epoch, abort, the batch's meta) + small region, job
// table, then the look-back area: tile descriptors for every tile of the
// batch and (tiles/64 + 1) group descriptors per column.
using PlLookback = Lookback<LB_WIN | LB_OVF | LB_SUM>;

static uint64_t pl_lb_bytes(uint64_t tiles, uint64_t groups)
{
    return tiles * PlLookback::TS * 8 + groups * PlLookback::GS * 8;
}

static uint64_t pl_jobs_bytes(uint32_t count) { return ((uint64_t)count * sizeof(PipeJob) + 255) / 256 * 256; }

extern "C" uint64_t am_pipeline_workspace_bytes(const uint64_t* n, uint32_t count)
{
    const uint32_t PL_TILE = pl_tile(8); // the smallest tile bounds the tile count
    uint64_t tiles = 1, groups = 1;
    for (uint32_t i = 0; i < count; i++) {
        const uint64_t t = n[i] ? (n[i] + PL_TILE - 1) / PL_TILE : 1;
        tiles += t;
        groups += t / 64 + 1;
    }
    return 256 + WS_SMALL_BYTES + pl_jobs_bytes(count) + pl_lb_bytes(tiles, groups);
}

// Host-side validation + job table upload; after this, am_pipeline_run
// launches one kernel only and can be captured in a graph.
extern "C" int am_pipeline_prepare(const am_pipeline_col* cols, uint32_t count, void* ws, uint64_t ws_bytes,
                                   uint32_t* total_tiles, void* stream)
{
    if (!cols || count == 0) return AM_E_INVALID;
    if (!ws) return AM_E_WORKSPACE;
    std::vector<PipeJob> jobs(count);
    uint64_t tiles = 0, groups = 0;
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
        if (c.pred.op > AM_PRED_NE) return AM_E_INVALID;
        uint32_t v = c.variant;
        if (v == 0) {
            v = am_tuned(AM_KNOB_PIPELINE_TILE, c.a.bits, c.a.format, c.s.format);
            if (v == 0) v = 16;
        }
        if (v != 8 && v != 16) return AM_E_INVALID;
        J.J = v;
        J.pr = make_pred(c.pred, J.a.mask);
        J.stats = c.stats;
        J.vbits = c.a.bits + 1;
        J.chk_ovf = c.s.bits < J.vbits ? 1 : 0;
        const uint32_t TL = pl_tile(v);
        J.T = J.a.n ? (uint32_t)((J.a.n + TL - 1) / TL) : 1;
        J.tile0 = (uint32_t)tiles;
        J.gbase = (uint32_t)groups;
        tiles += J.T;
        groups += J.T / 64 + 1;
        if (tiles >= (1ull << 31)) return AM_E_UNSUPPORTED_CONFIG;
    }
    const uint64_t need = 256 + WS_SMALL_BYTES + pl_jobs_bytes(count) + pl_lb_bytes(tiles, groups);
    if (ws_bytes < need) return AM_E_WORKSPACE;
    // the previous run on this stream may still read the job table
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return AM_E_HIP;
    uint8_t* jt = (uint8_t*)ws + 256 + WS_SMALL_BYTES;
    if (hipMemcpy(jt, jobs.data(), count * sizeof(PipeJob), hipMemcpyHostToDevice) != hipSuccess) return AM_E_HIP;
    const uint64_t off_t = 256 + WS_SMALL_BYTES + pl_jobs_bytes(count);
    PipeMeta m{PL_MAGIC, count, (uint32_t)tiles, 0, need, off_t, off_t + tiles * PlLookback::TS * 8};
    if (hipMemcpy(pl_meta(ws), &m, sizeof(m), hipMemcpyHostToDevice) != hipSuccess) return AM_E_HIP;
    if (total_tiles) *total_tiles = (uint32_t)tiles;
    return AM_OK;
}

extern "C" int am_pipeline_run(void* ws, uint64_t ws_bytes, uint32_t total_tiles, void* stream)
{
    if (!ws || ws_bytes < 256 + WS_SMALL_BYTES) return AM_E_WORKSPACE;
    if (total_tiles == 0) return AM_E_INVALID;
    // everything else -- job table, look-back areas, the tile count the grid
    // must have -- was recorded by am_pipeline_prepare in the workspace and
    // is checked by the kernel itself (the launch stays graph-capturable)
    PipeArgs a{};
    a.ws = (uint8_t*)ws;
    a.hdr = (WsHeader*)ws;
    a.meta = pl_meta(ws);
    a.jobs = (const PipeJob*)((uint8_t*)ws + 256 + WS_SMALL_BYTES);
    a.total_tiles = total_tiles;
    a.ws_bytes = ws_bytes;
    hipLaunchKernelGGL(k_pipeline, dim3(total_tiles), dim3(256), 0, (hipStream_t)stream, a);
    return check_launch();
}
