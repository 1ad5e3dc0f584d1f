// k_pipeline.hip -- the per-column operator chain
//     mask = filter(a, pred); s = compressstore(a, mask) + compressstore(b, mask);
//     total = sum of the selected a+b (exact, whatever width s is stored at)
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
// (J elements per lane: 2048, 4096 or 8192 elements) is part of its plan. A block
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
#include <type_traits>
#include <vector>

#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t PL_WAVES = 4;
__host__ __device__ constexpr uint32_t pl_tile(uint32_t J) { return PL_WAVES * J * 64; }
// LDS staging of a tile's compacted values, in bytes: with the block's other
// shared words just under 40 KB, so four blocks fill a CU's 160 KB
constexpr uint32_t PL_STAGE_BYTES = 40448;
constexpr uint32_t PL_TMAX = PL_STAGE_BYTES / 4; // u32 values
constexpr uint32_t PL_MAGIC = 0x41504C31u; // "APL1": the workspace holds a prepared batch

struct PipeJob {
    Col a, b;         // inputs: same n, bits, format
    Col s;            // output: n = capacity; kind SLAB or BITSTREAM
    Pred pr;
    uint64_t* stats;  // device: {K, overflow, sum, 0}
    uint32_t tile0;   // first global tile
    uint32_t T;       // tiles of this column (>= 1)
    uint32_t J;       // element path: elements per lane per tile (8..128)
    uint32_t vbits;   // width of a+b (w+1)
    uint32_t chk_ovf; // s narrower than vbits
    uint32_t gbase;   // first look-back group of this column
    uint32_t wsub;    // word path (slab a, b of 2..31 bits): 256-word sub-passes per tile; 0: element path
    uint32_t vsz;     // bytes per staged value: 1, 2 or 4 (s.bits rounded up)
    TopPred tp;       // word path: the predicate over top-aligned fields
    uint32_t pad1;
};

// A block's work: column `job`, tile t of it (job PL_IDLE: a padding block).
struct PlSlot {
    uint32_t job, t;
};
constexpr uint32_t PL_IDLE = 0xFFFFFFFFu;
// The GPU's XCDs (MI355X: 8, each with its own L2); workgroups go to them
// round-robin by index, in order within each.
constexpr uint32_t PL_XCDS = 8;

// Recorded by prepare in the workspace header's spare words.
struct PipeMeta {
    uint32_t magic, count, tiles, done; // done: columns finished in this launch
    uint32_t grid, xcd_local;           // blocks of a run; columns kept on one XCD each
    uint64_t need;                      // workspace bytes the batch needs
    uint64_t off_jobs;                  // byte offset of the job table
    uint64_t off_tiles, off_groups;     // byte offsets of the look-back areas
};
static_assert(sizeof(PipeMeta) <= sizeof(WsHeader) - 8, "meta fits the header");
__host__ __device__ inline PipeMeta* pl_meta(void* ws) { return (PipeMeta*)((uint8_t*)ws + 8); }

struct PipeArgs {
    uint8_t* ws;
    const PlSlot* map; // block -> (column, tile)
    const PipeJob* jobs;
    WsHeader* hdr;
    PipeMeta* meta;
    uint32_t total_tiles;
    uint64_t ws_bytes;
};

// One tile of one column: decode, select, add, count and sum; compact the
// selected values into LDS; look back within the column; write the output
// words the tile completes. IK is the column's input kind (block-uniform).
// Tile totals of phase 1 (every thread holds them after the stage).
struct TileTot {
    uint32_t ctile, otile;
    uint64_t stile;
};

// Phase 1: the tile's J*256 elements in J/8 sub-passes of 2048 (8 per
// lane); each lane fetches its elements of sub-pass sp+1 (all loads in
// flight) before it uses those of sp, so a tile waits for one memory round
// trip plus the sub-passes' arithmetic rather than one round trip each.
// Per sub-pass the wave64 ballot of the predicate selects 64 consecutive
// rows and the selected a+b go to s_vals at their tile ranks (sub-pass base
// + the waves before + the rank in the ballots).
constexpr uint32_t PL_SUB = 8; // elements per lane per sub-pass
template <int IK>
struct PlStage {
    Fetch<IK, false> fa[PL_SUB], fb[PL_SUB];
};
template <int IK, typename V>
__device__ inline TileTot pl_select_elems(const PipeJob& J, uint32_t t, V* s_vals, uint32_t* s_wcnt,
                                          uint32_t* s_wovf, uint64_t* s_wsum)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint64_t n = J.a.n;
    const uint64_t om = J.s.mask, vm = lowmask(J.vbits);
    const uint32_t nsub = J.J / PL_SUB;
    const uint64_t t0 = (uint64_t)t * pl_tile(J.J);
    uint64_t vsum = 0;
    uint32_t ovf = 0, base = 0;
    // this wave's first element of sub-pass sp, and how many of its 512 exist
    auto wbase = [&](uint32_t sp) { return t0 + sp * pl_tile(PL_SUB) + wave * (PL_SUB * 64); };
    auto wrem = [&](uint64_t wb) -> uint32_t {
        return wb < n ? (uint32_t)(n - wb < PL_SUB * 64 ? n - wb : PL_SUB * 64) : 0;
    };
    auto idx = [&](uint32_t j, uint32_t rem) { return j * 64 + lane < rem ? j * 64 + lane : (rem ? rem - 1 : 0); };
    auto load = [&](PlStage<IK>& S, uint32_t sp) {
        const uint64_t wb = wbase(sp);
        const uint32_t rem = wrem(wb);
        if (!rem) { // (past the column's end: nothing is selected)
#pragma unroll
            for (uint32_t j = 0; j < PL_SUB; j++) S.fa[j].raw = S.fb[j].raw = 0;
            return;
        }
        const TileReader<IK> ra(J.a, wb), rb(J.b, wb);
#pragma unroll
        for (uint32_t j = 0; j < PL_SUB; j++) {
            S.fa[j].load(ra, idx(j, rem));
            S.fb[j].load(rb, idx(j, rem));
        }
    };
    auto process = [&](const PlStage<IK>& S, uint32_t sp) {
        const uint64_t wb = wbase(sp);
        const uint32_t rem = wrem(wb);
        const TileReader<IK> ra(J.a, rem ? wb : 0), rb(J.b, rem ? wb : 0);
        uint64_t mine = 0;
        uint32_t cnt = 0;
        uint32_t val[PL_SUB];
#pragma unroll
        for (uint32_t j = 0; j < PL_SUB; j++) {
            const uint64_t x = S.fa[j].value(ra, idx(j, rem)), y = S.fb[j].value(rb, idx(j, rem));
            const bool fl = (j * 64 + lane < rem) && J.pr.sel32((uint32_t)x);
            val[j] = (uint32_t)((x + y) & vm);
            const uint64_t bl = ballot(fl);
            if (lane == j) mine = bl;
            cnt += (uint32_t)__popcll(bl);
            vsum += fl ? val[j] : 0;
            // (an output at least as wide as a+b never overflows: val & ~om is 0)
            ovf += (uint32_t)__popcll(ballot(fl && ((uint64_t)val[j] & ~om) != 0));
        }
        // (double-buffered counts: a wave reads sub-pass sp's before the
        // barrier of sp+1, so sp+2 may overwrite them)
        uint32_t* wc = s_wcnt + (sp & 1) * PL_WAVES;
        if (lane == 0) wc[wave] = cnt;
        lds_barrier();
        uint32_t run = base, tot = 0;
#pragma unroll
        for (uint32_t w = 0; w < PL_WAVES; w++) {
            run += w < wave ? wc[w] : 0;
            tot += wc[w];
        }
#pragma unroll
        for (uint32_t j = 0; j < PL_SUB; j++) {
            const uint64_t bl = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(mine >> 32), j) << 32) |
                                (uint32_t)__builtin_amdgcn_readlane((uint32_t)mine, j);
            if ((bl >> lane) & 1) s_vals[run + mbcnt(bl)] = (V)(val[j] & om);
            run += (uint32_t)__popcll(bl);
        }
        base += tot;
    };
    PlStage<IK> A, B;
    load(A, 0);
    for (uint32_t sp = 0; sp < nsub; sp += 2) {
        if (sp + 1 < nsub) load(B, sp + 1);
        process(A, sp);
        if (sp + 1 < nsub) {
            if (sp + 2 < nsub) load(A, sp + 2);
            process(B, sp + 1);
        }
    }
    vsum = wave_sum_u64(vsum);
    if (lane == 0) {
        s_wovf[wave] = ovf;
        s_wsum[wave] = vsum;
    }
    __syncthreads();
    uint32_t otile = 0;
    uint64_t stile = 0;
#pragma unroll
    for (uint32_t w = 0; w < PL_WAVES; w++) {
        otile += s_wovf[w];
        stile += s_wsum[w];
    }
    return TileTot{base, otile, stile};
}

// Phase 1 of the word path (slab a and b of 2..31 bits, eps = 64/bits
// fields per word): the tile is J.wsub sub-passes of 256 words, lane
// `tid` takes word tid of each (so a wave's words, and their rows, are
// consecutive); the words of the next eight sub-passes are in flight while
// one is used (four: 2 % slower on the 512-column batch). Per word: the selection bits of its eps fields at once
// (select_fields: one compare and one add-with-carry per field), the rank
// of the lane's first selected row by a DPP scan of the per-lane counts
// plus the waves before, then a+b of every field (SWAR) to s_vals.
constexpr uint32_t PL_RING = 8;
template <typename V>
__device__ inline TileTot pl_select_words(const PipeJob& J, uint32_t t, V* s_vals, uint32_t* s_wcnt,
                                          uint32_t* s_wovf, uint64_t* s_wsum)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t w = J.a.bits, eps = J.a.eps, nsub = J.wsub;
    const uint32_t m = (uint32_t)J.a.mask;
    const uint32_t om = J.s.bits >= 32 ? 0xFFFFFFFFu : (uint32_t)J.s.mask;
    const uint64_t words = J.a.words, n = J.a.n;
    const uint64_t tw0 = (uint64_t)t * nsub * 256; // first word of the tile
    const AM_GLOBAL uint64_t* pa = gmem(J.a.p);
    const AM_GLOBAL uint64_t* pb = gmem(J.b.p);
    // the tile reaches past the column's last element: mask fields >= n
    const bool tail = (tw0 + (uint64_t)nsub * 256) * eps > n;
    // SWAR sums: me covers the even fields 0, 2, 4, ... of a word, mo the
    // even fields f that have an odd field f+1 above them; a sum of two
    // w-bit fields takes w+1 bits, its carry landing in the (masked-out)
    // place of the field above
    uint64_t me = 0, mo = 0;
    for (uint32_t f = 0; f < eps; f += 2) {
        me |= (uint64_t)m << (f * w);
        if (f + 1 < eps) mo |= (uint64_t)m << (f * w);
    }
    const uint32_t m1 = w >= 31 ? 0xFFFFFFFFu : (2u << w) - 1; // w+1 bits
    // unselected fields are stored too, each lane to a slot of its own past
    // the staging area (no branch around the store)
    const uint32_t trash = PL_STAGE_BYTES / (uint32_t)sizeof(V) + lane * (4 / (uint32_t)sizeof(V));
    uint64_t X[PL_RING], Y[PL_RING];
    auto fetch = [&](uint32_t k, uint32_t sp) {
        const uint64_t wi = tw0 + sp * 256 + tid;
        X[k] = wi < words ? pa[wi] : 0;
        Y[k] = wi < words ? pb[wi] : 0;
    };
#pragma unroll
    for (uint32_t k = 0; k < PL_RING; k++)
        if (k < nsub) fetch(k, k);
    uint64_t vsum = 0;
    uint32_t ovf = 0, base = 0;
    for (uint32_t sp0 = 0; sp0 < nsub; sp0 += PL_RING) {
#pragma unroll
        for (uint32_t k = 0; k < PL_RING; k++) {
            const uint32_t sp = sp0 + k;
            if (sp >= nsub) break;
            const uint64_t x = X[k], y = Y[k];
            if (sp + PL_RING < nsub) fetch(k, sp + PL_RING);
            uint32_t gg = select_fields(x, eps, w, w, J.tp);
            if (tail) {
                const uint64_t e0 = (tw0 + sp * 256 + tid) * eps;
                const uint64_t left = e0 < n ? n - e0 : 0;
                gg &= left >= eps ? 0xFFFFFFFFu : (uint32_t)((1ull << left) - 1);
            }
            const uint32_t c = (uint32_t)__popc(gg);
            const uint32_t incl = wave_incl_scan(c);
            // (double-buffered counts: a wave reads sub-pass sp's before the
            // barrier of sp+1, so sp+2 may overwrite them)
            uint32_t* wc = s_wcnt + (sp & 1) * PL_WAVES;
            if (lane == 63) wc[wave] = incl;
            lds_barrier();
            uint32_t pos = base + incl - c, tot = 0;
#pragma unroll
            for (uint32_t v = 0; v < PL_WAVES; v++) {
                pos += v < wave ? wc[v] : 0;
                tot += wc[v];
            }
            // a+b of every field of the word: even fields in E, odd fields
            // (shifted down by w) in O, each at its even field's place
            const uint64_t E = (x & me) + (y & me);
            const uint64_t O = ((x >> w) & mo) + ((y >> w) & mo);
            // (an output at least w+1 bits wide holds every a+b: no masking,
            // no overflow count -- a separate instance of the field loop)
            auto fields = [&](auto chkc) {
                constexpr bool CHK = decltype(chkc)::value;
                auto one = [&](uint64_t S, uint32_t f, uint32_t sh) {
                    const uint32_t v = (uint32_t)(S >> sh) & m1; // < 2^32
                    const bool sel = __builtin_amdgcn_ubfe(gg, f, 1) != 0;
                    s_vals[sel ? pos : trash] = (V)(CHK ? (v & om) : v);
                    pos += sel ? 1u : 0u;
                    vsum += sel ? v : 0u;
                    if (CHK) ovf += (sel && (v & ~om) != 0) ? 1 : 0;
                };
                for (uint32_t f = 0, sh = 0; f < eps; f += 2, sh += 2 * w) {
                    one(E, f, sh);
                    if (f + 1 < eps) one(O, f + 1, sh);
                }
            };
            if (J.chk_ovf) fields(std::true_type{});
            else fields(std::false_type{});
            base += tot;
        }
    }
    vsum = wave_sum_u64(vsum);
    ovf = wave_sum32(ovf);
    if (lane == 0) {
        s_wovf[wave] = ovf;
        s_wsum[wave] = vsum;
    }
    __syncthreads();
    uint32_t otile = 0;
    uint64_t stile = 0;
#pragma unroll
    for (uint32_t v = 0; v < PL_WAVES; v++) {
        otile += s_wovf[v];
        stile += s_wsum[v];
    }
    return TileTot{base, otile, stile};
}

// Phases 2 and 3 of a tile: look back within the column (wave 0), then
// write the output words the tile completes.
template <typename V>
__device__ inline void pl_finish(const PipeArgs& a, const PipeMeta& m, const PipeJob& J, uint32_t t, uint32_t epoch,
                                 const V* s_vals,
                                 const TileTot& tt, uint64_t* s_P, uint64_t* s_carry)
{
    constexpr uint32_t F = LB_WIN | LB_OVF | LB_SUM;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t bo = J.s.bits;
    const uint32_t ctile = tt.ctile, otile = tt.otile;
    const uint64_t stile = tt.stile;

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
        L.tiles = (uint64_t*)(a.ws + m.off_tiles) + (uint64_t)J.tile0 * Lookback<F>::TS;
        L.groups = (uint64_t*)(a.ws + m.off_groups) + (uint64_t)J.gbase * Lookback<F>::GS;
        const Agg own{ctile, own_w, otile, stile};
        const Agg ex = L.run(t, own, lane);
        if (lane == 0) {
            *s_P = ex.c;
            *s_carry = ex.w;
            if (t == J.T - 1) {
                if (J.stats) {
                    AM_GLOBAL uint64_t* st = gmem(J.stats);
                    st[0] = L.aborted() ? STATS_INVALID : ex.c + ctile;
                    st[1] = ovf_add(ex.o, otile);
                    st[2] = ex.s + stile;
                    st[3] = 0;
                }
                // every tile of this column has published (so has read the
                // epoch); the last column to get here advances it
                if (add32(&a.meta->done, 1u) == m.count - 1) {
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
    // a+b of <= 31-bit inputs fits 32 bits; 64 more words take the word
    // path's stores of unselected fields
    __shared__ uint32_t s_vals[PL_TMAX + 64];
    __shared__ uint32_t s_wcnt[2 * PL_WAVES], s_wovf[PL_WAVES];
    __shared__ uint64_t s_wsum[PL_WAVES];
    __shared__ uint64_t s_P, s_carry;
    __shared__ uint32_t s_epoch, s_abort;

    const uint32_t tid = threadIdx.x;
    const uint32_t g = blockIdx.x;
    // ---- the launch must match the prepared batch (the block map lies
    // within the workspace whatever it holds: read with the meta at once)
    const PipeMeta m = ldk(a.meta);
    const PlSlot e = ldk(a.map + g);
    if (m.magic != PL_MAGIC) return; // never prepared: nothing to run or report
    if (m.grid != a.total_tiles || m.need > a.ws_bytes || (e.job != PL_IDLE && e.job >= m.count)) {
        // (the table as prepared -- its place depends on the prepared tile
        // count -- within the extent the prepare checked)
        const PipeJob* jobs = (const PipeJob*)(a.ws + m.off_jobs);
        if (g == 0)
            for (uint32_t c = tid; c < m.count; c += 256)
                if (jobs[c].stats) jobs[c].stats[0] = STATS_INVALID;
        return;
    }
    if (e.job == PL_IDLE) return; // (an XCD's list shorter than the longest)
    if (tid == 0) {
        s_epoch = ld32(&a.hdr->epoch);
        s_abort = ld32(&a.hdr->abort);
    }
    const PipeJob J = ldk(a.jobs + e.job); // (uniform: scalar registers for the whole tile)
    const uint32_t t = e.t;
    __syncthreads();
    if (s_abort) {
        if (t == J.T - 1 && tid == 0 && J.stats) gmem(J.stats)[0] = STATS_INVALID;
        return;
    }
    // the compacted values are staged at 1, 2 or 4 bytes (the output width
    // rounded up), so narrow outputs get larger tiles in the same LDS
    auto tile = [&](auto* sv) {
        TileTot tt;
        if (J.wsub) tt = pl_select_words(J, t, sv, s_wcnt, s_wovf, s_wsum);
        else switch (J.a.kind) { // (one kind per column; in practice one per batch)
            case K_SLAB: tt = pl_select_elems<K_SLAB>(J, t, sv, s_wcnt, s_wovf, s_wsum); break;
            case K_BITSTREAM: tt = pl_select_elems<K_BITSTREAM>(J, t, sv, s_wcnt, s_wovf, s_wsum); break;
            default: tt = pl_select_elems<K_STRIDED>(J, t, sv, s_wcnt, s_wovf, s_wsum); break;
        }
        pl_finish(a, m, J, t, s_epoch, sv, tt, &s_P, &s_carry);
    };
    if (J.vsz == 1) tile((uint8_t*)s_vals);
    else if (J.vsz == 2) tile((uint16_t*)s_vals);
    else tile(s_vals);
}

} // namespace am

using namespace am;

// Workspace: header (epoch, abort, the batch's meta) + small region, the
// tile map (global tile -> column, u32), the job table, then the look-back
// area: tile descriptors for every tile of the batch and (tiles/64 + 1)
// group descriptors per column.
using PlLookback = Lookback<LB_WIN | LB_OVF | LB_SUM>;

static uint64_t pl_lb_bytes(uint64_t tiles, uint64_t groups)
{
    return tiles * PlLookback::TS * 8 + groups * PlLookback::GS * 8;
}

static uint64_t pl_jobs_bytes(uint32_t count) { return ((uint64_t)count * sizeof(PipeJob) + 255) / 256 * 256; }
static uint64_t pl_map_bytes(uint64_t grid) { return (grid * sizeof(PlSlot) + 255) / 256 * 256; }
// the XCD-local layout is used only when its grid is at most this
static uint64_t pl_grid_cap(uint64_t tiles) { return tiles + tiles / 10 + PL_XCDS; }
constexpr uint64_t PL_OFF_MAP = 256 + WS_SMALL_BYTES;

extern "C" uint64_t am_pipeline_workspace_bytes(const uint64_t* n, uint32_t count)
{
    const uint32_t PL_TILE = pl_tile(8); // the smallest tile bounds the tile count
    uint64_t tiles = 1, groups = 1;
    for (uint32_t i = 0; i < count; i++) {
        const uint64_t t = n[i] ? (n[i] + PL_TILE - 1) / PL_TILE : 1;
        tiles += t;
        groups += t / 64 + 1;
    }
    return PL_OFF_MAP + pl_map_bytes(pl_grid_cap(tiles)) + pl_jobs_bytes(count) + pl_lb_bytes(tiles, groups);
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
            if (v == 0) v = 32;
        }
        if (v != 8 && v != 16 && v != 32) return AM_E_INVALID;
        // staged values: the output width rounded up to 1, 2 or 4 bytes, so
        // a tile holds up to cap = PL_STAGE_BYTES / vsz values; variant 32
        // takes the largest tile that fits, 16 and 8 that many elements per
        // lane times 4 / vsz
        J.vsz = c.s.bits <= 8 ? 1 : c.s.bits <= 16 ? 2 : 4;
        const uint32_t f = 4 / J.vsz, cap = PL_STAGE_BYTES / J.vsz;
        const uint32_t jmax = cap / pl_tile(PL_SUB) * PL_SUB; // element path: whole sub-passes
        const uint32_t vf = v == 32 ? jmax : v * f;
        J.J = vf;
        J.pr = make_pred(c.pred, J.a.mask);
        J.stats = c.stats;
        J.vbits = c.a.bits + 1;
        J.chk_ovf = c.s.bits < J.vbits ? 1 : 0;
        // word path: whole slab words per tile (256 per sub-pass), about as
        // many elements as the element path's tile: at least 2048 (the tile
        // count am_pipeline_workspace_bytes allows for), at most cap
        const uint32_t spe = 256 * J.a.eps; // elements per sub-pass
        J.wsub = (J.a.kind == K_SLAB && c.a.bits >= 2)
                     ? std::min(std::max((v == 32 ? cap : pl_tile(vf)) / spe, (2048 + spe - 1) / spe), cap / spe)
                     : 0;
        J.tp = J.wsub ? make_top_pred(J.pr, c.a.bits, J.a.eps) : TopPred{0, 0, 0};
        if (J.wsub) {
            const uint64_t TW = (uint64_t)J.wsub * 256;
            J.T = J.a.words ? (uint32_t)((J.a.words + TW - 1) / TW) : 1;
        }
        else {
            const uint32_t TL = pl_tile(vf);
            J.T = J.a.n ? (uint32_t)((J.a.n + TL - 1) / TL) : 1;
        }
        J.tile0 = (uint32_t)tiles;
        J.gbase = (uint32_t)groups;
        tiles += J.T;
        groups += J.T / 64 + 1;
        if (tiles >= (1ull << 31)) return AM_E_UNSUPPORTED_CONFIG;
    }
    // Block map. XCD-local when it balances: columns go whole to the XCD
    // with the fewest tiles so far (largest first), block i*8+x runs the
    // i-th tile of XCD x's list, so a column's look-back traffic stays in
    // one XCD's L2 and a tile still waits only on lower blocks of its own
    // XCD. Otherwise (a few columns hold most tiles) block g runs global
    // tile g, columns in order.
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return jobs[x].T > jobs[y].T; });
    std::vector<std::vector<uint32_t>> lists(PL_XCDS);
    std::vector<uint64_t> load(PL_XCDS, 0);
    for (uint32_t i : order) {
        const uint32_t x = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
        lists[x].push_back(i);
        load[x] += jobs[i].T;
    }
    const uint64_t maxl = *std::max_element(load.begin(), load.end());
    const bool xcd = PL_XCDS * maxl <= pl_grid_cap(tiles);
    const uint64_t grid = xcd ? PL_XCDS * maxl : tiles;
    if (grid >= (1ull << 31)) return AM_E_UNSUPPORTED_CONFIG;
    std::vector<PlSlot> map(grid, PlSlot{PL_IDLE, 0});
    if (xcd) {
        for (uint32_t x = 0; x < PL_XCDS; x++) {
            std::sort(lists[x].begin(), lists[x].end());
            uint64_t i = 0;
            for (uint32_t c : lists[x])
                for (uint32_t t = 0; t < jobs[c].T; t++, i++) map[i * PL_XCDS + x] = PlSlot{c, t};
        }
    }
    else {
        for (uint32_t c = 0; c < count; c++)
            for (uint32_t t = 0; t < jobs[c].T; t++) map[jobs[c].tile0 + t] = PlSlot{c, t};
    }
    const uint64_t off_j = PL_OFF_MAP + pl_map_bytes(grid), off_t = off_j + pl_jobs_bytes(count);
    const uint64_t need = off_t + pl_lb_bytes(tiles, groups);
    if (ws_bytes < need) return AM_E_WORKSPACE;
    // the previous run on this stream may still read the job table
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return AM_E_HIP;
    uint8_t* w8 = (uint8_t*)ws;
    if (hipMemcpy(w8 + PL_OFF_MAP, map.data(), grid * sizeof(PlSlot), hipMemcpyHostToDevice) != hipSuccess)
        return AM_E_HIP;
    if (hipMemcpy(w8 + off_j, jobs.data(), count * sizeof(PipeJob), hipMemcpyHostToDevice) != hipSuccess)
        return AM_E_HIP;
    PipeMeta m{PL_MAGIC, count, (uint32_t)tiles, 0,   (uint32_t)grid, xcd ? 1u : 0u,
               need,     off_j, off_t,               off_t + tiles * PlLookback::TS * 8};
    if (hipMemcpy(pl_meta(ws), &m, sizeof(m), hipMemcpyHostToDevice) != hipSuccess) return AM_E_HIP;
    if (total_tiles) *total_tiles = (uint32_t)grid;
    return AM_OK;
}

extern "C" int am_pipeline_run(void* ws, uint64_t ws_bytes, uint32_t total_tiles, void* stream)
{
    if (!ws || ws_bytes < PL_OFF_MAP) return AM_E_WORKSPACE;
    if (total_tiles == 0) return AM_E_INVALID;
    // every block reads its tile-map entry before it checks the recorded
    // batch, so the map alone must fit whatever the workspace holds
    if (ws_bytes < PL_OFF_MAP + pl_map_bytes(total_tiles)) return AM_E_WORKSPACE;
    // everything else -- job table, look-back areas, the tile count the grid
    // must have -- was recorded by am_pipeline_prepare in the workspace and
    // is checked by the kernel itself (the launch stays graph-capturable)
    PipeArgs a{};
    a.ws = (uint8_t*)ws;
    a.hdr = (WsHeader*)ws;
    a.meta = pl_meta(ws);
    a.map = (const PlSlot*)((uint8_t*)ws + PL_OFF_MAP);
    a.jobs = (const PipeJob*)((uint8_t*)ws + PL_OFF_MAP + pl_map_bytes(total_tiles));
    a.total_tiles = total_tiles;
    a.ws_bytes = ws_bytes;
    hipLaunchKernelGGL(k_pipeline, dim3(total_tiles), dim3(256), 0, (hipStream_t)stream, a);
    return check_launch();
}
