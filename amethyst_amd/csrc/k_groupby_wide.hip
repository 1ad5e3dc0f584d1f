// k_groupby_wide.hip -- group-by SUM/COUNT over keys of 13..24 bits.
//
// The reference caps its shared-memory group-by at 12 key bits
// (src/benchs/bench_groupby_sm_array.cu:124-132); its path for wider keys is
// a sort of all (key, value) pairs followed by a reduce-by-key
// (cub::DeviceRadixSort::SortPairs + cub::DeviceReduce::ReduceByKey,
// src/benchs/bench_groupby_sort_cub.cu:50-63), disabled in its driver
// (src/main.cu:316). Sorting moves every pair through several passes; a
// group-by needs only each key's rows together. Here the rows are
// partitioned once by the key's high bits and each partition is aggregated
// in an LDS table over the key's low LB bits:
//   1. hist      per row block, rows per partition (LDS histogram)
//   2. scan_part per partition: exclusive offsets over the row blocks, the
//                partition's size and aggregation chunks; a partition that
//                several chunks share has its table region zeroed here
//   3. scan_top  one block: partition bases, chunk bases, maxima reset
//   4. scatter   rows again, each (low key, value) pair packed into 32 or 64
//                bits at its partition's next slot (LDS counters)
//   5. agg       per chunk of one partition: LDS sums and counts over the
//                2^LB low keys, then stored (single-chunk partition) or
//                added with 64-bit atomics into the dense key table
//   6. count / scan_present / write / pack: present keys per partition,
//                their global ranks, key-sorted u64 staging, and the three
//                slab outputs at minimal width.
// Results are exact u64 sums and counts, in ascending key order, the same
// outputs as the small-key path; the integer sums do not depend on the
// order in which pairs land in a partition.
#include <type_traits>

#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t GW_TH = 256;          // threads of every wide-path block
constexpr uint32_t GW_ROWS = 8;          // rows per thread per step (partition passes)
constexpr uint32_t GW_MAXP = 4096;       // partitions
constexpr uint32_t GW_MAXLB = 12;        // low key bits per LDS table
constexpr uint32_t GW_LC = 15;           // log2 of the pairs per aggregation chunk
constexpr uint32_t GW_C = 1u << GW_LC;  // (a count fits 16 bits, a sum vbits + 15)
constexpr uint32_t GW_MAXG = 1024;       // row blocks of the partition passes
constexpr uint32_t GW_AGG_U = 8;         // pairs per thread in flight in the aggregation

struct WideArgs {
    Col keys, vals;
    Pred pr;
    uint32_t use_pred, mode;
    uint32_t kbits, LB, P, G; // partitions P = 2^(kbits-LB); G row blocks
    uint64_t rpb;             // rows per row block (multiple of GW_TH*GW_ROWS)
    uint32_t pair64, vbits;   // pair = (low key << vbits) | value, 32 or 64 bits
    uint32_t* hist;           // [P][G]: rows of partition p in block b, then exclusive offsets
    uint64_t* pbase;          // [P+1]: partition sizes, then exclusive bases (pbase[P] = pairs)
    uint32_t* cbase;          // [P+1]: partition chunks, then exclusive bases (cbase[P] = chunks)
    uint32_t* present;        // [P+1]: present keys, then exclusive bases (present[P] = groups)
    uint64_t* gmax;           // [2]: largest sum, largest count
    void* pairs;              // [rows]
    uint64_t* tsum;           // [2^kbits]
    uint64_t* tcnt;           // [2^kbits]
    uint64_t *skey, *ssum, *scnt; // key-sorted staging
    uint64_t *out_keys, *out_sums, *out_counts, *result;
};

__device__ inline uint32_t gw_width(uint64_t v) { return v ? 64 - (uint32_t)__clzll(v) : 1; }

// LDS counter increments where many lanes may hit one counter (a Zipf key's
// partition): the active lanes that share the first active lane's counter
// take one atomic for all of them, the others one each. lds_add_agg adds 1
// per lane; lds_inc_agg returns each lane's slot (the counter before its own
// increment, lanes of the shared counter in lane order).
__device__ inline void lds_add_agg(uint32_t* c, uint32_t p)
{
    const uint32_t p0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)p);
    const uint64_t same = ballot(p == p0);
    if (p == p0) {
        if (mbcnt(same) == 0) atomicAdd(&c[p0], (uint32_t)__popcll(same));
    }
    else atomicAdd(&c[p], 1u);
}
__device__ inline uint32_t lds_inc_agg(uint32_t* c, uint32_t p)
{
    const uint32_t p0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)p);
    const uint64_t same = ballot(p == p0);
    uint32_t r;
    if (p == p0) {
        uint32_t b = 0;
        if (mbcnt(same) == 0) b = atomicAdd(&c[p0], (uint32_t)__popcll(same));
        r = (uint32_t)__builtin_amdgcn_readfirstlane((int)b) + mbcnt(same); // (the first active lane leads)
    }
    else r = atomicAdd(&c[p], 1u);
    return r;
}

__device__ inline uint64_t gw_shfl_up(uint64_t v, uint32_t d)
{
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, d), hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), d);
    return ((uint64_t)hi << 32) | lo;
}

// Exclusive block scan of one u32 per thread (GW_TH threads); returns the
// thread's exclusive prefix, *total the block sum.
__device__ inline uint32_t block_excl_u32(uint32_t v, uint32_t* s_w, uint32_t* total)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t x = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d);
        if (lane >= d) x += y;
    }
    __syncthreads();
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (uint32_t w = 0; w < GW_TH / 64; w++) {
        pre += w < wave ? s_w[w] : 0;
        tot += s_w[w];
    }
    *total = tot;
    return pre + x - v;
}

__device__ inline uint64_t block_excl_u64(uint64_t v, uint64_t* s_w, uint64_t* total)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint64_t x = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint64_t y = gw_shfl_up(x, d);
        if (lane >= d) x += y;
    }
    __syncthreads();
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint64_t pre = 0, tot = 0;
    for (uint32_t w = 0; w < GW_TH / 64; w++) {
        pre += w < wave ? s_w[w] : 0;
        tot += s_w[w];
    }
    *total = tot;
    return pre + x - v;
}

// Rows [b*rpb, ...) of the block, GW_ROWS per thread per step; fn(key, value)
// for each row that passes the value predicate. Values are read only when
// the predicate or the caller needs them. The next step's rows are loaded
// before this step's are used (a block's steps would otherwise each wait a
// full memory round trip).
constexpr uint32_t TILE_ROWS = GW_TH * GW_ROWS; // rows per step of the partition passes

// BATCH: rows past the end read the last row, which the callers skip; with
// no branch around them the compiler issues all loads before the first is
// used (the histogram: 0.71 -> 0.3 ms at 24-bit keys). Without it each load
// sits in its own branch and is waited for there -- kept for the sorted
// scatter, whose registers the unconditional form raises past four waves
// per SIMD.
template <int KK, int VK, bool BATCH>
__device__ inline void gw_load(const WideArgs& a, uint64_t t0, uint64_t r1, bool wantv, uint64_t* kv, uint64_t* vv)
{
    const TileReader<KK> rk(a.keys, t0);
    const TileReader<VK> rv(a.vals, t0);
    const uint64_t rem = r1 - t0;
    if (!BATCH) {
#pragma unroll
        for (uint32_t r = 0; r < GW_ROWS; r++) {
            const uint32_t i = r * GW_TH + threadIdx.x;
            kv[r] = i < rem ? rk.get(i) : 0;
            vv[r] = (wantv && i < rem) ? rv.get(i) : 0;
        }
        return;
    }
    const uint32_t last = (uint32_t)(rem < TILE_ROWS ? rem - 1 : TILE_ROWS - 1);
    if (wantv) {
#pragma unroll
        for (uint32_t r = 0; r < GW_ROWS; r++) {
            const uint32_t i = r * GW_TH + threadIdx.x;
            kv[r] = rk.get(i < last ? i : last);
            vv[r] = rv.get(i < last ? i : last);
        }
    }
    else {
#pragma unroll
        for (uint32_t r = 0; r < GW_ROWS; r++) {
            const uint32_t i = r * GW_TH + threadIdx.x;
            kv[r] = rk.get(i < last ? i : last);
            vv[r] = 0;
        }
    }
}

template <int KK, int VK, bool NEEDV, typename Fn>
__device__ inline void gw_rows(const WideArgs& a, Fn fn)
{
    constexpr uint32_t TILE = GW_TH * GW_ROWS;
    const uint64_t n = a.keys.n;
    const uint64_t r0 = (uint64_t)blockIdx.x * a.rpb;
    const uint64_t r1 = r0 + a.rpb < n ? r0 + a.rpb : n;
    const bool wantv = NEEDV || a.use_pred;
    uint64_t kv[GW_ROWS], vv[GW_ROWS], kn[GW_ROWS], vn[GW_ROWS];
    if (r0 < r1) gw_load<KK, VK, true>(a, r0, r1, wantv, kv, vv);
    for (uint64_t t0 = r0; t0 < r1; t0 += TILE) {
        if (t0 + TILE < r1) gw_load<KK, VK, true>(a, t0 + TILE, r1, wantv, kn, vn);
        const uint64_t rem = r1 - t0;
#pragma unroll
        for (uint32_t r = 0; r < GW_ROWS; r++) {
            const uint32_t i = r * GW_TH + threadIdx.x;
            if (i < rem && (!a.use_pred || a.pr(vv[r]))) fn(kv[r], vv[r]);
        }
#pragma unroll
        for (uint32_t r = 0; r < GW_ROWS; r++) {
            kv[r] = kn[r];
            vv[r] = vn[r];
        }
    }
}

template <int KK, int VK>
__global__ __launch_bounds__(GW_TH) void k_gbw_hist(WideArgs a)
{
    __shared__ uint32_t h[GW_MAXP];
    for (uint32_t p = threadIdx.x; p < a.P; p += GW_TH) h[p] = 0;
    __syncthreads();
    const uint32_t LB = a.LB;
    gw_rows<KK, VK, false>(a, [&](uint64_t k, uint64_t) { lds_add_agg(h, (uint32_t)(k >> LB)); });
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < a.P; p += GW_TH) a.hist[(uint64_t)p * a.G + blockIdx.x] = h[p];
}

// one block per partition
__global__ __launch_bounds__(GW_TH) void k_gbw_scan_part(WideArgs a)
{
    __shared__ uint32_t s_w[GW_TH / 64];
    const uint32_t p = blockIdx.x;
    uint32_t* row = a.hist + (uint64_t)p * a.G;
    constexpr uint32_t PER = GW_MAXG / GW_TH; // entries per thread
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
        const uint32_t b = threadIdx.x * PER + k;
        v[k] = b < a.G ? row[b] : 0;
        sum += v[k];
    }
    uint32_t total;
    uint32_t pre = block_excl_u32(sum, s_w, &total);
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
        const uint32_t b = threadIdx.x * PER + k;
        if (b < a.G) row[b] = pre;
        pre += v[k];
    }
    const uint32_t chunks = (total + GW_C - 1) / GW_C;
    if (threadIdx.x == 0) {
        a.pbase[p] = total;
        a.cbase[p] = chunks;
    }
    if (chunks > 1) { // several chunks add into this partition's table region
        const uint64_t k0 = (uint64_t)p << a.LB;
        for (uint32_t k = threadIdx.x; k < (1u << a.LB); k += GW_TH) {
            a.tsum[k0 + k] = 0;
            a.tcnt[k0 + k] = 0;
        }
    }
}

// one block: exclusive bases of partitions and chunks
__global__ __launch_bounds__(GW_TH) void k_gbw_scan_top(WideArgs a)
{
    __shared__ uint64_t s_w64[GW_TH / 64];
    __shared__ uint32_t s_w32[GW_TH / 64];
    constexpr uint32_t PER = GW_MAXP / GW_TH;
    uint64_t pv[PER], ps = 0;
    uint32_t cv[PER], cs = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
        const uint32_t p = threadIdx.x * PER + k;
        pv[k] = p < a.P ? a.pbase[p] : 0;
        cv[k] = p < a.P ? a.cbase[p] : 0;
        ps += pv[k];
        cs += cv[k];
    }
    uint64_t ptot;
    uint32_t ctot;
    uint64_t pp = block_excl_u64(ps, s_w64, &ptot);
    uint32_t cp = block_excl_u32(cs, s_w32, &ctot);
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
        const uint32_t p = threadIdx.x * PER + k;
        if (p < a.P) {
            a.pbase[p] = pp;
            a.cbase[p] = cp;
        }
        pp += pv[k];
        cp += cv[k];
    }
    if (threadIdx.x == 0) {
        a.pbase[a.P] = ptot;
        a.cbase[a.P] = ctot;
        a.gmax[0] = 0;
        a.gmax[1] = 0;
    }
}

// Up to GW_SORTP partitions: each tile of rows is counting-sorted by
// partition in LDS first (local ranks from LDS counters, a block scan of the
// tile's partition counts), so every partition's pairs of the tile leave as
// one contiguous run -- whole lines instead of one line per pair.
constexpr uint32_t GW_SORTP = 256;
static_assert(GW_SORTP == GW_TH, "one tile partition count per thread");

template <int KK, int VK, bool SORT>
__device__ inline void gbw_scatter(const WideArgs& a)
{
    constexpr uint32_t TILE = GW_TH * GW_ROWS;
    constexpr uint32_t NP = SORT ? GW_SORTP : GW_MAXP;
    __shared__ uint32_t pos[NP];
    __shared__ uint64_t pb[NP];
    __shared__ uint32_t lh[SORT ? GW_SORTP : 1], loff[SORT ? GW_SORTP : 1];
    __shared__ uint64_t s_pair[SORT ? TILE : 1];
    __shared__ uint16_t s_part[SORT ? TILE : 1];
    __shared__ uint32_t s_w[GW_TH / 64];
    for (uint32_t p = threadIdx.x; p < a.P; p += GW_TH) {
        pos[p] = a.hist[(uint64_t)p * a.G + blockIdx.x];
        pb[p] = a.pbase[p];
    }
    __syncthreads();
    const uint32_t LB = a.LB, vb = a.vbits;
    const uint64_t lm = lowmask(LB);
    const uint64_t n = a.keys.n;
    const uint64_t r0 = (uint64_t)blockIdx.x * a.rpb;
    const uint64_t r1 = r0 + a.rpb < n ? r0 + a.rpb : n;
    uint64_t kv[GW_ROWS], vv[GW_ROWS], kn[GW_ROWS], vn[GW_ROWS];
    if (r0 < r1) gw_load<KK, VK, false>(a, r0, r1, true, kv, vv);
    for (uint64_t t0 = r0; t0 < r1; t0 += TILE) {
        if (t0 > r0) { // the rows loaded during the previous step
#pragma unroll
            for (uint32_t r = 0; r < GW_ROWS; r++) {
                kv[r] = kn[r];
                vv[r] = vn[r];
            }
        }
        if (t0 + TILE < r1) gw_load<KK, VK, false>(a, t0 + TILE, r1, true, kn, vn);
        const uint64_t rem = r1 - t0;
        if (!SORT) {
#pragma unroll
            for (uint32_t r = 0; r < GW_ROWS; r++) {
                const uint32_t i = r * GW_TH + threadIdx.x;
                if (i < rem && (!a.use_pred || a.pr(vv[r]))) {
                    const uint32_t p = (uint32_t)(kv[r] >> LB);
                    const uint32_t slot = atomicAdd(&pos[p], 1u);
                    const uint64_t x = ((kv[r] & lm) << vb) | vv[r];
                    if (a.pair64) ((uint64_t*)a.pairs)[pb[p] + slot] = x;
                    else ((uint32_t*)a.pairs)[pb[p] + slot] = (uint32_t)x;
                }
            }
            continue;
        }
        for (uint32_t p = threadIdx.x; p < a.P; p += GW_TH) lh[p] = 0;
        __syncthreads();
        uint32_t rk_[GW_ROWS];
#pragma unroll
        for (uint32_t r = 0; r < GW_ROWS; r++) {
            const uint32_t i = r * GW_TH + threadIdx.x;
            const bool ok = i < rem && (!a.use_pred || a.pr(vv[r]));
            rk_[r] = 0xFFFFFFFFu;
            if (ok) rk_[r] = lds_inc_agg(lh, (uint32_t)(kv[r] >> LB));
        }
        __syncthreads();
        // exclusive offsets of the tile's partitions (one per thread, GW_SORTP == GW_TH)
        const uint32_t c = threadIdx.x < a.P ? lh[threadIdx.x] : 0;
        uint32_t tot;
        const uint32_t off = block_excl_u32(c, s_w, &tot);
        if (threadIdx.x < a.P) loff[threadIdx.x] = off;
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < GW_ROWS; r++) {
            if (rk_[r] == 0xFFFFFFFFu) continue;
            const uint32_t p = (uint32_t)(kv[r] >> LB);
            const uint32_t d = loff[p] + rk_[r];
            s_pair[d] = ((kv[r] & lm) << vb) | vv[r];
            s_part[d] = (uint16_t)p;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < tot; i += GW_TH) {
            const uint32_t p = s_part[i];
            const uint64_t dst = pb[p] + pos[p] + (i - loff[p]);
            if (a.pair64) ((uint64_t*)a.pairs)[dst] = s_pair[i];
            else ((uint32_t*)a.pairs)[dst] = (uint32_t)s_pair[i];
        }
        __syncthreads();
        if (threadIdx.x < a.P) pos[threadIdx.x] += lh[threadIdx.x];
        __syncthreads();
    }
}

template <int KK, int VK>
__global__ __launch_bounds__(GW_TH) void k_gbw_scatter_sorted(WideArgs a) { gbw_scatter<KK, VK, true>(a); }
template <int KK, int VK>
__global__ __launch_bounds__(GW_TH) void k_gbw_scatter_direct(WideArgs a) { gbw_scatter<KK, VK, false>(a); }

// More than GW_SORTP partitions (keys of 21..24 bits, 512..4096
// partitions): a block of GW_BTH threads counting-sorts tiles of GW_BIG
// rows in LDS, so a partition's pairs of the tile leave as one run of about
// GW_BIG / P pairs -- one store per row was one memory transaction per row. The block covers the same rows as row
// block b of the histogram (blockIdx.x = b), so its starting places are
// hist[p][b].
constexpr uint32_t GW_BTH = 1024;
constexpr uint32_t GW_BIG = 8192; // 8 rows per thread (16 spill at 1024 threads)

__device__ inline uint32_t block_excl_u32_1024(uint32_t v, uint32_t* s_w, uint32_t* total)
{
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t x = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, d);
        if (lane >= d) x += y;
    }
    __syncthreads();
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (uint32_t w = 0; w < GW_BTH / 64; w++) {
        pre += w < wave ? s_w[w] : 0;
        tot += s_w[w];
    }
    *total = tot;
    return pre + x - v;
}

template <int KK, int VK, bool P64>
__global__ __launch_bounds__(GW_BTH) void k_gbw_scatter_big(WideArgs a)
{
    typedef typename std::conditional<P64, uint64_t, uint32_t>::type PairT;
    constexpr uint32_t TB = GW_BIG, RB = TB / GW_BTH;
    constexpr uint32_t PPT = GW_MAXP / GW_BTH; // partition counters per thread in the scan
    __shared__ PairT s_pair[TB];
    __shared__ uint16_t s_part[TB];
    __shared__ uint32_t cnt[GW_MAXP], loff[GW_MAXP], gpos[GW_MAXP];
    __shared__ uint32_t s_w[GW_BTH / 64];
    const uint32_t tid = threadIdx.x, P = a.P, LB = a.LB, vb = a.vbits;
    for (uint32_t p = tid; p < P; p += GW_BTH) gpos[p] = (uint32_t)a.pbase[p] + a.hist[(uint64_t)p * a.G + blockIdx.x];
    const uint64_t lm = lowmask(LB);
    const uint64_t n = a.keys.n;
    const uint64_t r0 = (uint64_t)blockIdx.x * a.rpb;
    const uint64_t r1 = r0 + a.rpb < n ? r0 + a.rpb : n;
    for (uint64_t t0 = r0; t0 < r1; t0 += TB) {
        const TileReader<KK> rk(a.keys, t0);
        const TileReader<VK> rv(a.vals, t0);
        const uint64_t rem = r1 - t0;
        // rows in groups of 8 (loads of a group in flight together), kept as
        // (partition, pair): 2 registers a row instead of 4
        uint32_t part[RB];
        PairT pv[RB];
#pragma unroll
        for (uint32_t h = 0; h < RB; h += 8) {
            uint64_t kv[8], vv[8];
            const uint32_t last = (uint32_t)(rem < TB ? rem - 1 : TB - 1); // (as in gw_load)
#pragma unroll
            for (uint32_t r = 0; r < 8; r++) {
                const uint32_t i = (h + r) * GW_BTH + tid;
                kv[r] = rk.get(i < last ? i : last);
                vv[r] = rv.get(i < last ? i : last);
            }
#pragma unroll
            for (uint32_t r = 0; r < 8; r++) {
                const uint32_t i = (h + r) * GW_BTH + tid;
                const bool ok = i < rem && (!a.use_pred || a.pr(vv[r]));
                part[h + r] = ok ? (uint32_t)(kv[r] >> LB) : 0xFFFFFFFFu;
                pv[h + r] = (PairT)(((kv[r] & lm) << vb) | vv[r]);
            }
            asm volatile("" ::: "memory"); // (the next group's loads after this group's registers are freed)
        }
        for (uint32_t p = tid; p < P; p += GW_BTH) cnt[p] = 0;
        __syncthreads();
        uint32_t rank[RB];
#pragma unroll
        for (uint32_t r = 0; r < RB; r++) rank[r] = part[r] != 0xFFFFFFFFu ? atomicAdd(&cnt[part[r]], 1u) : 0;
        __syncthreads();
        // exclusive offsets of the tile's partitions, PPT consecutive per thread
        uint32_t c[PPT], cs = 0;
#pragma unroll
        for (uint32_t q = 0; q < PPT; q++) {
            const uint32_t p = tid * PPT + q;
            c[q] = p < P ? cnt[p] : 0;
            cs += c[q];
        }
        uint32_t tot;
        uint32_t off = block_excl_u32_1024(cs, s_w, &tot);
#pragma unroll
        for (uint32_t q = 0; q < PPT; q++) {
            const uint32_t p = tid * PPT + q;
            if (p < P) loff[p] = off;
            off += c[q];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < RB; r++) {
            if (part[r] == 0xFFFFFFFFu) continue;
            const uint32_t d = loff[part[r]] + rank[r];
            s_pair[d] = pv[r];
            s_part[d] = (uint16_t)part[r];
        }
        __syncthreads();
        for (uint32_t i = tid; i < tot; i += GW_BTH) {
            const uint32_t p = s_part[i];
            ((PairT*)a.pairs)[(uint64_t)gpos[p] + (i - loff[p])] = s_pair[i];
        }
        __syncthreads();
        for (uint32_t p = tid; p < P; p += GW_BTH) gpos[p] += cnt[p];
        __syncthreads();
    }
}

// one block per chunk of GW_C pairs of one partition: one packed LDS word
// per low key, (count << SB) | sum with SB = vbits + 15 -- a chunk's sum
// fits SB bits and its count (<= 2^15) the 64 - SB >= 17 above, so a row
// costs one 64-bit LDS atomic and nothing overflows
__global__ __launch_bounds__(GW_TH) void k_gbw_agg(WideArgs a)
{
    __shared__ uint64_t s_t[1u << GW_MAXLB];
    const uint32_t i = blockIdx.x;
    if (i >= a.cbase[a.P]) return;
    uint32_t lo = 0, hi = a.P - 1; // partition: the last p with cbase[p] <= i
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (a.cbase[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const uint32_t p = lo;
    const uint32_t c = i - a.cbase[p], nch = a.cbase[p + 1] - a.cbase[p];
    const uint64_t q0 = a.pbase[p] + (uint64_t)c * GW_C;
    const uint64_t qe = a.pbase[p + 1];
    const uint32_t cnt = (uint32_t)(qe - q0 < GW_C ? qe - q0 : GW_C);
    const uint32_t LB = a.LB, K = 1u << LB, vb = a.vbits;
    const uint32_t SB = vb + GW_LC;
    const uint64_t ONE = 1ull << SB, SM = lowmask(SB);
    // small tables: one copy per wave (fewer lanes on one LDS word)
    const uint32_t copies = K <= (1u << GW_MAXLB) / 4 ? 4 : 1;
    const uint32_t cp = (threadIdx.x >> 6) % copies;
    for (uint32_t k = threadIdx.x; k < K * copies; k += GW_TH) s_t[k] = 0;
    __syncthreads();
    const uint64_t vm = lowmask(vb);
    // (GW_AGG_U pairs per thread loaded before any is added)
    constexpr uint32_t U = GW_AGG_U;
    if (a.pair64) {
        const uint64_t* pr = (const uint64_t*)a.pairs + q0;
        for (uint32_t j0 = 0; j0 < cnt; j0 += GW_TH * U) {
            uint64_t x[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t j = j0 + u * GW_TH + threadIdx.x;
                x[u] = j < cnt ? pr[j] : 0;
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                if (j0 + u * GW_TH + threadIdx.x < cnt)
                    atomicAdd((unsigned long long*)&s_t[cp * K + (uint32_t)(x[u] >> vb)],
                              (unsigned long long)(ONE | (x[u] & vm)));
        }
    }
    else {
        const uint32_t* pr = (const uint32_t*)a.pairs + q0;
        for (uint32_t j0 = 0; j0 < cnt; j0 += GW_TH * U) {
            uint32_t x[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t j = j0 + u * GW_TH + threadIdx.x;
                x[u] = j < cnt ? pr[j] : 0;
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                if (j0 + u * GW_TH + threadIdx.x < cnt) {
                    const uint32_t k = vb < 32 ? (x[u] >> vb) : 0;
                    atomicAdd((unsigned long long*)&s_t[cp * K + k], (unsigned long long)(ONE | (x[u] & (uint32_t)vm)));
                }
        }
    }
    __syncthreads();
    const uint64_t k0 = (uint64_t)p << LB;
    for (uint32_t k = threadIdx.x; k < K; k += GW_TH) {
        uint64_t s = 0, n = 0;
        for (uint32_t q = 0; q < copies; q++) {
            const uint64_t t = s_t[q * K + k];
            s += t & SM;
            n += t >> SB;
        }
        if (nch == 1) {
            a.tsum[k0 + k] = s;
            a.tcnt[k0 + k] = n;
        }
        else if (n) {
            add64(&a.tsum[k0 + k], s);
            add64(&a.tcnt[k0 + k], n);
        }
    }
}

// one block per partition: present keys and the maxima
__global__ __launch_bounds__(GW_TH) void k_gbw_count(WideArgs a)
{
    __shared__ uint32_t s_w[GW_TH / 64];
    __shared__ uint64_t s_m[2][GW_TH / 64];
    const uint32_t p = blockIdx.x;
    // a partition without rows never had its table region written
    const uint32_t K = a.pbase[p + 1] != a.pbase[p] ? 1u << a.LB : 0u;
    const uint64_t k0 = (uint64_t)p << a.LB;
    uint32_t pres = 0;
    uint64_t smax = 0, cmax = 0;
    for (uint32_t k = threadIdx.x; k < K; k += GW_TH) {
        const uint64_t n = a.tcnt[k0 + k];
        uint64_t s = a.tsum[k0 + k];
        if (a.mode == AM_GB_REF) s &= a.vals.mask;
        if (n) {
            pres++;
            smax = s > smax ? s : smax;
            cmax = n > cmax ? n : cmax;
        }
    }
    uint32_t total;
    (void)block_excl_u32(pres, s_w, &total);
    smax = wave_max_u64(smax);
    cmax = wave_max_u64(cmax);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_m[0][wave] = smax;
        s_m[1][wave] = cmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 0; w < GW_TH / 64; w++) {
            smax = s_m[0][w] > smax ? s_m[0][w] : smax;
            cmax = s_m[1][w] > cmax ? s_m[1][w] : cmax;
        }
        a.present[p] = total;
        if (smax) atomicMax((unsigned long long*)&a.gmax[0], (unsigned long long)smax);
        if (cmax) atomicMax((unsigned long long*)&a.gmax[1], (unsigned long long)cmax);
    }
}

// one block: group bases and the result record
__global__ __launch_bounds__(GW_TH) void k_gbw_scan_present(WideArgs a)
{
    __shared__ uint32_t s_w[GW_TH / 64];
    constexpr uint32_t PER = GW_MAXP / GW_TH;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
        const uint32_t p = threadIdx.x * PER + k;
        v[k] = p < a.P ? a.present[p] : 0;
        sum += v[k];
    }
    uint32_t total;
    uint32_t pre = block_excl_u32(sum, s_w, &total);
#pragma unroll
    for (uint32_t k = 0; k < PER; k++) {
        const uint32_t p = threadIdx.x * PER + k;
        if (p < a.P) a.present[p] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0) {
        a.present[a.P] = total;
        a.result[0] = total;
        a.result[1] = a.mode == AM_GB_REF ? a.vals.bits : gw_width(a.gmax[0]);
        a.result[2] = gw_width(a.gmax[1]);
        a.result[3] = a.pbase[a.P];
    }
}

// one block per partition: its present keys, in key order, into the staging arrays
__global__ __launch_bounds__(GW_TH) void k_gbw_write(WideArgs a)
{
    __shared__ uint32_t s_w[GW_TH / 64];
    const uint32_t p = blockIdx.x;
    if (a.pbase[p + 1] == a.pbase[p]) return; // no rows, no keys
    const uint32_t K = 1u << a.LB;
    const uint64_t k0 = (uint64_t)p << a.LB;
    uint64_t o = a.present[p];
    for (uint32_t kb = 0; kb < K; kb += GW_TH) {
        const uint32_t k = kb + threadIdx.x;
        const uint64_t n = k < K ? a.tcnt[k0 + k] : 0;
        uint32_t total;
        const uint32_t r = block_excl_u32(n ? 1u : 0u, s_w, &total);
        if (n) {
            uint64_t s = a.tsum[k0 + k];
            if (a.mode == AM_GB_REF) s &= a.vals.mask;
            a.skey[o + r] = k0 + k;
            a.ssum[o + r] = s;
            a.scnt[o + r] = n;
        }
        o += total;
    }
}

// one thread per output word of each output (slab at the result widths)
__global__ __launch_bounds__(GW_TH) void k_gbw_pack(WideArgs a)
{
    const uint64_t wd = (uint64_t)blockIdx.x * GW_TH + threadIdx.x;
    const uint64_t ng = a.result[0];
    const uint32_t wb[3] = {a.kbits, (uint32_t)a.result[1], (uint32_t)a.result[2]};
    const uint64_t* src[3] = {a.skey, a.ssum, a.scnt};
    uint64_t* dst[3] = {a.out_keys, a.out_sums, a.out_counts};
#pragma unroll
    for (uint32_t o = 0; o < 3; o++) {
        if (!dst[o]) continue;
        const uint32_t b = wb[o], e = 64 / b;
        const uint64_t words = (ng + e - 1) / e;
        if (wd >= words) continue;
        const uint64_t m = lowmask(b);
        uint64_t w = 0;
        for (uint32_t s = 0; s < e; s++) {
            const uint64_t i = wd * e + s;
            if (i < ng) w |= (src[o][i] & m) << (s * b);
        }
        dst[o][wd] = w;
    }
}

// Low key bits per LDS table: 256 partitions while the table can take the
// rest (keys <= 20 bits; the scatter then sorts each tile by partition),
// wider keys keep 4096-key tables.
static uint32_t gw_lb(uint32_t kbits) { return kbits >= 20 ? 12 : kbits - 8; }

struct WideLayout {
    uint64_t hist, pbase, cbase, present, gmax, pairs, tsum, tcnt, skey, ssum, scnt, end;
};

static WideLayout gw_layout(uint64_t n, uint32_t kbits, uint32_t vbits)
{
    const uint32_t LB = gw_lb(kbits), P = 1u << (kbits - LB);
    const uint64_t K = 1ull << kbits;
    const bool p64 = LB + vbits > 32;
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    WideLayout L{};
    uint64_t o = 0;
    L.hist = o, o = al(o + (uint64_t)P * GW_MAXG * 4);
    L.pbase = o, o = al(o + (P + 1) * 8ull);
    L.cbase = o, o = al(o + (P + 1) * 4ull);
    L.present = o, o = al(o + (P + 1) * 4ull);
    L.gmax = o, o = al(o + 16);
    L.pairs = o, o = al(o + n * (p64 ? 8 : 4));
    L.tsum = o, o = al(o + K * 8);
    L.tcnt = o, o = al(o + K * 8);
    L.skey = o, o = al(o + K * 8);
    L.ssum = o, o = al(o + K * 8);
    L.scnt = o, o = al(o + K * 8);
    L.end = o;
    return L;
}

} // namespace am

using namespace am;

uint64_t am_groupby_wide_bytes(uint64_t n, uint32_t kbits, uint32_t vbits)
{
    return WS_FIXED + gw_layout(n, kbits, vbits).end;
}

// Keys of 13..24 bits (validated by am_groupby).
int am_groupby_wide(const Col& ck, const Col& cv, const am_predicate* value_pred, uint32_t mode, uint64_t* out_keys,
                    uint64_t* out_sums, uint64_t* out_counts, uint64_t* result_dev, void* ws, uint64_t ws_bytes,
                    hipStream_t s)
{
    if (ck.n >= (1ull << 32)) return AM_E_UNSUPPORTED_CONFIG; // partition sizes are 32-bit
    if (ws_bytes < am_groupby_wide_bytes(ck.n, ck.bits, cv.bits)) return AM_E_WORKSPACE;
    const WideLayout L = gw_layout(ck.n, ck.bits, cv.bits);
    uint8_t* base = (uint8_t*)ws + WS_FIXED;
    WideArgs a{};
    a.keys = ck;
    a.vals = cv;
    a.use_pred = value_pred ? 1 : 0;
    if (value_pred) a.pr = make_pred(*value_pred, cv.mask);
    a.mode = mode;
    a.kbits = ck.bits;
    a.LB = gw_lb(ck.bits);
    a.P = 1u << (ck.bits - a.LB);
    a.vbits = cv.bits;
    a.pair64 = a.LB + cv.bits > 32 ? 1 : 0;
    const uint64_t step = (uint64_t)GW_TH * GW_ROWS;
    const uint64_t steps = (ck.n + step - 1) / step;
    uint64_t G = steps < 1 ? 1 : (steps < GW_MAXG ? steps : GW_MAXG);
    const uint64_t spb = steps ? (steps + G - 1) / G : 1;
    G = steps ? (steps + spb - 1) / spb : 1;
    a.G = (uint32_t)G;
    a.rpb = spb * step;
    a.hist = (uint32_t*)(base + L.hist);
    a.pbase = (uint64_t*)(base + L.pbase);
    a.cbase = (uint32_t*)(base + L.cbase);
    a.present = (uint32_t*)(base + L.present);
    a.gmax = (uint64_t*)(base + L.gmax);
    a.pairs = base + L.pairs;
    a.tsum = (uint64_t*)(base + L.tsum);
    a.tcnt = (uint64_t*)(base + L.tcnt);
    a.skey = (uint64_t*)(base + L.skey);
    a.ssum = (uint64_t*)(base + L.ssum);
    a.scnt = (uint64_t*)(base + L.scnt);
    a.out_keys = out_keys;
    a.out_sums = out_sums;
    a.out_counts = out_counts;
    a.result = result_dev;
    const uint64_t chunks_max = (ck.n + GW_C - 1) / GW_C + a.P;
    if (chunks_max >= (1ull << 31)) return AM_E_UNSUPPORTED_CONFIG;
#define GW_KV(KERNEL)                                                                                        \
    switch (ck.kind) {                                                                                       \
        case K_STRIDED: GW_V(KERNEL, K_STRIDED); break;                                                      \
        case K_SLAB: GW_V(KERNEL, K_SLAB); break;                                                            \
        default: GW_V(KERNEL, K_BITSTREAM); break;                                                           \
    }
#define GW_V(KERNEL, KK)                                                                                     \
    switch (cv.kind) {                                                                                       \
        case K_STRIDED: hipLaunchKernelGGL((KERNEL<KK, K_STRIDED>), dim3(a.G), dim3(GW_TH), 0, s, a); break; \
        case K_SLAB: hipLaunchKernelGGL((KERNEL<KK, K_SLAB>), dim3(a.G), dim3(GW_TH), 0, s, a); break;       \
        default: hipLaunchKernelGGL((KERNEL<KK, K_BITSTREAM>), dim3(a.G), dim3(GW_TH), 0, s, a); break;      \
    }
#define GW_KVB(P64)                                                                                          \
    switch (ck.kind) {                                                                                       \
        case K_STRIDED: GW_VB(K_STRIDED, P64); break;                                                        \
        case K_SLAB: GW_VB(K_SLAB, P64); break;                                                              \
        default: GW_VB(K_BITSTREAM, P64); break;                                                             \
    }
#define GW_VB(KK, P64)                                                                                       \
    switch (cv.kind) {                                                                                       \
        case K_STRIDED: hipLaunchKernelGGL((k_gbw_scatter_big<KK, K_STRIDED, P64>), dim3(a.G), dim3(GW_BTH), 0, s, a); break; \
        case K_SLAB: hipLaunchKernelGGL((k_gbw_scatter_big<KK, K_SLAB, P64>), dim3(a.G), dim3(GW_BTH), 0, s, a); break;       \
        default: hipLaunchKernelGGL((k_gbw_scatter_big<KK, K_BITSTREAM, P64>), dim3(a.G), dim3(GW_BTH), 0, s, a); break;      \
    }
    GW_KV(k_gbw_hist)
    hipLaunchKernelGGL(k_gbw_scan_part, dim3(a.P), dim3(GW_TH), 0, s, a);
    hipLaunchKernelGGL(k_gbw_scan_top, dim3(1), dim3(GW_TH), 0, s, a);
    if (a.P <= GW_SORTP) {
        GW_KV(k_gbw_scatter_sorted)
    }
    else if (a.pair64) {
        GW_KVB(true)
    }
    else {
        GW_KVB(false)
    }
#undef GW_VB
#undef GW_KVB
#undef GW_V
#undef GW_KV
    hipLaunchKernelGGL(k_gbw_agg, dim3((uint32_t)chunks_max), dim3(GW_TH), 0, s, a);
    hipLaunchKernelGGL(k_gbw_count, dim3(a.P), dim3(GW_TH), 0, s, a);
    hipLaunchKernelGGL(k_gbw_scan_present, dim3(1), dim3(GW_TH), 0, s, a);
    hipLaunchKernelGGL(k_gbw_write, dim3(a.P), dim3(GW_TH), 0, s, a);
    const uint64_t maxw = 1ull << ck.bits; // words of the widest output (64-bit fields, every key present)
    hipLaunchKernelGGL(k_gbw_pack, dim3((uint32_t)((maxw + GW_TH - 1) / GW_TH)), dim3(GW_TH), 0, s, a);
    return check_launch();
}
