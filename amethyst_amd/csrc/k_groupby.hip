// k_groupby.hip -- group-by SUM/COUNT over small key domains (keys <= 12 bits).
//
// Reference: kernel_groupby_sm_array_gather / _writeout
// (src/kernels/groupby_sm_array.cuh:4-81): a per-block shared-memory table of
// 2^bits u32 sums plus u32 occupancy flags, two shared atomics per row, merged
// into global memory with atomics, then written out through a global atomic
// counter -- so the output order is nondeterministic, sums are truncated to
// `bits` and there is no COUNT.
//
// Here, three stream-ordered kernels and no global atomics:
//   1. partial: each block keeps ONE packed LDS word per key,
//      (count << SB) | sum, so a row costs a single 64-bit LDS atomic add; SB
//      is chosen so neither field can overflow within a fold period of R rows
//      (vbits + log2 R <= SB, log2(R+1) <= 64 - SB). Every R rows the packed
//      table is folded into a wide LDS table (u64 sum, u64 count); at the end
//      the block stores its wide table as one row of a partial-table matrix.
//   2. reduce: column sums of that matrix (64 keys per block, coalesced rows).
//   3. writeout (one block): ascending key order via a block scan over
//      occupancy, minimal output widths from the maxima, slab-packed outputs.
// An optional value predicate fuses the chained filter -> group-by.
#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t GB_THREADS = 1024;
constexpr uint32_t GB_ROWS = 8; // rows per thread per step (loads in flight)

struct GroupArgs {
    Col keys, vals;
    Pred pr;
    uint32_t use_pred, pad;
    uint32_t kbits, mode;
    uint32_t SB;          // sum field bits of the packed LDS word
    uint32_t fold_tiles;  // tiles between packed -> wide folds
    uint64_t rows_per_block;
    uint64_t* part;       // [blocks][2G]: sums then counts
    uint64_t* total;      // [2G]
    uint32_t blocks;
    uint64_t* out_keys;
    uint64_t* out_sums;
    uint64_t* out_counts;
    uint64_t* result;
    uint32_t vmagic; // word path: floor(2^32 / eps of the values) + 1
};

__device__ inline uint32_t width_of(uint64_t v)
{
    return v ? 64 - (uint32_t)__clzll(v) : 1;
}

template <int KK, int VK, uint32_t TH>
__global__ __launch_bounds__(TH) void k_groupby_partial(GroupArgs a)
{
    constexpr uint32_t TILE = GB_ROWS * TH; // rows per step
    extern __shared__ uint64_t lds[];
    const uint32_t G = 1u << a.kbits;
    if (blockIdx.x == 0) // the reduce accumulates into the totals
        for (uint32_t c = threadIdx.x; c < 2 * G; c += TH) a.total[c] = 0;
    uint64_t* packed = lds;       // [G]
    uint64_t* wsum = lds + G;     // [G]
    uint64_t* wcnt = lds + 2 * G; // [G]
    const uint32_t tid = threadIdx.x;
    for (uint32_t g = tid; g < G; g += TH) {
        packed[g] = 0;
        wsum[g] = 0;
        wcnt[g] = 0;
    }
    __syncthreads();

    const uint64_t n = a.keys.n;
    const uint64_t r0 = (uint64_t)blockIdx.x * a.rows_per_block;
    uint64_t r1 = r0 + a.rows_per_block;
    if (r1 > n) r1 = n;
    const uint64_t SMASK = lowmask(a.SB);
    const uint64_t ONE = 1ull << a.SB;
    uint32_t since = 0;
    for (uint64_t t0 = r0; t0 < r1; t0 += TILE) {
        TileReader<KK> rk(a.keys, t0);
        TileReader<VK> rv(a.vals, t0);
        const uint64_t rem = r1 - t0;
        uint64_t kv[TILE / TH], vv[TILE / TH];
#pragma unroll
        for (uint32_t r = 0; r < TILE / TH; r++) {
            const uint32_t i = r * TH + tid;
            kv[r] = i < rem ? rk.get(i) : 0;
            vv[r] = i < rem ? rv.get(i) : 0;
        }
#pragma unroll
        for (uint32_t r = 0; r < TILE / TH; r++) {
            const uint32_t i = r * TH + tid;
            if (i < rem && (!a.use_pred || a.pr(vv[r])))
                atomicAdd((unsigned long long*)&packed[kv[r]], (unsigned long long)(ONE | vv[r]));
        }
        if (++since == a.fold_tiles) {
            since = 0;
            __syncthreads();
            for (uint32_t g = tid; g < G; g += TH) {
                const uint64_t p = packed[g];
                if (p) {
                    wsum[g] += p & SMASK;
                    wcnt[g] += p >> a.SB;
                    packed[g] = 0;
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    uint64_t* row = a.part + (uint64_t)blockIdx.x * 2 * G;
    for (uint32_t g = tid; g < G; g += TH) {
        const uint64_t p = packed[g];
        row[g] = wsum[g] + (p & SMASK);
        row[G + g] = wcnt[g] + (p >> a.SB);
    }
}

// Slab keys and slab values, word per lane: per step a thread takes
// GW_UNITS key words (ek rows each), their key fields at fixed shifts, and
// the 2..4 value words those rows span, walking the value fields with a
// small register window (shifted when a word is used up). Per row: two
// field extracts and the LDS atomic -- no per-row division or load. Same
// table, folds and partial rows as k_groupby_partial.
constexpr uint32_t GW_UNITS = 2;

__global__ __launch_bounds__(GB_THREADS) void k_groupby_words(GroupArgs a)
{
    extern __shared__ uint64_t lds[];
    const uint32_t G = 1u << a.kbits;
    if (blockIdx.x == 0) // the reduce accumulates into the totals
        for (uint32_t c = threadIdx.x; c < 2 * G; c += GB_THREADS) a.total[c] = 0;
    uint64_t* packed = lds;
    uint64_t* wsum = lds + G;
    uint64_t* wcnt = lds + 2 * G;
    const uint32_t tid = threadIdx.x;
    for (uint32_t g = tid; g < G; g += GB_THREADS) {
        packed[g] = 0;
        wsum[g] = 0;
        wcnt[g] = 0;
    }
    __syncthreads();
    const uint64_t n = a.keys.n;
    const uint32_t ek = a.keys.eps, ev = a.vals.eps, wk = a.keys.bits, wv = a.vals.bits;
    const uint64_t km = a.keys.mask, vm = a.vals.mask;
    const uint64_t vwords = a.vals.words, kwords = a.keys.words;
    const uint64_t r0 = (uint64_t)blockIdx.x * a.rows_per_block; // a multiple of ek
    uint64_t r1 = r0 + a.rows_per_block;
    if (r1 > n) r1 = n;
    const uint64_t SMASK = lowmask(a.SB);
    const uint64_t ONE = 1ull << a.SB;
    const uint32_t step = GB_THREADS * GW_UNITS * ek; // rows per step (<= one fold period)
    uint32_t since = 0;
    for (uint64_t t0 = r0; t0 < r1; t0 += step) {
        const uint64_t qv0 = t0 / ev; // value word and field of row t0 (uniform)
        const uint32_t sv0 = (uint32_t)(t0 - qv0 * ev);
        uint64_t K[GW_UNITS], V[GW_UNITS][4];
        uint32_t S[GW_UNITS];
#pragma unroll
        for (uint32_t u = 0; u < GW_UNITS; u++) {
            const uint32_t rel = (u * GB_THREADS + tid) * ek; // < 2^24
            const uint64_t kq = t0 / ek + u * GB_THREADS + tid;
            K[u] = kq < kwords ? a.keys.p[kq] : 0;
            const uint32_t x = sv0 + rel;
            const uint32_t dq = ev == 1 ? x : (uint32_t)(((uint64_t)x * a.vmagic) >> 32);
            S[u] = x - dq * ev;
            const uint64_t qv = qv0 + dq;
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) V[u][k] = qv + k < vwords ? a.vals.p[qv + k] : 0;
        }
#pragma unroll
        for (uint32_t u = 0; u < GW_UNITS; u++) {
            const uint64_t e0 = t0 + (u * GB_THREADS + tid) * ek;
            if (e0 >= r1) continue;
            const uint32_t cnt = (uint32_t)(r1 - e0 < ek ? r1 - e0 : ek);
            uint64_t v0 = V[u][0], v1 = V[u][1], v2 = V[u][2], v3 = V[u][3];
            uint32_t sv = S[u];
            for (uint32_t j = 0, sk = 0; j < cnt; j++, sk += wk) {
                const uint64_t key = (K[u] >> sk) & km;
                const uint64_t val = (v0 >> (sv * wv)) & vm;
                if (!a.use_pred || a.pr(val))
                    atomicAdd((unsigned long long*)&packed[key], (unsigned long long)(ONE | val));
                if (++sv == ev) {
                    sv = 0;
                    v0 = v1;
                    v1 = v2;
                    v2 = v3;
                }
            }
        }
        if (++since == a.fold_tiles) {
            since = 0;
            __syncthreads();
            for (uint32_t g = tid; g < G; g += GB_THREADS) {
                const uint64_t p = packed[g];
                if (p) {
                    wsum[g] += p & SMASK;
                    wcnt[g] += p >> a.SB;
                    packed[g] = 0;
                }
            }
            __syncthreads();
        }
    }
    __syncthreads();
    uint64_t* row = a.part + (uint64_t)blockIdx.x * 2 * G;
    for (uint32_t g = tid; g < G; g += GB_THREADS) {
        const uint64_t p = packed[g];
        row[g] = wsum[g] + (p & SMASK);
        row[G + g] = wcnt[g] + (p >> a.SB);
    }
}

// total[c] = sum over blocks of part[b][c], c < 2G; 64 columns per block,
// each wave sums every 4th row (coalesced 512-byte rows).
// Column sums of the partial tables: block (x, y) sums partial rows
// [y*R, (y+1)*R) of columns [64x, 64x+64) and adds them into the totals
// (zeroed by the partial kernel) with 64-bit integer atomics -- exact, and
// independent of the order the blocks finish in.
constexpr uint32_t GB_RED_ROWS = 16; // partial rows per reduce block
__global__ __launch_bounds__(256) void k_groupby_reduce(GroupArgs a)
{
    __shared__ uint64_t s_acc[4][64];
    const uint32_t G2 = 2u << a.kbits;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t c = blockIdx.x * 64 + lane;
    const uint32_t b0 = blockIdx.y * GB_RED_ROWS, b1 = min(b0 + GB_RED_ROWS, a.blocks);
    uint64_t acc = 0;
    if (c < G2)
        for (uint32_t b = b0 + wave; b < b1; b += 4) acc += a.part[(uint64_t)b * G2 + c];
    s_acc[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && c < G2) {
        const uint64_t sum = s_acc[0][lane] + s_acc[1][lane] + s_acc[2][lane] + s_acc[3][lane];
        if (sum) atomicAdd((unsigned long long*)&a.total[c], (unsigned long long)sum);
    }
}

__global__ __launch_bounds__(GB_THREADS) void k_groupby_writeout(GroupArgs a)
{
    __shared__ uint32_t s_rank_key[4096];
    __shared__ uint32_t s_wave[GB_THREADS / 64];
    __shared__ uint64_t s_red[3][GB_THREADS / 64];
    const uint32_t G = 1u << a.kbits;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t NW = GB_THREADS / 64;
    const uint64_t* sums = a.total;
    const uint64_t* counts = a.total + G;
    uint64_t smax = 0, cmax = 0, rows = 0;
    for (uint32_t g = tid; g < G; g += GB_THREADS) {
        const uint64_t c = counts[g];
        uint64_t s = sums[g];
        if (a.mode == AM_GB_REF) s &= a.vals.mask;
        if (c) {
            smax = s > smax ? s : smax;
            cmax = c > cmax ? c : cmax;
            rows += c;
        }
    }
    smax = wave_max_u64(smax);
    cmax = wave_max_u64(cmax);
    rows = wave_sum_u64(rows);
    if (lane == 0) {
        s_red[0][wave] = smax;
        s_red[1][wave] = cmax;
        s_red[2][wave] = rows;
    }
    __syncthreads();
    smax = 0;
    cmax = 0;
    rows = 0;
    for (uint32_t w = 0; w < NW; w++) {
        smax = s_red[0][w] > smax ? s_red[0][w] : smax;
        cmax = s_red[1][w] > cmax ? s_red[1][w] : cmax;
        rows += s_red[2][w];
    }
    const uint32_t sbits = a.mode == AM_GB_REF ? a.vals.bits : width_of(smax);
    const uint32_t cbits = width_of(cmax);
    // ranks of present keys in ascending key order
    uint32_t base = 0;
    for (uint32_t g0 = 0; g0 < G; g0 += GB_THREADS) {
        const uint32_t g = g0 + tid;
        const bool present = g < G && counts[g] != 0;
        const uint64_t b = ballot(present);
        if (lane == 0) s_wave[wave] = (uint32_t)__popcll(b);
        __syncthreads();
        uint32_t pre = base, tot = 0;
        for (uint32_t w = 0; w < NW; w++) {
            pre += w < wave ? s_wave[w] : 0;
            tot += s_wave[w];
        }
        if (present) s_rank_key[pre + mbcnt(b)] = g;
        base += tot;
        __syncthreads();
    }
    const uint32_t ng = base;
    const uint32_t ek = 64 / a.kbits, es = 64 / sbits, ec = 64 / cbits;
    const uint64_t smask = lowmask(sbits);
    for (uint32_t wd = tid; wd < (ng + ek - 1) / ek; wd += GB_THREADS) {
        uint64_t word = 0;
        for (uint32_t s = 0; s < ek && wd * ek + s < ng; s++) word |= (uint64_t)s_rank_key[wd * ek + s] << (s * a.kbits);
        a.out_keys[wd] = word;
    }
    for (uint32_t wd = tid; wd < (ng + es - 1) / es; wd += GB_THREADS) {
        uint64_t word = 0;
        for (uint32_t s = 0; s < es && wd * es + s < ng; s++)
            word |= (sums[s_rank_key[wd * es + s]] & smask) << (s * sbits);
        a.out_sums[wd] = word;
    }
    if (a.out_counts) {
        for (uint32_t wd = tid; wd < (ng + ec - 1) / ec; wd += GB_THREADS) {
            uint64_t word = 0;
            for (uint32_t s = 0; s < ec && wd * ec + s < ng; s++) word |= counts[s_rank_key[wd * ec + s]] << (s * cbits);
            a.out_counts[wd] = word;
        }
    }
    if (tid == 0) {
        a.result[0] = ng;
        a.result[1] = sbits;
        a.result[2] = cbits;
        a.result[3] = rows;
    }
}

} // namespace am

using namespace am;

extern "C" uint64_t am_groupby_workspace_bytes(uint64_t n, uint32_t key_bits, uint32_t value_bits)
{
    if (key_bits < 1 || key_bits > 24 || value_bits < 1 || value_bits > 32) return 0;
    return key_bits <= 12 ? WS_FIXED : am_groupby_wide_bytes(n, key_bits, value_bits);
}

extern "C" int am_groupby(const am_column* keys, const am_column* values, const am_predicate* value_pred, uint32_t mode,
                          uint64_t* out_keys, uint64_t* out_sums, uint64_t* out_counts, uint64_t* result_dev, void* ws,
                          uint64_t ws_bytes, void* stream)
{
    Col ck, cv;
    int st = make_col(keys, &ck);
    if (st) return st;
    if ((st = make_col(values, &cv))) return st;
    if (ck.n != cv.n || !out_keys || !out_sums || !result_dev) return AM_E_INVALID;
    if (mode > AM_GB_REF || (value_pred && value_pred->op > AM_PRED_NE)) return AM_E_INVALID;
    if (ck.bits > 24 || cv.bits > 32) return AM_E_UNSUPPORTED_CONFIG;
    if (!ws || ws_bytes < WS_FIXED) return AM_E_WORKSPACE;
    if (ck.bits > 12) // partitioned path
        return am_groupby_wide(ck, cv, value_pred, mode, out_keys, out_sums, out_counts, result_dev, ws, ws_bytes,
                               (hipStream_t)stream);
    GroupArgs a{};
    a.keys = ck;
    a.vals = cv;
    a.use_pred = value_pred ? 1 : 0;
    if (value_pred) a.pr = make_pred(*value_pred, cv.mask);
    a.kbits = ck.bits;
    a.mode = mode;
    // packed word: sum field SB bits, count field 64-SB bits; R = 2^r rows per fold
    const uint32_t r = (63 - cv.bits) / 2;
    a.SB = cv.bits + r;
    // up to 1024 keys: 4-wave blocks (24 KB of LDS, 6 per CU) and as many as the
    // partial-table region holds, so enough rows are in flight per CU; wider
    // keys: 16-wave blocks (LDS-bound), 256 of them
    const uint32_t G = 1u << ck.bits;
    const uint32_t TH = GB_THREADS;
    const uint32_t tile = GB_ROWS * TH;
    a.fold_tiles = (uint32_t)((1ull << r) / tile); // r >= 15 for vbits <= 32, so >= 8 steps
    uint64_t maxb = WS_GBPART_BYTES / (16ull * G);
    if (maxb > 2048) maxb = 2048;
    if (const uint32_t tb = am_tuned(AM_KNOB_GROUPBY_BLOCKS, keys->bits, keys->format, values->format))
        if (tb < maxb) maxb = tb < 64 ? 64 : tb;
    const uint64_t tiles = (ck.n + tile - 1) / tile;
    uint64_t blocks = tiles < maxb ? (tiles ? tiles : 1) : maxb;
    const uint64_t tiles_per_block = tiles ? (tiles + blocks - 1) / blocks : 1;
    blocks = tiles ? (tiles + tiles_per_block - 1) / tiles_per_block : 1;
    a.rows_per_block = tiles_per_block * tile;
    a.blocks = (uint32_t)blocks;
    a.part = ws_gbpart(ws);
    a.total = ws_small(ws);
    a.out_keys = out_keys;
    a.out_sums = out_sums;
    a.out_counts = out_counts;
    a.result = result_dev;
    const size_t lds = (size_t)G * 24;
    hipStream_t s = (hipStream_t)stream;
    // word path: slab keys and values, a key word's rows within 4 value words,
    // a step's rows within one fold period (the packed count cannot overflow)
    const uint32_t wstep = GB_THREADS * GW_UNITS * ck.eps;
    if (ck.kind == K_SLAB && cv.kind == K_SLAB && ck.eps + cv.eps - 1 <= 4 * cv.eps && wstep <= (1u << r) &&
        ck.n) {
        a.fold_tiles = (uint32_t)((1ull << r) / wstep);
        a.vmagic = (uint32_t)((1ull << 32) / cv.eps + 1);
        uint64_t steps = (ck.n + wstep - 1) / wstep;
        uint64_t wb = steps < maxb ? steps : maxb;
        const uint64_t spb = (steps + wb - 1) / wb;
        wb = (steps + spb - 1) / spb;
        a.rows_per_block = spb * wstep; // a multiple of the key word's rows
        a.blocks = (uint32_t)wb;
        if (lds > 65536) (void)hipFuncSetAttribute((const void*)k_groupby_words, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_groupby_words, dim3((uint32_t)wb), dim3(GB_THREADS), lds, s, a);
        hipLaunchKernelGGL(k_groupby_reduce, dim3((2 * G + 63) / 64, (a.blocks + GB_RED_ROWS - 1) / GB_RED_ROWS), dim3(256),
                           0, s, a);
        hipLaunchKernelGGL(k_groupby_writeout, dim3(1), dim3(GB_THREADS), 0, s, a);
        return check_launch();
    }
#define GB_GO(KK, VK, T)                                                                                       \
    {                                                                                                          \
        if (lds > 65536) (void)hipFuncSetAttribute((const void*)k_groupby_partial<KK, VK, T>,                  \
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);      \
        hipLaunchKernelGGL((k_groupby_partial<KK, VK, T>), dim3((uint32_t)blocks), dim3(T), lds, s, a);        \
    }
#define GB_LAUNCH(KK, VK)                      \
    {                                          \
        if (TH == 256) GB_GO(KK, VK, 256)      \
        else GB_GO(KK, VK, GB_THREADS)         \
    }
#define GB_V(KK)                                         \
    switch (cv.kind) {                                   \
        case K_STRIDED: GB_LAUNCH(KK, K_STRIDED); break; \
        case K_SLAB: GB_LAUNCH(KK, K_SLAB); break;       \
        default: GB_LAUNCH(KK, K_BITSTREAM); break;      \
    }
    switch (ck.kind) {
        case K_STRIDED: GB_V(K_STRIDED); break;
        case K_SLAB: GB_V(K_SLAB); break;
        default: GB_V(K_BITSTREAM); break;
    }
#undef GB_V
#undef GB_LAUNCH
#undef GB_GO
    hipLaunchKernelGGL(k_groupby_reduce, dim3((2 * G + 63) / 64, (a.blocks + GB_RED_ROWS - 1) / GB_RED_ROWS), dim3(256), 0,
                       s, a);
    hipLaunchKernelGGL(k_groupby_writeout, dim3(1), dim3(GB_THREADS), 0, s, a);
    return check_launch();
}
