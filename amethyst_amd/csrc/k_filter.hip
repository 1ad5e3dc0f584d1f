// k_filter.hip -- predicate filter to an LSB-first bitmask, and the u64 sum.
//
// Filter reference: kernel_filter_baseline / _nogaps / _slabs
// (src/kernels/filter_baseline.cuh:9-36, filter_nogaps.cuh:4-65,
// filter_slabs.cuh:4-57). There each thread decodes 32 consecutive elements
// serially and writes one u32 mask word; the baseline variant needs a 32x32
// shared-memory bit transpose.
// Here lane l of a wave decodes element base + 64j + l, so one wave64
// ballot (a single v_cmp into an SGPR pair) *is* mask word j -- no transpose,
// no shared memory. Lane j keeps ballot j and the wave writes its 16 words
// with one coalesced store.
//
// Sum reference: kernel_hop_slabs / kernel_hop_nogaps (src/kernels/hop.cuh:7-76),
// one global atomicAdd per thread. Here: per-block wave reductions, one
// agent-scope atomic per block into the workspace, and the last block to
// finish moves the total to the caller's pointer and re-zeroes the
// accumulator (no memset, deterministic: u64 addition wraps and commutes).
#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t FT_J = 16;                 // mask words per wave
constexpr uint32_t FT_TILE = 4 * FT_J * 64;   // elements per 256-thread block

template <int IK>
__global__ __launch_bounds__(256) void k_filter(Col in, uint32_t pop, uint64_t pc, uint64_t* mask_out)
{
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t wbase = (uint64_t)blockIdx.x * FT_TILE + wave * (FT_J * 64);
    if (wbase >= in.n) return;
    TileReader<IK> rd(in, wbase);
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t j = 0; j < FT_J; j++) {
        const uint64_t e = wbase + j * 64 + lane;
        bool f = false;
        if (e < in.n) f = pred_eval(pop, pc, rd.get(j * 64 + lane));
        uint64_t b = ballot(f);
        if (lane == j) mine = b;
    }
    const uint64_t words = (in.n + 63) >> 6;
    const uint64_t w = (wbase >> 6) + lane;
    if (lane < FT_J && w < words) mask_out[w] = mine;
}

template <int IK>
__global__ __launch_bounds__(256) void k_sum(Col in, WsHeader* hdr, uint64_t* out)
{
    __shared__ uint64_t s_part[4];
    uint64_t acc = 0;
    for (uint64_t t0 = (uint64_t)blockIdx.x * FT_TILE; t0 < in.n; t0 += (uint64_t)gridDim.x * FT_TILE) {
        TileReader<IK> rd(in, t0);
        const uint64_t rem = in.n - t0;
#pragma unroll 4
        for (uint32_t r = 0; r < FT_TILE / 256; r++) {
            uint32_t i = r * 256 + threadIdx.x;
            if (i < rem) acc += rd.get(i);
        }
    }
    acc = wave_sum_u64(acc);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t tot = s_part[0] + s_part[1] + s_part[2] + s_part[3];
        (void)add64(&hdr->accum, tot);
        drain();
        uint32_t d = add32(&hdr->done, 1);
        if (d == gridDim.x - 1) {
            uint64_t total = __hip_atomic_exchange(&hdr->accum, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            out[0] = total;
            st32(&hdr->done, 0);
        }
    }
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
    dim3 grid((uint32_t)((c.n + FT_TILE - 1) / FT_TILE));
    hipStream_t s = (hipStream_t)stream;
    switch (c.kind) {
        case K_STRIDED: hipLaunchKernelGGL(k_filter<K_STRIDED>, grid, dim3(256), 0, s, c, pred.op, pred.c, mask_out); break;
        case K_SLAB: hipLaunchKernelGGL(k_filter<K_SLAB>, grid, dim3(256), 0, s, c, pred.op, pred.c, mask_out); break;
        default: hipLaunchKernelGGL(k_filter<K_BITSTREAM>, grid, dim3(256), 0, s, c, pred.op, pred.c, mask_out); break;
    }
    return check_launch();
}

extern "C" int am_sum(const am_column* in, uint64_t* sum_dev, void* ws, uint64_t ws_bytes, void* stream)
{
    Col c;
    int st = make_col(in, &c);
    if (st) return st;
    if (!sum_dev) return AM_E_INVALID;
    if (!ws || ws_bytes < WS_FIXED) return AM_E_WORKSPACE;
    uint64_t tiles = (c.n + FT_TILE - 1) / FT_TILE;
    uint32_t blocks = (uint32_t)(tiles < 2048 ? (tiles ? tiles : 1) : 2048);
    hipStream_t s = (hipStream_t)stream;
    WsHeader* h = (WsHeader*)ws;
    switch (c.kind) {
        case K_STRIDED: hipLaunchKernelGGL(k_sum<K_STRIDED>, dim3(blocks), dim3(256), 0, s, c, h, sum_dev); break;
        case K_SLAB: hipLaunchKernelGGL(k_sum<K_SLAB>, dim3(blocks), dim3(256), 0, s, c, h, sum_dev); break;
        default: hipLaunchKernelGGL(k_sum<K_BITSTREAM>, dim3(blocks), dim3(256), 0, s, c, h, sum_dev); break;
    }
    return check_launch();
}
