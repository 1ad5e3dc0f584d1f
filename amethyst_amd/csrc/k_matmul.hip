// k_matmul.hip -- exact integer matmul of bit-packed <= 7-bit matrices on
// the int8 matrix cores (v_mfma_i32_16x16x64_i8, exact i32 accumulation).
//
// Reference: the packed-matrix layout helpers only -- matrix_mul_calc_rowcolcnt
// and get_matrix_element (src/common_code.cuh:65-86,118-134): row-major,
// each row padded to whole slabs, element (r, c) in word r*ceil(cols/eps) +
// c/eps at bit bits*(c%eps). The reference ships no matmul kernel
// (README.md:20), so parity is pinned by the int64 oracle only.
//
// Pipeline (stream-ordered):
//   1. unpack A (M x K) to int8 rows [M][Kp] and B (K x N) to int8 columns
//      [N][Kp] (Kp = K rounded up to 64, zero-padded): 7-bit values are
//      non-negative in int8, so the signed int8 MFMA is exact;
//   2. GEMM: 128x128 block tiles, K-steps of 64 staged through LDS with
//      16-byte loads and double buffering; each of 4 waves owns a 64x64
//      quadrant = 4x4 MFMA tiles of 16x16 (i32 accumulators).
//      Both operands use the same k-to-slot mapping per lane, so the sum over
//      k is exact whatever order the matrix core visits the slots in;
//   3. epilogue: C as row-major int32 and/or re-packed as a row-padded slab
//      matrix at `out_bits` (fused when floor(64/out_bits) divides 16,
//      otherwise through the int32 copy and a row packer).
// The batched entry point runs every phase for all matrices of a batch in
// one persistent launch per phase (work lists built on the host).
#include <algorithm>
#include <vector>

#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr uint32_t MM_BM = 128, MM_BN = 128, MM_BK = 64;
constexpr uint32_t MM_LDS_STRIDE = MM_BK + 16; // bytes per staged row (bank spread)

struct MatJob {
    const uint64_t* A; // packed M x K
    const uint64_t* B; // packed K x N
    uint64_t* Cp;      // packed C (row-padded slab at out_bits) or null
    int32_t* Ci;       // int32 C [M][N] or null
    int8_t* A8;        // [M][Kp]
    int8_t* B8;        // [N][Kp]
    uint32_t M, N, K, Kp;
    uint32_t out_bits;
    uint32_t tiles_n;  // ceil(N / 128)
    uint32_t tile0;    // first GEMM tile of this job in the batch
    uint32_t achunk0;  // first A-unpack work item
    uint32_t bchunk0;  // first B-unpack work item
};

__device__ inline uint32_t find_job(const MatJob* jobs, uint32_t njobs, uint32_t item, int which)
{
    // last job whose first item <= item (jobs sorted by first item)
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        const uint32_t f = which == 0 ? jobs[mid].achunk0 : (which == 1 ? jobs[mid].bchunk0 : jobs[mid].tile0);
        if (f <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline uint32_t get7(const uint64_t* row, uint32_t c, uint32_t bits, uint32_t eps)
{
    const uint64_t w = row[c / eps];
    return (uint32_t)((w >> ((c % eps) * bits)) & lowmask(bits));
}

// A: one work item = one row and 16 consecutive k (one 16-byte store).
__global__ __launch_bounds__(256) void k_mm_unpack_a(const MatJob* jobs, uint32_t njobs, uint32_t bits,
                                                     uint32_t total_items)
{
    const uint32_t eps = 64 / bits;
    for (uint32_t it = blockIdx.x * 256 + threadIdx.x; it < total_items; it += gridDim.x * 256) {
        const MatJob& J = jobs[find_job(jobs, njobs, it, 0)];
        const uint32_t local = it - J.achunk0;
        const uint32_t cpr = J.Kp / 16; // chunks per row
        const uint32_t r = local / cpr, k0 = (local % cpr) * 16;
        const uint64_t* row = J.A + (uint64_t)r * ((J.K + eps - 1) / eps);
        uint32_t v[4] = {0, 0, 0, 0};
        for (uint32_t s = 0; s < 16; s++) {
            const uint32_t k = k0 + s;
            const uint32_t x = k < J.K ? get7(row, k, bits, eps) : 0u;
            v[s >> 2] |= x << (8 * (s & 3));
        }
        *(uint4*)(J.A8 + (uint64_t)r * J.Kp + k0) = make_uint4(v[0], v[1], v[2], v[3]);
    }
}

// B^T: one work item = one column n and 64 consecutive k; neighbouring lanes
// take neighbouring columns, so every load instruction reads a few adjacent
// words of one packed row.
__global__ __launch_bounds__(256) void k_mm_unpack_bt(const MatJob* jobs, uint32_t njobs, uint32_t bits,
                                                      uint32_t total_items)
{
    const uint32_t eps = 64 / bits;
    for (uint32_t it = blockIdx.x * 256 + threadIdx.x; it < total_items; it += gridDim.x * 256) {
        const MatJob& J = jobs[find_job(jobs, njobs, it, 1)];
        const uint32_t local = it - J.bchunk0;
        const uint32_t kc = local / J.N, n = local % J.N;
        const uint32_t rw = (J.N + eps - 1) / eps;
        const uint32_t sh = (n % eps) * bits, wi = n / eps;
        uint32_t v[16];
#pragma unroll
        for (uint32_t q = 0; q < 16; q++) v[q] = 0;
#pragma unroll 8
        for (uint32_t s = 0; s < 64; s++) {
            const uint32_t k = kc * 64 + s;
            uint32_t x = 0;
            if (k < J.K) x = (uint32_t)((J.B[(uint64_t)k * rw + wi] >> sh) & lowmask(bits));
            v[s >> 2] |= x << (8 * (s & 3));
        }
        uint4* dst = (uint4*)(J.B8 + (uint64_t)n * J.Kp + kc * 64);
        dst[0] = make_uint4(v[0], v[1], v[2], v[3]);
        dst[1] = make_uint4(v[4], v[5], v[6], v[7]);
        dst[2] = make_uint4(v[8], v[9], v[10], v[11]);
        dst[3] = make_uint4(v[12], v[13], v[14], v[15]);
    }
}

// GEMM over the batch's tile list; one 128x128 output tile per iteration.
__global__ __launch_bounds__(256) void k_mm_gemm(const MatJob* jobs, uint32_t njobs, uint32_t total_tiles)
{
    __shared__ __attribute__((aligned(16))) int8_t sA[2][MM_BM * MM_LDS_STRIDE];
    __shared__ __attribute__((aligned(16))) int8_t sB[2][MM_BN * MM_LDS_STRIDE];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t wm = wave >> 1, wn = wave & 1;
    const uint32_t lr = lane & 15, lg = lane >> 4;

    for (uint32_t tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
        const MatJob& J = jobs[find_job(jobs, njobs, tile, 2)];
        const uint32_t local = tile - J.tile0;
        const uint32_t m0 = (local / J.tiles_n) * MM_BM, n0 = (local % J.tiles_n) * MM_BN;
        const uint32_t steps = J.Kp / MM_BK;

        // each thread stages 2 x 16 bytes of A and of B per K-step
        uint4 ra[2], rb[2];
        auto gload = [&](uint32_t ks) {
#pragma unroll
            for (uint32_t h = 0; h < 2; h++) {
                const uint32_t c = tid + h * 256; // 512 chunks of 16 B per operand
                const uint32_t row = c >> 2, kq = (c & 3) * 16;
                const uint32_t gm = m0 + row, gn = n0 + row;
                ra[h] = gm < J.M ? *(const uint4*)(J.A8 + (uint64_t)gm * J.Kp + ks * MM_BK + kq) : make_uint4(0, 0, 0, 0);
                rb[h] = gn < J.N ? *(const uint4*)(J.B8 + (uint64_t)gn * J.Kp + ks * MM_BK + kq) : make_uint4(0, 0, 0, 0);
            }
        };
        auto sstore = [&](uint32_t buf) {
#pragma unroll
            for (uint32_t h = 0; h < 2; h++) {
                const uint32_t c = tid + h * 256;
                const uint32_t row = c >> 2, kq = (c & 3) * 16;
                *(uint4*)(&sA[buf][row * MM_LDS_STRIDE + kq]) = ra[h];
                *(uint4*)(&sB[buf][row * MM_LDS_STRIDE + kq]) = rb[h];
            }
        };

        v4i acc[4][4];
#pragma unroll
        for (uint32_t i = 0; i < 4; i++)
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) acc[i][j] = v4i{0, 0, 0, 0};

        gload(0);
        __syncthreads(); // previous tile's readers are done with the buffers
        sstore(0);
        __syncthreads();
        for (uint32_t ks = 0; ks < steps; ks++) {
            const uint32_t buf = ks & 1;
            if (ks + 1 < steps) gload(ks + 1);
            v4i fa[4], fb[4];
#pragma unroll
            for (uint32_t i = 0; i < 4; i++) {
                fa[i] = *(const v4i*)(&sA[buf][(wm * 64 + i * 16 + lr) * MM_LDS_STRIDE + lg * 16]);
                fb[i] = *(const v4i*)(&sB[buf][(wn * 64 + i * 16 + lr) * MM_LDS_STRIDE + lg * 16]);
            }
#pragma unroll
            for (uint32_t i = 0; i < 4; i++)
#pragma unroll
                for (uint32_t j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
            if (ks + 1 < steps) {
                sstore(buf ^ 1);
                __syncthreads();
            }
        }

        // ---- epilogue: lane holds C[row0 + 4*lg + r][col0 + lr], r < 4
        const uint32_t ob = J.out_bits;
        const uint32_t eps_o = 64 / ob;
        const bool fused = J.Cp && (16 % eps_o == 0);
        const uint32_t rw_o = (J.N + eps_o - 1) / eps_o;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t col = n0 + wn * 64 + j * 16 + lr;
#pragma unroll
                for (uint32_t r = 0; r < 4; r++) {
                    const uint32_t row = m0 + wm * 64 + i * 16 + 4 * lg + r;
                    const int32_t val = acc[i][j][r];
                    if (J.Ci && row < J.M && col < J.N) J.Ci[(uint64_t)row * J.N + col] = val;
                    if (fused) {
                        uint64_t word = (col < J.N) ? ((uint64_t)(uint32_t)val & lowmask(ob)) : 0;
                        for (uint32_t s = 1; s < eps_o; s++) {
                            const uint32_t src = __shfl((int)val, (int)(lane + s), 64);
                            if (col + s < J.N) word |= ((uint64_t)src & lowmask(ob)) << (s * ob);
                        }
                        if (row < J.M && col < J.N && (lr % eps_o) == 0)
                            J.Cp[(uint64_t)row * rw_o + col / eps_o] = word;
                    }
                }
            }
        }
    }
}

// Row packer for widths whose slabs do not tile 16 columns: int32 C -> slab rows.
__global__ __launch_bounds__(256) void k_mm_pack_rows(const MatJob* jobs, uint32_t njobs, const uint32_t* word0,
                                                      uint32_t total_words)
{
    for (uint32_t it = blockIdx.x * 256 + threadIdx.x; it < total_words; it += gridDim.x * 256) {
        uint32_t lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (word0[mid] <= it) lo = mid;
            else hi = mid - 1;
        }
        const MatJob& J = jobs[lo];
        if (!J.Cp || !J.Ci) continue;
        const uint32_t ob = J.out_bits, eps_o = 64 / ob;
        const uint32_t rw_o = (J.N + eps_o - 1) / eps_o;
        const uint32_t local = it - word0[lo];
        const uint32_t row = local / rw_o, wc = local % rw_o;
        uint64_t word = 0;
        for (uint32_t s = 0; s < eps_o; s++) {
            const uint32_t col = wc * eps_o + s;
            if (col < J.N) word |= ((uint64_t)(uint32_t)J.Ci[(uint64_t)row * J.N + col] & lowmask(ob)) << (s * ob);
        }
        J.Cp[(uint64_t)row * rw_o + wc] = word;
    }
}

inline uint32_t width_of_host(uint64_t v)
{
    uint32_t w = 1;
    while (w < 64 && (v >> w)) w++;
    return w;
}

} // namespace am

using namespace am;

// Workspace bytes for one matrix product (int8 copies, int32 staging, job list).
static uint64_t mm_ws_one(uint64_t M, uint64_t N, uint64_t K)
{
    const uint64_t Kp = (K + 63) / 64 * 64;
    return (M * Kp + 255) / 256 * 256 + (N * Kp + 255) / 256 * 256 + (M * N * 4 + 255) / 256 * 256;
}

extern "C" uint64_t am_matmul_workspace_bytes(const uint32_t* M, const uint32_t* N, const uint32_t* K, uint32_t count)
{
    uint64_t b = 4096 + (uint64_t)count * (sizeof(MatJob) + 4) + 256;
    for (uint32_t i = 0; i < count; i++) b += mm_ws_one(M[i], N[i], K[i]);
    return b;
}

extern "C" uint32_t am_matmul_out_bits(uint32_t bits, uint32_t K)
{
    if (bits < 1 || bits > 7 || K == 0) return 1;
    const uint64_t vmax = (uint64_t)K * (((1ull << bits) - 1) * ((1ull << bits) - 1));
    return width_of_host(vmax);
}

extern "C" int am_matmul_packed_batched(uint32_t bits, const am_matmul_desc* descs, uint32_t count, void* ws,
                                        uint64_t ws_bytes, void* stream)
{
    if (bits < 1 || bits > 7) return AM_E_UNSUPPORTED_CONFIG;
    if (!descs || count == 0) return AM_E_INVALID;
    if (!ws) return AM_E_WORKSPACE;
    std::vector<uint32_t> Ms(count), Ns(count), Ks(count);
    for (uint32_t i = 0; i < count; i++) {
        const am_matmul_desc& d = descs[i];
        if (!d.A || !d.B || (!d.C_packed && !d.C_i32) || d.M == 0 || d.N == 0 || d.K == 0) return AM_E_INVALID;
        if (d.M > (1u << 24) || d.N > (1u << 24) || d.K > (1u << 24)) return AM_E_UNSUPPORTED_CONFIG;
        if ((uint64_t)d.K * 16129ull >= (1ull << 31)) return AM_E_UNSUPPORTED_CONFIG; // i32 exactness
        if (d.out_bits > 32) return AM_E_INVALID;
        Ms[i] = d.M;
        Ns[i] = d.N;
        Ks[i] = d.K;
    }
    if (ws_bytes < am_matmul_workspace_bytes(Ms.data(), Ns.data(), Ks.data(), count)) return AM_E_WORKSPACE;
    std::vector<MatJob> jobs(count);
    std::vector<uint32_t> word0(count);
    uint8_t* cur = (uint8_t*)ws + 4096 + (uint64_t)count * (sizeof(MatJob) + 4) + 256;
    uint64_t tiles = 0, ach = 0, bch = 0, pwords = 0;
    bool need_pack = false;
    for (uint32_t i = 0; i < count; i++) {
        const am_matmul_desc& d = descs[i];
        MatJob& J = jobs[i];
        J.A = d.A;
        J.B = d.B;
        J.M = d.M;
        J.N = d.N;
        J.K = d.K;
        J.Kp = (d.K + 63) / 64 * 64;
        J.out_bits = d.out_bits ? d.out_bits : am_matmul_out_bits(bits, d.K);
        J.A8 = (int8_t*)cur;
        cur += ((uint64_t)J.M * J.Kp + 255) / 256 * 256;
        J.B8 = (int8_t*)cur;
        cur += ((uint64_t)J.N * J.Kp + 255) / 256 * 256;
        J.Cp = d.C_packed;
        J.Ci = d.C_i32;
        const uint32_t eps_o = 64 / J.out_bits;
        if (J.Cp && (16 % eps_o) != 0) {
            need_pack = true;
            if (!J.Ci) J.Ci = (int32_t*)cur; // staging copy
        }
        cur += ((uint64_t)J.M * J.N * 4 + 255) / 256 * 256;
        J.tiles_n = (J.N + MM_BN - 1) / MM_BN;
        J.tile0 = (uint32_t)tiles;
        J.achunk0 = (uint32_t)ach;
        J.bchunk0 = (uint32_t)bch;
        word0[i] = (uint32_t)pwords;
        tiles += (uint64_t)((J.M + MM_BM - 1) / MM_BM) * J.tiles_n;
        ach += (uint64_t)J.M * (J.Kp / 16);
        bch += (uint64_t)J.N * (J.Kp / 64);
        if (J.Cp && (16 % eps_o) != 0) pwords += (uint64_t)J.M * ((J.N + eps_o - 1) / eps_o);
    }
    if (tiles >= (1ull << 32) || ach >= (1ull << 32) || bch >= (1ull << 32) || pwords >= (1ull << 32))
        return AM_E_UNSUPPORTED_CONFIG;
    hipStream_t s = (hipStream_t)stream;
    MatJob* djobs = (MatJob*)((uint8_t*)ws + 4096);
    uint32_t* dword0 = (uint32_t*)((uint8_t*)djobs + (uint64_t)count * sizeof(MatJob));
    if (hipMemcpyAsync(djobs, jobs.data(), count * sizeof(MatJob), hipMemcpyHostToDevice, s) != hipSuccess) return AM_E_HIP;
    if (hipMemcpyAsync(dword0, word0.data(), count * 4, hipMemcpyHostToDevice, s) != hipSuccess) return AM_E_HIP;
    auto grid_for = [](uint64_t items) { return (uint32_t)std::min<uint64_t>((items + 255) / 256, 65535); };
    hipLaunchKernelGGL(k_mm_unpack_a, dim3(grid_for(ach)), dim3(256), 0, s, djobs, count, bits, (uint32_t)ach);
    hipLaunchKernelGGL(k_mm_unpack_bt, dim3(grid_for(bch)), dim3(256), 0, s, djobs, count, bits, (uint32_t)bch);
    static uint32_t capb = 0;
    if (!capb) capb = resident_blocks(k_mm_gemm, 256, 0);
    hipLaunchKernelGGL(k_mm_gemm, dim3((uint32_t)std::min<uint64_t>(tiles, capb)), dim3(256), 0, s, djobs, count,
                       (uint32_t)tiles);
    if (need_pack)
        hipLaunchKernelGGL(k_mm_pack_rows, dim3(grid_for(pwords)), dim3(256), 0, s, djobs, count, dword0,
                           (uint32_t)pwords);
    return check_launch();
}

extern "C" int am_matmul_packed(uint32_t bits, const uint64_t* A, const uint64_t* B, uint32_t M, uint32_t N, uint32_t K,
                                uint32_t out_bits, uint64_t* C_packed, int32_t* C_i32, void* ws, uint64_t ws_bytes,
                                void* stream)
{
    am_matmul_desc d{A, B, C_packed, C_i32, M, N, K, out_bits};
    return am_matmul_packed_batched(bits, &d, 1, ws, ws_bytes, stream);
}
