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
//      [N][Kp] (Kp = K rounded up to 128, zero-padded): 7-bit values are
//      non-negative in int8, so the signed int8 MFMA is exact;
//   2. GEMM: 256x256 block tiles, K-steps of 128 staged through LDS with
//      16-byte loads and double buffering; each of 4 waves owns a 128x128
//      quadrant = 8x8 MFMA tiles of 16x16 (i32 accumulators);
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

constexpr uint32_t MM_BM = 256, MM_BN = 256, MM_BK = 128, MM_GROUP = 8;
constexpr int MM_RSRC_FLAGS = 0x00020000;        // buffer descriptor word 3: 32-bit data format
constexpr uint32_t MM_STAGE = (MM_BM + MM_BN) * MM_BK; // bytes of one K-step stage (unpadded rows)
constexpr uint32_t MM_C_STRIDE = 260;            // int32 per staged C row (4*stride = 16 mod 32 banks)
constexpr uint32_t MM_LDS = 128 * MM_C_STRIDE * 4 > 2 * MM_STAGE ? 128 * MM_C_STRIDE * 4 : 2 * MM_STAGE;
constexpr unsigned MM_WAIT_VM0 = 0x0F70;         // s_waitcnt vmcnt(0), other counters unconstrained

struct MatJob {
    const uint64_t* A; // packed M x K
    const uint64_t* B; // packed K x N
    uint64_t* Cp;      // packed C (row-padded slab at out_bits) or null
    int32_t* Ci;       // int32 C [M][N] or null
    int8_t* A8;        // [M][Kp]
    int8_t* B8;        // [N][Kp]
    uint32_t M, N, K, Kp;
    uint32_t out_bits;
    uint32_t tiles_n;  // ceil(N / 256)
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

// Tile id -> (job, m0, n0). Tiles of one job are taken in groups of
// MM_GROUP tile rows (column-major inside a group), and the persistent
// grid's blocks are renumbered so that the blocks of one XCD (block b runs
// on XCD b % 8) take consecutive tiles: at any time one XCD works on a few
// A row-tiles and B column-tiles, whose K-steps its L2 then serves once.
__device__ inline uint32_t mm_logical_tile(uint32_t tile, uint32_t grid, uint32_t total)
{
    // a permutation inside each full round of `grid` tiles; a partial last
    // round keeps the identity so every id stays below `total`
    const uint32_t round = tile / grid, b = tile % grid;
    if (grid % 8 != 0 || (uint64_t)(round + 1) * grid > total) return tile;
    return round * grid + (b % 8) * (grid / 8) + b / 8;
}

__device__ inline uint32_t rfl(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
template <typename T> __device__ inline T* rfl_ptr(T* p)
{
    const uint64_t v = (uint64_t)p;
    return (T*)(((uint64_t)rfl((uint32_t)(v >> 32)) << 32) | rfl((uint32_t)v));
}

// The job's fields as wave-uniform (scalar) values: the job list is read with
// vector loads, and buffer descriptors must be built from scalars.
__device__ inline MatJob uniform_job(const MatJob& j)
{
    MatJob u;
    u.A = rfl_ptr(j.A);
    u.B = rfl_ptr(j.B);
    u.Cp = rfl_ptr(j.Cp);
    u.Ci = rfl_ptr(j.Ci);
    u.A8 = rfl_ptr(j.A8);
    u.B8 = rfl_ptr(j.B8);
    u.M = rfl(j.M);
    u.N = rfl(j.N);
    u.K = rfl(j.K);
    u.Kp = rfl(j.Kp);
    u.out_bits = rfl(j.out_bits);
    u.tiles_n = rfl(j.tiles_n);
    u.tile0 = rfl(j.tile0);
    u.achunk0 = rfl(j.achunk0);
    u.bchunk0 = rfl(j.bchunk0);
    return u;
}

// GEMM over the batch's tile list; one 256x256 output tile per iteration.
// 4 waves in a 2x2 arrangement, each owning 128x128 = 8x8 MFMA tiles of
// 16x16 (256 i32 accumulators per lane). K-steps of 128 bytes are staged in
// LDS (double-buffered, 144 KB, one block per CU); the next step's global
// loads are in flight in registers while the current step computes.
__global__ __launch_bounds__(256, 1) void k_mm_gemm(const MatJob* jobs, uint32_t njobs, uint32_t total_tiles)
{
    // two K-step stages [A rows | B rows]; the epilogue reuses it as 128 int32 rows
    __shared__ __attribute__((aligned(1024))) int8_t smem[MM_LDS];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t wm = wave >> 1, wn = wave & 1;
    const uint32_t lr = lane & 15, lg = lane >> 4;
    auto sA = [&](uint32_t buf) { return smem + buf * MM_STAGE; };
    auto sB = [&](uint32_t buf) { return smem + buf * MM_STAGE + MM_BM * MM_BK; };

    for (uint32_t tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
        const uint32_t lt = mm_logical_tile(tile, gridDim.x, total_tiles);
        const MatJob J = uniform_job(jobs[find_job(jobs, njobs, lt, 2)]);
        const uint32_t local = lt - J.tile0;
        const uint32_t tiles_m = (J.M + MM_BM - 1) / MM_BM;
        const uint32_t per_group = MM_GROUP * J.tiles_n, grp = local / per_group;
        const uint32_t first = grp * MM_GROUP, gsz = min(tiles_m - first, MM_GROUP), r = local % per_group;
        const uint32_t m0 = (first + r % gsz) * MM_BM, n0 = (r / gsz) * MM_BN;
        const uint32_t steps = J.Kp / MM_BK;

        // K-steps land in LDS straight from global memory (buffer loads with
        // the LDS flag, 16 bytes per lane, no registers): each instruction
        // fills 1 KB = 8 rows of 128 bytes, lane l the 16-byte slot l%8 of row
        // l/8, and slot s of row R holds the row's chunk s ^ (R%8) (an XOR
        // swizzle: the fragment reads of 8 consecutive rows then hit 8
        // different slots). Wave w loads rows 64w..64w+63 of A and of B.
        // Loads bounded by the tile's valid rows read zeros past M (or N).
        const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(J.A8 + (uint64_t)m0 * J.Kp), (short)0, (int)(min(J.M - m0, MM_BM) * J.Kp), MM_RSRC_FLAGS);
        const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(J.B8 + (uint64_t)n0 * J.Kp), (short)0, (int)(min(J.N - n0, MM_BN) * J.Kp), MM_RSRC_FLAGS);
        const uint32_t lrow = lane >> 3;
        const uint32_t vo = (wave * 64 + lrow) * J.Kp + ((lane & 7) ^ lrow) * 16, rs = 8 * J.Kp;
        auto gload = [&](uint32_t ks, uint32_t buf) {
            const uint32_t ko = vo + ks * MM_BK;
#pragma unroll
            for (uint32_t h = 0; h < 8; h++) {
                const uint32_t lo = (wave * 8 + h) * 1024;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rA, (__attribute__((address_space(3))) void*)(sA(buf) + lo),
                                                         16, ko + h * rs, 0, 0, 0);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rB, (__attribute__((address_space(3))) void*)(sB(buf) + lo),
                                                         16, ko + h * rs, 0, 0, 0);
            }
        };
        // fragment reads: row R = 128*wm (or wn) + 16*i + lr, chunk 4*sub + lg
        const uint32_t fo0 = lr * MM_BK + ((lg ^ (lr & 7)) * 16), fo1 = lr * MM_BK + (((4 + lg) ^ (lr & 7)) * 16);

        v4i acc[8][8];
#pragma unroll
        for (uint32_t i = 0; i < 8; i++)
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) acc[i][j] = v4i{0, 0, 0, 0};

        __syncthreads(); // previous tile's epilogue is done with the LDS
        gload(0, 0);
        __builtin_amdgcn_s_waitcnt(MM_WAIT_VM0);
        __syncthreads();
        for (uint32_t ks = 0; ks < steps; ks++) {
            const uint32_t buf = ks & 1;
            if (ks + 1 < steps) gload(ks + 1, buf ^ 1); // read in step ks-1, released by its barrier
#pragma unroll
            for (uint32_t sub = 0; sub < MM_BK / 64; sub++) {
                const uint32_t fo = sub ? fo1 : fo0;
                v4i fa[8], fb[8];
#pragma unroll
                for (uint32_t i = 0; i < 8; i++) {
                    fa[i] = *(const v4i*)(sA(buf) + (wm * 128 + i * 16) * MM_BK + fo);
                    fb[i] = *(const v4i*)(sB(buf) + (wn * 128 + i * 16) * MM_BK + fo);
                }
#pragma unroll
                for (uint32_t i = 0; i < 8; i++)
#pragma unroll
                    for (uint32_t j = 0; j < 8; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
            }
            __builtin_amdgcn_s_waitcnt(MM_WAIT_VM0); // this wave's loads of the next step have landed
            __syncthreads();                         // ... and every wave's, and step ks is read
        }

        // ---- epilogue through LDS, one 128-row half of the tile at a time:
        // the two waves owning the half write their accumulators (lane holds
        // C[row0 + 4*lg + r][col0 + lr], r < 4) as int32 rows, then all
        // threads store whole rows: int32 C in 16-byte pieces and/or packed
        // words (when floor(64/out_bits) divides 16, words never straddle
        // tiles; other widths are packed later from the int32 copy)
        const uint32_t ob = J.out_bits;
        const uint32_t eps_o = 64 / ob;
        const bool fused = J.Cp && (16 % eps_o == 0);
        const uint32_t rw_o = (J.N + eps_o - 1) / eps_o;
        const uint64_t omask = lowmask(ob);
        int32_t* sc = (int32_t*)smem;
        __syncthreads(); // every wave is done reading the last K-step
#pragma unroll
        for (uint32_t half = 0; half < 2; half++) {
            if (wm == half) {
#pragma unroll
                for (uint32_t i = 0; i < 8; i++)
#pragma unroll
                    for (uint32_t j = 0; j < 8; j++)
#pragma unroll
                        for (uint32_t r = 0; r < 4; r++)
                            sc[(i * 16 + 4 * lg + r) * MM_C_STRIDE + wn * 128 + j * 16 + lr] = acc[i][j][r];
            }
            __syncthreads();
            const uint32_t rbase = m0 + half * 128;
            if (J.Ci) {
                for (uint32_t t = tid; t < 128 * 64; t += 256) {
                    const uint32_t row = t >> 6, c = n0 + (t & 63) * 4, gr = rbase + row;
                    if (gr >= J.M || c >= J.N) continue;
                    const int32_t* src = sc + row * MM_C_STRIDE + (t & 63) * 4;
                    int32_t* dst = J.Ci + (uint64_t)gr * J.N + c;
                    if (c + 4 <= J.N && ((uintptr_t)dst & 15) == 0) {
                        *(v4i*)dst = *(const v4i*)src;
                    } else {
                        for (uint32_t q = 0; q < 4 && c + q < J.N; q++) dst[q] = src[q];
                    }
                }
            }
            if (fused) {
                const uint32_t lw = 31 - __builtin_clz(256 / eps_o); // log2 words per tile row
                for (uint32_t t = tid; t < (128u << lw); t += 256) {
                    const uint32_t row = t >> lw, w = t & ((1u << lw) - 1), gr = rbase + row;
                    const uint32_t c0 = n0 + w * eps_o;
                    if (gr >= J.M || c0 >= J.N) continue;
                    const int32_t* src = sc + row * MM_C_STRIDE + w * eps_o;
                    uint64_t word = 0;
                    for (uint32_t q = 0; q < eps_o && c0 + q < J.N; q++)
                        word |= ((uint64_t)(uint32_t)src[q] & omask) << (q * ob);
                    J.Cp[(uint64_t)gr * rw_o + c0 / eps_o] = word;
                }
            }
            __syncthreads(); // the half's rows are stored before the next half lands
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
    const uint64_t Kp = (K + MM_BK - 1) / MM_BK * MM_BK;
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
        J.Kp = (d.K + MM_BK - 1) / MM_BK * MM_BK;
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
