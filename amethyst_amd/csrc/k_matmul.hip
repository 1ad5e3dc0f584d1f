// k_matmul.hip -- exact integer matmul of bit-packed <= 7-bit matrices on
// the int8 matrix cores (v_mfma_i32_32x32x32_i8, exact i32 accumulation).
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
//   2. GEMM: 256x256 block tiles, K-steps of 128 loaded straight into LDS
//      one step ahead; each of 8 waves owns 64x128 = 2x4 MFMA tiles of 32x32
//      (i32 accumulators).
//      Both operands use the same k-to-slot mapping per lane, so the sum over
//      k is exact whatever order the matrix core visits the slots in;
//   3. epilogue: C as row-major int32 and/or re-packed as a row-padded slab
//      matrix at `out_bits` (fused when floor(64/out_bits) divides 16,
//      otherwise through the int32 copy and a row packer).
// The batched entry point runs every phase for all matrices of a batch in
// one persistent launch per phase (work lists built on the host).
#include <algorithm>
#include <cstring>
#include <vector>

#include <new>

#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr uint32_t MM_BM = 256, MM_BN = 256, MM_BK = 128, MM_GROUP = 8;
constexpr uint32_t MM_KPAD = 128; // K padding of the int8 copies (whole K-steps and B-unpack tiles)
constexpr int MM_RSRC_FLAGS = 0x00020000;        // buffer descriptor word 3: 32-bit data format
constexpr uint32_t MM_STAGE = (MM_BM + MM_BN) * MM_BK; // bytes of one K-step stage (unpadded rows)
constexpr uint32_t MM_C_STRIDE = 260;            // int32 per staged C row (4*stride = 16 mod 32 banks)
constexpr uint32_t MM_SLOT = 64 * MM_C_STRIDE * 4; // one stage; in the epilogue 64 rows of C
static_assert(MM_SLOT >= MM_STAGE, "a stage fits its slot");
// s_waitcnt vmcnt(0) with the other counters unconstrained (vmcnt bits 3:0 and 15:14)
constexpr unsigned MM_WAIT_VM0 = 0x0F70;
constexpr unsigned MM_WAIT_VM0_LGKM0 = 0x0070; // vmcnt(0) lgkmcnt(0) (bits 11:8), expcnt unconstrained
constexpr uint32_t MM_THREADS = 512; // 8 waves: 4 x 64 rows by 2 x 128 columns of the tile

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
};

// One block of an unpack launch: its job and its first work unit there
// (host-built table, so blocks never search the job list).
struct MatSeg {
    uint32_t job, first;
};

constexpr uint32_t MU_A_ITEMS = 1024; // A-unpack items (one row, 16 k) per block
constexpr uint32_t MU_KT = 128;       // B-unpack tile: k rows
constexpr uint32_t MU_NT = 576;       // B-unpack tile: at most this many columns (whole words)

__device__ inline uint32_t find_job(const MatJob* jobs, uint32_t njobs, uint32_t tile)
{
    // last job whose first tile <= tile (jobs sorted by first tile)
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].tile0 <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// A -> int8 rows [M][Kp]: one item = one row and 16 consecutive k (one
// 16-byte store). The 16 fields span at most 3 packed words (>= 9 fields per
// word); neighbouring lanes take neighbouring items of a row, so the word
// loads of a wave cover a contiguous stretch of the row. A thread issues the
// loads of all its items before unpacking any (the int8 stores may alias the
// packed input as far as the compiler knows, so it would not hoist a later
// item's loads above an earlier item's store by itself).
template <uint32_t BITS>
__global__ __launch_bounds__(256) void k_mm_unpack_a(const MatJob* jobs, const MatSeg* segs)
{
    constexpr uint32_t EPS = 64 / BITS, Q = MU_A_ITEMS / 256;
    constexpr uint64_t MASK = (1ull << BITS) - 1;
    const MatSeg sg = segs[blockIdx.x];
    const MatJob& J = jobs[sg.job];
    const uint32_t cpr = J.Kp / 16, rw = (J.K + EPS - 1) / EPS;
    const uint32_t items = J.M * cpr; // < 2^32 (host-checked)
    uint64_t wv[Q][3];
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
        const uint32_t it = sg.first + q * 256 + threadIdx.x;
        wv[q][0] = wv[q][1] = wv[q][2] = 0;
        if (it >= items) continue;
        const uint32_t r = it / cpr, k0 = (it - r * cpr) * 16;
        if (k0 >= J.K) continue;
        const uint64_t* row = J.A + (uint64_t)r * rw;
        const uint32_t w = k0 / EPS;
        wv[q][0] = row[w];
        if (w + 1 < rw) wv[q][1] = row[w + 1];
        if (w + 2 < rw) wv[q][2] = row[w + 2];
    }
#pragma unroll
    for (uint32_t q = 0; q < Q; q++) {
        const uint32_t it = sg.first + q * 256 + threadIdx.x;
        if (it >= items) break;
        const uint32_t r = it / cpr, k0 = (it - r * cpr) * 16;
        uint32_t v[4] = {0, 0, 0, 0};
        if (k0 < J.K) {
            // the three words' payloads (dead top bits dropped) as one
            // 192-bit stream s0:s1:s2, shifted down to the item's first field:
            // field s is then bits [BITS*s, BITS*s + BITS) of t0:t1
            constexpr uint32_t L = EPS * BITS; // payload bits per word
            const uint32_t w = k0 / EPS, off = k0 - w * EPS;
            uint64_t s0 = wv[q][0], s1 = wv[q][1], s2 = wv[q][2];
            if constexpr (L < 64) {
                constexpr uint64_t PM = (1ull << L) - 1;
                s0 = (s0 & PM) | (s1 << L);
                s1 = ((s1 & PM) >> (64 - L)) | (s2 << (2 * L - 64));
                s2 = (s2 & PM) >> (128 - 2 * L);
            }
            const uint32_t sh = off * BITS; // < 64
            uint64_t t0 = (s0 >> sh) | ((s1 << 1) << (63 - sh));
            uint64_t t1 = (s1 >> sh) | ((s2 << 1) << (63 - sh));
            // fields past K (end of the row) read as 0
            const uint32_t nb = (J.K - k0 < 16 ? J.K - k0 : 16) * BITS; // valid bits, <= 112
            if (nb < 64) t0 &= (1ull << nb) - 1;
            t1 &= nb <= 64 ? 0ull : (1ull << (nb - 64)) - 1;
#pragma unroll
            for (uint32_t f = 0; f < 16; f++) {
                const uint32_t b = f * BITS;
                const uint64_t x = b + BITS <= 64 ? (t0 >> b) : (b >= 64 ? (t1 >> (b - 64)) : ((t0 >> b) | (t1 << (64 - b))));
                v[f >> 2] |= ((uint32_t)x & (uint32_t)MASK) << (8 * (f & 3));
            }
        }
        *(uint4*)(J.A8 + (uint64_t)r * J.Kp + k0) = make_uint4(v[0], v[1], v[2], v[3]);
    }
}

// B (K x N) -> int8 columns [N][Kp], through an LDS transpose. A block takes
// MU_KT rows of k and WN whole words of every row (WN*EPS <= MU_NT columns):
// lanes on neighbouring words load 4 rows each (coalesced row segments) and
// write, per field, the 4 k-bytes of that column as one LDS word; then 8
// lanes per column store its MU_KT contiguous bytes.
template <uint32_t BITS>
__global__ __launch_bounds__(256) void k_mm_unpack_bt(const MatJob* jobs, const MatSeg* segs)
{
    constexpr uint32_t EPS = 64 / BITS, WN = MU_NT / EPS, NT = WN * EPS;
    constexpr uint64_t MASK = (1ull << BITS) - 1;
    constexpr uint32_t LS = MU_KT / 4 + 1; // LDS words per column (odd: columns spread over banks)
    __shared__ uint32_t tr[NT * LS];
    const MatSeg sg = segs[blockIdx.x];
    const MatJob& J = jobs[sg.job];
    const uint32_t ntn = (J.N + NT - 1) / NT;
    const uint32_t kt = sg.first / ntn, nt = sg.first - kt * ntn;
    const uint32_t k0 = kt * MU_KT, wb = nt * WN, rw = (J.N + EPS - 1) / EPS;
    // every load of the thread's tasks first, then the transposes into LDS
    constexpr uint32_t NTASK = WN * (MU_KT / 4), TPT = (NTASK + 255) / 256;
    uint64_t x[TPT][4];
#pragma unroll
    for (uint32_t p = 0; p < TPT; p++) {
        const uint32_t task = p * 256 + threadIdx.x;
        const uint32_t q = task / WN, w = task - q * WN, gw = wb + w;
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            const uint32_t k = k0 + 4 * q + i;
            x[p][i] = (task < NTASK && gw < rw && k < J.K) ? J.B[(uint64_t)k * rw + gw] : 0;
        }
    }
#pragma unroll
    for (uint32_t p = 0; p < TPT; p++) {
        const uint32_t task = p * 256 + threadIdx.x;
        if (task >= NTASK) break;
        const uint32_t q = task / WN, w = task - q * WN;
#pragma unroll
        for (uint32_t f = 0; f < EPS; f++) {
            uint32_t d = 0;
#pragma unroll
            for (uint32_t i = 0; i < 4; i++) d |= (uint32_t)((x[p][i] >> (f * BITS)) & MASK) << (8 * i);
            tr[(w * EPS + f) * LS + q] = d;
        }
    }
    __syncthreads();
    for (uint32_t task = threadIdx.x; task < NT * (MU_KT / 16); task += 256) {
        const uint32_t nl = task / (MU_KT / 16), pc = task % (MU_KT / 16);
        const uint32_t n = wb * EPS + nl;
        if (n >= J.N) continue;
        const uint32_t* src = tr + nl * LS + pc * 4;
        *(uint4*)(J.B8 + (uint64_t)n * J.Kp + k0 + pc * 16) = make_uint4(src[0], src[1], src[2], src[3]);
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
    return u;
}

// GEMM over the batch's tile list; one 256x256 output tile per iteration.
// 8 waves, 4 bands of 64 rows by 2 halves of 128 columns, each owning
// 64x128 = 2x4 MFMA tiles of 32x32 (v_mfma_i32_32x32x32_i8; 16 i32
// accumulators per tile and lane, 128 in all); one block per CU, 2 waves per
// SIMD.
//
// K-steps of 128 bytes land in LDS straight from global memory (buffer loads
// with the LDS flag, 16 bytes per lane, no registers), one step ahead, in two
// stage arrays used alternately. A stage holds 256 rows of A then 256 of B,
// 128 bytes each; chunk c (16 bytes) of row R sits in slot c ^ (R % 8) ^ (R / 8 % 2)
// (an XOR swizzle: any 16 consecutive rows, 2 KB, put their chunk c in 16
// distinct 16-byte bank groups, so the fragment reads meet no bank twice).
// One load instruction fills 1 KB = 8 rows: lane l takes row l/8, slot l%8.
// Wave w loads rows 32w..32w+31 of A and of B. Loads bounded by the tile's
// valid rows read zeros past M (or N).
//
// The stages are distinct LDS objects and the K loop is unrolled by 2, so the
// compiler knows the fragment reads of one stage do not alias the loads in
// flight into the other and inserts no wait for them.
__global__ __launch_bounds__(MM_THREADS, 1) void k_mm_gemm(const MatJob* jobs, uint32_t njobs, uint32_t total_tiles)
{
    // two K-step stages [A rows | B rows]; the epilogue reuses them as 2 x 64
    // rows of int32 C
    __shared__ __attribute__((aligned(1024))) int8_t st0[MM_SLOT], st1[MM_SLOT];
    // wave index as a scalar: the LDS destinations of the loads are then
    // scalar arithmetic (M0) rather than per-lane values
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = rfl(tid >> 6);
    const uint32_t wm = wave >> 1, wn = wave & 1; // wm: 64-row band (0..3), wn: 128-column half
    const uint32_t l32 = lane & 31, lh = lane >> 5;

    for (uint32_t tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
        const uint32_t lt = mm_logical_tile(tile, gridDim.x, total_tiles);
        const MatJob J = uniform_job(jobs[find_job(jobs, njobs, lt)]);
        const uint32_t local = lt - J.tile0;
        const uint32_t tiles_m = (J.M + MM_BM - 1) / MM_BM;
        const uint32_t per_group = MM_GROUP * J.tiles_n, grp = local / per_group;
        const uint32_t first = grp * MM_GROUP, gsz = min(tiles_m - first, MM_GROUP), r = local % per_group;
        const uint32_t m0 = (first + r % gsz) * MM_BM, n0 = (r / gsz) * MM_BN;
        const uint32_t steps = J.Kp / MM_BK;

        const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(J.A8 + (uint64_t)m0 * J.Kp), (short)0, (int)(min(J.M - m0, MM_BM) * J.Kp), MM_RSRC_FLAGS);
        const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(J.B8 + (uint64_t)n0 * J.Kp), (short)0, (int)(min(J.N - n0, MM_BN) * J.Kp), MM_RSRC_FLAGS);
        const uint32_t lrow = lane >> 3;
        const uint32_t vo = (wave * 32 + lrow) * J.Kp + ((lane & 7) ^ lrow) * 16, rs = 8 * J.Kp;
        auto gload = [&](uint32_t ks, int8_t* st) {
#pragma unroll
            for (uint32_t h = 0; h < 4; h++) {
                const uint32_t ko = (vo ^ ((h & 1) * 16)) + ks * MM_BK; // rows 8h..: slot ^ 1 when h is odd
                const uint32_t lo = (wave * 4 + h) * 1024;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rA, (__attribute__((address_space(3))) void*)(st + lo), 16,
                                                         ko + h * rs, 0, 0, 0);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rB, (__attribute__((address_space(3))) void*)(st + MM_BM * MM_BK + lo), 16, ko + h * rs, 0, 0, 0);
            }
        };
        // fragment reads: row R = 128*wm (or wn) + 32*i + l32, k chunk 2*sub + lh
        // (k = 32*sub .. 32*sub+31 of the step; both operands map k to lanes alike)
        const uint32_t fbase = l32 * MM_BK, fx = (l32 ^ ((l32 >> 3) & 1)) & 7;

        v16i acc[2][4];
#pragma unroll
        for (uint32_t i = 0; i < 2; i++)
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) acc[i][j] = v16i{};

        // two sets of fragment registers: the reads of sub-step s+1 go out
        // between the matrix ops of sub-step s, so the matrix core is not left
        // waiting for LDS at the start of each sub-step
        v4i fa[2][2], fb[2][4];
        auto fread = [&](const int8_t* st, uint32_t sub, uint32_t b) {
            const uint32_t fo = fbase + (((2 * sub + lh) ^ fx) * 16);
#pragma unroll
            for (uint32_t i = 0; i < 2; i++) fa[b][i] = *(const v4i*)(st + (wm * 64 + i * 32) * MM_BK + fo);
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) fb[b][j] = *(const v4i*)(st + MM_BM * MM_BK + (wn * 128 + j * 32) * MM_BK + fo);
        };
        auto mfmas = [&](uint32_t b) {
#pragma unroll
            for (uint32_t i = 0; i < 2; i++)
#pragma unroll
                for (uint32_t j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[b][i], fb[b][j], acc[i][j], 0, 0, 0);
        };
        // schedule of one sub-step: matrix op, LDS read, ... (6 reads), then
        // the last two matrix ops; the region ends at the sched_barrier
        auto interleave = [&]() {
#pragma unroll
            for (int k = 0; k < 6; k++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); // one MFMA
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0); // one LDS read
            }
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
            __builtin_amdgcn_sched_barrier(0);
        };

        // one K-step from `cur` (sub-step 0's fragments already read); the
        // other stage `nxt` holds or receives step ks+1, loaded since the
        // previous step's barrier. Sub-steps 0..2 read the next sub-step's
        // fragments; then the wave waits for its loads of step ks+1 and its
        // reads of `cur`, meets the other waves (after which every wave is
        // done with `cur` and step ks+1 is in LDS), refills `cur` with step
        // ks+2 and reads step ks+1's first fragments during sub-step 3. A bare
        // s_barrier (no fences; the waits before it are the ones needed); the
        // empty asm keeps the compiler from moving LDS accesses across it.
        auto step = [&](uint32_t ks, int8_t* cur, const int8_t* nxt) {
#pragma unroll
            for (uint32_t sub = 0; sub < 3; sub++) {
                fread(cur, sub + 1, (sub + 1) & 1);
                mfmas(sub & 1);
                interleave();
            }
            __builtin_amdgcn_s_waitcnt(MM_WAIT_VM0_LGKM0);
            asm volatile("" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            if (ks + 2 < steps) gload(ks + 2, cur);
            fread(nxt, 0, 0); // (after the last step: stale data, never used)
            mfmas(1);
            interleave();
        };

        __builtin_amdgcn_s_waitcnt(MM_WAIT_VM0); // the previous tile's stores are out of the count
        __syncthreads();                         // ... and its epilogue is done with the LDS
        gload(0, st0);
        __builtin_amdgcn_s_waitcnt(MM_WAIT_VM0);
        __syncthreads();
        if (1 < steps) gload(1, st1);
        fread(st0, 0, 0);
        for (uint32_t ks = 0; ks < steps; ks += 2) {
            step(ks, st0, st1);
            if (ks + 1 < steps) step(ks + 1, st1, st0);
        }
        // (the fragment reads after the last barrier fetch data nobody uses;
        // the epilogue's LDS writes may overtake them)

        // ---- epilogue through LDS, one 128-row half of the tile at a time:
        // the four waves owning the half write their accumulators (lane holds
        // C[row0 + 8*(q/4) + 4*(lane/32) + q%4][col0 + lane%32], q < 16) as
        // int32 rows, then all threads store whole rows: int32 C in 16-byte
        // pieces and/or packed words (when floor(64/out_bits) divides 16,
        // words never straddle tiles; other widths are packed later from the
        // int32 copy)
        const uint32_t ob = J.out_bits;
        const uint32_t eps_o = 64 / ob;
        const bool fused = J.Cp && (16 % eps_o == 0);
        const uint32_t rw_o = (J.N + eps_o - 1) / eps_o;
        const uint64_t omask = lowmask(ob);
        auto slot = [&](uint32_t q) { return (int32_t*)(q == 0 ? st0 : st1); };
#pragma unroll
        for (uint32_t half = 0; half < 2; half++) {
            if ((wm >> 1) == half) {
#pragma unroll
                for (uint32_t i = 0; i < 2; i++)
#pragma unroll
                    for (uint32_t j = 0; j < 4; j++)
#pragma unroll
                        for (uint32_t q = 0; q < 16; q++)
                            slot(wm & 1)[(i * 32 + 8 * (q >> 2) + 4 * lh + (q & 3)) * MM_C_STRIDE + wn * 128 + j * 32 +
                                         l32] = acc[i][j][q];
            }
            __syncthreads();
            const uint32_t rbase = m0 + half * 128;
            if (J.Ci) {
                for (uint32_t t = tid; t < 128 * 64; t += MM_THREADS) {
                    const uint32_t row = t >> 6, c = n0 + (t & 63) * 4, gr = rbase + row;
                    if (gr >= J.M || c >= J.N) continue;
                    const int32_t* src = slot(row >> 6) + (row & 63) * MM_C_STRIDE + (t & 63) * 4;
                    int32_t* dst = J.Ci + (uint64_t)gr * J.N + c;
                    if (c + 4 <= J.N && ((uintptr_t)dst & 15) == 0) {
                        st_nt((v4i*)dst, *(const v4i*)src);
                    } else {
                        for (uint32_t q = 0; q < 4 && c + q < J.N; q++) st_nt(dst + q, src[q]);
                    }
                }
            }
            if (fused) {
                const uint32_t lw = 31 - __builtin_clz(256 / eps_o); // log2 words per tile row
                for (uint32_t t = tid; t < (128u << lw); t += MM_THREADS) {
                    const uint32_t row = t >> lw, w = t & ((1u << lw) - 1), gr = rbase + row;
                    const uint32_t c0 = n0 + w * eps_o;
                    if (gr >= J.M || c0 >= J.N) continue;
                    const int32_t* src = slot(row >> 6) + (row & 63) * MM_C_STRIDE + w * eps_o;
                    uint64_t word = 0;
                    for (uint32_t q = 0; q < eps_o && c0 + q < J.N; q++)
                        word |= ((uint64_t)(uint32_t)src[q] & omask) << (q * ob);
                    st_nt(J.Cp + (uint64_t)gr * rw_o + c0 / eps_o, word);
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
    const uint64_t Kp = (K + MM_KPAD - 1) / MM_KPAD * MM_KPAD;
    return (M * Kp + 255) / 256 * 256 + (N * Kp + 255) / 256 * 256 + (M * N * 4 + 255) / 256 * 256;
}

// Unpack blocks of one product: A items / MU_A_ITEMS, plus B tiles (at least
// 512 columns per tile for every width, so ceil(N/512) bounds the column tiles).
static uint64_t mm_segs_bound(uint64_t M, uint64_t N, uint64_t K)
{
    const uint64_t Kp = (K + MM_KPAD - 1) / MM_KPAD * MM_KPAD;
    return (M * (Kp / 16) + MU_A_ITEMS - 1) / MU_A_ITEMS + (N + 511) / 512 * (Kp / MU_KT);
}

// Header: job list, packer word offsets, unpack block table.
static uint64_t mm_header_bytes(const uint32_t* M, const uint32_t* N, const uint32_t* K, uint32_t count)
{
    uint64_t segs = 0;
    for (uint32_t i = 0; i < count; i++) segs += mm_segs_bound(M[i], N[i], K[i]);
    return (4096 + (uint64_t)count * (sizeof(MatJob) + 4) + 8 + segs * sizeof(MatSeg) + 255) / 256 * 256;
}

extern "C" uint64_t am_matmul_workspace_bytes(const uint32_t* M, const uint32_t* N, const uint32_t* K, uint32_t count)
{
    uint64_t b = mm_header_bytes(M, N, K, count);
    for (uint32_t i = 0; i < count; i++) b += mm_ws_one(M[i], N[i], K[i]);
    return b;
}

template <uint32_t BITS>
static void mm_launch_unpack(const MatJob* djobs, const MatSeg* dsegs, uint32_t na, uint32_t nb, hipStream_t s)
{
    if (na) hipLaunchKernelGGL(k_mm_unpack_a<BITS>, dim3(na), dim3(256), 0, s, djobs, dsegs);
    if (nb) hipLaunchKernelGGL(k_mm_unpack_bt<BITS>, dim3(nb), dim3(256), 0, s, djobs, dsegs + na);
}

extern "C" uint32_t am_matmul_out_bits(uint32_t bits, uint32_t K)
{
    if (bits < 1 || bits > 7 || K == 0) return 1;
    const uint64_t vmax = (uint64_t)K * (((1ull << bits) - 1) * ((1ull << bits) - 1));
    return width_of_host(vmax);
}

// Launch parameters of a prepared batch; the header (job list, packer word
// offsets, unpack block table) lives in the workspace.
struct am_matmul_plan {
    uint32_t bits, count, na, nb;
    uint64_t tiles, pwords;
    uint32_t capb; // resident GEMM blocks (the persistent grid)
    bool need_pack;
    MatJob* djobs;
    uint32_t* dword0;
    MatSeg* dsegs;
};

// Validate, lay out the workspace and build the header on the host; the
// header is copied to the device with a stream-ordered copy from `hdr`
// (host memory the caller keeps until the copy has been issued: a copy from
// pageable memory returns once the source has been read).
static int mm_prepare(uint32_t bits, const am_matmul_desc* descs, uint32_t count, void* ws, uint64_t ws_bytes,
                      am_matmul_plan& P, hipStream_t s)
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
    uint8_t* cur = (uint8_t*)ws + mm_header_bytes(Ms.data(), Ns.data(), Ks.data(), count);
    std::vector<MatSeg> segs_a, segs_b;
    const uint32_t eps_i = 64 / bits, nt_b = MU_NT / eps_i * eps_i; // B-unpack tile columns
    uint64_t tiles = 0, pwords = 0;
    bool need_pack = false;
    // jobs in order of decreasing K (the work per GEMM tile): the persistent
    // grid takes tiles in job order, so the longest tiles start first and
    // the short ones fill the tail (results are per job, order-independent)
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return descs[x].K > descs[y].K; });
    for (uint32_t i = 0; i < count; i++) {
        const am_matmul_desc& d = descs[order[i]];
        MatJob& J = jobs[i];
        J.A = d.A;
        J.B = d.B;
        J.M = d.M;
        J.N = d.N;
        J.K = d.K;
        J.Kp = (d.K + MM_KPAD - 1) / MM_KPAD * MM_KPAD;
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
        word0[i] = (uint32_t)pwords;
        tiles += (uint64_t)((J.M + MM_BM - 1) / MM_BM) * J.tiles_n;
        const uint64_t items_a = (uint64_t)J.M * (J.Kp / 16);
        if (items_a >= (1ull << 32)) return AM_E_UNSUPPORTED_CONFIG;
        for (uint64_t f = 0; f < items_a; f += MU_A_ITEMS) segs_a.push_back(MatSeg{i, (uint32_t)f});
        const uint32_t tiles_b = (J.N + nt_b - 1) / nt_b * (J.Kp / MU_KT);
        for (uint32_t t = 0; t < tiles_b; t++) segs_b.push_back(MatSeg{i, t});
        if (J.Cp && (16 % eps_o) != 0) pwords += (uint64_t)J.M * ((J.N + eps_o - 1) / eps_o);
    }
    if (tiles >= (1ull << 32) || pwords >= (1ull << 32) || segs_a.size() + segs_b.size() >= (1ull << 31))
        return AM_E_UNSUPPORTED_CONFIG;
    P.na = (uint32_t)segs_a.size();
    P.nb = (uint32_t)segs_b.size();
    segs_a.insert(segs_a.end(), segs_b.begin(), segs_b.end());
    P.djobs = (MatJob*)((uint8_t*)ws + 4096);
    P.dword0 = (uint32_t*)((uint8_t*)P.djobs + (uint64_t)count * sizeof(MatJob));
    P.dsegs = (MatSeg*)(((uintptr_t)(P.dword0 + count) + 7) & ~(uintptr_t)7);
    const size_t o_w0 = (uint8_t*)P.dword0 - (uint8_t*)P.djobs, o_seg = (uint8_t*)P.dsegs - (uint8_t*)P.djobs;
    std::vector<uint8_t> hdr(o_seg + segs_a.size() * sizeof(MatSeg));
    memcpy(hdr.data(), jobs.data(), count * sizeof(MatJob));
    memcpy(hdr.data() + o_w0, word0.data(), count * 4);
    memcpy(hdr.data() + o_seg, segs_a.data(), segs_a.size() * sizeof(MatSeg));
    // (stream-ordered: earlier work on `s` that reads this workspace's
    // header finishes first; `hdr` may be released once the call returns)
    if (hipMemcpyAsync(P.djobs, hdr.data(), hdr.size(), hipMemcpyHostToDevice, s) != hipSuccess) return AM_E_HIP;
    P.capb = resident_blocks(k_mm_gemm, MM_THREADS, 0);
    P.bits = bits;
    P.count = count;
    P.tiles = tiles;
    P.pwords = pwords;
    P.need_pack = need_pack;
    return AM_OK;
}

static int mm_launch(const am_matmul_plan& P, hipStream_t s)
{
    switch (P.bits) {
        case 1: mm_launch_unpack<1>(P.djobs, P.dsegs, P.na, P.nb, s); break;
        case 2: mm_launch_unpack<2>(P.djobs, P.dsegs, P.na, P.nb, s); break;
        case 3: mm_launch_unpack<3>(P.djobs, P.dsegs, P.na, P.nb, s); break;
        case 4: mm_launch_unpack<4>(P.djobs, P.dsegs, P.na, P.nb, s); break;
        case 5: mm_launch_unpack<5>(P.djobs, P.dsegs, P.na, P.nb, s); break;
        case 6: mm_launch_unpack<6>(P.djobs, P.dsegs, P.na, P.nb, s); break;
        default: mm_launch_unpack<7>(P.djobs, P.dsegs, P.na, P.nb, s); break;
    }
    auto grid_for = [](uint64_t items) { return (uint32_t)std::min<uint64_t>((items + 255) / 256, 65535); };
    hipLaunchKernelGGL(k_mm_gemm, dim3((uint32_t)std::min<uint64_t>(P.tiles, P.capb)), dim3(MM_THREADS), 0, s, P.djobs,
                       P.count, (uint32_t)P.tiles);
    if (P.need_pack)
        hipLaunchKernelGGL(k_mm_pack_rows, dim3(grid_for(P.pwords)), dim3(256), 0, s, P.djobs, P.count, P.dword0,
                           (uint32_t)P.pwords);
    return check_launch();
}

extern "C" int am_matmul_packed_batched(uint32_t bits, const am_matmul_desc* descs, uint32_t count, void* ws,
                                        uint64_t ws_bytes, void* stream)
{
    am_matmul_plan P{};
    const int st = mm_prepare(bits, descs, count, ws, ws_bytes, P, (hipStream_t)stream);
    return st ? st : mm_launch(P, (hipStream_t)stream);
}

extern "C" int am_matmul_plan_create(uint32_t bits, const am_matmul_desc* descs, uint32_t count, void* ws,
                                     uint64_t ws_bytes, am_matmul_plan** plan, void* stream)
{
    if (!plan) return AM_E_INVALID;
    *plan = nullptr;
    am_matmul_plan* P = new (std::nothrow) am_matmul_plan{};
    if (!P) return AM_E_INVALID;
    const int st = mm_prepare(bits, descs, count, ws, ws_bytes, *P, (hipStream_t)stream);
    if (st) {
        delete P;
        return st;
    }
    *plan = P;
    return AM_OK;
}

extern "C" int am_matmul_plan_run(const am_matmul_plan* plan, void* stream)
{
    if (!plan) return AM_E_INVALID;
    return mm_launch(*plan, (hipStream_t)stream);
}

extern "C" void am_matmul_plan_destroy(am_matmul_plan* plan) { delete plan; }

extern "C" int am_matmul_packed(uint32_t bits, const uint64_t* A, const uint64_t* B, uint32_t M, uint32_t N, uint32_t K,
                                uint32_t out_bits, uint64_t* C_packed, int32_t* C_i32, void* ws, uint64_t ws_bytes,
                                void* stream)
{
    am_matmul_desc d{A, B, C_packed, C_i32, M, N, K, out_bits};
    return am_matmul_packed_batched(bits, &d, 1, ws, ws_bytes, stream);
}
