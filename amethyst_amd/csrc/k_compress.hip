// k_compress.hip -- single-pass filter + compressstore with decoupled look-back.
//
// Reference (src/kernels/compressstore_slabs.cuh:4-206,
// src/benchs/bench_compressstore_slabs.cu:38-56): three launches -- a
// per-1024-element popcount of the mask, cub::DeviceScan::ExclusiveSum over
// the counts, then a writer that serially walks 32 mask words per warp,
// re-reads an input word from global memory for every output element, and
// merges the slab words shared by two warps with atomicAnd/atomicOr.
//
// Here it is one launch. Each 256-thread block takes a 4096-element tile
// from a ticket counter (tiles start in ticket order, so look-back always
// makes progress), and:
//   1. decodes element base + 64j + lane per lane; the selection flag (mask
//      bit or the predicate evaluated in place) goes through one wave64
//      ballot, which is also the output mask word and, via v_mbcnt, each
//      lane's rank among the wave's selected elements;
//   2. publishes its aggregate and walks back over predecessors 64 at a
//      time (one descriptor per lane), combining (count, window) pairs with
//      an ordered butterfly; the window is the last 63 bits of the packed
//      output stream, so the tile that completes an output word can build it
//      from its predecessors' values without atomics;
//   3. writes its selected values: byte-aligned outputs (u8/u16/u32/u64
//      containers) are scattered directly at their rank; slab and bitstream
//      outputs are assembled one whole word per lane from the LDS-compacted
//      tile plus the carried window, so every word is written exactly once,
//      dead and padding bits zero, no read-modify-write.
// The descriptors live in a caller workspace; each launch clears the other
// parity's descriptors and the last block to finish resets the counters, so
// repeated launches (and graph replays) need no memset.
#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t CS_WAVES = 4;
constexpr uint32_t CS_J = 16;
constexpr uint32_t CS_TILE = CS_WAVES * CS_J * 64; // 4096

struct CompressArgs {
    Col in;
    const uint64_t* mask_in;
    uint64_t pc;
    uint32_t pop;
    uint32_t T; // tiles in this launch
    Col out;
    uint64_t cap; // output capacity in elements
    uint64_t* mask_out;
    uint64_t* stats;
    WsHeader* hdr;
    Lookback lb;
};

__device__ inline void store_container(const Col& o, uint64_t r, uint64_t v)
{
    switch (o.cbytes) {
        case 1: ((uint8_t*)o.p)[r] = (uint8_t)v; break;
        case 2: ((uint16_t*)o.p)[r] = (uint16_t)v; break;
        case 4: ((uint32_t*)o.p)[r] = (uint32_t)v; break;
        default: o.p[r] = v;
    }
}

template <int IK, int OK, bool PRED>
__global__ __launch_bounds__(256) void k_compress(CompressArgs a)
{
    __shared__ uint32_t s_tile, s_par, s_lastT;
    __shared__ uint32_t s_wcnt[CS_WAVES];
    __shared__ uint64_t s_P, s_carry;
    __shared__ uint64_t s_vals[OK == K_STRIDED ? 1 : CS_TILE];

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t bo = a.out.bits;
    const uint64_t om = a.out.mask;

    if (tid == 0) {
        s_tile = add32(&a.hdr->ticket, 1);
        uint32_t ep = ld32(&a.hdr->epoch);
        s_par = ep & 1;
        s_lastT = ld32(&a.hdr->lastT[(ep & 1) ^ 1]);
    }
    __syncthreads();
    const uint32_t t = s_tile, par = s_par;
    const uint64_t cap = a.lb.cap;
    uint64_t* st = a.lb.status + par * cap;
    uint64_t* tA = a.lb.tailA + par * cap;
    uint64_t* tP = a.lb.tailP + par * cap;
    {
        // clear the other parity's descriptors used by the previous launch
        uint64_t* ost = a.lb.status + (par ^ 1) * cap;
        uint64_t* otA = a.lb.tailA + (par ^ 1) * cap;
        uint64_t* otP = a.lb.tailP + (par ^ 1) * cap;
        for (uint64_t i = t + (uint64_t)tid * a.T; i < s_lastT; i += 256ull * a.T) {
            st_rel(&ost[i], 0);
            st_rel(&otA[i], 0);
            st_rel(&otP[i], 0);
        }
    }

    // ---- 1. decode, select, rank within the wave
    const uint64_t n = a.in.n;
    const uint64_t wbase = (uint64_t)t * CS_TILE + wave * (CS_J * 64);
    TileReader<IK> rd(a.in, wbase < n ? wbase : 0);
    uint64_t v[CS_J];
    uint64_t bal[CS_J];
    uint32_t cnt = 0, ovf = 0;
#pragma unroll
    for (uint32_t j = 0; j < CS_J; j++) {
        const uint64_t e = wbase + j * 64 + lane;
        const bool inb = e < n;
        v[j] = inb ? rd.get(j * 64 + lane) : 0;
        bool f;
        if (PRED) {
            f = inb && pred_eval(a.pop, a.pc, v[j]);
        }
        else {
            const uint64_t mw = (wbase + j * 64 < n) ? a.mask_in[(wbase >> 6) + j] : 0;
            f = inb && ((mw >> lane) & 1);
        }
        bal[j] = ballot(f);
        cnt += (uint32_t)__popcll(bal[j]);
        ovf += (uint32_t)__popcll(ballot(f && (v[j] & ~om) != 0));
    }
    if (a.mask_out) {
        uint64_t mine = 0;
#pragma unroll
        for (uint32_t j = 0; j < CS_J; j++)
            if (lane == j) mine = bal[j];
        const uint64_t w = (wbase >> 6) + lane;
        if (lane < CS_J && w < ((n + 63) >> 6)) a.mask_out[w] = mine;
    }
    if (lane == 0) s_wcnt[wave] = cnt;
    __syncthreads();
    uint32_t wpre = 0, ctile = 0;
#pragma unroll
    for (uint32_t w = 0; w < CS_WAVES; w++) {
        uint32_t c = s_wcnt[w];
        wpre += w < wave ? c : 0;
        ctile += c;
    }

    if (OK != K_STRIDED) {
        uint32_t run = wpre;
#pragma unroll
        for (uint32_t j = 0; j < CS_J; j++) {
            if ((bal[j] >> lane) & 1) s_vals[run + mbcnt(bal[j])] = v[j] & om;
            run += (uint32_t)__popcll(bal[j]);
        }
        __syncthreads();
    }

    // ---- 2. look-back (wave 0)
    if (wave == 0) {
        // own window: last 63 bits of this tile's packed stream
        uint64_t own_w = 0;
        if (OK != K_STRIDED && bo < 64) {
            const uint32_t k = (63 + bo - 1) / bo;
            const uint32_t take = ctile < k ? ctile : k;
            if (lane < take) {
                const uint32_t jx = ctile - take + lane;
                const int64_t shift = (int64_t)jx * bo - ((int64_t)ctile * bo - 63);
                const uint64_t val = s_vals[jx];
                own_w = shift >= 0 ? (val << shift) : (val >> (-shift));
            }
            own_w = wave_or_u64(own_w) & TAIL_BITS;
        }
        uint64_t excl_c = 0, excl_w = 0;
        if (t == 0) {
            if (lane == 0) {
                st_rel(&tP[0], TAIL_VALID | own_w);
                st_rel(&st[0], ST_P | ctile);
            }
        }
        else {
            if (lane == 0) {
                st_rel(&tA[t], TAIL_VALID | own_w);
                st_rel(&st[t], ST_A | ctile);
            }
            uint64_t acc_c = 0, acc_w = 0;
            int64_t pos = (int64_t)t - 1;
            while (true) {
                const int64_t idx = pos - (int64_t)lane;
                uint64_t s = ST_P;
                if (idx >= 0) {
                    while ((s = ld_acq(&st[idx])) == 0) __builtin_amdgcn_s_sleep(1);
                }
                const bool isP = (s >> 62) == 2;
                const uint64_t pm = ballot(isP);
                const uint32_t stop = pm ? (uint32_t)__builtin_ctzll(pm) : 64;
                const bool incl = lane <= stop;
                uint64_t c = incl ? (s & ST_COUNT) : 0;
                uint64_t w = 0;
                if (incl && idx >= 0) {
                    const uint64_t* tp = isP ? &tP[idx] : &tA[idx];
                    uint64_t x;
                    while (!((x = ld_acq(tp)) & TAIL_VALID)) __builtin_amdgcn_s_sleep(1);
                    w = x & TAIL_BITS;
                }
                // ordered butterfly: higher lanes hold older tiles
                for (int m = 1; m < 64; m <<= 1) {
                    const uint64_t oc = shfl_xor_u64(c, m), ow = shfl_xor_u64(w, m);
                    if (lane & m) w = win_combine(w, ow, oc, bo);
                    else w = win_combine(ow, w, c, bo);
                    c += oc;
                }
                acc_w = win_combine(w, acc_w, acc_c, bo);
                acc_c += c;
                if (pm) break;
                pos -= 64;
            }
            excl_c = acc_c;
            excl_w = acc_w;
            if (lane == 0) {
                st_rel(&tP[t], TAIL_VALID | win_combine(excl_w, own_w, ctile, bo));
                st_rel(&st[t], ST_P | (excl_c + ctile));
            }
        }
        if (lane == 0) {
            s_P = excl_c;
            s_carry = excl_w;
        }
    }
    __syncthreads();
    const uint64_t P = s_P, carry = s_carry;
    const bool last = (t == a.T - 1);
    const uint64_t ocap = a.cap;

    // ---- 3. write
    if (OK == K_STRIDED) {
        uint32_t run = wpre;
#pragma unroll
        for (uint32_t j = 0; j < CS_J; j++) {
            if ((bal[j] >> lane) & 1) {
                const uint64_t r = P + run + mbcnt(bal[j]);
                if (r < ocap) store_container(a.out, r, v[j]);
            }
            run += (uint32_t)__popcll(bal[j]);
        }
        if (last && tid == 0) {
            // zero the padding bytes of the final word
            uint64_t K = P + ctile;
            if (K > ocap) K = ocap;
            const uint64_t b0 = K * a.out.cbytes, b1 = (b0 + 7) & ~7ull;
            for (uint64_t b = b0; b < b1; b++) ((uint8_t*)a.out.p)[b] = 0;
        }
    }
    else if (OK == K_SLAB) {
        const uint64_t eps = a.out.eps;
        const uint64_t kf = P / eps;
        uint64_t kl = (P + ctile) / eps;
        if (last && ((P + ctile) % eps)) kl += 1;
        const uint64_t wcap = (ocap + eps - 1) / eps;
        if (kl > wcap) kl = wcap;
        for (uint64_t k = kf + tid; k < kl; k += 256) {
            uint64_t word = 0;
            const uint64_t e0 = k * eps;
            for (uint32_t s = 0; s < eps; s++) {
                const uint64_t e = e0 + s;
                if (e >= ocap) break;
                uint64_t val;
                if (e < P) val = (carry >> (63 - (P - e) * bo)) & om;
                else if (e < P + ctile) val = s_vals[e - P];
                else break;
                word |= val << (s * bo);
            }
            a.out.p[k] = word;
        }
    }
    else {
        const uint64_t B0 = P * bo, B1 = (P + ctile) * bo, Bcap = ocap * bo;
        const uint64_t kf = B0 >> 6;
        uint64_t kl = B1 >> 6;
        if (last && (B1 & 63)) kl += 1;
        const uint64_t wcap = (Bcap + 63) >> 6;
        if (kl > wcap) kl = wcap;
        for (uint64_t k = kf + tid; k < kl; k += 256) {
            const uint64_t b0 = k * 64;
            uint64_t word = 0;
            if (b0 < B0) word = carry >> (b0 + 63 - B0);
            uint64_t e0 = b0 / bo;
            if (e0 < P) e0 = P;
            uint64_t e1 = (b0 + 63) / bo + 1;
            if (e1 > P + ctile) e1 = P + ctile;
            for (uint64_t e = e0; e < e1; e++) {
                const int64_t pos = (int64_t)(e * bo) - (int64_t)b0;
                const uint64_t val = s_vals[e - P];
                word |= pos >= 0 ? (val << pos) : (val >> (-pos));
            }
            if (b0 + 64 > Bcap) word &= lowmask((uint32_t)(Bcap - b0));
            a.out.p[k] = word;
        }
    }

    // ---- finish: overflow count, last-block bookkeeping
    if (lane == 0 && ovf) (void)add64(&a.hdr->overflow, ovf);
    drain();
    __syncthreads();
    if (tid == 0) {
        const uint32_t d = add32(&a.hdr->done, 1);
        if (d == a.T - 1) {
            uint64_t s;
            while (((s = ld_acq(&st[a.T - 1])) >> 62) != 2) __builtin_amdgcn_s_sleep(1);
            const uint64_t ov = __hip_atomic_exchange(&a.hdr->overflow, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (a.stats) {
                a.stats[0] = s & ST_COUNT;
                a.stats[1] = ov;
            }
            st32(&a.hdr->lastT[par], a.T);
            st32(&a.hdr->ticket, 0);
            st32(&a.hdr->done, 0);
            st32(&a.hdr->epoch, ld32(&a.hdr->epoch) + 1);
        }
    }
}

} // namespace am

using namespace am;

extern "C" int am_compressstore(const am_column* in, const uint64_t* mask, const am_predicate* pred, am_column* out,
                                uint64_t* mask_out, uint64_t* stats_dev, void* ws, uint64_t ws_bytes, void* stream)
{
    Col ci, co;
    int st = make_col(in, &ci);
    if (st) return st;
    if ((st = make_col(out, &co))) return st;
    if (!mask && !pred) return AM_E_INVALID;
    if (pred && pred->op > AM_PRED_NE) return AM_E_INVALID;
    if (!ws || ws_bytes < WS_FIXED) return AM_E_WORKSPACE;
    if (out->n && !out->data) return AM_E_INVALID;
    const uint64_t T = ci.n ? (ci.n + CS_TILE - 1) / CS_TILE : 1;
    Lookback lb = lookback_of(ws, ws_bytes);
    if (T > lb.cap || T >= (1ull << 31)) return AM_E_WORKSPACE;
    CompressArgs a{};
    a.in = ci;
    a.mask_in = mask;
    a.pc = pred ? pred->c : 0;
    a.pop = pred ? pred->op : 0;
    a.T = (uint32_t)T;
    a.out = co;
    a.cap = out->n;
    a.mask_out = mask_out;
    a.stats = stats_dev;
    a.hdr = (WsHeader*)ws;
    a.lb = lb;
    hipStream_t s = (hipStream_t)stream;
    const bool use_pred = (mask == nullptr);
#define CS_LAUNCH(IK, OK)                                                                              \
    if (use_pred) hipLaunchKernelGGL((k_compress<IK, OK, true>), dim3((uint32_t)T), dim3(256), 0, s, a); \
    else hipLaunchKernelGGL((k_compress<IK, OK, false>), dim3((uint32_t)T), dim3(256), 0, s, a);
#define CS_OUT(IK)                                          \
    switch (co.kind) {                                      \
        case K_STRIDED: CS_LAUNCH(IK, K_STRIDED); break;    \
        case K_SLAB: CS_LAUNCH(IK, K_SLAB); break;          \
        default: CS_LAUNCH(IK, K_BITSTREAM); break;         \
    }
    switch (ci.kind) {
        case K_STRIDED: CS_OUT(K_STRIDED); break;
        case K_SLAB: CS_OUT(K_SLAB); break;
        default: CS_OUT(K_BITSTREAM); break;
    }
#undef CS_OUT
#undef CS_LAUNCH
    return check_launch();
}
