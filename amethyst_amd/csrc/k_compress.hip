// k_compress.hip -- single-pass filter + compressstore with decoupled look-back.
//
// Reference (src/kernels/compressstore_slabs.cuh:4-206,
// src/benchs/bench_compressstore_slabs.cu:38-56): three launches -- a
// per-1024-element popcount of the mask, cub::DeviceScan::ExclusiveSum over
// the counts, then a writer that serially walks 32 mask words per warp,
// re-reads an input word from global memory for every output element, and
// merges the slab words shared by two warps with atomicAnd/atomicOr.
//
// Here it is one launch and it uses no contended global atomics. A grid of
// at most as many 256-thread blocks as can be resident at once walks the
// 4096-element tiles in order (block b takes tiles b, b+G, ...; a tile only
// waits on lower tiles, whose blocks are running, so look-back always makes
// progress). Each block loads the next tile while it works on the current
// one, and per tile:
//   1. decodes element base + 64j + lane per lane; the selection flag (mask
//      bit or the predicate evaluated in place) goes through one wave64
//      ballot, which is also the output mask word and, via v_mbcnt, each
//      lane's rank among the wave's selected elements;
//   2. publishes its aggregate and walks back over predecessors 64 at a
//      time (one descriptor per lane), combining (count, window, overflow)
//      with an ordered butterfly; the window is the last 63 bits of the
//      packed output stream, so the tile that completes an output word can
//      build it from its predecessors' values without atomics;
//   3. writes its selected values: byte-aligned outputs (u8/u16/u32/u64
//      containers) are scattered directly at their rank; slab and bitstream
//      outputs are assembled one whole word per lane from the LDS-compacted
//      tile plus the carried window, so every word is written exactly once,
//      dead and padding bits zero, no read-modify-write.
// Descriptor words carry the launch epoch (bumped by the last tile), so the
// workspace needs no clearing between launches or graph replays.
#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t CS_WAVES = 4;
constexpr uint32_t CS_J = 16;
constexpr uint32_t CS_TILE = CS_WAVES * CS_J * 64; // 4096

struct CompressArgs {
    Col in;
    const uint64_t* mask_in;
    Pred pr;
    uint32_t chk_ovf; // output narrower than input: count values that do not fit
    uint32_t T;       // tiles
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

template <int IK, int OK, bool PRED, bool WIDE>
__global__ __launch_bounds__(256) void k_compress(CompressArgs a)
{
    __shared__ uint32_t s_wcnt[CS_WAVES];
    __shared__ uint32_t s_wovf[CS_WAVES];
    __shared__ uint64_t s_P, s_carry;
    __shared__ uint32_t s_epoch;
    __shared__ uint64_t s_vals[OK == K_STRIDED ? 1 : CS_TILE];

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t bo = a.out.bits;
    const uint64_t om = a.out.mask;
    const uint64_t n = a.in.n;
    const uint32_t T = a.T;

    if (tid == 0) s_epoch = ld32(&a.hdr->epoch);
    __syncthreads();
    const uint32_t epoch = s_epoch;
    {
        // zero one slice of the descriptors this launch does not use
        const uint64_t cap = a.lb.cap;
        const uint64_t per = (cap + LB_CLEAR_SLICES - 1) / LB_CLEAR_SLICES;
        uint64_t lo = (uint64_t)(epoch % LB_CLEAR_SLICES) * per;
        uint64_t hi = lo + per < cap ? lo + per : cap;
        if (lo < T) lo = T;
        for (uint64_t i = lo + (uint64_t)blockIdx.x * 256 + tid; i < hi; i += (uint64_t)gridDim.x * 256) {
            st_rel(&a.lb.status[i], 0);
            st_rel(&a.lb.winA[2 * i], 0);
            st_rel(&a.lb.winA[2 * i + 1], 0);
            st_rel(&a.lb.winP[2 * i], 0);
            st_rel(&a.lb.winP[2 * i + 1], 0);
        }
    }
    uint32_t t = blockIdx.x;
    if (t >= T) return;

    Fetch<IK, WIDE> f[CS_J];
    uint64_t mk = 0; // lane j < CS_J: mask word j of this wave's chunk (mask mode)
    auto fetch_tile = [&](uint32_t tt, Fetch<IK, WIDE>* ff, uint64_t& mm) {
        const uint64_t wb = (uint64_t)tt * CS_TILE + wave * (CS_J * 64);
        if (wb < n) {
            const uint32_t rem = (uint32_t)(n - wb < CS_J * 64 ? n - wb : CS_J * 64);
            TileReader<IK> rd(a.in, wb);
#pragma unroll
            for (uint32_t j = 0; j < CS_J; j++) {
                const uint32_t i = j * 64 + lane;
                ff[j].load(rd, i < rem ? i : rem - 1);
            }
        }
        else {
#pragma unroll
            for (uint32_t j = 0; j < CS_J; j++) ff[j].raw = 0, ff[j].sh = 0;
        }
        if (!PRED) {
            const uint64_t w = (wb >> 6) + lane;
            mm = (lane < CS_J && w < ((n + 63) >> 6)) ? a.mask_in[w] : 0;
        }
    };
    fetch_tile(t, f, mk);

    for (; t < T; t += gridDim.x) {
        // ---- prefetch the next tile of this block
        Fetch<IK, WIDE> nf[CS_J];
        uint64_t nmk = 0;
        const uint32_t tn = t + gridDim.x;
        if (tn < T) fetch_tile(tn, nf, nmk);

        // ---- 1. decode, select, count (lane j keeps ballot j in `mine`)
        const uint64_t wbase = (uint64_t)t * CS_TILE + wave * (CS_J * 64);
        const uint32_t wrem = wbase < n ? (uint32_t)(n - wbase < CS_J * 64 ? n - wbase : CS_J * 64) : 0;
        uint64_t mine = 0;
        uint32_t cnt = 0, ovf = 0;
#pragma unroll
        for (uint32_t j = 0; j < CS_J; j++) {
            const bool inb = j * 64 + lane < wrem;
            const uint64_t v = f[j].value(a.in.mask);
            bool fl;
            if (PRED) {
                fl = inb && (WIDE ? a.pr(v) : a.pr.sel32((uint32_t)v));
            }
            else {
                const uint64_t mw = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(mk >> 32), j) << 32) |
                                    (uint32_t)__builtin_amdgcn_readlane((uint32_t)mk, j);
                fl = inb && ((mw >> lane) & 1);
            }
            const uint64_t b = ballot(fl);
            if (lane == j) mine = b;
            cnt += (uint32_t)__popcll(b);
            if (a.chk_ovf) ovf += (uint32_t)__popcll(ballot(fl && (v & ~om) != 0));
        }
        if (a.mask_out) {
            const uint64_t w = (wbase >> 6) + lane;
            if (lane < CS_J && w < ((n + 63) >> 6)) a.mask_out[w] = mine;
        }
        if (lane == 0) {
            s_wcnt[wave] = cnt;
            s_wovf[wave] = ovf;
        }
        __syncthreads();
        uint32_t wpre = 0, ctile = 0, otile = 0;
#pragma unroll
        for (uint32_t w = 0; w < CS_WAVES; w++) {
            const uint32_t c = s_wcnt[w];
            wpre += w < wave ? c : 0;
            ctile += c;
            otile += s_wovf[w];
        }
        auto bal = [&](uint32_t j) -> uint64_t {
            return ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(mine >> 32), j) << 32) |
                   (uint32_t)__builtin_amdgcn_readlane((uint32_t)mine, j);
        };

        if (OK != K_STRIDED) {
            uint32_t run = wpre;
#pragma unroll
            for (uint32_t j = 0; j < CS_J; j++) {
                const uint64_t b = bal(j);
                if ((b >> lane) & 1) s_vals[run + mbcnt(b)] = f[j].value(a.in.mask) & om;
                run += (uint32_t)__popcll(b);
            }
            __syncthreads();
        }

        // ---- 2. look-back (wave 0)
        if (wave == 0) {
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
                own_w = wave_or_u64(own_w) & WIN_BITS;
            }
            uint64_t excl_c = 0, excl_w = 0, excl_o = 0;
            if (t > 0) {
                if (lane == 0) {
                    st_rel(&a.lb.winA[2 * t], lb_word(1, epoch, own_w));
                    st_rel(&a.lb.winA[2 * t + 1], lb_word(1, epoch, (own_w >> 46) | ((uint64_t)otile << 17)));
                    st_rel(&a.lb.status[t], lb_status(1, epoch, ctile));
                }
                int64_t pos = (int64_t)t - 1;
                while (true) {
                    const int64_t idx = pos - (int64_t)lane;
                    uint64_t s = 0;
                    uint32_t lvl = 2;
                    if (idx >= 0) {
                        while (!lb_status_ok(s = ld_acq(&a.lb.status[idx]), epoch)) __builtin_amdgcn_s_sleep(1);
                        lvl = (uint32_t)(s >> 62);
                    }
                    const uint64_t pm = ballot(lvl == 2);
                    const uint32_t stop = pm ? (uint32_t)__builtin_ctzll(pm) : 64;
                    const bool incl = lane <= stop && idx >= 0;
                    uint64_t c = 0, w = 0, o = 0;
                    if (incl) {
                        c = s & CNT_MASK;
                        const uint64_t* wp = (lvl == 2 ? a.lb.winP : a.lb.winA) + 2 * idx;
                        uint64_t w0, w1;
                        while (!lb_word_ok(w0 = ld_acq(wp), lvl, epoch)) __builtin_amdgcn_s_sleep(1);
                        while (!lb_word_ok(w1 = ld_acq(wp + 1), lvl, epoch)) __builtin_amdgcn_s_sleep(1);
                        w = (w0 & P46) | ((w1 & 0x1FFFF) << 46);
                        o = (w1 & P46) >> 17;
                    }
                    // ordered butterfly: higher lanes hold older tiles
                    for (int m = 1; m < 64; m <<= 1) {
                        const uint64_t oc = shfl_xor_u64(c, m), ow = shfl_xor_u64(w, m), oo = shfl_xor_u64(o, m);
                        if (lane & m) w = win_combine(w, ow, oc, bo);
                        else w = win_combine(ow, w, c, bo);
                        c += oc;
                        o = ovf_add(o, oo);
                    }
                    excl_w = win_combine(w, excl_w, excl_c, bo);
                    excl_c += c;
                    excl_o = ovf_add(excl_o, o);
                    if (pm) break;
                    pos -= 64;
                }
            }
            const uint64_t incl_c = excl_c + ctile;
            const uint64_t incl_w = win_combine(excl_w, own_w, ctile, bo);
            const uint64_t incl_o = ovf_add(excl_o, otile);
            if (lane == 0) {
                st_rel(&a.lb.winP[2 * t], lb_word(2, epoch, incl_w));
                st_rel(&a.lb.winP[2 * t + 1], lb_word(2, epoch, (incl_w >> 46) | (incl_o << 17)));
                st_rel(&a.lb.status[t], lb_status(2, epoch, incl_c));
                s_P = excl_c;
                s_carry = excl_w;
                if (t == T - 1) {
                    if (a.stats) {
                        a.stats[0] = incl_c;
                        a.stats[1] = incl_o;
                    }
                    st32(&a.hdr->epoch, epoch + 1);
                }
            }
        }
        __syncthreads();
        const uint64_t P = s_P, carry = s_carry;
        const bool last = (t == T - 1);
        const uint64_t ocap = a.cap;

        // ---- 3. write
        if (OK == K_STRIDED) {
            uint32_t run = wpre;
#pragma unroll
            for (uint32_t j = 0; j < CS_J; j++) {
                const uint64_t b = bal(j);
                if ((b >> lane) & 1) {
                    const uint64_t r = P + run + mbcnt(b);
                    if (r < ocap) store_container(a.out, r, f[j].value(a.in.mask));
                }
                run += (uint32_t)__popcll(b);
            }
            if (last && tid == 0) {
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
        __syncthreads(); // LDS reuse by the next tile
#pragma unroll
        for (uint32_t j = 0; j < CS_J; j++) f[j] = nf[j];
        mk = nmk;
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
    if (pred) a.pr = make_pred(*pred, ci.mask);
    a.chk_ovf = co.bits < ci.bits ? 1 : 0;
    a.T = (uint32_t)T;
    a.out = co;
    a.cap = out->n;
    a.mask_out = mask_out;
    a.stats = stats_dev;
    a.hdr = (WsHeader*)ws;
    a.lb = lb;
    hipStream_t s = (hipStream_t)stream;
    const bool use_pred = (mask == nullptr);
    const bool wide = ci.bits > 32;
#define CS_GO(IK, OK, P, W)                                                                       \
    {                                                                                             \
        static uint32_t cap_blocks = 0;                                                           \
        if (!cap_blocks) cap_blocks = resident_blocks(k_compress<IK, OK, P, W>, 256, 0);          \
        const uint32_t g = (uint32_t)(T < cap_blocks ? T : cap_blocks);                           \
        hipLaunchKernelGGL((k_compress<IK, OK, P, W>), dim3(g), dim3(256), 0, s, a);              \
    }
#define CS_LAUNCH(IK, OK)                                    \
    if (use_pred) {                                          \
        if (wide) CS_GO(IK, OK, true, true) else CS_GO(IK, OK, true, false)     \
    } else {                                                 \
        if (wide) CS_GO(IK, OK, false, true) else CS_GO(IK, OK, false, false)   \
    }
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
#undef CS_GO
    return check_launch();
}
