// k_compress.hip -- single-pass filter + compressstore with decoupled look-back.
//
// Reference (src/kernels/compressstore_slabs.cuh:4-206,
// src/benchs/bench_compressstore_slabs.cu:38-56): three launches -- a
// per-1024-element popcount of the mask, cub::DeviceScan::ExclusiveSum over
// the counts, then a writer that serially walks 32 mask words per warp,
// re-reads an input word from global memory for every output element, and
// merges the slab words shared by two warps with atomicAnd/atomicOr.
//
// Here it is one launch without contended global atomics: one 256-thread
// block per tile, in tile order, each
//   1. selecting: the word kernel (slab and 1/2/4/8-byte container inputs)
//      gives every lane whole 64-bit input words, decodes all fields of a
//      word in registers and forms the word's selection bits (predicate
//      evaluated in place, or the bits of the input mask); per 64-word chunk
//      a DPP wave scan gives each lane its rank, and the lanes' bit groups
//      are OR-ed into the tile's output mask words in LDS. The element
//      kernel (bitstream inputs) decodes element base + 64j + lane per lane
//      and uses wave64 ballots, which are the mask words themselves;
//   2. staging its selected values, compacted, in LDS at the narrowest
//      container that holds them;
//   3. publishing its aggregate and looking back (one wave) over its
//      predecessors, two-level: group-mates, then whole 64-tile groups. The
//      scan element carries (count, window[, overflow]); the window is the
//      last 63 bits of the packed output stream, so the tile that completes
//      an output word builds it from its predecessors' values;
//   4. writing: containers are copied out contiguously at the tile's offset;
//      slab and bitstream outputs are assembled one whole word per lane, so
//      every word is written exactly once, dead and padding bits zero, with
//      no read-modify-write.
// Descriptor words carry the launch epoch (bumped by the last tile), so the
// workspace needs no clearing between launches or graph replays.
#include <atomic>
#include <mutex>
#include <type_traits>

#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t CS_THREADS = 256;          // 4 waves
constexpr uint32_t CE_J = 16;                 // element kernel: elements per lane
constexpr uint32_t CE_TILE = 4 * CE_J * 64;   // element kernel: 4096 elements per tile

template <int VB> struct VT;
template <> struct VT<1> { using T = uint8_t; };
template <> struct VT<2> { using T = uint16_t; };
template <> struct VT<4> { using T = uint32_t; };
template <> struct VT<8> { using T = uint64_t; };

struct CompressArgs {
    Col in;
    const uint64_t* mask_in;
    Pred pr;
    uint32_t fw, eps, R; // word kernel: field width, fields per word, chunks per wave
    uint32_t T;          // tiles
    uint64_t in_words;   // word kernel: words holding the n elements
    Col out;
    uint64_t cap; // output capacity in elements
    uint64_t* mask_out;
    uint64_t* stats;
    WsHeader* hdr;
    uint64_t* lb_tiles;
    uint64_t* lb_groups;
    uint64_t lb_cap;
    uint32_t lds16;   // input pointer 16-byte aligned: 16-byte global->LDS copies
    uint32_t out_vec; // output pointer aligned for 4-element vector stores
    TopPred tp;       // predicate over top-aligned fields (direct kernel)
    uint32_t swar, swinv;   // 2/4/8-bit fields: select_fields_pow2 with these masks
    uint64_t swE, swK, swG;
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

__device__ inline uint64_t readlane64(uint64_t v, uint32_t j)
{
    return ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), j) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((uint32_t)v, j);
}

// Window of a tile: the last 63 bits of its packed output stream.
template <typename V>
__device__ inline uint64_t tile_window(const V* s_vals, uint32_t ctile, uint32_t bo, uint32_t lane)
{
    uint64_t w = 0;
    if (bo < 64) {
        const uint32_t k = (63 + bo - 1) / bo;
        const uint32_t take = ctile < k ? ctile : k;
        if (lane < take) {
            const uint32_t jx = ctile - take + lane;
            const int64_t shift = (int64_t)jx * bo - ((int64_t)ctile * bo - 63);
            const uint64_t v = s_vals[jx];
            w = shift >= 0 ? (v << shift) : (v >> (-shift));
        }
        w = wave_or_u64(w) & WIN_BITS;
    }
    return w;
}

// Look-back for tile t (one wave; PUBLISHED: its aggregate, and for a
// group's last tile the group's, are already out);
// the last tile writes the statistics and advances the epoch.
template <uint32_t F, bool PUBLISHED = false>
__device__ inline void tile_lookback(const CompressArgs& a, const Lookback<F>& L, uint64_t t, uint32_t ctile,
                                     uint32_t otile, uint64_t own_w, uint32_t lane, uint64_t* s_P, uint64_t* s_carry)
{
    const Agg own{ctile, own_w, otile, 0};
    const Agg ex = PUBLISHED ? L.walk(t, own, lane, true) : L.run(t, own, lane);
    if (lane == 0) {
        *s_P = ex.c;
        if (s_carry) *s_carry = ex.w;
        if (t == a.T - 1) {
            const bool bad = L.aborted();
            if (a.stats) {
                a.stats[0] = bad ? STATS_INVALID : ex.c + ctile;
                a.stats[1] = ovf_add(ex.o, otile);
            }
            st32(&a.hdr->epoch, L.epoch + 1);
        }
    }
}

// Write the tile's ctile staged values at output position P.
template <int OK, typename V>
__device__ inline void tile_write(const CompressArgs& a, const V* s_vals, uint64_t P, uint64_t carry,
                                  uint32_t ctile, bool last, uint32_t tid)
{
    const uint64_t ocap = a.cap;
    if (OK == K_STRIDED) {
        for (uint32_t i = tid; i < ctile; i += CS_THREADS) {
            const uint64_t r = P + i;
            if (r < ocap) store_container(a.out, r, s_vals[i]);
        }
        if (last && tid == 0) {
            uint64_t K = P + ctile;
            if (K > ocap) K = ocap;
            const uint64_t b0 = K * a.out.cbytes, b1 = (b0 + 7) & ~7ull;
            for (uint64_t b = b0; b < b1; b++) ((uint8_t*)a.out.p)[b] = 0;
        }
    }
    else write_packed_tile<OK>(a.out, ocap, s_vals, P, carry, ctile, last, tid, CS_THREADS);
}

template <uint32_t F>
__device__ inline Lookback<F> make_lookback(const CompressArgs& a, uint32_t epoch)
{
    Lookback<F> L{};
    L.tiles = a.lb_tiles;
    L.groups = a.lb_groups;
    L.cap = a.lb_cap;
    L.epoch = epoch;
    L.bits = a.out.bits;
    L.hdr = a.hdr;
    return L;
}

template <int OK, bool OVF>
struct CsFlags {
    static constexpr uint32_t F = OK != K_STRIDED ? (LB_WIN | LB_OVF) : (OVF ? LB_OVF : 0);
};

// ------------------------------------------------------------ direct kernel
//
// Container outputs (1/2/4/8-byte) from slab / container inputs. A block is
// 4 waves (the largest shape of which 7-8 blocks stay resident per CU). The
// input tile (256*R words, R <= 8) is copied into LDS by the load unit
// itself (global_load_lds: no registers are held while it is in flight);
// each wave then takes its words into registers and forms their selection
// bits, a DPP scan ranks them, and -- while wave 0 looks back over the
// predecessors -- the other waves compact their selected values into the
// same LDS (now free) at the narrowest width that holds them; wave 0 does
// its own share after the look-back. Once the tile's output offset P is
// known, the block copies the compacted run to the output with vector stores.
constexpr uint32_t CW_MASKW = 256; // output mask words per tile (4*R*eps)
constexpr uint32_t CD_RMAX = 8;
constexpr uint32_t CD_WAVES = 4;
constexpr uint32_t CD_THREADS = CD_WAVES * 64;
constexpr uint32_t CD_STAGE = 19456; // LDS bytes: input words, then compacted values (8 blocks per CU)
constexpr uint32_t CD_STAGE_C = 16384; // container outputs: the same, compacted values in passes (8 blocks per CU)

// Copy values s_vals[0, cnt) >> k to out[P, P + cnt) (capacity-checked),
// four per thread step with one vector store where the output is aligned.
// (k drops the bits beneath a value staged top-aligned in its slot.)
template <typename C, typename V>
__device__ inline void copy_out(const CompressArgs& a, const V* s_vals, uint64_t P, uint32_t cnt, uint32_t tid, uint32_t k)
{
    C* out = (C*)a.out.p;
    const uint64_t cap = a.cap;
    uint32_t head = (uint32_t)((4 - (P & 3)) & 3);
    if (!a.out_vec) head = cnt; // unaligned output buffer: element stores only
    if (head > cnt) head = cnt;
    for (uint32_t i = tid; i < head; i += CD_THREADS)
        if (P + i < cap) st_nt(out + P + i, (C)(s_vals[i] >> k));
    const uint32_t nq = (cnt - head) / 4;
    for (uint32_t q = tid; q < nq; q += CD_THREADS) {
        const uint32_t i0 = head + 4 * q;
        const uint64_t r = P + i0;
        const C v0 = (C)(s_vals[i0] >> k), v1 = (C)(s_vals[i0 + 1] >> k), v2 = (C)(s_vals[i0 + 2] >> k),
                v3 = (C)(s_vals[i0 + 3] >> k);
        if (r + 4 <= cap) {
            if constexpr (sizeof(C) == 4) st_nt((u32x4*)(out + r), u32x4{(uint32_t)v0, (uint32_t)v1, (uint32_t)v2, (uint32_t)v3});
            else if constexpr (sizeof(C) == 2)
                st_nt((u32x2*)(out + r), u32x2{(uint32_t)v0 | ((uint32_t)v1 << 16), (uint32_t)v2 | ((uint32_t)v3 << 16)});
            else if constexpr (sizeof(C) == 1)
                st_nt((uint32_t*)(out + r), (uint32_t)v0 | ((uint32_t)v1 << 8) | ((uint32_t)v2 << 16) | ((uint32_t)v3 << 24));
            else {
                st_nt((u64x2*)(out + r), u64x2{(uint64_t)v0, (uint64_t)v1});
                st_nt((u64x2*)(out + r + 2), u64x2{(uint64_t)v2, (uint64_t)v3});
            }
        }
        else {
            if (r < cap) out[r] = v0;
            if (r + 1 < cap) out[r + 1] = v1;
            if (r + 2 < cap) out[r + 2] = v2;
            if (r + 3 < cap) out[r + 3] = v3;
        }
    }
    for (uint32_t i = head + 4 * nq + tid; i < cnt; i += CD_THREADS)
        if (P + i < cap) st_nt(out + P + i, (C)(s_vals[i] >> k));
}

// container outputs of a count-only scan and packed outputs fit 8 waves per
// SIMD (8 blocks per CU, the most 4-wave blocks a CU holds; packed outputs
// spill a few registers for it and still gain 3-8 %): ask the compiler to
// keep them there
template <int OK, bool OVF>
struct CdWaves {
    static constexpr int V = (OK != K_STRIDED || !OVF) ? 8 : 1;
};

template <int OK, bool PRED, bool OVF, int VB, bool WIDE>
__global__ __launch_bounds__(CD_THREADS, (CdWaves<OK, OVF>::V)) void k_compress_direct(CompressArgs a)
{
    constexpr uint32_t F = CsFlags<OK, OVF>::F;
    using V = typename VT<VB>::T;
    typedef __attribute__((address_space(3))) void* lds_ptr;
    constexpr uint32_t STAGE = OK == K_STRIDED ? CD_STAGE_C : CD_STAGE;
    __shared__ uint64_t s_buf[STAGE / 8]; // input words, then compacted values
    __shared__ uint64_t s_mask[PRED ? 1 : CW_MASKW]; // input mask words
    __shared__ uint32_t s_ccnt[4 * CD_RMAX], s_covf[4 * CD_RMAX];
    __shared__ uint64_t s_P, s_carry, s_win[CD_WAVES];
    __shared__ uint32_t s_epoch, s_abort;
    uint64_t* const s_x = s_buf;
    V* const s_vals = (V*)s_buf;

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool lead = wave == 0; // also looks back
    const uint64_t t = blockIdx.x;
    const uint32_t fw = a.fw, eps = a.eps, R = a.R, bw = a.in.bits;
    const uint64_t n = a.in.n, nw = a.in_words;
    const uint64_t w0 = t * (uint64_t)(CD_WAVES * 64 * R); // first word of the tile
    const uint64_t nmw = (n + 63) >> 6;
    const uint32_t NMW = CD_WAVES * R * eps;                 // mask words per tile (<= CW_MASKW)
    const uint64_t mw0 = (w0 * eps) >> 6;
    const bool last = t == a.T - 1;
    const bool pow2 = (eps & (eps - 1)) == 0;

    // ---- loads: the tile's words -> LDS (wave w: local words [64Rw, 64R(w+1)))
    uint32_t ep = 0, ab = 0;
    if (lead && lane == 0) {
        ep = ld32(&a.hdr->epoch); // needed after the first barrier
        ab = ld32(&a.hdr->abort);
    }
    {
        const uint32_t wl0 = wave * 64 * R, wend = wl0 + 64 * R;
        // Mask input: chunk r of this wave (64 words, 64*eps elements) is
        // exactly mask words [r*eps, (r+1)*eps) of the wave's R*eps, so the
        // wave reads its mask words first and loads only the input words of
        // chunks that select something (empty regions are never read: the
        // reference skips empty 1024-element chunks the same way,
        // src/kernels/compressstore_slabs.cuh:75-85).
        uint32_t live = 0xFFFFFFFFu;
        if (!PRED) {
            const uint32_t nwm = R * eps; // <= 64
            const uint64_t mi = mw0 + (uint64_t)wave * nwm + lane;
            uint64_t m = 0;
            if (lane < nwm && mi < nmw) m = a.mask_in[mi];
            if (lane < nwm) s_mask[wave * nwm + lane] = m;
            const uint64_t nz = ballot(m != 0);
            live = 0;
            for (uint32_t r = 0; r < R; r++)
                if ((nz >> (r * eps)) & lowmask(eps)) live |= 1u << r;
        }
        if (a.lds16) {
            for (uint32_t j = 0; j < R; j += 2) { // 128 words per instruction, 16 bytes per lane
                if (!((live >> j) & 3u)) continue; // (wave-uniform)
                const uint32_t wl = wl0 + j * 64 + 2 * lane;
                const uint64_t gw = w0 + wl;
                if (wl < wend) {
                    if (gw + 2 <= nw) __builtin_amdgcn_global_load_lds(a.in.p + gw, (lds_ptr)(s_x + wl0 + j * 64), 16, 0, LD_NT);
                    else if (gw < nw) s_x[wl] = a.in.p[gw];
                }
            }
        }
        else {
            const uint32_t* src = (const uint32_t*)a.in.p;
            for (uint32_t j = 0; j < 2 * R; j++) { // 32 words per instruction, 4 bytes per lane
                if (!((live >> (j >> 1)) & 1u)) continue; // (wave-uniform)
                const uint32_t dl = 2 * wl0 + j * 64 + lane;
                if (2 * w0 + dl < 2 * nw)
                    __builtin_amdgcn_global_load_lds(src + 2 * w0 + dl, (lds_ptr)((uint32_t*)s_x + 2 * wl0 + j * 64), 4, 0, LD_NT);
            }
        }
        __builtin_amdgcn_s_waitcnt(0); // this wave's global->LDS copies
    }
    __syncthreads();

    // ---- data waves: words -> registers, selection bits, ranks
    uint64_t x[CD_RMAX];
    uint32_t g[CD_RMAX], lp[CD_RMAX];
    const uint32_t om32 = (uint32_t)a.out.mask;
    const uint32_t leps = __umul24(lane, eps);
    // selection bits of the lane's word in chunk ci (word xr)
    auto select_word = [&](uint32_t ci, uint64_t xr) -> uint32_t {
        uint32_t gg;
        if (!PRED) {
            const uint32_t p = ci * 64 * eps + leps; // first bit of the word's elements in the tile mask
            const uint32_t w = p >> 6, off = p & 63;
            uint64_t m = s_mask[w] >> off;
            if (off && off + eps > 64) m |= s_mask[w + 1] << (64 - off);
            gg = (uint32_t)m & (uint32_t)lowmask(eps);
        }
        else if (WIDE)
            gg = a.pr(xr & a.in.mask);
        else if (VB == 1 && a.swar) // (2/4/8-bit fields stage 1-byte values; other instances keep
                                    // their code as it was: the branch had cost w=11 1-2 %)
            gg = select_fields_pow2(xr, fw, a.swE, a.swK, a.swG, a.swinv);
        else
            gg = select_fields(xr, eps, fw, bw, a.tp);
        if (last) { // elements past n (only in the last tile)
            const uint64_t e0 = (w0 + ci * 64 + lane) * eps;
            const uint32_t nv = e0 >= n ? 0 : (n - e0 < eps ? (uint32_t)(n - e0) : eps);
            gg &= (uint32_t)lowmask(nv);
        }
        return gg;
    };
    {
#pragma unroll
        for (uint32_t r = 0; r < CD_RMAX; r++) {
            x[r] = 0;
            g[r] = 0;
            lp[r] = 0;
            if (r >= R) continue;
            const uint32_t ci = wave * R + r;
            x[r] = s_x[ci * 64 + lane];
            uint32_t go = 0;
            const uint32_t gg = select_word(ci, x[r]);
            if (OVF) {
                if (WIDE) go = (gg & 1) && ((x[r] & a.in.mask) & ~a.out.mask) != 0;
                else
                    for (uint32_t s = 0, sh = 0; s < eps; s++, sh += fw)
                        go += ((gg >> s) & 1) && (((uint32_t)(x[r] >> sh) & (uint32_t)a.in.mask) & ~om32) != 0;
                const uint32_t oc = wave_sum32(go);
                if (lane == 0) s_covf[ci] = oc;
            }
            g[r] = gg;
            const uint32_t c = (uint32_t)__popc(gg);
            const uint32_t inc = wave_incl_scan(c);
            lp[r] = inc - c;
            if (lane == 63) s_ccnt[ci] = inc;
            if (PRED && a.mask_out) {
                if (pow2) { // the lanes of one mask word are an aligned segment: DPP OR
                    // (into this wave's consumed input words, stored together below)
                    const uint32_t S = 64 / eps;
                    const uint64_t m = eps == 1 ? ballot(gg & 1)
                                                : segment_or((uint64_t)gg << __umul24(lane & (S - 1), eps), S);
                    if ((lane & (S - 1)) == S - 1) s_x[wave * 64 * R + r * eps + lane / S] = m;
                }
                else // the selection bits, into this wave's own (consumed) input words
                    ((uint32_t*)(s_x + wave * 64 * R))[r * 64 + lane] = gg;
            }
        }
        if (PRED && a.mask_out && pow2) {
            // the wave's R*eps mask words are consecutive: one store of them
            // all (per-chunk stores were 8-16 byte pieces of a line at eps = 1, 2)
            const uint32_t nwm = R * eps;
            for (uint32_t k = lane; k < nwm; k += 64) {
                const uint64_t mi = mw0 + (uint64_t)wave * nwm + k;
                if (mi < nmw) st_nt(a.mask_out + mi, s_x[wave * 64 * R + k]);
            }
        }
        if (PRED && a.mask_out && !pow2) {
            // the wave's R*eps (<= 64) mask words: lane m gathers word m from
            // the selection bits of the (at most ceil(64/eps)+1) words it spans
            const uint32_t* sg = (const uint32_t*)(s_x + wave * 64 * R);
            const uint32_t nwm = R * eps;
            if (lane < nwm) {
                const uint32_t wa = (lane * 64) / eps;       // the word holding the mask word's first element
                const uint32_t sk = lane * 64 - wa * eps;    // its elements before that one
                uint64_t m = sg[wa] >> sk;
                uint32_t w = wa + 1, sh = eps - sk;
                // two words per step (independent loads: half the LDS round
                // trips); both start inside the mask word, so both exist
                for (; sh + eps < 64; w += 2, sh += 2 * eps) {
                    const uint32_t g0 = sg[w], g1 = sg[w + 1];
                    m |= ((uint64_t)g0 << sh) | ((uint64_t)g1 << (sh + eps));
                }
                if (sh < 64) m |= (uint64_t)sg[w] << sh;
                const uint64_t mi = mw0 + (uint64_t)wave * nwm + lane;
                if (mi < nmw) st_nt(a.mask_out + mi, m); // bits past n are zero (cleared above)
            }
        }
    }
    if (lead && lane == 0) {
        s_epoch = ep;
        s_abort = ab;
    }
    lds_barrier(); // every word is in registers: s_buf is free for the compacted values
    if (s_abort) {
        if (last && tid == 0 && a.stats) a.stats[0] = STATS_INVALID;
        return;
    }
    const uint32_t NC = CD_WAVES * R;
    const uint32_t cc = lane < NC ? s_ccnt[lane] : 0;
    const uint32_t cinc = wave_incl_scan(cc);
    const uint32_t ctile = __builtin_amdgcn_readlane(cinc, 63);
    if (!PRED && a.mask_out && tid < NMW && mw0 + tid < nmw) {
        uint64_t m = s_mask[tid];
        if (mw0 + tid == nmw - 1 && (n & 63)) m &= lowmask((uint32_t)(n & 63)); // input mask bits past n
        st_nt(a.mask_out + mw0 + tid, m);
    }
    // wave 0 publishes the tile's aggregate before the block compacts, and
    // looks back after. Packed outputs carry the tile's window (the last 63
    // bits of its output stream): each wave ORs in its values among the
    // last ceil(63/bits) ranks, taken from registers (only chunks reaching
    // into that tail do any work).
    const uint32_t otile = (OVF && lead) ? wave_sum32(lane < NC ? s_covf[lane] : 0) : 0;
    uint64_t own_w = 0;
    if (OK != K_STRIDED) {
        const uint32_t bo = a.out.bits;
        uint64_t wp = 0;
        if (bo < 64) {
            const uint32_t kq = (63 + bo - 1) / bo;
            const uint32_t q0 = ctile > kq ? ctile - kq : 0;
            const int32_t org = (int32_t)(ctile * bo) - 63; // stream bit of window bit 0
            const uint32_t vmw = (uint32_t)(a.in.mask & a.out.mask);
#pragma unroll
            for (uint32_t r = 0; r < CD_RMAX; r++) {
                if (r >= R) continue;
                const uint32_t ci = wave * R + r;
                if ((uint32_t)__builtin_amdgcn_readlane(cinc, ci) <= q0) continue; // chunk ends before the tail
                uint32_t q = __builtin_amdgcn_readlane(cinc, ci) - __builtin_amdgcn_readlane(cc, ci) + lp[r];
                const uint32_t gg = g[r];
                for (uint32_t s = 0, sh = 0; s < eps; s++, sh += fw)
                    if (__builtin_amdgcn_ubfe(gg, s, 1)) {
                        if (q >= q0) {
                            const uint64_t v = WIDE ? (x[r] & a.in.mask & a.out.mask) : (uint64_t)((uint32_t)(x[r] >> sh) & vmw);
                            const int32_t pos = (int32_t)(q * bo) - org;
                            wp |= pos >= 0 ? v << pos : v >> (-pos);
                        }
                        q++;
                    }
            }
            wp = wave_or_u64(wp);
        }
        if (lane == 0) s_win[wave] = wp;
        lds_barrier();
        own_w = (s_win[0] | s_win[1] | s_win[2] | s_win[3]) & WIN_BITS;
    }
    if (lead) {
        const Lookback<F> L = make_lookback<F>(a, s_epoch);
        const Agg own{ctile, own_w, otile, 0};
        L.publish_agg(t, own, lane);
        if ((t & 63) == 63) L.publish_group_early(t, own, lane);
    }
    if (ctile == 0 && !last && (t & 63) != 63) {
        // A tile that selects nothing writes nothing and needs no prefix:
        // with its (empty) aggregate out, successors walk past it. (A
        // group's last tile still walks: later walks stop at its inclusive
        // value; the last tile writes the statistics and the tail.)
        if (lead) make_lookback<F>(a, s_epoch).clear_slice(a.T, t * 64 + lane, (uint64_t)a.T * 64);
        return;
    }
    // (every LDS operation so far has completed; saying so keeps the compiler
    // from re-waiting on them inside the staging loops below, where each such
    // wait would also drain the previous value store)
    __builtin_amdgcn_s_waitcnt(0xC07F); // lgkmcnt(0)
    // compact the selected values into LDS at their tile ranks; container
    // outputs stage CAP values per pass (one pass unless the tile selects
    // more), packed outputs always fit
    constexpr uint32_t CAP = STAGE / sizeof(V);
    const uint32_t vm32 = (uint32_t)(a.in.mask & a.out.mask);
    // Container outputs stage a field's low vbits bits top-aligned in its
    // slot (one shift of the word, no mask; the bits beneath are neighbouring
    // fields', dropped by copy_out's shift by ksh) at a running byte offset.
    constexpr bool TOP = OK == K_STRIDED && !WIDE;
    const uint32_t vbits = bw < a.out.bits ? bw : a.out.bits;
    const uint32_t ksh = TOP ? 8 * (uint32_t)sizeof(V) - vbits : 0;
    auto stage_top = [&](uint64_t xr, uint32_t gg, uint32_t d, bool chk) {
        uint32_t da = d * (uint32_t)sizeof(V);
        for (uint32_t s = 0, sh = 64 - vbits; s < eps; s++, sh -= fw)
            if (__builtin_amdgcn_ubfe(gg, s, 1)) {
                if (!chk || da < CAP * (uint32_t)sizeof(V))
                    *(V*)((uint8_t*)s_vals + da) = (V)((xr << sh) >> (64 - 8 * sizeof(V)));
                da += (uint32_t)sizeof(V);
            }
    };
    auto stage = [&](uint32_t base, auto checked) {
        constexpr bool CHK = decltype(checked)::value;
#pragma unroll
        for (uint32_t r = 0; r < CD_RMAX; r++) {
            if (r >= R || !g[r]) continue;
            const uint32_t ci = wave * R + r;
            uint32_t d = __builtin_amdgcn_readlane(cinc, ci) - __builtin_amdgcn_readlane(cc, ci) + lp[r] - base;
            const uint32_t gg = g[r];
            if (WIDE) {
                if (!CHK || d < CAP) s_vals[d] = (V)(x[r] & a.in.mask & a.out.mask);
            }
            else if (TOP)
                stage_top(x[r], gg, d, CHK);
            else // (a branch-free form storing every field, unselected ones to a
                 // trash slot, measured slower: the extra LDS stores cost more)
                for (uint32_t s = 0, sh = 0; s < eps; s++, sh += fw)
                    if (__builtin_amdgcn_ubfe(gg, s, 1)) {
                        if (!CHK || d < CAP) s_vals[d] = (V)((uint32_t)(x[r] >> sh) & vm32);
                        d++;
                    }
        }
    };
    if (OK != K_STRIDED || ctile <= CAP) stage(0u, std::false_type{});
    else stage(0u, std::true_type{});
    if (OK == K_STRIDED) {
        if (lead) {
            const Lookback<F> L = make_lookback<F>(a, s_epoch);
            L.clear_slice(a.T, t * 64 + lane, (uint64_t)a.T * 64);
            tile_lookback<F, true>(a, L, t, ctile, otile, 0, lane, &s_P, nullptr);
        }
        lds_barrier();
        const uint64_t P = s_P;
        for (uint32_t base = 0;;) {
            const uint32_t cnt = ctile - base < CAP ? ctile - base : CAP;
            switch (a.out.cbytes) {
                case 1: copy_out<uint8_t>(a, s_vals, P + base, cnt, tid, ksh); break;
                case 2: copy_out<uint16_t>(a, s_vals, P + base, cnt, tid, ksh); break;
                case 4: copy_out<uint32_t>(a, s_vals, P + base, cnt, tid, ksh); break;
                default: copy_out<uint64_t>(a, s_vals, P + base, cnt, tid, ksh);
            }
            base += CAP;
            if (base >= ctile) break;
            lds_barrier(); // the pass is out: stage the next
            // (chunk by chunk from the tile's words again: the words, bits
            // and ranks are not kept live across the look-back, which would
            // cost occupancy)
            for (uint32_t r = 0; r < R; r++) {
                const uint32_t ci = wave * R + r;
                const uint64_t gw = w0 + ci * 64 + lane;
                const uint64_t xr = gw < nw ? a.in.p[gw] : 0;
                const uint32_t gg = select_word(ci, xr);
                const uint32_t c = (uint32_t)__popc(gg);
                uint32_t d = __builtin_amdgcn_readlane(cinc, ci) - __builtin_amdgcn_readlane(cc, ci) + wave_incl_scan(c) - c - base;
                if (WIDE) {
                    if (gg && d < CAP) s_vals[d] = (V)(xr & a.in.mask & a.out.mask);
                }
                else if (TOP)
                    stage_top(xr, gg, d, true);
                else
                    for (uint32_t s = 0, sh = 0; s < eps; s++, sh += fw)
                        if (__builtin_amdgcn_ubfe(gg, s, 1)) {
                            if (d < CAP) s_vals[d] = (V)((uint32_t)(xr >> sh) & vm32);
                            d++;
                        }
            }
            lds_barrier();
        }
        if (last && tid == 0) {
            uint64_t K = P + ctile;
            if (K > a.cap) K = a.cap;
            const uint64_t b0 = K * a.out.cbytes, b1 = (b0 + 7) & ~7ull;
            for (uint64_t b = b0; b < b1; b++) ((uint8_t*)a.out.p)[b] = 0;
        }
    }
    else {
        // packed outputs: output words are assembled from the staged values
        // and the carried window of the predecessors
        lds_barrier();
        if (lead) {
            const Lookback<F> L = make_lookback<F>(a, s_epoch);
            L.clear_slice(a.T, t * 64 + lane, (uint64_t)a.T * 64);
            tile_lookback<F, true>(a, L, t, ctile, otile, own_w, lane, &s_P, &s_carry);
        }
        lds_barrier();
        tile_write<OK>(a, s_vals, s_P, s_carry, ctile, last, tid);
    }
}

// ------------------------------------------------------------ element kernel
//
// Bitstream inputs: 4096-element tiles, element base + 64j + lane per lane.
template <int OK, bool PRED, bool OVF, bool WIDE>
__global__ __launch_bounds__(CS_THREADS) void k_compress_elems(CompressArgs a)
{
    constexpr uint32_t F = CsFlags<OK, OVF>::F;
    constexpr int IK = K_BITSTREAM;
    __shared__ uint64_t s_vals[CE_TILE];
    __shared__ uint32_t s_wcnt[4], s_wovf[4];
    __shared__ uint64_t s_P, s_carry;
    __shared__ uint32_t s_epoch, s_abort;

    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t t = blockIdx.x;
    if (tid == 0) {
        s_epoch = ld32(&a.hdr->epoch);
        s_abort = ld32(&a.hdr->abort);
    }
    __syncthreads();
    if (s_abort) {
        if (t == a.T - 1 && tid == 0 && a.stats) a.stats[0] = STATS_INVALID;
        return;
    }
    const Lookback<F> L = make_lookback<F>(a, s_epoch);
    L.clear_slice(a.T, t * CS_THREADS + tid, (uint64_t)a.T * CS_THREADS);

    const uint64_t n = a.in.n;
    const uint64_t om = a.out.mask;
    const uint64_t nmw = (n + 63) >> 6;
    const uint64_t wb = t * CE_TILE + (uint64_t)wave * (CE_J * 64);
    const uint32_t wrem = wb < n ? (uint32_t)(n - wb < CE_J * 64 ? n - wb : CE_J * 64) : 0;
    const TileReader<IK> rd(a.in, wb < n ? wb : 0);
    auto idx = [&](uint32_t j) { return j * 64 + lane < wrem ? j * 64 + lane : (wrem ? wrem - 1 : 0); };
    Fetch<IK, WIDE> f[CE_J];
    uint64_t mk = 0;
    if (wrem) {
        if (!PRED) { // mask input: 64-element chunks with an empty mask word are not read
            const uint64_t w = (wb >> 6) + lane;
            mk = (lane < CE_J && w < nmw) ? a.mask_in[w] : 0;
        }
#pragma unroll
        for (uint32_t j = 0; j < CE_J; j++) {
            f[j].raw = 0;
            if (PRED || readlane64(mk, j) != 0) f[j].load(rd, idx(j));
        }
    }
    uint64_t mine = 0;
    uint32_t cnt = 0, ovf = 0;
#pragma unroll
    for (uint32_t j = 0; j < CE_J; j++) {
        const bool inb = j * 64 + lane < wrem;
        const uint64_t v = wrem ? f[j].value(rd, idx(j)) : 0;
        bool fl;
        if (PRED) fl = inb && (WIDE ? a.pr(v) : a.pr.sel32((uint32_t)v));
        else fl = inb && ((readlane64(mk, j) >> lane) & 1);
        const uint64_t b = ballot(fl);
        if (lane == j) mine = b;
        cnt += (uint32_t)__popcll(b);
        if (OVF) ovf += (uint32_t)__popcll(ballot(fl && (v & ~om) != 0));
    }
    if (a.mask_out) {
        const uint64_t w = (wb >> 6) + lane;
        if (lane < CE_J && w < nmw) st_nt(a.mask_out + w, mine);
    }
    if (lane == 0) {
        s_wcnt[wave] = cnt;
        s_wovf[wave] = ovf;
    }
    lds_barrier();
    uint32_t wpre = 0, ctile = 0, otile = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; w++) {
        const uint32_t c = s_wcnt[w];
        wpre += w < wave ? c : 0;
        ctile += c;
        otile += s_wovf[w];
    }
    if (ctile == 0 && t != a.T - 1 && (t & 63) != 63) { // selects nothing: publish the empty aggregate, done
        if (wave == 0) L.publish_agg(t, Agg{0, 0, 0, 0}, lane);
        return;
    }
    if (OK == K_STRIDED && wave == 0) tile_lookback<F>(a, L, t, ctile, otile, 0, lane, &s_P, &s_carry);
    uint32_t run = wpre;
#pragma unroll
    for (uint32_t j = 0; j < CE_J; j++) {
        const uint64_t b = readlane64(mine, j);
        if ((b >> lane) & 1) s_vals[run + mbcnt(b)] = f[j].value(rd, idx(j)) & om;
        run += (uint32_t)__popcll(b);
    }
    lds_barrier();
    if (OK != K_STRIDED) {
        if (wave == 0) {
            const uint64_t w = tile_window(s_vals, ctile, a.out.bits, lane);
            tile_lookback<F>(a, L, t, ctile, otile, w, lane, &s_P, &s_carry);
        }
        lds_barrier();
    }
    tile_write<OK>(a, s_vals, s_P, s_carry, ctile, t == a.T - 1, tid);
}

// ------------------------------------------------------------ host dispatch

template <int OK, bool PRED, bool OVF>
static void launch_words(const CompressArgs& a, uint32_t vb, bool wide, hipStream_t s)
{
    const dim3 grid(a.T), block(CD_THREADS);
    if (wide) {
        switch (vb) {
            case 1: hipLaunchKernelGGL((k_compress_direct<OK, PRED, OVF, 1, true>), grid, block, 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_compress_direct<OK, PRED, OVF, 2, true>), grid, block, 0, s, a); break;
            case 4: hipLaunchKernelGGL((k_compress_direct<OK, PRED, OVF, 4, true>), grid, block, 0, s, a); break;
            default: hipLaunchKernelGGL((k_compress_direct<OK, PRED, OVF, 8, true>), grid, block, 0, s, a);
        }
    }
    else {
        switch (vb) {
            case 1: hipLaunchKernelGGL((k_compress_direct<OK, PRED, OVF, 1, false>), grid, block, 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_compress_direct<OK, PRED, OVF, 2, false>), grid, block, 0, s, a); break;
            default: hipLaunchKernelGGL((k_compress_direct<OK, PRED, OVF, 4, false>), grid, block, 0, s, a);
        }
    }
}

template <int OK, bool PRED, bool OVF>
static void launch_elems(const CompressArgs& a, bool wide, hipStream_t s)
{
    const dim3 grid(a.T), block(CS_THREADS);
    if (wide) hipLaunchKernelGGL((k_compress_elems<OK, PRED, OVF, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_compress_elems<OK, PRED, OVF, false>), grid, block, 0, s, a);
}

template <int OK>
static void launch_out(const CompressArgs& a, bool words, bool pred, bool ovf, uint32_t vb, bool wide, hipStream_t s)
{
    if (words) {
        if (pred) {
            if (ovf) launch_words<OK, true, true>(a, vb, wide, s);
            else launch_words<OK, true, false>(a, vb, wide, s);
        }
        else {
            if (ovf) launch_words<OK, false, true>(a, vb, wide, s);
            else launch_words<OK, false, false>(a, vb, wide, s);
        }
    }
    else {
        if (pred) {
            if (ovf) launch_elems<OK, true, true>(a, wide, s);
            else launch_elems<OK, true, false>(a, wide, s);
        }
        else {
            if (ovf) launch_elems<OK, false, true>(a, wide, s);
            else launch_elems<OK, false, false>(a, wide, s);
        }
    }
}

static inline uint32_t value_bytes(uint32_t bits) { return bits <= 8 ? 1 : (bits <= 16 ? 2 : (bits <= 32 ? 4 : 8)); }

// look-back flags of a launch (descriptor layout)
static inline uint32_t cs_flags(uint32_t out_kind, bool ovf)
{
    return out_kind != K_STRIDED ? (LB_WIN | LB_OVF) : (ovf ? LB_OVF : 0);
}

} // namespace am

using namespace am;

// Largest descriptor footprint of any compaction launch over n elements.
uint64_t am_compress_desc_bytes(uint64_t n)
{
    const uint64_t tiles = n / 1024 + 2; // tiles hold at least 1024 elements
    return Lookback<LB_WIN | LB_OVF>::bytes_for(tiles);
}

// One compaction launch with its look-back descriptors at `desc` (`desc_bytes`
// bytes) and the launch epoch / abort word at `hdr`.
static int compress_launch(const am_column* in, const uint64_t* mask, const am_predicate* pred, am_column* out,
                           uint64_t* mask_out, uint64_t* stats_dev, WsHeader* hdr, uint8_t* desc, uint64_t desc_bytes,
                           hipStream_t s)
{
    Col ci, co;
    int st = make_col(in, &ci);
    if (st) return st;
    if ((st = make_col(out, &co))) return st;
    if (!mask && !pred) return AM_E_INVALID;
    if (pred && pred->op > AM_PRED_NE) return AM_E_INVALID;
    if (out->n && !out->data) return AM_E_INVALID;
    CompressArgs a{};
    a.in = ci;
    a.mask_in = mask;
    if (pred) a.pr = make_pred(*pred, ci.mask);
    a.out = co;
    a.cap = out->n;
    a.mask_out = mask_out;
    a.stats = stats_dev;
    a.hdr = hdr;
    if (ci.kind == K_SLAB && ci.bits == 1) ci.kind = K_BITSTREAM; // 64 one-bit fields per word = dense stream
    a.in = ci;
    const bool words = ci.kind != K_BITSTREAM; // word kernel: at most 32 fields per word
    uint64_t T;
    if (words) {
        a.fw = ci.kind == K_SLAB ? ci.bits : 8 * ci.cbytes;
        a.eps = 64 / a.fw;
        {
            // mask words per tile 4*R*eps <= 256; packed outputs: compacted values 256*R*eps*VB <= CD_STAGE
            const uint32_t vbh = value_bytes(ci.bits < co.bits ? ci.bits : co.bits);
            a.R = 64 / a.eps;
            const uint32_t rs = CD_STAGE / (CD_WAVES * 64 * a.eps * vbh);
            if (co.kind != K_STRIDED && a.R > rs) a.R = rs; // packed outputs stage a whole tile
            static_assert(CD_STAGE_C >= CD_WAVES * 64 * CD_RMAX * 8, "container tile words fit the stage");
            if (a.R > CD_RMAX) a.R = CD_RMAX;
            // a tuned chunk count, within the stage limits above and keeping
            // tiles of >= 1024 elements (the descriptor area is sized for that)
            const uint32_t rmin = (4 + a.eps - 1) / a.eps;
            if (const uint32_t tr = am_tuned(AM_KNOB_COMPRESS_CHUNKS, in->bits, in->format, out->format))
                if (tr < a.R) a.R = tr < rmin ? rmin : tr;
            if (a.R < 1) a.R = 1;
            a.lds16 = ((uintptr_t)ci.p & 15) == 0;
            a.out_vec = ((uintptr_t)co.p & (4 * co.cbytes - 1)) == 0;
            if (pred && !mask && ci.bits <= 32) {
                a.tp = make_top_pred(a.pr, ci.bits, a.eps);
                SwarSel sw;
                if (swar_select(a.pr, a.tp, ci.bits, a.fw, a.eps, &sw)) {
                    a.swar = 1;
                    a.swinv = sw.inv;
                    a.swE = sw.E;
                    a.swK = sw.K;
                    a.swG = sw.G;
                }
            }
        }
        a.in_words = (ci.n + a.eps - 1) / a.eps;
        const uint64_t TW = (uint64_t)256 * a.R; // 4 (data) waves x R chunks x 64 words
        T = a.in_words ? (a.in_words + TW - 1) / TW : 1;
    }
    else {
        T = ci.n ? (ci.n + CE_TILE - 1) / CE_TILE : 1;
    }
    const bool ovf = co.bits < ci.bits;
    const uint32_t F = cs_flags(co.kind, ovf);
    uint64_t cap;
    if (F == 0) {
        const auto L = Lookback<0>::at(desc, desc_bytes);
        a.lb_tiles = L.tiles, a.lb_groups = L.groups, cap = L.cap;
    }
    else if (F == LB_OVF) {
        const auto L = Lookback<LB_OVF>::at(desc, desc_bytes);
        a.lb_tiles = L.tiles, a.lb_groups = L.groups, cap = L.cap;
    }
    else {
        const auto L = Lookback<LB_WIN | LB_OVF>::at(desc, desc_bytes);
        a.lb_tiles = L.tiles, a.lb_groups = L.groups, cap = L.cap;
    }
    if (T > cap || T >= (1ull << 31)) return AM_E_WORKSPACE;
    a.lb_cap = cap;
    a.T = (uint32_t)T;
    const uint32_t vb = value_bytes(ci.bits < co.bits ? ci.bits : co.bits);
    const bool wide = ci.bits > 32;
    switch (co.kind) {
        case K_STRIDED: launch_out<K_STRIDED>(a, words, pred != nullptr && !mask, ovf, vb, wide, s); break;
        case K_SLAB: launch_out<K_SLAB>(a, words, pred != nullptr && !mask, ovf, vb, wide, s); break;
        default: launch_out<K_BITSTREAM>(a, words, pred != nullptr && !mask, ovf, vb, wide, s); break;
    }
    return check_launch();
}

extern "C" int am_compressstore(const am_column* in, const uint64_t* mask, const am_predicate* pred, am_column* out,
                                uint64_t* mask_out, uint64_t* stats_dev, void* ws, uint64_t ws_bytes, void* stream)
{
    if (!ws || ws_bytes < WS_FIXED) return AM_E_WORKSPACE;
    return compress_launch(in, mask, pred, out, mask_out, stats_dev, (WsHeader*)ws, (uint8_t*)ws + WS_FIXED,
                           ws_bytes - WS_FIXED, (hipStream_t)stream);
}

// ------------------------------------------------------------ no-allocation entry point
//
// am_compressstore_noalloc keeps its launch epoch and look-back descriptors
// in `scratch`, a file-scope device array (one copy per GPU, alive until the
// process ends). Host code never passes the array's name to a kernel: its
// device address comes from hipGetSymbolAddress, once per device, and is
// kept in na_ptr[device].
namespace am {

constexpr uint64_t NA_TILES = AM_NOALLOC_MAX_ELEMENTS / 1024 + 2; // tiles hold at least 1024 elements
constexpr uint64_t NA_DESC_BYTES = Lookback<LB_WIN | LB_OVF>::bytes_for(NA_TILES);
constexpr uint64_t NA_BYTES = sizeof(WsHeader) + NA_DESC_BYTES;
static_assert(NA_BYTES <= AM_NOALLOC_SCRATCH_BYTES, "scratch size stated in amethyst.h");
constexpr uint64_t NA_WORDS = AM_NOALLOC_SCRATCH_BYTES / 8;
constexpr int NA_MAX_DEVICES = 64;

static __device__ __align__(256) uint64_t scratch[NA_WORDS];

static std::atomic<uint64_t*> na_ptr[NA_MAX_DEVICES];
static std::mutex na_mu;

__global__ __launch_bounds__(256) void k_na_clear(uint64_t* p, uint64_t words)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// One block: if an earlier call gave up waiting (abort word set), zero the
// whole scratch -- header included -- so this call starts from a clean
// state; otherwise do nothing.
__global__ __launch_bounds__(256) void k_na_heal(uint64_t* p, uint64_t words)
{
    __shared__ uint32_t s_ab;
    if (threadIdx.x == 0) s_ab = ld32(&((WsHeader*)p)->abort);
    __syncthreads();
    if (!s_ab) return;
    for (uint64_t i = threadIdx.x; i < words; i += 256) p[i] = 0;
}

// Device address of this device's scratch; the first call on a device looks
// it up and queues its zeroing on `s`.
static int na_scratch(hipStream_t s, uint64_t** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return AM_E_HIP;
    if (dev < 0 || dev >= NA_MAX_DEVICES) return AM_E_UNSUPPORTED_CONFIG;
    uint64_t* p = na_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> g(na_mu);
        p = na_ptr[dev].load(std::memory_order_relaxed);
        if (!p) {
            void* q = nullptr;
            if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return AM_E_HIP;
            p = (uint64_t*)q;
            hipLaunchKernelGGL(k_na_clear, dim3(1024), dim3(256), 0, s, p, NA_WORDS);
            if (hipGetLastError() != hipSuccess) return AM_E_HIP;
            na_ptr[dev].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return AM_OK;
}

} // namespace am

extern "C" int am_compressstore_noalloc(const am_column* in, const uint64_t* mask, const am_predicate* pred,
                                        am_column* out, uint64_t* mask_out, uint64_t* stats_dev, void* stream)
{
    if (!in || !out) return AM_E_INVALID;
    if (in->n > AM_NOALLOC_MAX_ELEMENTS) return AM_E_SCRATCH_LIMIT;
    hipStream_t s = (hipStream_t)stream;
    uint64_t* p = nullptr;
    int st = na_scratch(s, &p);
    if (st) return st;
    hipLaunchKernelGGL(k_na_heal, dim3(1), dim3(256), 0, s, p, NA_WORDS);
    if (hipGetLastError() != hipSuccess) return AM_E_HIP;
    st = compress_launch(in, mask, pred, out, mask_out, stats_dev, (WsHeader*)p, (uint8_t*)p + sizeof(WsHeader),
                         AM_NOALLOC_SCRATCH_BYTES - sizeof(WsHeader), s);
    return st == AM_E_WORKSPACE ? AM_E_SCRATCH_LIMIT : st;
}
