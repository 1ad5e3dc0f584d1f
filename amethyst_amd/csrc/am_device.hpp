// am_device.hpp -- device-side building blocks shared by the gfx950 kernels:
// column-format decoders, wave64 primitives, and the single-pass look-back
// state used by compaction.
//
// Formats follow the reference layouts (src/common_code.cuh:11-19,43-55,
// 100-116). Internally the five public formats reduce to three "kinds":
// STRIDED (one element per 1/2/4/8-byte container: BASELINE, TIGHT, U32),
// SLAB (floor(64/bits) elements per uint64, dead bits on top) and
// BITSTREAM (dense).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "../../include/amethyst.h"

namespace am {

enum Kind : int { K_STRIDED = 0, K_SLAB = 1, K_BITSTREAM = 2 };

__host__ __device__ inline uint64_t lowmask(uint32_t bits)
{
    return bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
}

// Column parameters resolved on the host (see capi.cpp: make_col).
struct Col {
    uint64_t* p;        // packed data (uint64 words)
    uint64_t n;         // elements
    uint64_t mask;      // lowmask(bits)
    uint32_t bits;
    uint32_t kind;      // Kind
    uint32_t cbytes;    // STRIDED container bytes (1, 2, 4, 8)
    uint32_t eps;       // SLAB elements per word
    uint32_t eps_magic; // SLAB: floor(2^32/eps)+1, exact quotient for idx < 2^24
    uint32_t pad;
    uint64_t words;     // uint64 words of the buffer
};

// A predicate normalised on the host to a branch-free range test over the
// column's value range: selected = ((v - lo) <= span) != inv.
struct Pred {
    uint64_t lo, span;
    uint32_t inv, pad;
    __device__ inline bool operator()(uint64_t v) const { return ((v - lo) <= span) != (inv != 0); }
    __device__ inline bool sel32(uint32_t v) const { return ((v - (uint32_t)lo) <= (uint32_t)span) != (inv != 0); }
};

// Streaming stores for outputs written once and not read again by the
// kernel (the non-temporal hint keeps them from displacing cached inputs and
// look-back state; 4 % on the 1 GB compaction). LD_NT is the matching cache
// policy for global->LDS loads of inputs read once.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
template <typename T> __device__ inline void st_nt(T* p, T v) { __builtin_nontemporal_store(v, p); }
template <typename T> __device__ inline void st_nt(__attribute__((address_space(1))) T* p, T v)
{
    __builtin_nontemporal_store(v, p);
}

// Pointers that kernels load from memory (job tables) are generic to the
// compiler, whose flat accesses also count against the LDS wait counter, so
// an LDS barrier would wait for every global load in flight. Device buffers
// are global memory: these casts give global_load / global_store.
#define AM_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ inline AM_GLOBAL T* gmem(T* p) { return (AM_GLOBAL T*)p; }
template <typename T> __device__ inline const AM_GLOBAL T* gmem(const T* p) { return (const AM_GLOBAL T*)p; }
// A load through the constant address space (scalar, kept in SGPRs): for
// tables no kernel writes while it runs (job tables, recorded batch meta).
template <typename T> __device__ inline T ldk(const T* p)
{
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    const __attribute__((address_space(4))) uint32_t* q = (const __attribute__((address_space(4))) uint32_t*)p;
    T v;
    uint32_t* d = (uint32_t*)&v;
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; i++) d[i] = q[i];
    return v;
}
constexpr int LD_NT = 2; // cache-policy operand: non-temporal

// 32-bit quotient idx/eps for idx < 2^24 and eps in 2..64 (multiply-high by
// floor(2^32/eps)+1: the error term stays below 2^-8 < 1/eps).
__device__ inline uint32_t div_eps(uint32_t idx, const Col& c)
{
    return c.eps == 1 ? idx : __umulhi(idx, c.eps_magic);
}

// A decoder for elements [t0, t0 + 2^24) of a column: per-tile constants are
// formed once (scalar 64-bit math), per-element work is 32-bit.
template <int KIND>
struct TileReader {
    const AM_GLOBAL uint64_t* base; // first word touched by the tile
    uint32_t r0;                    // SLAB: t0 % eps ; BITSTREAM: (t0*bits) % 64
    uint64_t mask;
    uint32_t bits, cbytes, eps, eps_magic;
    uint32_t last32; // BITSTREAM: last valid u32 index relative to base (clamp)
    Col c;

    __device__ inline TileReader(const Col& col, uint64_t t0) : c(col)
    {
        mask = col.mask;
        bits = col.bits;
        cbytes = col.cbytes;
        eps = col.eps;
        eps_magic = col.eps_magic;
        if (KIND == K_STRIDED) {
            base = (const AM_GLOBAL uint64_t*)((const AM_GLOBAL uint8_t*)gmem(col.p) + t0 * col.cbytes);
            r0 = 0;
        }
        else if (KIND == K_SLAB) {
            uint64_t q = t0 / col.eps;
            base = gmem(col.p) + q;
            r0 = (uint32_t)(t0 - q * col.eps);
        }
        else {
            uint64_t bit = t0 * (uint64_t)col.bits;
            base = gmem(col.p) + (bit >> 6);
            r0 = (uint32_t)(bit & 63);
        }
        const uint64_t rel = col.words - (uint64_t)(base - gmem(col.p)); // words from base to the end
        last32 = (uint32_t)(rel * 2 > 0xFFFFFFFFull ? 0xFFFFFFFFull : (rel ? rel * 2 - 1 : 0));
    }

    // element t0 + i (caller guarantees it is < n)
    __device__ inline uint64_t get(uint32_t i) const
    {
        if (KIND == K_STRIDED) {
            switch (cbytes) {
                case 1: return ((const AM_GLOBAL uint8_t*)base)[i] & mask;
                case 2: return ((const AM_GLOBAL uint16_t*)base)[i] & mask;
                case 4: return ((const AM_GLOBAL uint32_t*)base)[i] & mask;
                default: return base[i] & mask;
            }
        }
        else if (KIND == K_SLAB) {
            uint32_t idx = r0 + i;
            uint32_t q = eps == 1 ? idx : __umulhi(idx, eps_magic);
            uint32_t s = idx - q * eps;
            uint64_t w = base[q];
            return (w >> (s * bits)) & mask;
        }
        else {
            uint32_t lb = r0 + i * bits;
            const AM_GLOBAL uint64_t* p = base + (lb >> 6);
            uint32_t o = lb & 63;
            uint64_t v = p[0] >> o;
            if (o + bits > 64) v |= p[1] << (64 - o);
            return v & mask;
        }
    }
};

// Raw fetch of one element for the lane: the loaded word(s) plus the shift
// that extracts the value, so loads can be issued a whole tile ahead.
template <int IK, bool WIDE>
struct Fetch {
    uint64_t raw;
    // Branch-free: callers clamp `i` to a valid element and mask the result.
    __device__ inline void load(const TileReader<IK>& rd, uint32_t i)
    {
        if (IK == K_STRIDED) {
            switch (rd.cbytes) { // uniform
                case 1: raw = ((const AM_GLOBAL uint8_t*)rd.base)[i]; break;
                case 2: raw = ((const AM_GLOBAL uint16_t*)rd.base)[i]; break;
                case 4: raw = ((const AM_GLOBAL uint32_t*)rd.base)[i]; break;
                default: raw = rd.base[i];
            }
        }
        else if (IK == K_SLAB) {
            const uint32_t idx = rd.r0 + i;
            const uint32_t q = rd.eps == 1 ? idx : __umulhi(idx, rd.eps_magic);
            raw = rd.base[q];
        }
        else if (!WIDE) {
            const uint32_t lb = rd.r0 + i * rd.bits;
            const AM_GLOBAL uint32_t* p32 = (const AM_GLOBAL uint32_t*)rd.base;
            const uint32_t i0 = lb >> 5;
            const uint32_t i1 = i0 + 1 <= rd.last32 ? i0 + 1 : rd.last32;
            raw = ((uint64_t)p32[i1] << 32) | p32[i0];
        }
        else {
            raw = rd.get(i);
        }
    }
    // the same (clamped) element index as load(); the shift is recomputed
    __device__ inline uint64_t value(const TileReader<IK>& rd, uint32_t i) const
    {
        if (IK == K_SLAB) {
            const uint32_t idx = rd.r0 + i;
            const uint32_t q = rd.eps == 1 ? idx : __umulhi(idx, rd.eps_magic);
            return (raw >> ((idx - q * rd.eps) * rd.bits)) & rd.mask;
        }
        if (IK == K_BITSTREAM && !WIDE) return (raw >> ((rd.r0 + i * rd.bits) & 31)) & rd.mask;
        return raw & rd.mask;
    }
};

// A predicate (values of `bits` <= 32 bits) over fields read top-aligned in a
// 32-bit register with arbitrary bits beneath: see select_fields.
struct TopPred {
    uint32_t lo, rg, inv;
};

// Selection bits of a word's eps fields (field s -> bit s; field s is bits
// [s*fw, s*fw + bw)). Shifting the word left by 64 - s*fw - bw puts field s
// at the top of the high half with lower fields' bits beneath it; the range
// test is scaled by 2^(32-bw) so those bits cannot change its outcome:
//   v in [lo, lo + span]  <=>  (y - lo*2^k) <= (span + 1)*2^k - 1,  k = 32 - bw.
// Compare and accumulate are one v_cmp + one v_addc (carry = the compare).
__device__ inline uint32_t select_fields(uint64_t x, uint32_t eps, uint32_t fw, uint32_t bw, const TopPred& tp)
{
    const uint32_t LO = tp.lo, RG = tp.rg;
    uint32_t gr = 0; // bits in reverse field order
    if (LO == 0) { // (wave-uniform) ranges from 0, e.g. v < c: no subtraction
        for (uint32_t s = 0, sh = 64 - bw; s < eps; s++, sh -= fw) {
            const uint32_t t = (uint32_t)((x << sh) >> 32);
            uint32_t ng;
            asm("v_cmp_ge_u32 vcc, %2, %1\n\tv_addc_co_u32 %0, vcc, %3, %3, vcc" : "=v"(ng) : "v"(t), "s"(RG), "v"(gr) : "vcc");
            gr = ng;
        }
    }
    else {
        for (uint32_t s = 0, sh = 64 - bw; s < eps; s++, sh -= fw) {
            const uint32_t t = (uint32_t)((x << sh) >> 32) - LO;
            uint32_t ng;
            asm("v_cmp_ge_u32 vcc, %2, %1\n\tv_addc_co_u32 %0, vcc, %3, %3, vcc" : "=v"(ng) : "v"(t), "s"(RG), "v"(gr) : "vcc");
            gr = ng;
        }
    }
    return (__builtin_bitreverse32(gr) >> (32 - eps)) ^ tp.inv;
}

// Selection bits of a word of eps = 64/fw fields, fw = 2, 4 or 8 (no dead
// bits), for a range from 0 (v < sp1): even and odd fields in 2fw-bit slots
// (E: their value bits), v + (2^fw - sp1) (K, per slot) carries into the
// slot's bit fw (G) iff v >= sp1, two 64-bit adds for all fields instead of
// a compare per field; the not-selected bits then sit in 2-bit groups at a
// stride of 2fw (even field low, odd field high), gathered by log steps.
__device__ inline uint32_t select_fields_pow2(uint64_t x, uint32_t fw, uint64_t E, uint64_t K, uint64_t G, uint32_t inv)
{
    const uint64_t ne = ((x & E) + K) & G;
    const uint64_t no = (((x >> fw) & E) + K) & G;
    uint64_t c = (ne >> fw) | (no >> (fw - 1));
    uint32_t all;
    if (fw == 2) {
        c = (c | (c >> 2)) & 0x0F0F0F0F0F0F0F0Full;
        c = (c | (c >> 4)) & 0x00FF00FF00FF00FFull;
        c = (c | (c >> 8)) & 0x0000FFFF0000FFFFull;
        c = (c | (c >> 16)) & 0xFFFFFFFFull;
        all = 0xFFFFFFFFu;
    }
    else if (fw == 4) {
        c = (c | (c >> 6)) & 0x000F000F000F000Full;
        c = (c | (c >> 12)) & 0x000000FF000000FFull;
        c = (c | (c >> 24)) & 0xFFFFull;
        all = 0xFFFFu;
    }
    else {
        c = (c | (c >> 14)) & 0x0000000F0000000Full;
        c = (c | (c >> 28)) & 0xFFull;
        all = 0xFFu;
    }
    return ((uint32_t)c ^ all) ^ inv;
}

// ------------------------------------------------------------ wave64 helpers

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Workgroup barrier over LDS only: waits for this wave's LDS operations
// (lgkmcnt(0)), then s_barrier. __syncthreads() also drains the wave's
// global stores (its fence waits vmcnt(0)), a full memory round trip that
// nothing after these barriers needs: no wave reads global memory that
// another wave of the block wrote before them.
__device__ inline void lds_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F); // lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Inclusive wave64 prefix sum with DPP: row_shr 1/2/4/8 within rows of 16
// lanes (out-of-row sources read 0), then row_bcast 15 / 31 carry the row
// totals into the following rows.
template <int CTRL, int ROWM, bool BC>
__device__ inline uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWM, 0xf, BC);
}
__device__ inline uint32_t wave_incl_scan(uint32_t v)
{
    v += dpp_u32<0x111, 0xf, true>(v);  // row_shr:1
    v += dpp_u32<0x112, 0xf, true>(v);  // row_shr:2
    v += dpp_u32<0x114, 0xf, true>(v);  // row_shr:4
    v += dpp_u32<0x118, 0xf, true>(v);  // row_shr:8
    v += dpp_u32<0x142, 0xa, false>(v); // row_bcast:15 -> rows 1, 3
    v += dpp_u32<0x143, 0xc, false>(v); // row_bcast:31 -> rows 2, 3
    return v;
}

__device__ inline uint32_t wave_sum32(uint32_t v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += (uint32_t)__shfl_xor((int)v, m);
    return v;
}

// number of set bits of m below this lane (v_mbcnt_lo/hi)
__device__ inline uint32_t mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

// f(integral_constant<J>) for J in the sequence, unrolled at compile time
template <uint32_t... J, typename Fn>
__device__ inline void static_for(std::integer_sequence<uint32_t, J...>, Fn&& f)
{
    (f(std::integral_constant<uint32_t, J>{}), ...);
}

// Lane J of (lo, hi) := the uniform 64-bit value v (two v_writelane_b32 with
// an immediate lane select). gfx950 needs 2 wait states between a VALU write
// of an SGPR (such as the compare that produced v) and a VALU read of it; the
// compiler cannot see inside the asm, so the s_nop 1 is explicit.
template <uint32_t J>
__device__ inline void writelane_imm64(uint32_t& lo, uint32_t& hi, uint64_t v)
{
    asm volatile("s_nop 1\n\tv_writelane_b32 %0, %2, %4\n\tv_writelane_b32 %1, %3, %4"
                 : "+v"(lo), "+v"(hi)
                 : "s"((uint32_t)v), "s"((uint32_t)(v >> 32)), "i"(J));
}

__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m)
{
    uint32_t lo = __shfl_xor((unsigned)(uint32_t)v, m);
    uint32_t hi = __shfl_xor((unsigned)(uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint64_t wave_sum_u64(uint64_t v)
{
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_u64(v, m);
    return v;
}

__device__ inline uint64_t wave_or_u64(uint64_t v)
{
    for (int m = 32; m >= 1; m >>= 1) v |= shfl_xor_u64(v, m);
    return v;
}

__device__ inline uint64_t wave_max_u64(uint64_t v)
{
    for (int m = 32; m >= 1; m >>= 1) {
        uint64_t o = shfl_xor_u64(v, m);
        v = o > v ? o : v;
    }
    return v;
}

__device__ inline bool pred_eval(uint32_t op, uint64_t c, uint64_t v)
{
    switch (op) {
        case AM_PRED_LT: return v < c;
        case AM_PRED_LE: return v <= c;
        case AM_PRED_GT: return v > c;
        case AM_PRED_GE: return v >= c;
        case AM_PRED_EQ: return v == c;
        default: return v != c;
    }
}

// OR over aligned segments of S = 64/eps lanes (S a power of two, 2..32):
// after the call the last lane of each segment holds the OR of the segment
// (Hillis-Steele with DPP row_shr 1/2/4/8 inside rows of 16, then row_bcast
// 15 carries row 0/2 into row 1/3 for S = 32).
template <int CTRL, int ROWM, bool BC>
__device__ inline uint64_t dpp_or64(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, ROWM, 0xf, BC);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, ROWM, 0xf, BC);
    return v | (((uint64_t)hi << 32) | lo);
}
__device__ inline uint64_t segment_or(uint64_t v, uint32_t S)
{
    if (S >= 2) v = dpp_or64<0x111, 0xf, true>(v);
    if (S >= 4) v = dpp_or64<0x112, 0xf, true>(v);
    if (S >= 8) v = dpp_or64<0x114, 0xf, true>(v);
    if (S >= 16) v = dpp_or64<0x118, 0xf, true>(v);
    if (S >= 32) v = dpp_or64<0x142, 0xa, false>(v);
    return v;
}

// ------------------------------------------------------------ agent-scope atomics

__device__ inline uint64_t ld_acq(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_rel(uint64_t* p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t ld32(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st32(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t add32(uint32_t* p, uint32_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t add64(uint64_t* p, uint64_t v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wait until every memory operation this wave issued has completed
__device__ inline void drain() { __builtin_amdgcn_s_waitcnt(0); }

// ------------------------------------------------------------ workspace

// Workspace: header, a small scratch region (sum partials, group-by totals),
// the group-by partial tables, then the look-back descriptors.
struct WsHeader {
    uint32_t epoch; // launch counter of the look-back kernels
    uint32_t abort; // nonzero once a look-back wait gave up (poisoned until re-zeroed)
    uint32_t pad0[62];
};
static_assert(sizeof(WsHeader) == 256, "header size");

constexpr uint64_t WS_SMALL_BYTES = 65536;                  // 8192 u64
constexpr uint32_t GB_MAX_BLOCKS = 256;                      // group-by partial tables
constexpr uint64_t GB_MAX_KEYS = 4096;                       // keys <= 12 bits
constexpr uint64_t WS_GBPART_BYTES = GB_MAX_BLOCKS * 2 * GB_MAX_KEYS * 8; // 16 MiB
constexpr uint64_t WS_FIXED = 256 + WS_SMALL_BYTES + WS_GBPART_BYTES;

__host__ __device__ inline uint64_t* ws_small(void* ws) { return (uint64_t*)((uint8_t*)ws + 256); }
__host__ __device__ inline uint64_t* ws_gbpart(void* ws) { return (uint64_t*)((uint8_t*)ws + 256 + WS_SMALL_BYTES); }

// ------------------------------------------------------------ two-level look-back
//
// Scan element of the single-pass compaction kernels: (count, window,
// overflow[, sum]). The window is the last 63 bits of the dense stream of
// selected values at the output width, right-aligned (bit 62 = last stream
// bit); appending y (m = y.count*bits stream bits) to x gives
// (x.win >> m) | y.win when m < 63, else y.win -- so the tile that completes
// an output word can build it from its predecessors' values.
//
// One block per tile, tiles in launch order: a tile waits only on lower
// tiles, whose blocks were dispatched before it, so the look-back always
// makes progress without a tile counter.
//
// Descriptors carry the launch epoch so stale words from earlier launches
// are recognised without clearing: a status word
//   [63:62] level (1 aggregate, 2 inclusive) [61:40] epoch&(2^22-1) [39:0] count
// and, when the scan carries more than a count, payload words
//   [63:48] epoch&0xffff [47:46] level [45:0] payload
// (window low 46 bits; window high 17 bits | overflow << 17; sum low 46
// bits; sum high 18 bits). Per tile: status, aggregate words, inclusive
// words. Per group of 64 tiles: status and aggregate words of the whole
// group, published by the group's last tile as soon as its 63 group-mates'
// aggregates are in (or during its own walk); the group's inclusive value is
// that tile's. A look-back reads the tiles from its predecessor down to the
// first tile of the previous group (so it never waits for that group's
// words) and, for the 64 groups older than that, each group's aggregate and
// its last tile's status -- in one round trip for a count-only scan, in two
// when payload words come along (holding both sets would cost occupancy) --
// and stops at the nearest tile or group whose inclusive value is known.
// Descriptor reads are
// device-coherent loads that each move a whole memory line; words that many
// walks poll are written once per launch (the group aggregates), since
// rewriting polled lines slows the memory path that serves them.
//
// Every wait is bounded: after LB_WAIT_TICKS of the 100 MHz wall clock the
// waiting wave sets the workspace's abort word and all waiters stop. An
// aborted workspace stays poisoned (launches that see it do nothing except
// marking their statistics invalid) until it is zero-filled again.
constexpr uint32_t LB_CLEAR_SLICES = 64;
constexpr uint64_t CNT_MASK = (1ull << 40) - 1;
constexpr uint64_t P46 = (1ull << 46) - 1;
constexpr uint64_t OVF_MAX = (1ull << 29) - 1;
constexpr uint64_t WIN_BITS = (1ull << 63) - 1;
constexpr uint64_t LB_WAIT_TICKS = 200000000ull; // 2 s at 100 MHz
constexpr uint64_t STATS_INVALID = ~0ull;        // statistics of an aborted launch

enum LbFlags : uint32_t { LB_WIN = 1, LB_OVF = 2, LB_SUM = 4 };

__device__ inline uint64_t lb_status(uint32_t level, uint32_t epoch, uint64_t count)
{
    return ((uint64_t)level << 62) | ((uint64_t)(epoch & 0x3FFFFF) << 40) | (count & CNT_MASK);
}
__device__ inline bool lb_status_ok(uint64_t s, uint32_t epoch)
{
    return (s >> 62) != 0 && ((s >> 40) & 0x3FFFFF) == (epoch & 0x3FFFFF);
}
__device__ inline uint64_t lb_word(uint32_t level, uint32_t epoch, uint64_t payload)
{
    return ((uint64_t)(epoch & 0xFFFF) << 48) | ((uint64_t)level << 46) | (payload & P46);
}
__device__ inline bool lb_word_ok(uint64_t w, uint32_t level, uint32_t epoch)
{
    return (w >> 48) == (epoch & 0xFFFF) && ((w >> 46) & 3) == level;
}
__device__ inline uint64_t win_combine(uint64_t xw, uint64_t yw, uint64_t ycount, uint32_t bits)
{
    uint64_t m = ycount * bits;
    if (m >= 63) return yw & WIN_BITS;
    return ((xw & WIN_BITS) >> m) | (yw & WIN_BITS);
}
__device__ inline uint64_t ovf_add(uint64_t a, uint64_t b)
{
    uint64_t s = a + b;
    return s > OVF_MAX ? OVF_MAX : s;
}

struct Agg {
    uint64_t c, w, o, s;
};

// x (older) then y (newer); only the fields a scan with flags F carries
// are combined (the others stay zero)
template <uint32_t F>
__device__ inline Agg agg_combine(const Agg& x, const Agg& y, uint32_t bits)
{
    Agg r{x.c + y.c, 0, 0, 0};
    if (F & LB_WIN) r.w = win_combine(x.w, y.w, y.c, bits);
    if (F & LB_OVF) r.o = ovf_add(x.o, y.o);
    if (F & LB_SUM) r.s = x.s + y.s;
    return r;
}

template <uint32_t F>
__device__ inline Agg shfl_agg(const Agg& a, int m)
{
    Agg r{shfl_xor_u64(a.c, m), 0, 0, 0};
    if (F & LB_WIN) r.w = shfl_xor_u64(a.w, m);
    if (F & LB_OVF) r.o = shfl_xor_u64(a.o, m);
    if (F & LB_SUM) r.s = shfl_xor_u64(a.s, m);
    return r;
}

// Ordered wave reduction: lane j holds element j, higher lanes are OLDER;
// every lane returns the combination oldest..newest.
template <uint32_t F>
__device__ inline Agg wave_combine_ordered(Agg v, uint32_t lane, uint32_t bits)
{
    for (int m = 1; m < 64; m <<= 1) {
        const Agg o = shfl_agg<F>(v, m);
        v = (lane & m) ? agg_combine<F>(v, o, bits) : agg_combine<F>(o, v, bits);
    }
    return v;
}

template <uint32_t F>
struct Lookback {
    static constexpr uint32_t NX = ((F & (LB_WIN | LB_OVF)) ? 2 : 0) + ((F & LB_SUM) ? 2 : 0); // payload words
    static constexpr uint32_t NXA = NX ? NX : 1;                                           // array size
    static constexpr uint32_t TS = 1 + 2 * NX; // tile stride: status, aggregate words, inclusive words
    static constexpr uint32_t GS = 1 + NX;     // group stride: status, aggregate words
    uint64_t* tiles;  // [cap][TS]
    uint64_t* groups; // [cap/64 + 2][GS]
    uint64_t cap;     // tiles
    uint32_t epoch, bits;
    WsHeader* hdr;

    __host__ __device__ static constexpr uint64_t bytes_for(uint64_t tiles_cap)
    {
        return tiles_cap * TS * 8 + (tiles_cap / 64 + 2) * GS * 8;
    }
    __host__ __device__ static Lookback at(void* base, uint64_t bytes)
    {
        Lookback L{};
        const uint64_t fixed = 2 * GS * 8;
        uint64_t c = bytes > fixed ? ((bytes - fixed) * 64) / (TS * 8 * 64 + GS * 8) : 0;
        while (c && bytes_for(c) > bytes) c--;
        while (bytes_for(c + 1) <= bytes) c++;
        L.cap = c;
        L.tiles = (uint64_t*)base;
        L.groups = L.tiles + L.cap * TS;
        return L;
    }
    __device__ uint64_t* tile(uint64_t t) const { return tiles + t * TS; }
    __device__ uint64_t* group(uint64_t g) const { return groups + g * GS; }

    // ---- bounded waiting
    __device__ bool aborted() const { return ld32(&hdr->abort) != 0; }
    // one round of a wait loop (wave-uniform); false once the launch gave up
    __device__ bool keep_waiting(uint32_t& it, uint64_t& t0) const
    {
        if (it++ == 0) t0 = wall_clock64();
        __builtin_amdgcn_s_sleep(2);
        if ((it & 63) == 0) {
            if (aborted()) return false;
            if (wall_clock64() - t0 > LB_WAIT_TICKS) {
                st32(&hdr->abort, 1u);
                return false;
            }
        }
        return true;
    }

    // ---- payload words
    __device__ void put_x(uint64_t* p, uint32_t level, const Agg& a) const
    {
        uint32_t k = 0;
        if (F & (LB_WIN | LB_OVF)) {
            st_rel(&p[k++], lb_word(level, epoch, a.w));
            st_rel(&p[k++], lb_word(level, epoch, (a.w >> 46) | (a.o << 17)));
        }
        if (F & LB_SUM) {
            st_rel(&p[k++], lb_word(level, epoch, a.s));
            st_rel(&p[k++], lb_word(level, epoch, a.s >> 46));
        }
    }
    __device__ void load_x(const uint64_t* p, uint64_t* x) const
    {
        for (uint32_t k = 0; k < NX; k++) x[k] = ld_acq(&p[k]);
    }
    // decode payload words (true when every tag is current)
    __device__ bool decode(const uint64_t* x, uint32_t level, uint64_t count, Agg& a) const
    {
        bool ok = true;
        for (uint32_t k = 0; k < NX; k++) ok = ok && lb_word_ok(x[k], level, epoch);
        a.c = count;
        a.w = 0;
        a.o = 0;
        a.s = 0;
        uint32_t k = 0;
        if (F & (LB_WIN | LB_OVF)) {
            a.w = (x[0] & P46) | ((x[1] & 0x1FFFF) << 46);
            a.o = (x[1] & P46) >> 17;
            k = 2;
        }
        if (F & LB_SUM) a.s = (x[k] & P46) | ((x[k + 1] & P46) << 46);
        return ok;
    }

    // Wave-level: publish tile t's aggregate, look back, publish its
    // inclusive value. Returns the exclusive prefix (all lanes).
    //
    // One round trip in the common case: the window of tiles from t-1 down
    // to the first tile of the previous group (at most 127, two per lane)
    // and the groups older than the window (64, one per lane) are read
    // together; the walk stops at the nearest inclusive value in the window,
    // else combines the window's aggregates with whole groups.
    __device__ Agg run(uint64_t t, const Agg& own, uint32_t lane) const
    {
        publish_agg(t, own, lane);
        return walk(t, own, lane);
    }
    // tile t's aggregate (tile 0: its inclusive value)
    __device__ void publish_agg(uint64_t t, const Agg& own, uint32_t lane) const
    {
        if (lane == 0) {
            uint64_t* D = tile(t);
            const uint32_t lvl = t == 0 ? 2 : 1;
            if (NX) put_x(D + 1 + (lvl - 1) * NX, lvl, own);
            st_rel(&D[0], lb_status(lvl, epoch, own.c));
        }
    }
    // The last tile of a group (t = 64g+63), its aggregate published: once
    // its 63 group-mates have published theirs, publish the group's
    // aggregate, so later look-backs need not wait for this tile's own
    // look-back. Skipped when a mate holds an inclusive value (only tile 0
    // can: later walks then stop at this tile's inclusive value instead).
    __device__ void publish_group_early(uint64_t t, const Agg& own, uint32_t lane) const
    {
        const Agg zero{0, 0, 0, 0};
        const bool act = lane < 63;
        const uint64_t* q = tile(act ? t - 1 - lane : t);
        uint64_t st = 0, xa[NXA];
        if (act) {
            st = ld_acq(q);
            load_x(q + 1, xa);
        }
        uint32_t it = 0;
        uint64_t t0 = 0;
        Agg v = zero;
        for (;;) {
            const bool ok = act && lb_status_ok(st, epoch);
            if (ballot(ok && (st >> 62) == 2)) return;
            const bool dec = !act || (ok && decode(xa, 1, st & CNT_MASK, v));
            if (!ballot(!dec)) break;
            if (!keep_waiting(it, t0)) return;
            if (!dec) {
                if (!ok) st = ld_acq(q);
                load_x(q + 1, xa);
            }
        }
        const Agg ga = agg_combine<F>(wave_combine_ordered<F>(act ? v : zero, lane, bits), own, bits);
        if (lane == 0) {
            if (NX) put_x(group(t >> 6) + 1, 1, ga);
            st_rel(group(t >> 6), lb_status(1, epoch, ga.c));
        }
    }
    // The look-back of a tile whose aggregate is published; returns the
    // exclusive prefix (all lanes). Only a group's last tile publishes its
    // inclusive value (the one later walks stop at; the others would mostly
    // arrive too late to shorten a walk and cost a second write into a line
    // that walks are polling). group_out: the group aggregate is already out.
    __device__ Agg walk(uint64_t t, const Agg& own, uint32_t lane, bool group_out = false) const
    {
        const Agg zero{0, 0, 0, 0};
        uint64_t* D = tile(t);
        if (t == 0) return zero;
        const uint64_t g = t >> 6;
        const uint32_t below = (uint32_t)(t & 63);
        const uint64_t lo = g ? (g - 1) * 64 : 0;
        const uint32_t nw = (uint32_t)(t - lo); // window tiles [lo, t): 1..127
        // window index i: near i = lane (tile t-1-lane), far i = 64 + lane (tile t-65-lane)
        const bool an = lane < nw, af = 64 + lane < nw;
        const uint64_t* qn = tile(an ? t - 1 - lane : t);
        const uint64_t* qf = tile(af ? t - 65 - lane : t);
        // groups older than the window: lane j -> group g-2-j
        const int64_t h0 = (int64_t)g - 2 - (int64_t)lane;
        const uint64_t* gq = group(h0 >= 0 ? (uint64_t)h0 : 0);
        const uint64_t* lt = tile(h0 >= 0 ? (uint64_t)h0 * 64 + 63 : 0);
        // aggregate payloads come with the statuses; the one inclusive
        // payload a walk uses (at its stop) is fetched once the stop is known
        uint64_t sn = 0, sf = 0, gs = 0, ls = 0;
        uint64_t xan[NXA], xaf[NXA], xg[NXA];
        if (an) {
            sn = ld_acq(qn);
            load_x(qn + 1, xan);
        }
        if (af) {
            sf = ld_acq(qf);
            load_x(qf + 1, xaf);
        }
        if (NX == 0 && h0 >= 0) { // (with payload words, holding these too costs occupancy)
            ls = ld_acq(lt);
            gs = ld_acq(gq);
            load_x(gq + 1, xg);
        }
        bool gave_up = false;
        uint32_t it = 0, stop = nw;
        uint64_t t0 = 0;
        Agg vn = zero, vf = zero;
        bool needn = false, needf = false;
        for (;;) {
            const bool okn = an && lb_status_ok(sn, epoch), okf = af && lb_status_ok(sf, epoch);
            const uint64_t pn = ballot(okn && (sn >> 62) == 2), pf = ballot(okf && (sf >> 62) == 2);
            stop = pn ? (uint32_t)__builtin_ctzll(pn) : (pf ? 64 + (uint32_t)__builtin_ctzll(pf) : nw);
            const uint32_t fi = 64 + lane;
            needn = an && (lane < stop || (lane == stop && stop < nw));
            needf = af && (fi < stop || (fi == stop && stop < nw));
            bool dn = true, df = true;
            if (needn) {
                if (lane < stop) dn = okn && decode(xan, 1, sn & CNT_MASK, vn);
                else {
                    uint64_t xp[NXA];
                    load_x(qn + 1 + NX, xp);
                    dn = decode(xp, 2, sn & CNT_MASK, vn);
                }
            }
            if (needf) {
                if (fi < stop) df = okf && decode(xaf, 1, sf & CNT_MASK, vf);
                else {
                    uint64_t xp[NXA];
                    load_x(qf + 1 + NX, xp);
                    df = decode(xp, 2, sf & CNT_MASK, vf);
                }
            }
            if (!ballot((needn && !dn) || (needf && !df))) break;
            if (!keep_waiting(it, t0)) {
                gave_up = true;
                break;
            }
            if (needn && !dn && lane < stop) {
                if (!okn) sn = ld_acq(qn);
                load_x(qn + 1, xan);
            }
            if (needf && !df && fi < stop) {
                if (!okf) sf = ld_acq(qf);
                load_x(qf + 1, xaf);
            }
        }
        if (!needn) vn = zero;
        if (!needf) vf = zero;
        Agg excl = wave_combine_ordered<F>(vn, lane, bits);
        if (nw > 64 && stop >= 64) excl = agg_combine<F>(wave_combine_ordered<F>(vf, lane, bits), excl, bits);
        bool done = stop < nw || gave_up;
        if (below == 63 && stop >= 63 && !gave_up && !group_out) {
            // group aggregate of group g: its other 63 tiles are near lanes 0..62
            const Agg ga = agg_combine<F>(wave_combine_ordered<F>(lane < 63 ? vn : zero, lane, bits), own, bits);
            if (lane == 0) {
                if (NX) put_x(group(g) + 1, 1, ga);
                st_rel(group(g), lb_status(1, epoch, ga.c));
            }
        }
        // whole groups, 64 per step (lane j -> group gb-1-j); a count-only
        // scan has the first step's words in hand
        for (uint64_t gb = g - 1; !done;) {
            const int64_t h = (int64_t)gb - 1 - (int64_t)lane;
            const bool act = h >= 0;
            if (NX != 0 || gb != g - 1) {
                gq = group(act ? (uint64_t)h : 0);
                lt = tile(act ? (uint64_t)h * 64 + 63 : 0);
                if (act) {
                    ls = ld_acq(lt);
                    gs = ld_acq(gq);
                    load_x(gq + 1, xg);
                }
            }
            uint32_t it2 = 0, gstop = 64;
            uint64_t t2 = 0, p2 = 0;
            Agg vp = zero, vg = zero;
            for (;;) {
                // stop at the nearest group whose last tile has its inclusive value
                p2 = ballot(!act || (lb_status_ok(ls, epoch) && (ls >> 62) == 2));
                gstop = p2 ? (uint32_t)__builtin_ctzll(p2) : 64;
                bool hasP = false;
                if (lane == gstop && act) {
                    uint64_t xl[NXA];
                    load_x(lt + 1 + NX, xl);
                    hasP = decode(xl, 2, ls & CNT_MASK, vp);
                }
                const bool hasG = act && lb_status_ok(gs, epoch) && decode(xg, 1, gs & CNT_MASK, vg);
                const bool need = (lane < gstop && !hasG) || (lane == gstop && act && !hasP);
                if (!ballot(need)) break;
                if (!keep_waiting(it2, t2)) {
                    gave_up = true;
                    break;
                }
                if (need && lane < gstop) {
                    ls = ld_acq(lt);
                    gs = ld_acq(gq);
                    load_x(gq + 1, xg);
                }
            }
            const Agg v = lane < gstop ? vg : ((lane == gstop && act) ? vp : zero);
            const Agg part = wave_combine_ordered<F>(v, lane, bits);
            excl = agg_combine<F>(part, excl, bits);
            done = p2 != 0 || gave_up;
            gb = gb >= 64 ? gb - 64 : 0;
        }
        if (lane == 0 && !gave_up && below == 63) {
            const Agg incl = agg_combine<F>(excl, own, bits);
            if (NX) put_x(D + 1 + NX, 2, incl);
            st_rel(&D[0], lb_status(2, epoch, incl.c));
        }
        return excl;
    }

    // zero one slice of the tile and group descriptors not used by this
    // launch (tiles >= T, groups >= T/64), spread over the launch's threads
    __device__ void clear_slice(uint64_t T, uint64_t tid, uint64_t nthreads) const
    {
        const uint64_t per = (cap + LB_CLEAR_SLICES - 1) / LB_CLEAR_SLICES;
        uint64_t lo = (uint64_t)(epoch % LB_CLEAR_SLICES) * per;
        const uint64_t hi = lo + per < cap ? lo + per : cap;
        const uint64_t glo = lo / 64 < T / 64 ? T / 64 : lo / 64, ghi = (hi + 63) / 64;
        if (lo < T) lo = T;
        for (uint64_t i = lo + tid; i < hi; i += nthreads)
            for (uint32_t k = 0; k < TS; k++) st_rel(&tiles[i * TS + k], 0);
        for (uint64_t i = glo + tid; i < ghi && i < cap / 64 + 2; i += nthreads)
            for (uint32_t k = 0; k < GS; k++) st_rel(&groups[i * GS + k], 0);
    }
};


// Write the output words a tile of compacted values completes: ctile values
// s_vals[0, ctile) (LDS) at element P of a packed column `o` (slab or
// bitstream, capacity ocap elements; values past it stay unwritten). Values
// before P come from `carry`, the 63 output-stream bits before the tile
// (bit 62 = the bit just before P's first bit). The last tile also writes
// its final partial word. Tile-relative indices fit 32 bits (a tile holds
// < 2^17 values); nthr threads, this one tid.
template <int OK, typename V>
__device__ inline void write_packed_tile(const Col& o, uint64_t ocap, const V* s_vals, uint64_t P, uint64_t carry,
                                         uint32_t ctile, bool last, uint32_t tid, uint32_t nthr)
{
    const uint32_t bo = o.bits;
    const uint64_t om = o.mask;
    if (OK == K_SLAB) {
        // word kf + j holds tile values [j*eps - q0, (j+1)*eps - q0); values
        // before the tile come from the carried window. Tile-relative
        // indices fit 32 bits (a tile holds < 2^17 values).
        const uint32_t eps = o.eps;
        const uint64_t kf = P / eps;
        const int32_t q0 = (int32_t)(P - kf * eps); // carried values in the first word
        uint64_t kl = (P + ctile) / eps;
        if (last && ((P + ctile) % eps)) kl += 1;
        const uint64_t wcap = (ocap + eps - 1) / eps;
        if (kl > wcap) kl = wcap;
        const uint32_t nk = kl > kf ? (uint32_t)(kl - kf) : 0;
        // values at or past min(ctile, ocap - P) stay zero
        const int64_t room = (int64_t)ocap - (int64_t)P;
        const int32_t lim = room < (int64_t)ctile ? (int32_t)(room < -(int64_t)eps ? -(int64_t)eps : room) : (int32_t)ctile;
        for (uint32_t j = tid; j < nk; j += nthr) {
            const int32_t i0 = (int32_t)(j * eps) - q0;
            uint64_t word = 0;
            if (i0 >= 0 && i0 + (int32_t)eps <= lim) { // whole word from the staged values
                for (uint32_t s = 0, sh = 0; s < eps; s++, sh += bo) word |= (uint64_t)s_vals[i0 + s] << sh;
            }
            else {
                for (uint32_t s = 0, sh = 0; s < eps; s++, sh += bo) {
                    const int32_t i = i0 + (int32_t)s;
                    if (i >= lim) break;
                    const uint64_t v = i < 0 ? (carry >> (63 - (uint32_t)(-i) * bo)) & om : (uint64_t)s_vals[i];
                    word |= v << sh;
                }
            }
            st_nt(gmem(o.p) + kf + j, word);
        }
    }
    else {
        // word kf + j covers tile stream bits [j*64 - o0, j*64 - o0 + 64);
        // bits before the tile come from the carried window
        const uint64_t B0 = P * bo, B1 = (P + ctile) * bo, Bcap = ocap * bo;
        const uint64_t kf = B0 >> 6;
        const int32_t o0 = (int32_t)(B0 & 63);
        uint64_t kl = B1 >> 6;
        if (last && (B1 & 63)) kl += 1;
        const uint64_t wcap = (Bcap + 63) >> 6;
        if (kl > wcap) kl = wcap;
        const uint32_t nk = kl > kf ? (uint32_t)(kl - kf) : 0;
        // x / bo = (x * (floor(2^32/bo) + 1)) >> 32, exact for x < 2^24 (the
        // multiplier needs 33 bits when bo = 1)
        const uint64_t bmagic = (1ull << 32) / bo + 1;
        auto div_bo = [&](uint32_t x) { return (uint32_t)(((uint64_t)x * bmagic) >> 32); };
        for (uint32_t j = tid; j < nk; j += nthr) {
            const int32_t rb = (int32_t)(j * 64) - o0; // tile stream bit of the word's bit 0
            uint64_t word = 0;
            uint32_t i = 0;
            if (rb < 0) word = carry >> (63 + rb); // carry bit 62 = stream bit B0 - 1
            else i = div_bo((uint32_t)rb);
            uint32_t ie = div_bo((uint32_t)(rb + 63 + (int32_t)bo)); // one past the last value touching the word
            if (ie > ctile) ie = ctile;
            for (; i < ie; i++) {
                const int32_t pos = (int32_t)(i * bo) - rb;
                const uint64_t v = s_vals[i];
                word |= pos >= 0 ? (v << pos) : (v >> (-pos));
            }
            const uint64_t b0 = (kf + j) * 64;
            if (b0 + 64 > Bcap) word &= lowmask((uint32_t)(Bcap - b0));
            st_nt(gmem(o.p) + kf + j, word);
        }
    }
}

} // namespace am
