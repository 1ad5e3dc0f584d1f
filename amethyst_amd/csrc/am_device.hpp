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
    const uint64_t* base; // first word touched by the tile
    uint32_t r0;          // SLAB: t0 % eps ; BITSTREAM: (t0*bits) % 64
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
            base = (const uint64_t*)((const uint8_t*)col.p + t0 * col.cbytes);
            r0 = 0;
        }
        else if (KIND == K_SLAB) {
            uint64_t q = t0 / col.eps;
            base = col.p + q;
            r0 = (uint32_t)(t0 - q * col.eps);
        }
        else {
            uint64_t bit = t0 * (uint64_t)col.bits;
            base = col.p + (bit >> 6);
            r0 = (uint32_t)(bit & 63);
        }
        const uint64_t rel = col.words - (uint64_t)(base - col.p); // words from base to the end
        last32 = (uint32_t)(rel * 2 > 0xFFFFFFFFull ? 0xFFFFFFFFull : (rel ? rel * 2 - 1 : 0));
    }

    // element t0 + i (caller guarantees it is < n)
    __device__ inline uint64_t get(uint32_t i) const
    {
        if (KIND == K_STRIDED) {
            switch (cbytes) {
                case 1: return ((const uint8_t*)base)[i] & mask;
                case 2: return ((const uint16_t*)base)[i] & mask;
                case 4: return ((const uint32_t*)base)[i] & mask;
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
            const uint64_t* p = base + (lb >> 6);
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
    uint32_t sh;
    // Branch-free: callers clamp `i` to a valid element and mask the result.
    __device__ inline void load(const TileReader<IK>& rd, uint32_t i)
    {
        sh = 0;
        if (IK == K_STRIDED) {
            switch (rd.cbytes) { // uniform
                case 1: raw = ((const uint8_t*)rd.base)[i]; break;
                case 2: raw = ((const uint16_t*)rd.base)[i]; break;
                case 4: raw = ((const uint32_t*)rd.base)[i]; break;
                default: raw = rd.base[i];
            }
        }
        else if (IK == K_SLAB) {
            const uint32_t idx = rd.r0 + i;
            const uint32_t q = rd.eps == 1 ? idx : __umulhi(idx, rd.eps_magic);
            raw = rd.base[q];
            sh = (idx - q * rd.eps) * rd.bits;
        }
        else if (!WIDE) {
            const uint32_t lb = rd.r0 + i * rd.bits;
            const uint32_t* p32 = (const uint32_t*)rd.base;
            const uint32_t i0 = lb >> 5;
            const uint32_t i1 = i0 + 1 <= rd.last32 ? i0 + 1 : rd.last32;
            raw = ((uint64_t)p32[i1] << 32) | p32[i0];
            sh = lb & 31;
        }
        else {
            raw = rd.get(i);
        }
    }
    __device__ inline uint64_t value(uint64_t mask) const { return (raw >> sh) & mask; }
};

// ------------------------------------------------------------ wave64 helpers

__device__ inline uint32_t lane_id() { return __lane_id(); }

// number of set bits of m below this lane (v_mbcnt_lo/hi)
__device__ inline uint32_t mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

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
    uint32_t pad0[63];
};
static_assert(sizeof(WsHeader) == 256, "header size");

constexpr uint64_t WS_SMALL_BYTES = 65536;                  // 8192 u64
constexpr uint32_t GB_MAX_BLOCKS = 256;                      // group-by partial tables
constexpr uint64_t GB_MAX_KEYS = 4096;                       // keys <= 12 bits
constexpr uint64_t WS_GBPART_BYTES = GB_MAX_BLOCKS * 2 * GB_MAX_KEYS * 8; // 16 MiB
constexpr uint64_t WS_FIXED = 256 + WS_SMALL_BYTES + WS_GBPART_BYTES;
constexpr uint64_t WS_DESC_BYTES_PER_TILE = 5 * 8; // status + 2 window words per level

__host__ __device__ inline uint64_t* ws_small(void* ws) { return (uint64_t*)((uint8_t*)ws + 256); }
__host__ __device__ inline uint64_t* ws_gbpart(void* ws) { return (uint64_t*)((uint8_t*)ws + 256 + WS_SMALL_BYTES); }

// Look-back descriptors. Every word carries the launch epoch, so stale words
// from earlier launches are recognised without ever clearing the arrays:
//   status[t]    [63:62] level (1 aggregate, 2 inclusive) [61:40] epoch&(2^22-1) [39:0] count
//   winA/winP[2t..2t+1]  [63:48] epoch&0xffff [47:46] level [45:0] payload
// The two payload words hold a 63-bit window (46 + 17 bits) and a 29-bit
// saturating overflow count. Each launch also zeroes one slice of the
// indices it does not use, so no word can outlive 2^16 launches unrewritten.
struct Lookback {
    uint64_t* status; // [cap]
    uint64_t* winA;   // [2*cap]
    uint64_t* winP;   // [2*cap]
    uint64_t cap;
};

__host__ __device__ inline Lookback lookback_of(void* ws, uint64_t ws_bytes)
{
    Lookback L;
    uint64_t cap = ws_bytes > WS_FIXED ? (ws_bytes - WS_FIXED) / WS_DESC_BYTES_PER_TILE : 0;
    uint64_t* d = (uint64_t*)((uint8_t*)ws + WS_FIXED);
    L.status = d;
    L.winA = d + cap;
    L.winP = d + 3 * cap;
    L.cap = cap;
    return L;
}

constexpr uint32_t LB_CLEAR_SLICES = 64;
constexpr uint64_t CNT_MASK = (1ull << 40) - 1;
constexpr uint64_t P46 = (1ull << 46) - 1;
constexpr uint64_t OVF_MAX = (1ull << 29) - 1;
constexpr uint64_t WIN_BITS = (1ull << 63) - 1;

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

// Scan element of the compaction: (count, window, overflow). The window is
// the last 63 bits of the dense stream of selected values at the output
// width, right-aligned (bit 62 = last stream bit). Appending y
// (m = y.count*bits stream bits) to x: (x.win >> m) | y.win if m < 63, else y.win.
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

} // namespace am
