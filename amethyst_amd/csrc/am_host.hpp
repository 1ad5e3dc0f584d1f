// am_host.hpp -- host-side helpers shared by the C-ABI entry points:
// column validation and resolution into the device-side Col descriptor.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/amethyst.h"
#include "am_device.hpp"

// descriptor bytes of the largest compaction launch over n elements (k_compress.hip)
uint64_t am_compress_desc_bytes(uint64_t n);

namespace am { struct Col; }
// group-by over keys of 13..24 bits (k_groupby_wide.hip): workspace bytes and launch
uint64_t am_groupby_wide_bytes(uint64_t n, uint32_t kbits, uint32_t vbits);
int am_groupby_wide(const am::Col& ck, const am::Col& cv, const am_predicate* value_pred, uint32_t mode,
                    uint64_t* out_keys, uint64_t* out_sums, uint64_t* out_counts, uint64_t* result_dev, void* ws,
                    uint64_t ws_bytes, hipStream_t s);

// autotuned launch shape for (knob, bits, in format, out format), or 0 (capi.hip)
uint32_t am_tuned(uint32_t knob, uint32_t bits, uint32_t in_format, uint32_t out_format);

namespace am {

inline uint32_t tight_bytes(uint32_t bits)
{
    // tight_type<bits>::type, src/common_code.cuh:11-19
    uint32_t bc = (bits + 7) / 8;
    return bc > 4 ? 8 : (bc > 2 ? 4 : (bc > 1 ? 2 : 1));
}

inline bool valid_format(uint32_t fmt, uint32_t bits)
{
    if (bits < 1 || bits > 64) return false;
    if (fmt == AM_FMT_U32) return bits <= 32;
    return fmt <= AM_FMT_U32;
}

// Resolve a public column into the device descriptor (AM_OK or an error).
inline int make_col(const am_column* c, Col* out)
{
    if (!c) return AM_E_INVALID;
    if (!valid_format(c->format, c->bits)) return AM_E_INVALID;
    if (c->n && !c->data) return AM_E_INVALID;
    if (c->n >= (1ull << 40)) return AM_E_UNSUPPORTED_CONFIG;
    Col d{};
    d.p = (uint64_t*)c->data;
    d.n = c->n;
    d.bits = c->bits;
    d.mask = lowmask(c->bits);
    d.eps = 64 / c->bits;
    d.eps_magic = d.eps > 1 ? (uint32_t)((1ull << 32) / d.eps + 1) : 0;
    d.words = am_packed_words(c->format, c->bits, c->n);
    switch (c->format) {
        case AM_FMT_BASELINE: d.kind = K_STRIDED; d.cbytes = 8; break;
        case AM_FMT_TIGHT: d.kind = K_STRIDED; d.cbytes = tight_bytes(c->bits); break;
        case AM_FMT_U32: d.kind = K_STRIDED; d.cbytes = 4; break;
        case AM_FMT_SLAB: d.kind = K_SLAB; d.cbytes = 8; break;
        default:
            // a bitstream whose width divides 64 (2..64) is byte for byte the
            // slab of that width (no straddles, no dead bits): the word kernels
            // take it; 1-bit columns stay streams
            d.kind = (c->bits >= 2 && 64 % c->bits == 0) ? K_SLAB : K_BITSTREAM;
            d.cbytes = 8;
            break;
    }
    *out = d;
    return AM_OK;
}

// Normalise `v OP c` over values in [0, vmax] to selected = ((v-lo) <= span) != inv.
inline Pred make_pred(const am_predicate& p, uint64_t vmax)
{
    // the selected set as a closed range [a, b] (or its complement for NE)
    uint64_t a = 0, b = vmax;
    bool empty = false, inv = false;
    switch (p.op) {
        case AM_PRED_LT: if (p.c == 0) empty = true; else b = p.c - 1 < vmax ? p.c - 1 : vmax; break;
        case AM_PRED_LE: b = p.c < vmax ? p.c : vmax; break;
        case AM_PRED_GT: if (p.c >= vmax) empty = true; else a = p.c + 1; break;
        case AM_PRED_GE: if (p.c > vmax) empty = true; else a = p.c; break;
        case AM_PRED_EQ: if (p.c > vmax) empty = true; else a = b = p.c; break;
        default: // NE
            if (p.c > vmax) break; // everything
            a = b = p.c;
            inv = true;
            break;
    }
    Pred r{};
    if (empty) { r.lo = 0; r.span = ~0ull; r.inv = 1; return r; }
    r.lo = a;
    r.span = b - a;
    r.inv = inv ? 1 : 0;
    return r;
}

// The predicate scaled to top-aligned fields: (v - lo) <= span over values
// of `bits` <= 32 bits becomes (y - lo*2^k) <= (span+1)*2^k - 1, k = 32 - bits,
// for y = v*2^k + (anything < 2^k); inv flips the eps selection bits.
inline TopPred make_top_pred(const Pred& p, uint32_t bits, uint32_t eps)
{
    const uint32_t k = 32 - bits;
    const uint64_t vmask = lowmask(bits);
    const uint64_t sp1 = p.span >= vmask ? (vmask + 1) : p.span + 1;
    TopPred t;
    t.lo = (uint32_t)(p.lo << k);
    t.rg = (uint32_t)((sp1 << k) - 1);
    t.inv = p.inv ? (uint32_t)lowmask(eps) : 0;
    return t;
}

// SWAR selection (select_fields_pow2) applies to fields of fw = 2, 4 or 8
// bits with no dead bits (fw == bits) under a range from 0 or one ending at
// the top value (the complement of v < lo): fills its slot masks and the
// flip of the result; false when it does not apply.
struct SwarSel {
    uint64_t E, K, G;
    uint32_t inv;
};
inline bool swar_select(const Pred& pr, const TopPred& tp, uint32_t bits, uint32_t fw, uint32_t eps, SwarSel* o)
{
    if (bits > 32 || bits != fw || (fw != 2 && fw != 4 && fw != 8)) return false;
    const uint64_t vmax = lowmask(bits);
    const bool from0 = tp.lo == 0, to_top = pr.lo && pr.span >= vmax - pr.lo; // [0, ..) or [lo, vmax]
    if (!from0 && !to_top) return false;
    const uint64_t sp1 = from0 ? ((uint64_t)tp.rg + 1) >> (32 - bits) : pr.lo; // 1 .. 2^fw: v < sp1
    *o = SwarSel{0, 0, 0, from0 ? tp.inv : tp.inv ^ (uint32_t)lowmask(eps)};
    for (uint32_t k = 0; k < eps / 2; k++) {
        o->E |= lowmask(fw) << (2 * fw * k);
        o->K |= ((1ull << fw) - sp1) << (2 * fw * k);
        o->G |= (1ull << fw) << (2 * fw * k);
    }
    return true;
}

// The selected values of a predicate over `bits` <= 32-bit values as one
// cyclic range [c0, c0 + m) (mod 2^bits), scaled to values top-aligned in 32
// bits with arbitrary bits beneath: selected <=> (y - lo) mod 2^32 <= rg.
// One range covers every operator, NE included (its complement of one value
// wraps around). Returns false when nothing is selected.
inline bool cyclic_top_range(const Pred& p, uint32_t bits, uint32_t* lo, uint32_t* rg)
{
    const uint64_t vmax = lowmask(bits);
    uint64_t c0 = p.lo, m = p.span >= vmax ? vmax + 1 : p.span + 1;
    if (p.inv) {
        c0 = (p.lo + m) & vmax;
        m = vmax + 1 - m;
    }
    const uint32_t k = 32 - bits;
    *lo = (uint32_t)(c0 << k);
    *rg = (uint32_t)((m << k) - 1);
    return m != 0;
}

inline int check_launch()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? AM_OK : AM_E_HIP;
}

// Largest grid whose blocks are all resident at once (the tiles of a block
// wait only on lower tiles, which then always have a running owner).
template <typename K>
static uint32_t resident_blocks(K kernel, int threads, size_t lds)
{
    int dev = 0, cus = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, lds) != hipSuccess || per < 1) per = 1;
    if (cus < 1) cus = 1;
    return (uint32_t)(per * cus);
}

} // namespace am
