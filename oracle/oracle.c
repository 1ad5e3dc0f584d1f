/*
 * oracle.c -- CPU restatement of the reference's packed-column semantics.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (the amethyst_amd package
 * and its HIP library) may load or call this file; only tests/, the smoke
 * check and the CPU-baseline leg of bench.py use it, as the checker.
 *
 * Every function restates the reference (yogi-tud/Amethyst) from the cited
 * file:line. The reference keeps its "CPU implementation" in the validation
 * loops of its bench files; those loops decode each element through an
 * accessor (get_element / get_elem_nogaps / get_element_baseline) and compare
 * against the GPU result. The functions here do the same per-element work.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   - prng: against the reference's own fast_prng.cuh compiled into
 *     oracle/_ref (when built) and a committed golden fixture, plus the
 *     published PCG32 / splitmix64 algorithm.
 *   - formats: closed forms (bits in {8,16,32} slab == plain array, bits=64
 *     identity, bitstream == numpy bit unpacking), round trips.
 *   - filter / compressstore / sum / binary / group-by: numpy vectorised
 *     restatements (np.packbits, boolean indexing, np.bincount, uint64
 *     wrapping arithmetic) and invariants (count == popcount, Σcount == n).
 *   - mask generators: byte-level fixtures of the reference's
 *     generate_mask_uniform / create_bitmask, cross-checked against
 *     oracle/_ref where it builds.
 *   - matmul: brute-force int64 triple loop vs numpy int64 matmul.
 * Parity unpinned: none of the functions below (see tests/test_oracle.py).
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdlib.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/amethyst.h"

/* ------------------------------------------------------------------ prng */

/* fast_prng: PCG32 (XSH-RR) seeded via splitmix64, inc = 1.
 * src/fast_prng.cuh:11-19 (srand), :21-30 (rand). */
typedef struct {
    uint64_t state;
    uint64_t inc;
} or_rng;

static void or_srand(or_rng* r, uint64_t seed)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    r->state = z ^ (z >> 31);
    r->inc = 1;
}

static uint32_t or_rand(or_rng* r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ull + (r->inc | 1);
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

/* out[i] = draw (skip + i) of the stream seeded with `seed` (serial). */
void or_prng_fill(uint64_t seed, uint64_t skip, uint32_t* out, size_t n)
{
    or_rng r;
    or_srand(&r, seed);
    for (uint64_t i = 0; i < skip; i++) or_rand(&r);
    for (size_t i = 0; i < n; i++) out[i] = or_rand(&r);
}

/* Input data of the reference driver, src/main.cu:239-251: both buffers are
 * filled from one generator seeded 42, alternating, each word = hi<<32 | lo,
 * a zero word replaced by 1. */
void or_gen_main_words(uint64_t seed, uint64_t* d1, uint64_t* d2, size_t n)
{
    or_rng r;
    or_srand(&r, seed);
    for (size_t i = 0; i < n; i++) {
        uint64_t hi = or_rand(&r);
        d1[i] = (hi << 32) | (uint64_t)or_rand(&r);
        if (d1[i] == 0) d1[i] = 1;
        hi = or_rand(&r);
        d2[i] = (hi << 32) | (uint64_t)or_rand(&r);
        if (d2[i] == 0) d2[i] = 1;
    }
}

/* generate_mask_uniform, src/main.cu:33-46: its own generator seeded 42;
 * bit j of each byte, j from 7 down to 0 (MSB-first), is set iff
 * rand() < (uint64)(selectivity * UINT32_MAX). */
void or_mask_uniform_ref(uint8_t* buf, uint64_t offset, uint64_t length, double selectivity)
{
    uint64_t p_adjusted = (uint64_t)(selectivity * (double)UINT32_MAX);
    or_rng r;
    or_srand(&r, 42);
    for (uint64_t i = offset; i < offset + length; i++) {
        uint8_t acc = 0;
        for (int j = 7; j >= 0; j--) {
            if ((uint64_t)or_rand(&r) < p_adjusted) acc |= (uint8_t)(1u << j);
        }
        buf[i] = acc;
    }
}

/* create_bitmask, src/main.cu:59-108: `cluster_count` runs of
 * (size_t)(selectivity*total)/cluster_count ones, run i starting at
 * i*(total/cluster_count); bytes are MSB-first; total/8 bytes are produced.
 * selectivity is a float and the product is formed in float, as there. */
void or_mask_clusters_ref(uint8_t* out, float selectivity, size_t cluster_count, size_t total)
{
    size_t total_set_one = (size_t)(selectivity * (float)total);
    size_t cluster_size = total_set_one / cluster_count;
    size_t slice = total / cluster_count;
    memset(out, 0, total / 8);
    for (size_t i = 0; i < cluster_count; i++) {
        for (size_t k = 0; k < cluster_size; k++) {
            size_t bit = k + i * slice;
            if (bit / 8 < total / 8) out[bit / 8] |= (uint8_t)(1u << (7 - bit % 8));
        }
    }
}

/* ------------------------------------------------ this build's generators
 *
 * The synthetic inputs of the benchmark configurations that the reference
 * does not generate (its generators are the two above). Each restates, one
 * draw at a time from the serial fast_prng stream, the algorithm that the
 * matching am_gen_* entry point documents (include/amethyst.h); the device
 * code evaluates the same draws in parallel with jump-ahead. */

uint64_t or_mask(int bits);

static uint64_t thr32(double p) /* round(p * 2^32), clamped to [0, 2^32] */
{
    double t = p * 4294967296.0;
    return t <= 0 ? 0 : (t >= 4294967296.0 ? (1ull << 32) : (uint64_t)(t + 0.5));
}

/* am_gen_mask_bernoulli: bit i (LSB-first) = draw(i) < round(s * 2^32). */
void or_gen_mask_bernoulli(uint64_t seed, double s, uint64_t n, uint64_t* out)
{
    or_rng r;
    or_srand(&r, seed);
    uint64_t thr = thr32(s);
    memset(out, 0, ((n + 63) / 64) * 8);
    for (uint64_t i = 0; i < n; i++)
        if ((uint64_t)or_rand(&r) < thr) out[i / 64] |= 1ull << (i % 64);
}

/* am_gen_mask_runs: the bits are cut into chunks of 4096; chunk c uses draws
 * c*4097 .. : its first draw starts the chain in state 1 iff draw <
 * round(s*2^32); then per bit the current state is written and one draw
 * moves the chain (state 1 leaves with probability 1/mean_run, state 0
 * with s/(mean_run*(1-s)), thresholds rounded to 32-bit fractions). */
void or_gen_mask_runs(uint64_t seed, double s, double mean_run, uint64_t n, uint64_t* out)
{
    double p10 = 1.0 / mean_run, p01 = s / (mean_run * (1.0 - s));
    if (p01 > 1.0) p01 = 1.0;
    uint64_t t10 = (uint64_t)(p10 * 4294967296.0 + 0.5), t01 = (uint64_t)(p01 * 4294967296.0 + 0.5);
    uint64_t tpi = (uint64_t)(s * 4294967296.0 + 0.5);
    memset(out, 0, ((n + 63) / 64) * 8);
    or_rng r;
    or_srand(&r, seed);
    int on = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (i % 4096 == 0) on = (uint64_t)or_rand(&r) < tpi;
        if (on) out[i / 64] |= 1ull << (i % 64);
        uint64_t d = or_rand(&r);
        on = on ? !(d < t10) : (d < t01);
    }
}

/* am_gen_from_cdf: v_i = the first b with draw(i) < thr[b] (thr ascending,
 * the last 2^32-1 so every draw but 2^32-1 itself maps below nb), masked. */
void or_gen_from_cdf(uint64_t seed, const uint32_t* thr, uint32_t nb, int bits, uint64_t n, uint64_t* vals)
{
    or_rng r;
    or_srand(&r, seed);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t d = or_rand(&r), b = 0;
        while (b + 1 < nb && !(d < thr[b])) b++;
        vals[i] = (uint64_t)b & or_mask(bits);
    }
}

/* am_gen_sorted_noise: v_i = min(mask, floor((i / n) * mask) + (draw(i) &
 * (2^noise_bits - 1))), the quotient and product in IEEE double. */
void or_gen_sorted_noise(uint64_t seed, int bits, int noise_bits, uint64_t n, uint64_t* vals)
{
    uint64_t m = or_mask(bits);
    uint64_t nm = noise_bits <= 0 ? 0 : (noise_bits >= 32 ? 0xFFFFFFFFull : ((1ull << noise_bits) - 1));
    or_rng r;
    or_srand(&r, seed);
    for (uint64_t i = 0; i < n; i++) {
        volatile double frac = (double)i / (double)n; /* two separately rounded operations */
        uint64_t base = (uint64_t)(frac * (double)m);
        uint64_t v = base + ((uint64_t)or_rand(&r) & nm);
        vals[i] = v > m ? m : v;
    }
}

/* --------------------------------------------------------------- formats */

/* gen_bitmask, src/common_code.cuh:38-41 (bits in 1..64). */
uint64_t or_mask(int bits)
{
    return (~(uint64_t)0) >> (64 - bits);
}

static int tight_bytes(int bits)
{
    /* tight_type<bits>::type, src/common_code.cuh:11-19 */
    int bc = (bits + 7) / 8;
    return bc > 4 ? 8 : (bc > 2 ? 4 : (bc > 1 ? 2 : 1));
}

uint64_t or_packed_words(int fmt, int bits, uint64_t n)
{
    switch (fmt) {
        case AM_FMT_BASELINE: return n;
        case AM_FMT_SLAB: {
            uint64_t eps = 64 / (uint64_t)bits;
            return (n + eps - 1) / eps;
        }
        case AM_FMT_BITSTREAM: return (n * (uint64_t)bits + 63) / 64;
        case AM_FMT_TIGHT: return (n * (uint64_t)tight_bytes(bits) + 7) / 8;
        case AM_FMT_U32: return (n + 1) / 2;
    }
    return 0;
}

/* Element accessors:
 *   baseline  get_element_baseline  src/common_code.cuh:51-55
 *   slab      get_element<bits>      src/common_code.cuh:43-50
 *   bitstream get_elem_nogaps<bits>  src/common_code.cuh:100-116
 *   tight     sentinel read          src/slab_iterator.cuh:83-91 */
uint64_t or_get(int fmt, int bits, const uint64_t* data, uint64_t i)
{
    uint64_t m = or_mask(bits);
    switch (fmt) {
        case AM_FMT_BASELINE: return data[i] & m;
        case AM_FMT_SLAB: {
            uint64_t eps = 64 / (uint64_t)bits;
            uint64_t slab = i / eps;
            uint64_t off = (i - slab * eps) * (uint64_t)bits;
            return (data[slab] >> off) & m;
        }
        case AM_FMT_BITSTREAM: {
            uint64_t slab = (i * (uint64_t)bits) / 64;
            uint64_t off = (i * (uint64_t)bits) % 64;
            uint64_t remaining = 64 - off;
            uint64_t e;
            if (remaining < (uint64_t)bits) {
                e = (data[slab] >> off) & or_mask((int)remaining);
                uint64_t rest = (uint64_t)bits - remaining;
                e |= (data[slab + 1] & or_mask((int)rest)) << remaining;
            }
            else {
                e = (data[slab] >> off) & m;
            }
            return e;
        }
        case AM_FMT_TIGHT: {
            switch (tight_bytes(bits)) {
                case 1: return ((const uint8_t*)data)[i] & m;
                case 2: return ((const uint16_t*)data)[i] & m;
                case 4: return ((const uint32_t*)data)[i] & m;
                default: return data[i] & m;
            }
        }
        case AM_FMT_U32: return ((const uint32_t*)data)[i] & m;
    }
    return 0;
}

/* Store v & mask(bits) as element i (the buffer's other bits untouched).
 * Slab placement as in kernel_binary_op_slabs (src/kernels/binary_op_slabs.cuh:62-63),
 * bitstream as in kernel_binary_op_nogaps (src/kernels/binary_op_nogaps.cuh:69-83). */
void or_set(int fmt, int bits, uint64_t* data, uint64_t i, uint64_t v)
{
    uint64_t m = or_mask(bits);
    v &= m;
    switch (fmt) {
        case AM_FMT_BASELINE: data[i] = v; return;
        case AM_FMT_SLAB: {
            uint64_t eps = 64 / (uint64_t)bits;
            uint64_t slab = i / eps;
            uint64_t off = (i - slab * eps) * (uint64_t)bits;
            data[slab] = (data[slab] & ~(m << off)) | (v << off);
            return;
        }
        case AM_FMT_BITSTREAM: {
            uint64_t bit = i * (uint64_t)bits;
            uint64_t slab = bit / 64, off = bit % 64;
            data[slab] = (data[slab] & ~(m << off)) | (v << off);
            if (off + (uint64_t)bits > 64) {
                uint64_t done = 64 - off;
                uint64_t rest_mask = or_mask((int)((uint64_t)bits - done));
                data[slab + 1] = (data[slab + 1] & ~rest_mask) | (v >> done);
            }
            return;
        }
        case AM_FMT_TIGHT: {
            switch (tight_bytes(bits)) {
                case 1: ((uint8_t*)data)[i] = (uint8_t)v; return;
                case 2: ((uint16_t*)data)[i] = (uint16_t)v; return;
                case 4: ((uint32_t*)data)[i] = (uint32_t)v; return;
                default: data[i] = v; return;
            }
        }
        case AM_FMT_U32: ((uint32_t*)data)[i] = (uint32_t)v; return;
    }
}

/* Pack n values into a zero-filled buffer of or_packed_words words. */
void or_pack(int fmt, int bits, const uint64_t* vals, uint64_t n, uint64_t* out)
{
    memset(out, 0, or_packed_words(fmt, bits, n) * 8);
    for (uint64_t i = 0; i < n; i++) or_set(fmt, bits, out, i, vals[i]);
}

void or_unpack(int fmt, int bits, const uint64_t* data, uint64_t n, uint64_t* vals)
{
    for (uint64_t i = 0; i < n; i++) vals[i] = or_get(fmt, bits, data, i);
}

/* ------------------------------------------------------------- operators */

static uint64_t udiv(uint64_t l, uint64_t r) { return r ? l / r : ~(uint64_t)0; }
static uint64_t umod(uint64_t l, uint64_t r) { return r ? l % r : l; }

/* The operator functors of src/config.cuh:5-42, each masked to `outbits`. */
uint64_t or_binop(int op, int outbits, uint64_t l, uint64_t r)
{
    uint64_t v;
    switch (op) {
        case AM_OP_ADD: v = l + r; break;                              /* :5-9   */
        case AM_OP_MUL: v = l * r; break;                              /* :27-31 */
        case AM_OP_MOD: v = umod(l, r); break;                         /* :17-21 */
        case AM_OP_DIV: v = udiv(l, r); break;                         /* :22-26 */
        case AM_OP_MATH:                                               /* :11-15 */
            v = udiv(l * l + r * r + 4 * l, l * l + r * r + 5 * l);
            break;
        case AM_OP_BOOL_SIMPLE: v = (l || r); break;                   /* :32-36 */
        case AM_OP_BOOL_COMPLEX: v = ((l && !r) || (!l && r)); break;  /* :38-42 */
        case AM_OP_AND: v = l & r; break;
        case AM_OP_OR: v = l | r; break;
        case AM_OP_XOR: v = l ^ r; break;
        case AM_OP_SUB: v = l - r; break;
        default: v = 0;
    }
    return v & or_mask(outbits);
}

/* op_add_outbits / op_mul_outbits / op_mod_outbits, src/config.cuh:44-55. */
int or_outbits(int op, int bits)
{
    if (op == AM_OP_ADD) return bits == 64 ? 64 : bits + 1;
    if (op == AM_OP_MUL) return bits * 2 >= 64 ? 64 : bits * 2;
    return bits;
}

int or_pred(int op, uint64_t c, uint64_t v)
{
    switch (op) {
        case AM_PRED_LT: return v < c;
        case AM_PRED_LE: return v <= c;
        case AM_PRED_GT: return v > c;
        case AM_PRED_GE: return v >= c; /* FILTER_PREDICATE_OP, src/config.cuh:107-113 */
        case AM_PRED_EQ: return v == c;
        case AM_PRED_NE: return v != c;
    }
    return 0;
}

/* Filter -> LSB-first u64 mask words, the check loop of
 * src/benchs/bench_filter_slabs.cu:38-46 / bench_filter_nogaps.cu:31-40
 * turned into the producer: bit i = pred(element i). */
void or_filter(int fmt, int bits, const uint64_t* in, uint64_t n, int pop, uint64_t c, uint64_t* mask)
{
    memset(mask, 0, ((n + 63) / 64) * 8);
    for (uint64_t i = 0; i < n; i++) {
        if (or_pred(pop, c, or_get(fmt, bits, in, i))) mask[i / 64] |= 1ull << (i % 64);
    }
}

/* Compressstore, the validation loop src/benchs/bench_compressstore_slabs.cu:74-95:
 * walk the LSB-first mask (bit_iterator, src/bit_iterator.cuh:18-21), append
 * each selected element in order. The output is written into a zero-filled
 * buffer in (ofmt, obits); out_cap limits how many are stored. Returns the
 * selected count; *overflow counts selected values wider than obits. */
uint64_t or_compressstore(int ifmt, int ibits, const uint64_t* in, uint64_t n, const uint64_t* mask,
                          int ofmt, int obits, uint64_t* out, uint64_t out_cap, uint64_t* overflow)
{
    memset(out, 0, or_packed_words(ofmt, obits, out_cap) * 8);
    uint64_t k = 0, ovf = 0, om = or_mask(obits);
    for (uint64_t i = 0; i < n; i++) {
        if ((mask[i / 64] >> (i % 64)) & 1) {
            uint64_t v = or_get(ifmt, ibits, in, i);
            if (v & ~om) ovf++;
            if (k < out_cap) or_set(ofmt, obits, out, k, v);
            k++;
        }
    }
    if (overflow) *overflow = ovf;
    return k;
}

/* Binary op, src/benchs/bench_binary_op_slabs.cu:32-41 /
 * bench_binary_op_nogaps.cu:32-45: out[i] = op(a[i], b[i]) at outbits. */
void or_binary(int op, int fmt, int bits, const uint64_t* a, const uint64_t* b, uint64_t n,
               int ofmt, int obits, uint64_t* out)
{
    memset(out, 0, or_packed_words(ofmt, obits, n) * 8);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t r = or_binop(op, obits, or_get(fmt, bits, a, i), or_get(fmt, bits, b, i));
        or_set(ofmt, obits, out, i, r);
    }
}

/* Sum with HOP_INIT = 0, HOP_OP = + (src/config.cuh:63-71), the loop of
 * src/benchs/bench_hop_slabs.cu:38-51. */
uint64_t or_sum(int fmt, int bits, const uint64_t* in, uint64_t n)
{
    uint64_t s = 0;
    for (uint64_t i = 0; i < n; i++) s += or_get(fmt, bits, in, i);
    return s;
}

/* Group-by, the map of src/benchs/bench_groupby_sm_array.cu:85-118 made
 * dense and key-ordered: sums[k] (u64, wrapping) and counts[k] for every key
 * k < 2^kbits; rows whose value fails the optional predicate are skipped.
 * Returns the number of rows aggregated. */
uint64_t or_groupby(int kfmt, int kbits, const uint64_t* keys, int vfmt, int vbits, const uint64_t* vals,
                    uint64_t n, int use_pred, int pop, uint64_t pc, uint64_t* sums, uint64_t* counts)
{
    uint64_t groups = 1ull << kbits, rows = 0;
    memset(sums, 0, groups * 8);
    memset(counts, 0, groups * 8);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t v = or_get(vfmt, vbits, vals, i);
        if (use_pred && !or_pred(pop, pc, v)) continue;
        uint64_t k = or_get(kfmt, kbits, keys, i);
        sums[k] += v;
        counts[k] += 1;
        rows++;
    }
    return rows;
}

/* C[M x N] = A[M x K] * B[K x N], exact int64, values read through
 * get_matrix_element's row-padded slab layout (src/common_code.cuh:124-129). */
void or_matmul_packed(int bits, const uint64_t* A, const uint64_t* B, uint64_t M, uint64_t N, uint64_t K, int64_t* C)
{
    uint64_t eps = 64 / (uint64_t)bits, m = or_mask(bits);
    uint64_t ra = (K + eps - 1) / eps, rb = (N + eps - 1) / eps;
    for (uint64_t i = 0; i < M; i++) {
        for (uint64_t j = 0; j < N; j++) {
            int64_t acc = 0;
            for (uint64_t k = 0; k < K; k++) {
                uint64_t a = (A[i * ra + k / eps] >> (bits * (k % eps))) & m;
                uint64_t b = (B[k * rb + j / eps] >> (bits * (j % eps))) & m;
                acc += (int64_t)(a * b);
            }
            C[i * N + j] = acc;
        }
    }
}

/* --------------------------------------------------- CPU baseline timing */

/* Faithful single-thread form of the benchmark-1 path: predicate
 * filter + compressstore of a slab column into uint32 + LSB-first mask
 * (per-element accessor decode, as the reference's loops do). */
uint64_t or_cpu_filter_compress_u32(int bits, const uint64_t* in, uint64_t n, int pop, uint64_t c,
                                    uint32_t* out, uint64_t* mask)
{
    uint64_t k = 0;
    uint64_t words = (n + 63) / 64;
    for (uint64_t w = 0; w < words; w++) {
        uint64_t mw = 0;
        uint64_t end = (w + 1) * 64 < n ? (w + 1) * 64 : n;
        for (uint64_t i = w * 64; i < end; i++) {
            uint64_t v = or_get(AM_FMT_SLAB, bits, in, i);
            if (or_pred(pop, c, v)) {
                mw |= 1ull << (i - w * 64);
                out[k++] = (uint32_t)v;
            }
        }
        mask[w] = mw;
    }
    return k;
}

/* Parallel form with identical arithmetic: per-chunk count, exclusive scan
 * over chunks, then write (OpenMP over contiguous chunks of 64-element
 * words). Returns the selected count. `threads` <= 0 uses all cores. */
uint64_t or_cpu_filter_compress_u32_mt(int bits, const uint64_t* in, uint64_t n, int pop, uint64_t c,
                                       uint32_t* out, uint64_t* mask, int threads)
{
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    int nt = 1;
    (void)threads;
#endif
    uint64_t words = (n + 63) / 64;
    uint64_t* counts = (uint64_t*)calloc((size_t)nt + 1, sizeof(uint64_t));
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        uint64_t w0 = words * (uint64_t)t / (uint64_t)nt, w1 = words * (uint64_t)(t + 1) / (uint64_t)nt;
        uint64_t cnt = 0;
        for (uint64_t w = w0; w < w1; w++) {
            uint64_t mw = 0;
            uint64_t end = (w + 1) * 64 < n ? (w + 1) * 64 : n;
            for (uint64_t i = w * 64; i < end; i++) {
                if (or_pred(pop, c, or_get(AM_FMT_SLAB, bits, in, i))) mw |= 1ull << (i - w * 64);
            }
            mask[w] = mw;
            cnt += (uint64_t)__builtin_popcountll(mw);
        }
        counts[t + 1] = cnt;
#pragma omp barrier
#pragma omp single
        for (int i = 0; i < nt; i++) counts[i + 1] += counts[i];
        uint64_t k = counts[t];
        for (uint64_t w = w0; w < w1; w++) {
            uint64_t mw = mask[w];
            while (mw) {
                int b = __builtin_ctzll(mw);
                out[k++] = (uint32_t)or_get(AM_FMT_SLAB, bits, in, w * 64 + (uint64_t)b);
                mw &= mw - 1;
            }
        }
    }
    uint64_t total = counts[nt];
    free(counts);
    return total;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
