// k_gen.hip -- seeded input generators.
//
// The reference draws its inputs from fast_prng, PCG32 (XSH-RR output) with a
// splitmix64 pre-seed and inc = 1 (src/fast_prng.cuh:11-30), serially on the
// host (src/main.cu:239-251). Here every lane jumps straight to its draw index
// with the LCG jump-ahead of Brown, "Random Number Generation with Arbitrary
// Strides" (1994) -- the pcg_advance method of the PCG family -- so a 2^28
// element column is generated in parallel and draw i is bit-identical to the
// i-th rand() of the reference generator.
#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint64_t PCG_MULT = 6364136223846793005ull;
constexpr uint64_t PCG_INC = 1ull; // inc | 1 with inc = 1

__host__ __device__ inline uint64_t splitmix_seed(uint64_t seed)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// state after `delta` LCG steps (Brown 1994)
__host__ __device__ inline void lcg_jump(uint64_t delta, uint64_t& mult_out, uint64_t& plus_out)
{
    uint64_t acc_mult = 1, acc_plus = 0, mult = PCG_MULT, plus = PCG_INC;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult *= mult;
            acc_plus = acc_plus * mult + plus;
        }
        plus = (mult + 1) * plus;
        mult *= mult;
        delta >>= 1;
    }
    mult_out = acc_mult;
    plus_out = acc_plus;
}

__device__ inline uint64_t state_at(uint64_t s0, uint64_t i)
{
    uint64_t m, p;
    lcg_jump(i, m, p);
    return m * s0 + p;
}

__device__ inline uint32_t pcg_out(uint64_t old)
{
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

// out[i] = draw(offset + i); lane-strided so stores coalesce: each lane steps
// its state by `stride` draws with the precomputed jump (smul, splus).
__global__ __launch_bounds__(256) void k_prng_fill(uint64_t s0, uint64_t offset, uint32_t* out, uint64_t n,
                                                    uint64_t smul, uint64_t splus)
{
    uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    if (gid >= n) return;
    uint64_t st = state_at(s0, offset + gid);
    for (uint64_t i = gid; i < n; i += stride) {
        out[i] = pcg_out(st);
        st = smul * st + splus;
    }
}

// Value of element i of a uniform column: draw(i) & mask for bits <= 32,
// (draw(2i) << 32 | draw(2i+1)) & mask for wider columns.
__device__ inline uint64_t uniform_value(uint64_t s0, uint64_t i, uint32_t bits, uint64_t mask)
{
    if (bits <= 32) return pcg_out(state_at(s0, i)) & mask;
    uint64_t st = state_at(s0, 2 * i);
    uint64_t hi = pcg_out(st);
    st = PCG_MULT * st + PCG_INC;
    return ((hi << 32) | pcg_out(st)) & mask;
}

// sequential next value after uniform_value(i) (state passed by reference)
struct UniformStream {
    uint64_t st;
    uint32_t bits;
    uint64_t mask;
    __device__ UniformStream(uint64_t s0, uint64_t i, uint32_t b, uint64_t m) : bits(b), mask(m)
    {
        st = state_at(s0, b <= 32 ? i : 2 * i);
    }
    __device__ inline uint64_t next()
    {
        uint64_t v = pcg_out(st);
        st = PCG_MULT * st + PCG_INC;
        if (bits > 32) {
            v = (v << 32) | pcg_out(st);
            st = PCG_MULT * st + PCG_INC;
        }
        return v & mask;
    }
};

// Optional inverse-CDF mapping of each draw: value = first b with draw < thr[b].
struct CdfMap {
    const uint32_t* thr;
    uint32_t nb;
    __device__ inline uint64_t map(uint32_t d) const
    {
        uint32_t lo = 0, hi = nb - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (d < thr[mid]) hi = mid;
            else lo = mid + 1;
        }
        return lo;
    }
};

// One lane per output word (packed kinds) or per element (strided).
template <int OK, bool CDF>
__global__ __launch_bounds__(256) void k_gen_column(uint64_t s0, Col out, CdfMap cdf)
{
    uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (OK == K_STRIDED) {
        if (gid >= out.n) return;
        uint64_t v;
        if (CDF) v = cdf.map(pcg_out(state_at(s0, gid))) & out.mask;
        else v = uniform_value(s0, gid, out.bits, out.mask);
        switch (out.cbytes) {
            case 1: ((uint8_t*)out.p)[gid] = (uint8_t)v; break;
            case 2: ((uint16_t*)out.p)[gid] = (uint16_t)v; break;
            case 4: ((uint32_t*)out.p)[gid] = (uint32_t)v; break;
            default: out.p[gid] = v;
        }
        return;
    }
    uint64_t words = OK == K_SLAB ? (out.n + out.eps - 1) / out.eps : (out.n * out.bits + 63) / 64;
    if (gid >= words) return;
    uint64_t e0, e1; // element range overlapping word gid
    if (OK == K_SLAB) {
        e0 = gid * out.eps;
        e1 = e0 + out.eps;
    }
    else {
        e0 = (gid * 64) / out.bits;
        e1 = (gid * 64 + 63) / out.bits + 1;
    }
    if (e1 > out.n) e1 = out.n;
    UniformStream us(s0, e0, out.bits, out.mask);
    uint64_t st = CDF ? state_at(s0, e0) : 0;
    uint64_t word = 0;
    for (uint64_t e = e0; e < e1; e++) {
        uint64_t v;
        if (CDF) {
            v = cdf.map(pcg_out(st)) & out.mask;
            st = PCG_MULT * st + PCG_INC;
        }
        else v = us.next();
        if (OK == K_SLAB) word |= v << ((e - e0) * out.bits);
        else {
            int64_t pos = (int64_t)(e * out.bits) - (int64_t)(gid * 64);
            if (pos >= 0) word |= v << pos;
            else word |= v >> (-pos);
        }
    }
    out.p[gid] = word;
}

// Sorted with noise: v_i = min(mask, floor(i * 2^bits / n) + (draw(i) & noise_mask)).
template <int OK>
__global__ __launch_bounds__(256) void k_gen_sorted(uint64_t s0, Col out, uint64_t noise_mask)
{
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    auto val = [&](uint64_t e, uint32_t d) -> uint64_t {
        const double frac = (double)e / (double)out.n;
        uint64_t base = (uint64_t)(frac * (double)(out.mask)) ;
        uint64_t v = base + (d & noise_mask);
        return v > out.mask ? out.mask : v;
    };
    if (OK == K_STRIDED) {
        if (gid >= out.n) return;
        const uint64_t v = val(gid, pcg_out(state_at(s0, gid)));
        switch (out.cbytes) {
            case 1: ((uint8_t*)out.p)[gid] = (uint8_t)v; break;
            case 2: ((uint16_t*)out.p)[gid] = (uint16_t)v; break;
            case 4: ((uint32_t*)out.p)[gid] = (uint32_t)v; break;
            default: out.p[gid] = v;
        }
        return;
    }
    const uint64_t words = OK == K_SLAB ? (out.n + out.eps - 1) / out.eps : (out.n * out.bits + 63) / 64;
    if (gid >= words) return;
    uint64_t e0, e1;
    if (OK == K_SLAB) {
        e0 = gid * out.eps;
        e1 = e0 + out.eps;
    }
    else {
        e0 = (gid * 64) / out.bits;
        e1 = (gid * 64 + 63) / out.bits + 1;
    }
    if (e1 > out.n) e1 = out.n;
    uint64_t st = state_at(s0, e0), word = 0;
    for (uint64_t e = e0; e < e1; e++) {
        const uint64_t v = val(e, pcg_out(st));
        st = PCG_MULT * st + PCG_INC;
        if (OK == K_SLAB) word |= v << ((e - e0) * out.bits);
        else {
            const int64_t pos = (int64_t)(e * out.bits) - (int64_t)(gid * 64);
            word |= pos >= 0 ? (v << pos) : (v >> (-pos));
        }
    }
    out.p[gid] = word;
}

// Bernoulli mask: bit i = draw(i) < thr (thr = round(s * 2^32)); lane per word.
__global__ __launch_bounds__(256) void k_gen_mask_bernoulli(uint64_t s0, uint64_t thr, uint64_t* mask, uint64_t n)
{
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t words = (n + 63) / 64;
    if (w >= words) return;
    uint64_t st = state_at(s0, w * 64), m = 0;
    uint64_t cnt = n - w * 64 < 64 ? n - w * 64 : 64;
    for (uint64_t b = 0; b < cnt; b++) {
        if ((uint64_t)pcg_out(st) < thr) m |= 1ull << b;
        st = PCG_MULT * st + PCG_INC;
    }
    mask[w] = m;
}

// Clustered mask: a two-state Markov chain (leave a 1-run with probability
// p10, leave a 0-run with p01) gives geometric run lengths with means 1/p10
// and 1/p01. Each lane runs the chain over a chunk of `chunk_words` words,
// starting from the stationary distribution (draw < pi1).
__global__ __launch_bounds__(256) void k_gen_mask_runs(uint64_t s0, uint64_t p10, uint64_t p01, uint64_t pi1,
                                                       uint64_t* mask, uint64_t n, uint32_t chunk_words)
{
    uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t words = (n + 63) / 64;
    uint64_t w0 = c * chunk_words;
    if (w0 >= words) return;
    uint64_t st = state_at(s0, w0 * 64 + c); // one extra draw per chunk for the start state
    bool on = (uint64_t)pcg_out(st) < pi1;
    st = PCG_MULT * st + PCG_INC;
    uint64_t w1 = w0 + chunk_words < words ? w0 + chunk_words : words;
    for (uint64_t w = w0; w < w1; w++) {
        uint64_t m = 0;
        uint64_t cnt = n - w * 64 < 64 ? n - w * 64 : 64;
        for (uint64_t b = 0; b < cnt; b++) {
            if (on) m |= 1ull << b;
            uint64_t d = pcg_out(st);
            st = PCG_MULT * st + PCG_INC;
            if (on) on = !(d < p10);
            else on = d < p01;
        }
        mask[w] = m;
    }
}

// generate_mask_uniform (src/main.cu:33-46), one lane per byte: byte i of
// the run gets bit j (7 down to 0, MSB first) set iff draw 8i + (7-j) <
// p_adj, p_adj = (uint64)(selectivity * UINT32_MAX).
__global__ __launch_bounds__(256) void k_gen_mask_uniform_ref(uint64_t s0, uint64_t p_adj, uint8_t* buf,
                                                              uint64_t length)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= length) return;
    uint64_t st = state_at(s0, 8 * i);
    uint32_t acc = 0;
    for (int j = 7; j >= 0; j--) {
        if ((uint64_t)pcg_out(st) < p_adj) acc |= 1u << j;
        st = PCG_MULT * st + PCG_INC;
    }
    buf[i] = (uint8_t)acc;
}

// create_bitmask (src/main.cu:59-108), one lane per byte: bit b is set iff
// it lies in the first `cluster` bits of one of the first `count` slices of
// `slice` bits; bit b is bit 7 - b%8 of byte b/8.
__global__ __launch_bounds__(256) void k_gen_mask_clusters_ref(uint64_t cluster, uint64_t slice, uint64_t count,
                                                               uint8_t* out, uint64_t bytes)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bytes) return;
    uint32_t acc = 0;
    for (uint32_t k = 0; k < 8; k++) {
        const uint64_t b = 8 * i + k;
        if (b / slice < count && b % slice < cluster) acc |= 1u << (7 - k);
    }
    out[i] = (uint8_t)acc;
}

} // namespace am

using namespace am;

extern "C" int am_prng_fill_u32(uint64_t seed, uint64_t offset, uint32_t* out, uint64_t n, void* stream)
{
    if (!out && n) return AM_E_INVALID;
    if (n == 0) return AM_OK;
    uint64_t threads = 256;
    uint64_t blocks = (n + threads * 64 - 1) / (threads * 64);
    if (blocks > 65535) blocks = 65535;
    uint64_t smul, splus;
    lcg_jump(blocks * threads, smul, splus);
    hipLaunchKernelGGL(k_prng_fill, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream, splitmix_seed(seed), offset,
                       out, n, smul, splus);
    return check_launch();
}

static int gen_column(uint64_t seed, am_column* out, const uint32_t* thr, uint32_t nb, void* stream)
{
    Col c;
    int st = make_col(out, &c);
    if (st) return st;
    if (c.n == 0) return AM_OK;
    uint64_t units = c.kind == K_STRIDED ? c.n : am_packed_words(out->format, out->bits, out->n);
    uint32_t blocks = (uint32_t)((units + 255) / 256);
    CdfMap cm{thr, nb};
    uint64_t s0 = splitmix_seed(seed);
    hipStream_t s = (hipStream_t)stream;
#define GEN_CASE(KD)                                                                                   \
    if (thr) hipLaunchKernelGGL((k_gen_column<KD, true>), dim3(blocks), dim3(256), 0, s, s0, c, cm);  \
    else hipLaunchKernelGGL((k_gen_column<KD, false>), dim3(blocks), dim3(256), 0, s, s0, c, cm);
    switch (c.kind) {
        case K_STRIDED: GEN_CASE(K_STRIDED); break;
        case K_SLAB: GEN_CASE(K_SLAB); break;
        default: GEN_CASE(K_BITSTREAM); break;
    }
#undef GEN_CASE
    return check_launch();
}

extern "C" int am_gen_uniform(uint64_t seed, am_column* out, void* stream)
{
    return gen_column(seed, out, nullptr, 0, stream);
}

extern "C" int am_gen_from_cdf(uint64_t seed, const uint32_t* thresholds_dev, uint32_t nbins, am_column* out,
                               void* stream)
{
    if (!thresholds_dev || nbins == 0) return AM_E_INVALID;
    return gen_column(seed, out, thresholds_dev, nbins, stream);
}

extern "C" int am_gen_sorted_noise(uint64_t seed, uint32_t noise_bits, am_column* out, void* stream)
{
    Col c;
    int st = make_col(out, &c);
    if (st) return st;
    if (c.n == 0) return AM_OK;
    const uint64_t units = c.kind == K_STRIDED ? c.n : am_packed_words(out->format, out->bits, out->n);
    const uint32_t blocks = (uint32_t)((units + 255) / 256);
    const uint64_t nm = lowmask(noise_bits > 32 ? 32 : noise_bits) & (noise_bits ? ~0ull : 0ull);
    hipStream_t s = (hipStream_t)stream;
    const uint64_t s0 = splitmix_seed(seed);
    switch (c.kind) {
        case K_STRIDED: hipLaunchKernelGGL(k_gen_sorted<K_STRIDED>, dim3(blocks), dim3(256), 0, s, s0, c, nm); break;
        case K_SLAB: hipLaunchKernelGGL(k_gen_sorted<K_SLAB>, dim3(blocks), dim3(256), 0, s, s0, c, nm); break;
        default: hipLaunchKernelGGL(k_gen_sorted<K_BITSTREAM>, dim3(blocks), dim3(256), 0, s, s0, c, nm); break;
    }
    return check_launch();
}

extern "C" int am_gen_mask_bernoulli(uint64_t seed, double selectivity, uint64_t* mask, uint64_t n, void* stream)
{
    if (!mask && n) return AM_E_INVALID;
    if (n == 0) return AM_OK;
    double t = selectivity * 4294967296.0;
    uint64_t thr = t <= 0 ? 0 : (t >= 4294967296.0 ? (1ull << 32) : (uint64_t)(t + 0.5));
    uint64_t words = (n + 63) / 64;
    hipLaunchKernelGGL(k_gen_mask_bernoulli, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       splitmix_seed(seed), thr, mask, n);
    return check_launch();
}

extern "C" int am_gen_mask_runs(uint64_t seed, double selectivity, double mean_run, uint64_t* mask, uint64_t n,
                                void* stream)
{
    if ((!mask && n) || selectivity <= 0.0 || selectivity >= 1.0 || mean_run < 1.0) return AM_E_INVALID;
    if (n == 0) return AM_OK;
    double p10 = 1.0 / mean_run;                                // 1-runs: mean `mean_run`
    double p01 = selectivity / (mean_run * (1.0 - selectivity)); // 0-runs: mean mean_run*(1-s)/s
    if (p01 > 1.0) p01 = 1.0;
    auto thr = [](double p) { return (uint64_t)(p * 4294967296.0 + 0.5); };
    uint64_t words = (n + 63) / 64;
    const uint32_t chunk_words = 64;
    uint64_t chunks = (words + chunk_words - 1) / chunk_words;
    hipLaunchKernelGGL(k_gen_mask_runs, dim3((uint32_t)((chunks + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       splitmix_seed(seed), thr(p10), thr(p01), thr(selectivity), mask, n, chunk_words);
    return check_launch();
}

extern "C" int am_gen_mask_uniform_ref(uint64_t seed, double selectivity, uint8_t* buf, uint64_t offset,
                                       uint64_t length, void* stream)
{
    if (!buf && length) return AM_E_INVALID;
    if (length == 0) return AM_OK;
    const uint64_t p_adj = (uint64_t)(selectivity * (double)UINT32_MAX);
    hipLaunchKernelGGL(k_gen_mask_uniform_ref, dim3((uint32_t)((length + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, splitmix_seed(seed), p_adj, buf + offset, length);
    return check_launch();
}

extern "C" int am_gen_mask_clusters_ref(float selectivity, uint64_t cluster_count, uint64_t total, uint8_t* out,
                                        void* stream)
{
    const uint64_t bytes = total / 8;
    if ((!out && bytes) || cluster_count == 0) return AM_E_INVALID;
    if (bytes == 0) return AM_OK;
    // the reference forms selectivity * total in float
    const uint64_t set_one = (uint64_t)(selectivity * (float)total);
    uint64_t cluster = set_one / cluster_count, slice = total / cluster_count, count = cluster_count;
    if (slice == 0) slice = 1, count = 0; // fewer bits than clusters: the reference sets none
    hipLaunchKernelGGL(k_gen_mask_clusters_ref, dim3((uint32_t)((bytes + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, cluster, slice, count, out, bytes);
    return check_launch();
}

// ------------------------------------------------------------ stream copy
//
// Measurement harness: a plain device copy with 16-byte non-temporal loads
// and stores, 4 per lane in flight -- the sustained read+write rate that the
// operators' HBM fractions are compared against.
__global__ __launch_bounds__(256) void k_stream_copy(const u64x2* src, u64x2* dst, uint64_t n2)
{
    constexpr uint32_t U = 4;
    const uint64_t stride = (uint64_t)gridDim.x * 256 * U;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 * U + threadIdx.x; i < n2; i += stride) {
        u64x2 v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint64_t j = i + u * 256;
            v[u] = j < n2 ? __builtin_nontemporal_load(src + j) : u64x2{0, 0};
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint64_t j = i + u * 256;
            if (j < n2) st_nt(dst + j, v[u]);
        }
    }
}

extern "C" int am_stream_copy(const uint64_t* src, uint64_t* dst, uint64_t nwords, void* stream)
{
    if (nwords && (!src || !dst)) return AM_E_INVALID;
    if ((((uintptr_t)src | (uintptr_t)dst) & 15) != 0) return AM_E_INVALID; // 16-byte aligned buffers
    if (nwords == 0) return AM_OK;
    hipStream_t s = (hipStream_t)stream;
    const uint64_t n2 = nwords / 2;
    if (n2) {
        const uint64_t blocks = (n2 + 1023) / 1024;
        hipLaunchKernelGGL(k_stream_copy, dim3((uint32_t)(blocks < 16384 ? blocks : 16384)), dim3(256), 0, s,
                           (const u64x2*)src, (u64x2*)dst, n2);
    }
    if (nwords & 1 && hipMemcpyAsync(dst + nwords - 1, src + nwords - 1, 8, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return AM_E_HIP;
    return check_launch();
}
