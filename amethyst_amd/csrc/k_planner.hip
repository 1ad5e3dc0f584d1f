// k_planner.hip -- device side of the learned plan selector for batched
// operator chains: per-column statistics (32 features from a strided and a
// contiguous sample) and a 3-layer MLP (32 -> 64 -> 64 -> P, fp32) whose
// outputs are the predicted log-time of each plan relative to the best.
// The plan with the lowest prediction is taken when its margin over the
// runner-up is at least `min_margin`; otherwise the column falls back to the
// analytic plan supplied by the host. No reference counterpart: the
// reference picks its operator per compile-time configuration
// (src/config.cuh:74-113) and its launch shape by exhaustive sweep
// (src/common_code.cuh:179-241).
#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

constexpr uint32_t PLAN_FEATURES = 32;
constexpr uint32_t PLAN_HIDDEN = 64;
constexpr uint32_t PLAN_SAMPLE = 4096; // strided sample and contiguous run, each

struct StatCol {
    Col a, b;
    Pred pr;
    uint64_t c; // predicate constant (raw, for the c/2^w feature)
};

// One 256-thread block per column.
__global__ __launch_bounds__(256) void k_plan_stats(const StatCol* cols, uint32_t ncols, float* feat)
{
    __shared__ uint32_t s_hist[16];
    __shared__ uint32_t s_red[8][4];
    __shared__ uint64_t s_max[4];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t ci = blockIdx.x;
    if (ci >= ncols) return;
    const StatCol& C = cols[ci];
    const uint64_t n = C.a.n;
    const uint32_t w = C.a.bits;
    if (tid < 16) s_hist[tid] = 0;
    __syncthreads();
    uint32_t sel = 0, nondec = 0, flips = 0, runsel = 0;
    uint64_t vmax = 0;
    const uint32_t S = n < PLAN_SAMPLE ? (uint32_t)n : PLAN_SAMPLE;
    // strided sample: selectivity, histogram, max of selected a+b
    if (S) {
        const uint64_t stride = n / S;
        for (uint32_t i = tid; i < S; i += 256) {
            const uint64_t e = (uint64_t)i * stride;
            uint64_t x, y;
            if (C.a.kind == K_SLAB) {
                const uint64_t q = e / C.a.eps, s = e - q * C.a.eps;
                x = (C.a.p[q] >> (s * w)) & C.a.mask;
                y = (C.b.p[q] >> (s * w)) & C.b.mask;
            }
            else {
                TileReader<K_BITSTREAM> xa(C.a, e), xb(C.b, e);
                x = xa.get(0);
                y = xb.get(0);
            }
            const bool f = C.pr(x);
            sel += f;
            atomicAdd(&s_hist[(uint32_t)(x >> (w > 4 ? w - 4 : 0)) & 15], 1u);
            if (f) vmax = (x + y) > vmax ? (x + y) : vmax;
        }
        // contiguous run: sortedness and selection run structure
        const uint64_t base = n > S ? (n - S) / 2 : 0;
        for (uint32_t i = tid; i + 1 < S; i += 256) {
            const uint64_t e = base + i;
            uint64_t x0, x1;
            if (C.a.kind == K_SLAB) {
                uint64_t q = e / C.a.eps, s = e - q * C.a.eps;
                x0 = (C.a.p[q] >> (s * w)) & C.a.mask;
                q = (e + 1) / C.a.eps;
                s = (e + 1) - q * C.a.eps;
                x1 = (C.a.p[q] >> (s * w)) & C.a.mask;
            }
            else {
                TileReader<K_BITSTREAM> xa(C.a, e);
                x0 = xa.get(0);
                x1 = xa.get(1);
            }
            nondec += x1 >= x0;
            const bool f0 = C.pr(x0), f1 = C.pr(x1);
            flips += f0 != f1;
            runsel += f0;
        }
    }
    auto red = [&](uint32_t v) { for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m); return v; };
    sel = red(sel);
    nondec = red(nondec);
    flips = red(flips);
    runsel = red(runsel);
    vmax = wave_max_u64(vmax);
    if (lane == 0) {
        s_red[wave][0] = sel;
        s_red[wave][1] = nondec;
        s_red[wave][2] = flips;
        s_red[wave][3] = runsel;
        s_max[wave] = vmax;
    }
    __syncthreads();
    if (tid != 0) return;
    uint32_t tsel = 0, tnd = 0, tfl = 0, trs = 0;
    uint64_t tmax = 0;
    for (int q = 0; q < 4; q++) {
        tsel += s_red[q][0];
        tnd += s_red[q][1];
        tfl += s_red[q][2];
        trs += s_red[q][3];
        tmax = s_max[q] > tmax ? s_max[q] : tmax;
    }
    float* F = feat + (uint64_t)ci * PLAN_FEATURES;
    const float Sf = S ? (float)S : 1.0f;
    const float s_frac = tsel / Sf;
    const uint32_t est_bits = tmax ? 64 - __clzll(tmax) : 1;
    const float logn = log2f((float)n + 1.0f);
    F[0] = w / 32.0f;
    F[1] = logn / 27.0f;
    F[2] = s_frac;
    for (int b = 0; b < 16; b++) F[3 + b] = s_hist[b] / Sf;
    F[19] = est_bits / 33.0f;
    F[20] = ((float)(w + 1) - (float)est_bits) / 8.0f;
    F[21] = S > 1 ? tnd / (float)(S - 1) : 1.0f;
    F[22] = S > 1 ? (trs + 1.0f) / (tfl + 1.0f) / 256.0f : 0.0f; // mean selected-run length
    F[23] = (float)((double)C.c / (double)(1ull << w));
    F[24] = n < (1ull << 20) ? 1.0f : 0.0f;
    F[25] = log2f((float)((n + 4095) / 4096) + 1.0f) / 15.0f;
    const float eps_in = (float)(64 / w);
    F[26] = (float)n / eps_in * 8.0f / 1e9f;                            // slab input bytes (GB)
    F[27] = s_frac * (float)n / (float)(64 / (w + 1)) * 8.0f / 1e9f;   // slab out, width w+1
    F[28] = s_frac * (float)n * (w + 1) / 8.0f / 1e9f;                 // bitstream out, w+1
    F[29] = s_frac * (float)n * (est_bits + 1) / 8.0f / 1e9f;          // bitstream out, est+1
    F[30] = s_frac * (1.0f - s_frac) * 4.0f;
    F[31] = 1.0f;
}

// MLP (ReLU hidden layers); weights in one flat fp32 array:
// W1[64][32] b1[64] W2[64][64] b2[64] W3[P][64] b3[P]. One thread per column.
__global__ __launch_bounds__(256) void k_plan_mlp(const float* feat, uint32_t ncols, const float* wts, uint32_t P,
                                                  const float* fmean, const float* fstd, float min_margin,
                                                  const int32_t* fallback, int32_t* plan, float* score)
{
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncols) return;
    float x[PLAN_FEATURES], h1[PLAN_HIDDEN], h2[PLAN_HIDDEN];
    for (uint32_t i = 0; i < PLAN_FEATURES; i++) x[i] = (feat[(uint64_t)c * PLAN_FEATURES + i] - fmean[i]) / fstd[i];
    const float* W1 = wts;
    const float* b1 = W1 + PLAN_HIDDEN * PLAN_FEATURES;
    const float* W2 = b1 + PLAN_HIDDEN;
    const float* b2 = W2 + PLAN_HIDDEN * PLAN_HIDDEN;
    const float* W3 = b2 + PLAN_HIDDEN;
    const float* b3 = W3 + P * PLAN_HIDDEN;
    for (uint32_t o = 0; o < PLAN_HIDDEN; o++) {
        float acc = b1[o];
        for (uint32_t i = 0; i < PLAN_FEATURES; i++) acc = fmaf(W1[o * PLAN_FEATURES + i], x[i], acc);
        h1[o] = acc > 0.0f ? acc : 0.0f;
    }
    for (uint32_t o = 0; o < PLAN_HIDDEN; o++) {
        float acc = b2[o];
        for (uint32_t i = 0; i < PLAN_HIDDEN; i++) acc = fmaf(W2[o * PLAN_HIDDEN + i], h1[i], acc);
        h2[o] = acc > 0.0f ? acc : 0.0f;
    }
    float best = 3.4e38f, second = 3.4e38f;
    int32_t arg = 0;
    for (uint32_t p = 0; p < P; p++) {
        float acc = b3[p];
        for (uint32_t i = 0; i < PLAN_HIDDEN; i++) acc = fmaf(W3[p * PLAN_HIDDEN + i], h2[i], acc);
        if (score) score[(uint64_t)c * P + p] = acc;
        if (acc < best) {
            second = best;
            best = acc;
            arg = (int32_t)p;
        }
        else if (acc < second) second = acc;
    }
    plan[c] = (second - best >= min_margin || !fallback) ? arg : fallback[c];
}

} // namespace am

using namespace am;

extern "C" int am_plan_features(const am_pipeline_col* cols, uint32_t count, float* feat_dev, void* scratch_dev,
                                void* stream)
{
    if (!cols || !feat_dev || !scratch_dev || count == 0) return AM_E_INVALID;
    StatCol* host = new StatCol[count];
    for (uint32_t i = 0; i < count; i++) {
        int st = make_col(&cols[i].a, &host[i].a);
        if (!st) st = make_col(&cols[i].b, &host[i].b);
        if (st || host[i].a.bits > 31 || (host[i].a.kind != K_SLAB && host[i].a.kind != K_BITSTREAM)) {
            delete[] host;
            return st ? st : AM_E_UNSUPPORTED_CONFIG;
        }
        host[i].pr = make_pred(cols[i].pred, host[i].a.mask);
        host[i].c = cols[i].pred.c;
    }
    // stream-ordered: a previous call's k_plan_stats on this stream has read
    // scratch_dev before it is overwritten (a copy from pageable memory
    // returns once the source has been read, so `host` can go)
    hipError_t e = hipMemcpyAsync(scratch_dev, host, count * sizeof(StatCol), hipMemcpyHostToDevice, (hipStream_t)stream);
    delete[] host;
    if (e != hipSuccess) return AM_E_HIP;
    hipLaunchKernelGGL(k_plan_stats, dim3(count), dim3(256), 0, (hipStream_t)stream, (const StatCol*)scratch_dev, count,
                       feat_dev);
    return check_launch();
}

extern "C" uint64_t am_plan_scratch_bytes(uint32_t count) { return (uint64_t)count * sizeof(StatCol); }

extern "C" int am_plan_infer(const float* feat_dev, uint32_t count, const float* weights_dev, uint32_t nplans,
                             const float* fmean_dev, const float* fstd_dev, float min_margin,
                             const int32_t* fallback_dev, int32_t* plan_dev, float* score_dev, void* stream)
{
    if (!feat_dev || !weights_dev || !fmean_dev || !fstd_dev || !plan_dev || count == 0 || nplans == 0 || nplans > 64)
        return AM_E_INVALID;
    hipLaunchKernelGGL(k_plan_mlp, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream, feat_dev, count,
                       weights_dev, nplans, fmean_dev, fstd_dev, min_margin, fallback_dev, plan_dev, score_dev);
    return check_launch();
}
