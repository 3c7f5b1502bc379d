// denoise.hip — the reference's film denoiser on gfx950 and its C ABI
// entry point rt_denoise (include/rt3d.h), replacing NforDenoiser.Denoise
// (src/RayTracer/Denoising/NforDenoiser.cs:29-57).
//
// Stages (each a few kernels on the context's stream):
//   depth      RenderBuffers.CombineDepthHalves, the 99th percentile of the
//              hit depths (PlaneOps.PercentileOfNonNegative: an exact radix
//              select, two 16-bit histogram passes), normalisation to [0, 1]
//              (NforDenoiser.NormaliseDepth, :92-112)
//   variance   VarianceEstimator.FullVariance / HalfFromFull (:24-60): the raw
//              dual-buffer estimate, the 7x7 binomial smooth (PlaneOps.cs:18-48)
//              and the floors
//   NL-means   NlMeansCore.Filter (NlMeansCore.cs:147-283) for the feature
//              prefilter (FeaturePrefilter.cs:19-28) and the joint filter
//              (NforDenoiser.JointNlm, :66-84)
//
// NL-means, offset-decomposed as the reference does it, in two kernels per
// chunk of search offsets: k_nlm_rows gives each (row, offset) pair one
// thread that walks the row with the reference's running patch sum (its
// rounding depends on the whole walk, so a row cannot be split) and writes
// one horizontally averaged plane per offset; k_nlm_accum gives each pixel
// one thread that takes the offsets in the reference's order (dy major),
// sums the vertical patch, forms the weight and accumulates.  The planes of
// a chunk of up to 128 offsets (1 GB at 1920x1080) live in HBM between the
// two kernels.  Every fp32 operation is the reference's, in its order; its
// Vector<float> body (8 wide on x64) and scalar tail differ in one rounding
// (reciprocal multiply vs division of the vertical patch sum), and the
// same split is made per row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rt3d.h"

namespace rt {

hipStream_t ctx_stream(rt_ctx* ctx);
int ctx_device(rt_ctx* ctx);
int ctx_fail(rt_ctx* ctx, int code, const std::string& msg);
int ctx_dn_buffer(rt_ctx* ctx, size_t bytes, void** out);

namespace dn {

constexpr int kVecW = 8;            // Vector<float>.Count on x64
constexpr float kEps = 1e-10f;      // NlMeansCore.Epsilon
constexpr float kCutoff = 10.0f;    // NlMeansCore.WeightCutoff
constexpr int kPatch = 3;           // DenoiserOptions.PatchRadius
constexpr int kChunk = 128;         // offsets per k_nlm_rows / k_nlm_accum pair
constexpr int kBlk = 256;

__device__ inline float fast_exp_neg(float d) {
    // NlMeansCore.FastExpNegScalar (Schraudolph's bit trick)
    const int i = (int)(-12102203.16f * d + 1064866805.0f);
    return __int_as_float(i);
}

__device__ inline float maxz(float a) { return a > 0.0f ? a : 0.0f; }

// DistanceRow (NlMeansCore.cs:103-139) at one pixel pair
__device__ inline float dpt_at(const float* __restrict__ g, const float* __restrict__ v, int64_t n, int gc,
                               float invC, int64_t p, int64_t q) {
    float d = 0.0f;
    for (int c = 0; c < gc; c++) {
        const int64_t cb = (int64_t)c * n;
        const float du = g[cb + p] - g[cb + q];
        const float vp = v[cb + p], vq = v[cb + q];
        const float vmin = vp < vq ? vp : vq;
        d += (du * du - (vp + vmin)) / (kEps + vp + vq);
    }
    return d * invC;
}

struct Geom {
    int w, h, R, side;
    int o0, no;   // offsets [o0, o0 + no) of the search window (o = (dy + R) * side + dx + R)
};

__device__ inline void offset_region(const Geom& G, int o, int& dx, int& dy, int& x0, int& x1, int& y0, int& y1) {
    dx = o % G.side - G.R;
    dy = o / G.side - G.R;
    x0 = max(0, -dx);
    x1 = G.w - 1 - max(0, dx);
    y0 = max(0, -dy);
    y1 = G.h - 1 - max(0, dy);
}

// Pass 1 of NlMeansCore.Filter for every (row, offset) of a chunk: pointwise
// distances and the running horizontal patch average (NlMeansCore.cs:177-195).
__global__ void __launch_bounds__(kBlk) k_nlm_rows(const float* __restrict__ g, const float* __restrict__ v, int gc,
                                                   Geom G, float* __restrict__ dh) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)G.h * G.no) return;
    const int ol = (int)(tid % G.no);
    const int y = (int)(tid / G.no);
    int dx, dy, x0, x1, y0, y1;
    offset_region(G, G.o0 + ol, dx, dy, x0, x1, y0, y1);
    if (x0 > x1 || y < y0 || y > y1) return;
    const int64_t n = (int64_t)G.w * G.h;
    const int64_t row = (int64_t)y * G.w;
    const int64_t qOff = (int64_t)dy * G.w + dx;
    const float invC = 1.0f / gc;
    float* out = dh + (int64_t)ol * n + row;
    constexpr int P = kPatch;
    float run = 0.0f;
    const int warm = min(x0 + P, x1);
    for (int xx = x0; xx <= warm; xx++) run += dpt_at(g, v, n, gc, invC, row + xx, row + xx + qOff);
    for (int x = x0; x <= x1; x++) {
        const int lo = max(x0, x - P), hi = min(x1, x + P);
        out[x] = run / (float)(hi - lo + 1);
        if (x + P + 1 <= x1) run += dpt_at(g, v, n, gc, invC, row + x + P + 1, row + x + P + 1 + qOff);
        if (x - P >= x0) run -= dpt_at(g, v, n, gc, invC, row + x - P, row + x - P + qOff);
    }
}

// Pass 2 of NlMeansCore.Filter for one chunk of offsets (NlMeansCore.cs:199-263),
// per pixel in offset order; num/den carry over between chunks in memory, the
// last chunk normalises (:268-281).  tc <= 3.
__global__ void __launch_bounds__(kBlk) k_nlm_accum(Geom G, const float* __restrict__ dh, int tc,
                                                    const float* __restrict__ t, float invK2, int nf,
                                                    const float* __restrict__ feat, const float* __restrict__ fbw,
                                                    float* __restrict__ num, float* __restrict__ den, int first,
                                                    int last, float* __restrict__ out) {
    const int64_t n = (int64_t)G.w * G.h;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int x = (int)(p % G.w), y = (int)(p / G.w);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    float dsum = 0.0f;
    if (!first) {
        for (int c = 0; c < tc; c++) acc[c] = num[(int64_t)c * n + p];
        dsum = den[p];
    }
    constexpr int P = kPatch;
    for (int ol = 0; ol < G.no; ol++) {
        int dx, dy, x0, x1, y0, y1;
        offset_region(G, G.o0 + ol, dx, dy, x0, x1, y0, y1);
        if (x0 > x1 || y0 > y1 || x < x0 || x > x1 || y < y0 || y > y1) continue;
        const int vlo = max(y0, y - P), vhi = min(y1, y + P);
        const int cnt = vhi - vlo + 1;
        const float* col = dh + (int64_t)ol * n + x;
        float sum = 0.0f;
        for (int yy = vlo; yy <= vhi; yy++) sum += col[(int64_t)yy * G.w];
        const int nvec = ((x1 - x0 + 1) / kVecW) * kVecW;
        float d = (x - x0 < nvec) ? maxz(sum * (1.0f / cnt)) * invK2 : maxz(sum / (float)cnt) * invK2;
        const int64_t q = p + (int64_t)dy * G.w + dx;
        for (int f = 0; f < nf; f++) {
            const float df = feat[(int64_t)f * n + p] - feat[(int64_t)f * n + q];
            d += df * df * fbw[f];
        }
        if (!(d < kCutoff)) continue;
        const float wgt = fast_exp_neg(d);
        dsum += wgt;
        for (int c = 0; c < tc; c++) acc[c] += wgt * t[(int64_t)c * n + q];
    }
    if (last) {
        const float invDen = dsum > 0.0f ? 1.0f / dsum : 0.0f;
        for (int c = 0; c < tc; c++) out[(int64_t)c * n + p] = acc[c] * invDen;
    } else {
        for (int c = 0; c < tc; c++) num[(int64_t)c * n + p] = acc[c];
        den[p] = dsum;
    }
}

// PlaneOps.Combine (PlaneOps.cs:104-118): dst = wa*a + wb*b
__global__ void __launch_bounds__(kBlk) k_combine(const float* a, const float* b, int64_t len, float wa, float wb,
                                                  float* dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) dst[i] = wa * a[i] + wb * b[i];
}

// RenderBuffers.CombineDepthHalves (RenderBuffers.cs:130-151)
__global__ void __launch_bounds__(kBlk) k_depth_combine(const float* da, const float* db, int64_t n, float wa,
                                                        float wb, int nB, float* dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float va = da[i], vb = db[i];
    const bool ma = va < 0.0f, mb = vb < 0.0f;
    dst[i] = (ma && mb) ? -1.0f : ma ? vb : mb ? va : (nB == 0 ? va : wa * va + wb * vb);
}

// non-negative float -> its order key (-0 is 0)
__device__ inline uint32_t nonneg_key(float v) {
    const uint32_t b = __float_as_uint(v);
    return b == 0x80000000u ? 0u : b;
}

__global__ void __launch_bounds__(kBlk) k_pct_hist_hi(const float* x, int64_t n, int32_t* hist) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = x[i];
    if (v >= 0.0f) atomicAdd(&hist[nonneg_key(v) >> 16], 1);
}

__global__ void __launch_bounds__(kBlk) k_pct_hist_lo(const float* x, int64_t n, uint32_t hi, int32_t* hist) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = x[i];
    if (!(v >= 0.0f)) return;
    const uint32_t k = nonneg_key(v);
    if ((k >> 16) == hi) atomicAdd(&hist[k & 0xffffu], 1);
}

// NforDenoiser.NormaliseDepth: sky (-1) at the far plane, else min(v/far, 1)
__global__ void __launch_bounds__(kBlk) k_depth_norm(const float* da, const float* db, int64_t n, float invFar,
                                                     float* oa, float* ob) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = da[i], b = db[i];
    oa[i] = a < 0.0f ? 1.0f : fminf(a * invFar, 1.0f);
    ob[i] = b < 0.0f ? 1.0f : fminf(b * invFar, 1.0f);
}

// raw dual-buffer variance of one channel: (0.5 (a - b))^2
__global__ void __launch_bounds__(kBlk) k_var_raw(const float* a, const float* b, int64_t n, float* raw) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = 0.5f * (a[i] - b[i]);
    raw[i] = d * d;
}

// one pass of PlaneOps.SeparablePass with the binomial [1,6,15,20,15,6,1]
// kernel, truncated and renormalised at the borders (PlaneOps.cs:25-48)
__global__ void __launch_bounds__(kBlk) k_binomial(const float* src, float* dst, int w, int h, int horizontal) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    const float k[7] = {1.0f, 6.0f, 15.0f, 20.0f, 15.0f, 6.0f, 1.0f};
    float sum = 0.0f, wSum = 0.0f;
    for (int j = -3; j <= 3; j++) {
        const int xx = horizontal ? x + j : x, yy = horizontal ? y : y + j;
        if (xx < 0 || xx >= w || yy < 0 || yy >= h) continue;
        const float kw = k[j + 3];
        sum += kw * src[(int64_t)yy * w + xx];
        wSum += kw;
    }
    dst[i] = sum / wSum;
}

// variance floors (VarianceEstimator.cs:42-45) and the half-buffer factor 2
__global__ void __launch_bounds__(kBlk) k_var_floor(const float* sm, const float* mean, int64_t n, float* half) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float m = mean[i];
    const float fl = fmaxf(1e-5f * m * m, 1e-12f);
    const float s = sm[i];
    half[i] = 2.0f * (s > fl ? s : fl);
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kBlk - 1) / kBlk); }

}  // namespace dn
}  // namespace rt

using namespace rt;
using namespace rt::dn;

namespace {

#define DN_TRY(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) return ctx_fail(ctx, RT_ERR_DEVICE, hipGetErrorString(e_)); \
    } while (0)

struct Arena {
    char* base;
    size_t off = 0;
    float* f(int64_t count) {
        char* r = base ? base + off : nullptr;
        off += ((size_t)count * 4 + 255) & ~(size_t)255;
        return (float*)r;
    }
};

struct Pipe {
    rt_ctx* ctx;
    hipStream_t st;
    int w, h;
    int64_t n;
    float* dh;       // kChunk planes
    float* num;      // 3 planes
    float* den;      // 1 plane
    int32_t* hist;   // 65536 bins
    float* bw;       // 8 inverse bandwidths (device)
};

// PlaneOps.Combine of two buffers of len floats (nB <= 0: copy of a)
int combine(Pipe& Q, const float* a, const float* b, int64_t len, int nA, int nB, float* dst) {
    rt_ctx* ctx = Q.ctx;
    if (nB <= 0) {
        DN_TRY(hipMemcpyAsync(dst, a, (size_t)len * 4, hipMemcpyDeviceToDevice, Q.st));
        return RT_OK;
    }
    const float wa = (float)nA / (float)(nA + nB), wb = 1.0f - wa;
    hipLaunchKernelGGL(k_combine, dim3(blocks(len)), dim3(kBlk), 0, Q.st, a, b, len, wa, wb, dst);
    return RT_OK;
}

// VarianceEstimator.FullVariance + HalfFromFull over c channels; scratch 2 planes
int half_variance(Pipe& Q, const float* A, const float* B, const float* mean, int c, float* s0, float* s1,
                  float* half) {
    for (int ch = 0; ch < c; ch++) {
        const int64_t o = (int64_t)ch * Q.n;
        hipLaunchKernelGGL(k_var_raw, dim3(blocks(Q.n)), dim3(kBlk), 0, Q.st, A + o, B + o, Q.n, s0);
        hipLaunchKernelGGL(k_binomial, dim3(blocks(Q.n)), dim3(kBlk), 0, Q.st, s0, s1, Q.w, Q.h, 1);
        hipLaunchKernelGGL(k_binomial, dim3(blocks(Q.n)), dim3(kBlk), 0, Q.st, s1, s0, Q.w, Q.h, 0);
        hipLaunchKernelGGL(k_var_floor, dim3(blocks(Q.n)), dim3(kBlk), 0, Q.st, s0, mean + o, Q.n, half + o);
    }
    return RT_OK;
}

// NlMeansCore.Filter: guide g/var v (gc planes), target t (tc planes) -> out
int nlm(Pipe& Q, const float* g, const float* v, int gc, const float* t, int tc, int R, float K, int nf,
        const float* feat, float* out) {
    const int side = 2 * R + 1, offsets = side * side;
    const float invK2 = 1.0f / (K * K);
    for (int o0 = 0; o0 < offsets; o0 += kChunk) {
        Geom G{Q.w, Q.h, R, side, o0, std::min(kChunk, offsets - o0)};
        const int64_t rows = (int64_t)Q.h * G.no;
        hipLaunchKernelGGL(k_nlm_rows, dim3(blocks(rows)), dim3(kBlk), 0, Q.st, g, v, gc, G, Q.dh);
        hipLaunchKernelGGL(k_nlm_accum, dim3(blocks(Q.n)), dim3(kBlk), 0, Q.st, G, (const float*)Q.dh, tc, t, invK2,
                           nf, feat, (const float*)Q.bw, Q.num, Q.den, o0 == 0 ? 1 : 0,
                           o0 + kChunk >= offsets ? 1 : 0, out);
    }
    return RT_OK;
}

// FeaturePrefilter.Prefilter: cross NL-means of the halves (R 5, k 1)
int prefilter(Pipe& Q, const float* A, const float* B, int c, int nA, int nB, float* mean, float* hv, float* s0,
              float* s1, float* fa, float* fb, float* out) {
    const int64_t len = (int64_t)c * Q.n;
    int rc = combine(Q, A, B, len, nA, nB, mean);
    if (!rc) rc = half_variance(Q, A, B, mean, c, s0, s1, hv);
    if (!rc) rc = nlm(Q, B, hv, c, A, c, 5, 1.0f, 0, nullptr, fa);
    if (!rc) rc = nlm(Q, A, hv, c, B, c, 5, 1.0f, 0, nullptr, fb);
    if (!rc) rc = combine(Q, fa, fb, len, nA, nB, out);
    return rc;
}

// PlaneOps.PercentileOfNonNegative (0.99) by a two-pass radix select
int percentile99(Pipe& Q, const float* x, float* result) {
    rt_ctx* ctx = Q.ctx;
    std::vector<int32_t> hist(65536);
    DN_TRY(hipMemsetAsync(Q.hist, 0, 65536 * 4, Q.st));
    hipLaunchKernelGGL(k_pct_hist_hi, dim3(blocks(Q.n)), dim3(kBlk), 0, Q.st, x, Q.n, Q.hist);
    DN_TRY(hipMemcpyAsync(hist.data(), Q.hist, 65536 * 4, hipMemcpyDeviceToHost, Q.st));
    DN_TRY(hipStreamSynchronize(Q.st));
    int64_t count = 0;
    for (int i = 0; i < 65536; i++) count += hist[i];
    if (count == 0) {
        *result = 1.0f;
        return RT_OK;
    }
    int64_t k = (int64_t)(0.99f * (float)(count - 1));
    k = std::min<int64_t>(std::max<int64_t>(k, 0), count - 1);
    uint32_t hi = 0;
    int64_t before = 0;
    for (; hi < 65536; hi++) {
        if (before + hist[hi] > k) break;
        before += hist[hi];
    }
    DN_TRY(hipMemsetAsync(Q.hist, 0, 65536 * 4, Q.st));
    hipLaunchKernelGGL(k_pct_hist_lo, dim3(blocks(Q.n)), dim3(kBlk), 0, Q.st, x, Q.n, hi, Q.hist);
    DN_TRY(hipMemcpyAsync(hist.data(), Q.hist, 65536 * 4, hipMemcpyDeviceToHost, Q.st));
    DN_TRY(hipStreamSynchronize(Q.st));
    uint32_t lo = 0;
    for (; lo < 65536; lo++) {
        if (before + hist[lo] > k) break;
        before += hist[lo];
    }
    const uint32_t bits = (hi << 16) | lo;
    memcpy(result, &bits, 4);
    return RT_OK;
}

}  // namespace

extern "C" int rt_denoise(rt_ctx* ctx, int32_t width, int32_t height, const rt_film* film,
                          const rt_denoise_options* opts, float* beauty_out) {
    if (!ctx || !film || !opts || !beauty_out) return RT_ERR_INVALID_ARGUMENT;
    if (width <= 0 || height <= 0) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "invalid film dimensions");
    if (!film->beauty_a || !film->beauty_b)
        return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "denoising needs the beauty halves");
    if (!film->albedo_a || !film->albedo_b || !film->normal_a || !film->normal_b || !film->depth_a || !film->depth_b)
        return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "Denoising requires AOV capture (RenderCaptureOptions.Full).");
    const int kind = opts->kind, quality = opts->quality;
    if (kind != RT_DENOISE_NLM && kind != RT_DENOISE_NFOR)
        return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "unexpected denoiser kind");
    if (quality != RT_DENOISE_FAST && quality != RT_DENOISE_HIGH)
        return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "unexpected denoise quality");
    if (kind == RT_DENOISE_NFOR) return ctx_fail(ctx, RT_ERR_UNSUPPORTED, "the NFOR regression is not built yet");
    const int nA = opts->samples_a, nB = opts->samples_b;
    if (nA <= 0 || nB < 0) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "samples_a must be > 0 and samples_b >= 0");
    DN_TRY(hipSetDevice(ctx_device(ctx)));

    Pipe Q{ctx, ctx_stream(ctx), width, height, (int64_t)width * height};
    const int64_t n = Q.n;
    // device layout (floats): inputs 20 n, features 7 n, temporaries, chunk planes
    float *bA, *bB, *aA, *aB, *nmA, *nmB, *dA, *dB, *comb, *dNa, *dNb, *feat, *mean, *hv, *s0, *s1, *fa, *fb, *bmean,
        *bhv, *oa, *ob, *res;
    size_t total = 0;
    for (int pass = 0; pass < 2; pass++) {
        char* base = nullptr;
        if (pass == 1) {
            void* mem = nullptr;
            int rc = ctx_dn_buffer(ctx, total, &mem);
            if (rc) return rc;
            base = (char*)mem;
        }
        Arena A{base};
        bA = A.f(3 * n), bB = A.f(3 * n), aA = A.f(3 * n), aB = A.f(3 * n), nmA = A.f(3 * n), nmB = A.f(3 * n);
        dA = A.f(n), dB = A.f(n), comb = A.f(n), dNa = A.f(n), dNb = A.f(n), feat = A.f(7 * n);
        mean = A.f(3 * n), hv = A.f(3 * n), s0 = A.f(n), s1 = A.f(n), fa = A.f(3 * n), fb = A.f(3 * n);
        bmean = A.f(3 * n), bhv = A.f(3 * n), oa = A.f(3 * n), ob = A.f(3 * n), res = A.f(3 * n);
        Q.dh = A.f((int64_t)kChunk * n);
        Q.num = A.f(3 * n);
        Q.den = A.f(n);
        Q.hist = (int32_t*)A.f(65536);
        Q.bw = A.f(8);
        total = A.off + 256;
    }
    const size_t p3 = (size_t)n * 12, p1 = (size_t)n * 4;
    DN_TRY(hipMemcpyAsync(bA, film->beauty_a, p3, hipMemcpyHostToDevice, Q.st));
    DN_TRY(hipMemcpyAsync(bB, film->beauty_b, p3, hipMemcpyHostToDevice, Q.st));
    DN_TRY(hipMemcpyAsync(aA, film->albedo_a, p3, hipMemcpyHostToDevice, Q.st));
    DN_TRY(hipMemcpyAsync(aB, film->albedo_b, p3, hipMemcpyHostToDevice, Q.st));
    DN_TRY(hipMemcpyAsync(nmA, film->normal_a, p3, hipMemcpyHostToDevice, Q.st));
    DN_TRY(hipMemcpyAsync(nmB, film->normal_b, p3, hipMemcpyHostToDevice, Q.st));
    DN_TRY(hipMemcpyAsync(dA, film->depth_a, p1, hipMemcpyHostToDevice, Q.st));
    DN_TRY(hipMemcpyAsync(dB, film->depth_b, p1, hipMemcpyHostToDevice, Q.st));
    const float bw[8] = {1.0f / 2e-3f, 1.0f / 2e-3f, 1.0f / 2e-3f, 1.0f / 1e-2f,
                         1.0f / 1e-2f, 1.0f / 1e-2f, 1.0f / 2e-3f, 0.0f};
    DN_TRY(hipMemcpyAsync(Q.bw, bw, sizeof(bw), hipMemcpyHostToDevice, Q.st));

    // 1. depth -> [0, 1]
    {
        const int tot = nA + nB;
        const float wa = tot > 0 ? (float)nA / (float)tot : 1.0f, wb = 1.0f - wa;
        hipLaunchKernelGGL(k_depth_combine, dim3(blocks(n)), dim3(kBlk), 0, Q.st, dA, dB, n, wa, wb, nB, comb);
        float p99 = 1.0f;
        int rc = percentile99(Q, comb, &p99);
        if (rc) return rc;
        const float far = std::max(1.05f * p99, 1e-6f), invFar = 1.0f / far;
        hipLaunchKernelGGL(k_depth_norm, dim3(blocks(n)), dim3(kBlk), 0, Q.st, dA, dB, n, invFar, dNa, dNb);
    }
    // 2. prefiltered features: albedo, normal, depth planes back to back
    int rc = prefilter(Q, aA, aB, 3, nA, nB, mean, hv, s0, s1, fa, fb, feat);
    if (!rc) rc = prefilter(Q, nmA, nmB, 3, nA, nB, mean, hv, s0, s1, fa, fb, feat + 3 * n);
    if (!rc) rc = prefilter(Q, dNa, dNb, 1, nA, nB, mean, hv, s0, s1, fa, fb, feat + 6 * n);
    // 3. beauty variance
    if (!rc) rc = combine(Q, bA, bB, 3 * n, nA, nB, bmean);
    if (!rc) rc = half_variance(Q, bA, bB, bmean, 3, s0, s1, bhv);
    // 4. joint NL-means
    if (!rc) {
        const float k = quality == RT_DENOISE_HIGH ? 1.0f : 0.7f;
        const int R = quality == RT_DENOISE_HIGH ? 9 : 7;
        rc = nlm(Q, bB, bhv, 3, bA, 3, R, k, 7, feat, oa);
        if (!rc) rc = nlm(Q, bA, bhv, 3, bB, 3, R, k, 7, feat, ob);
        if (!rc) rc = combine(Q, oa, ob, 3 * n, nA, nB, res);
    }
    if (rc) return rc;
    DN_TRY(hipGetLastError());
    DN_TRY(hipMemcpyAsync(beauty_out, res, p3, hipMemcpyDeviceToHost, Q.st));
    DN_TRY(hipStreamSynchronize(Q.st));
    return RT_OK;
}
