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
// batch of filterings (the six feature-prefilter passes run as one batch,
// the two halves of the main filter as another): k_nlm_rows gives each
// (filtering, row, offset) one thread that walks the row with the
// reference's running patch sum (its rounding depends on the whole walk, so
// a row cannot be split) and writes one horizontally averaged plane per
// offset; k_nlm_accum gives each (filtering, pixel) one thread that takes
// the offsets in the reference's order (dy major), sums the vertical patch,
// forms the weight and accumulates.  The planes of a batch (up to 768, 6.4
// GB at 1920x1080) live in HBM between the two kernels.  Every fp32 operation is the reference's, in its order; its
// Vector<float> body (8 wide on x64) and scalar tail differ in one rounding
// (reciprocal multiply vs division of the vertical patch sum), and the
// same split is made per row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
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

// One NL-means filtering of a batch: guide g / variance v (gc planes), target
// t (tc <= 3 planes), output out; the dh planes of the batch's job j start at
// dh + j * no * n.
struct NlmJob {
    const float* g;
    const float* v;
    const float* t;
    float* out;
    int gc, tc, pad0, pad1;
};

// Pass 1 of NlMeansCore.Filter for every (job, row, offset) of a chunk:
// pointwise distances and the running horizontal patch average
// (NlMeansCore.cs:177-195).  Each distance is evaluated once; the running
// sum subtracts it 2P+1 steps later from a per-thread ring of 8 in LDS.
__global__ void __launch_bounds__(kBlk) k_nlm_rows(const NlmJob* __restrict__ jobs, Geom G, float* __restrict__ dh) {
    __shared__ float ring[kBlk * 8];
    const NlmJob J = jobs[blockIdx.y];
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
    const float invC = 1.0f / J.gc;
    float* out = dh + ((int64_t)blockIdx.y * G.no + ol) * n + row;
    float* rg = ring + threadIdx.x * 8;
    constexpr int P = kPatch;
    static_assert(2 * P + 1 < 8, "ring of 8");
    float run = 0.0f;
    const int warm = min(x0 + P, x1);
    for (int xx = x0; xx <= warm; xx++) {
        const float d = dpt_at(J.g, J.v, n, J.gc, invC, row + xx, row + xx + qOff);
        rg[xx & 7] = d;
        run += d;
    }
    for (int x = x0; x <= x1; x++) {
        const int lo = max(x0, x - P), hi = min(x1, x + P);
        out[x] = run / (float)(hi - lo + 1);
        if (x + P + 1 <= x1) {
            const int m = x + P + 1;
            const float d = dpt_at(J.g, J.v, n, J.gc, invC, row + m, row + m + qOff);
            rg[m & 7] = d;
            run += d;
        }
        if (x - P >= x0) run -= rg[(x - P) & 7];
    }
}

// Pass 2 of NlMeansCore.Filter for one chunk of offsets (NlMeansCore.cs:199-263),
// per (job, pixel) in offset order; num/den carry over between chunks in
// memory, the last chunk normalises (:268-281).
__global__ void __launch_bounds__(kBlk) k_nlm_accum(const NlmJob* __restrict__ jobs, Geom G,
                                                    const float* __restrict__ dh, float invK2, int nf,
                                                    const float* __restrict__ feat, const float* __restrict__ fbw,
                                                    float* __restrict__ num, float* __restrict__ den, int first,
                                                    int last) {
    const NlmJob J = jobs[blockIdx.y];
    const int64_t n = (int64_t)G.w * G.h;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int tc = J.tc;
    const float* __restrict__ t = J.t;
    float* nm = num + (int64_t)blockIdx.y * 3 * n;
    float* dn = den + (int64_t)blockIdx.y * n;
    const float* dhj = dh + (int64_t)blockIdx.y * G.no * n;
    const int x = (int)(p % G.w), y = (int)(p / G.w);
    float acc[3] = {0.0f, 0.0f, 0.0f};
    float dsum = 0.0f;
    if (!first) {
        for (int c = 0; c < tc; c++) acc[c] = nm[(int64_t)c * n + p];
        dsum = dn[p];
    }
    constexpr int P = kPatch;
    for (int ol = 0; ol < G.no; ol++) {
        int dx, dy, x0, x1, y0, y1;
        offset_region(G, G.o0 + ol, dx, dy, x0, x1, y0, y1);
        if (x0 > x1 || y0 > y1 || x < x0 || x > x1 || y < y0 || y > y1) continue;
        const int vlo = max(y0, y - P), vhi = min(y1, y + P);
        const int cnt = vhi - vlo + 1;
        const float* col = dhj + (int64_t)ol * n + x;
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
        for (int c = 0; c < tc; c++) J.out[(int64_t)c * n + p] = acc[c] * invDen;
    } else {
        for (int c = 0; c < tc; c++) nm[(int64_t)c * n + p] = acc[c];
        dn[p] = dsum;
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

// histogram of one digit of the order keys of the non-negative values whose
// higher bits equal `prefix` (under `mask`); per-block counts in LDS
constexpr int kDigitBins = 2048;
__global__ void __launch_bounds__(kBlk) k_pct_hist(const float* x, int64_t n, int shift, uint32_t dmask,
                                                   uint32_t mask, uint32_t prefix, int32_t* hist) {
    __shared__ int32_t lh[kDigitBins];
    for (int i = threadIdx.x; i < kDigitBins; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = x[i];
        if (!(v >= 0.0f)) continue;
        const uint32_t k = nonneg_key(v);
        if ((k & mask) == prefix) atomicAdd(&lh[(k >> shift) & dmask], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kDigitBins; i += blockDim.x)
        if (lh[i]) atomicAdd(&hist[i], lh[i]);
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

// ── NFOR (NforRegression.cs) ──────────────────────────────────────────────
constexpr int kDim = 8, kFeat = 7;                 // [1, albedo 3, normal 3, depth]
constexpr int kTileCX = 16, kTileCY = 8;           // centres per k_nfor_dist workgroup
constexpr int kBW = 2 * kTileCX - 1 + 2 * kPatch;  // its distance block
constexpr int kBH = 2 * kTileCY - 1 + 2 * kPatch;
constexpr int kSplatTile = 64;                     // NforRegression.TileSize

struct NforGeom {
    int w, h, R, side, offsets;
    int cxn, cyn;      // centres: every (even x, even y) pixel (the stride-2 grid)
};

// Patch distance of every (centre, offset) pair (ProcessTile, NforRegression.cs:345-381):
// per offset, pointwise distances of the block in LDS, then each centre's
// clipped patch summed row by row and divided by its pixel count; INF when
// the centre has no partner inside the image.  dist is offset-major.
__global__ void __launch_bounds__(kBlk) k_nfor_dist(const float* __restrict__ g, const float* __restrict__ v, NforGeom G,
                                                    float* __restrict__ dist) {
    __shared__ float blk[kBW * kBH];
    const int64_t n = (int64_t)G.w * G.h, nc = (int64_t)G.cxn * G.cyn;
    const int tilesX = (G.cxn + kTileCX - 1) / kTileCX;
    const int tcx = blockIdx.x % tilesX, tcy = blockIdx.x / tilesX;
    const int px0 = 2 * kTileCX * tcx, py0 = 2 * kTileCY * tcy;
    const int bx0 = px0 - kPatch, by0 = py0 - kPatch;
    const int lcx = threadIdx.x % kTileCX, lcy = threadIdx.x / kTileCX;
    const bool is_c = threadIdx.x < kTileCX * kTileCY;
    const int gx = px0 + 2 * lcx, gy = py0 + 2 * lcy;
    const bool has_c = is_c && gx < G.w && gy < G.h;
    const int64_t ci = (int64_t)(gy / 2) * G.cxn + gx / 2;
    const float invC = 1.0f / 3.0f;
    constexpr int P = kPatch;
    for (int o = 0; o < G.offsets; o++) {
        const int dx = o % G.side - G.R, dy = o / G.side - G.R;
        const int pxLo = max(0, -dx), pxHi = G.w - 1 - max(0, dx);
        const int pyLo = max(0, -dy), pyHi = G.h - 1 - max(0, dy);
        const int64_t qOff = (int64_t)dy * G.w + dx;
        for (int i = threadIdx.x; i < kBW * kBH; i += blockDim.x) {
            const int xx = bx0 + i % kBW, yy = by0 + i / kBW;
            float d = 0.0f;
            if (xx >= pxLo && xx <= pxHi && yy >= pyLo && yy <= pyHi) {
                const int64_t p = (int64_t)yy * G.w + xx;
                d = dpt_at(g, v, n, 3, invC, p, p + qOff);
            }
            blk[i] = d;
        }
        __syncthreads();
        if (has_c) {
            float r;
            if (gx < pxLo || gx > pxHi || gy < pyLo || gy > pyHi) {
                r = INFINITY;
            } else {
                const int xLo = max(gx - P, pxLo), xHi = min(gx + P, pxHi);
                const int yLo = max(gy - P, pyLo), yHi = min(gy + P, pyHi);
                float sum = 0.0f;
                for (int yy = yLo; yy <= yHi; yy++)
                    for (int xx = xLo; xx <= xHi; xx++) sum += blk[(yy - by0) * kBW + (xx - bx0)];
                r = sum / (float)((xHi - xLo + 1) * (yHi - yLo + 1));
            }
            dist[(int64_t)o * nc + ci] = r;
        }
        __syncthreads();
    }
}

// Per centre and candidate: weights, the weighted normal equations over the
// window in offset order, the regularised Cholesky solve in double
// (RegressionSolver.cs:19-66) and the coefficients (NforRegression.cs:383-461).
// Only the lower triangle of the Gram matrix is formed: the solver reads no
// other element.  Weights go to W (offset-major like dist, so a wave's
// stores are contiguous; 0 = no contribution).
__global__ void __launch_bounds__(kBlk) k_nfor_solve(const float* __restrict__ dist, const float* __restrict__ feat,
                                                     const float* __restrict__ t, NforGeom G, float invK2,
                                                     float* __restrict__ W, float* __restrict__ beta) {
    const int64_t nc = (int64_t)G.cxn * G.cyn, n = (int64_t)G.w * G.h;
    const int64_t ci = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= nc) return;
    const int gx = 2 * (int)(ci % G.cxn), gy = 2 * (int)(ci / G.cxn);
    const int64_t p = (int64_t)gy * G.w + gx;
    float fp[kFeat];
    for (int i = 0; i < kFeat; i++) fp[i] = feat[(int64_t)i * n + p];
    float gram[kDim * (kDim + 1) / 2];   // lower triangle, row by row
    float rhs[3 * kDim];
    for (int i = 0; i < kDim * (kDim + 1) / 2; i++) gram[i] = 0.0f;
    for (int i = 0; i < 3 * kDim; i++) rhs[i] = 0.0f;
    for (int o = 0; o < G.offsets; o++) {
        const float dd = dist[(int64_t)o * nc + ci];
        float wgt = 0.0f;
        if (!(isinf(dd) && dd > 0.0f)) {
            const float d = maxz(dd) * invK2;
            if (d < kCutoff) wgt = fast_exp_neg(d);
        }
        W[(int64_t)o * nc + ci] = wgt;
        if (wgt == 0.0f) continue;
        const int dx = o % G.side - G.R, dy = o / G.side - G.R;
        const int64_t q = p + (int64_t)dy * G.w + dx;
        float f[kDim];
        f[0] = 1.0f;
        for (int i = 0; i < kFeat; i++) f[i + 1] = feat[(int64_t)i * n + q] - fp[i];
        const float wy0 = wgt * t[q], wy1 = wgt * t[n + q], wy2 = wgt * t[2 * n + q];
        int k = 0;
        for (int i = 0; i < kDim; i++) {
            const float wf = wgt * f[i];
            for (int j = 0; j <= i; j++) gram[k++] += wf * f[j];
        }
        for (int j = 0; j < kDim; j++) {
            rhs[j] += wy0 * f[j];
            rhs[kDim + j] += wy1 * f[j];
            rhs[2 * kDim + j] += wy2 * f[j];
        }
    }
    float* b = beta + ci * 3 * kDim;
    const float sumW = gram[0];
    if (sumW <= 0.0f) {
        for (int i = 0; i < 3 * kDim; i++) b[i] = 0.0f;
        return;
    }
    double L[kDim * (kDim + 1) / 2], B[3 * kDim];
    for (int i = 0; i < kDim * (kDim + 1) / 2; i++) L[i] = gram[i];
    for (int i = 0; i < 3 * kDim; i++) B[i] = rhs[i];
    double mf[3];
    for (int ch = 0; ch < 3; ch++) mf[ch] = B[ch * kDim] / (double)sumW;
    auto at = [](int i, int j) { return i * (i + 1) / 2 + j; };   // j <= i
    double trace = 0;
    for (int i = 0; i < kDim; i++) trace += L[at(i, i)];
    const double lambda = 1e-3 * trace / kDim + 1e-12;
    for (int i = 0; i < kDim; i++) L[at(i, i)] += lambda;
    bool ok = true;
    for (int i = 0; i < kDim && ok; i++)
        for (int j = 0; j <= i; j++) {
            double sum = L[at(i, j)];
            for (int k = 0; k < j; k++) sum -= L[at(i, k)] * L[at(j, k)];
            if (i == j) {
                if (sum <= 0) {
                    ok = false;
                    break;
                }
                L[at(i, i)] = sqrt(sum);
            } else {
                L[at(i, j)] = sum / L[at(j, j)];
            }
        }
    if (ok) {
        for (int r = 0; r < 3; r++) {
            double* x = B + r * kDim;
            for (int i = 0; i < kDim; i++) {
                double sum = x[i];
                for (int k = 0; k < i; k++) sum -= L[at(i, k)] * x[k];
                x[i] = sum / L[at(i, i)];
            }
            for (int i = kDim - 1; i >= 0; i--) {
                double sum = x[i];
                for (int k = i + 1; k < kDim; k++) sum -= L[at(k, i)] * x[k];
                x[i] = sum / L[at(i, i)];
            }
        }
        for (int i = 0; i < 3 * kDim; i++) b[i] = (float)B[i];
    } else {
        for (int i = 0; i < 3 * kDim; i++) b[i] = 0.0f;
        for (int ch = 0; ch < 3; ch++) b[ch * kDim] = (float)mf[ch];
    }
}

// The collaborative splat as a gather (NforRegression.cs:463-516 and the
// normalisation :288-311): pixel q takes the predictions of the centres
// whose window holds it in the order the reference adds them — tiles by
// checkerboard pass, centres row by row inside a tile — then divides.
__global__ void __launch_bounds__(kBlk) k_nfor_gather(const float* __restrict__ W, const float* __restrict__ beta,
                                                      const float* __restrict__ feat, const float* __restrict__ t,
                                                      NforGeom G, float* __restrict__ out) {
    const int64_t n = (int64_t)G.w * G.h, nc = (int64_t)G.cxn * G.cyn;
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int x = (int)(q % G.w), y = (int)(q / G.w);
    float fq[kFeat];
    for (int i = 0; i < kFeat; i++) fq[i] = feat[(int64_t)i * n + q];
    float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, den = 0.0f;
    const int xa = max(0, x - G.R), xb = min(G.w - 1, x + G.R);
    const int ya = max(0, y - G.R), yb = min(G.h - 1, y + G.R);
    const int txa = xa / kSplatTile, txb = xb / kSplatTile, tya = ya / kSplatTile, tyb = yb / kSplatTile;
    for (int pass = 0; pass < 4; pass++)
        for (int ty = tya; ty <= tyb; ty++) {
            if ((ty & 1) != (pass >> 1)) continue;
            for (int tx = txa; tx <= txb; tx++) {
                if ((tx & 1) != (pass & 1)) continue;
                const int gy0 = max(ty * kSplatTile, ya), gy1 = min(ty * kSplatTile + kSplatTile - 1, yb);
                const int gx0 = max(tx * kSplatTile, xa), gx1 = min(tx * kSplatTile + kSplatTile - 1, xb);
                for (int gy = gy0 + (gy0 & 1); gy <= gy1; gy += 2)
                    for (int gx = gx0 + (gx0 & 1); gx <= gx1; gx += 2) {
                        const int64_t ci = (int64_t)(gy / 2) * G.cxn + gx / 2;
                        const int o = (y - gy + G.R) * G.side + (x - gx + G.R);
                        const float wgt = W[(int64_t)o * nc + ci];
                        if (wgt <= 0.0f) continue;
                        const int64_t pc = (int64_t)gy * G.w + gx;
                        float f[kDim];
                        f[0] = 1.0f;
                        for (int i = 0; i < kFeat; i++) f[i + 1] = fq[i] - feat[(int64_t)i * n + pc];
                        const float* b = beta + ci * 3 * kDim;
                        float pr[3];
                        for (int ch = 0; ch < 3; ch++) {
                            float e[kDim];
                            for (int i = 0; i < kDim; i++) e[i] = b[ch * kDim + i] * f[i];
                            pr[ch] = ((e[0] + e[1]) + (e[2] + e[3])) + ((e[4] + e[5]) + (e[6] + e[7]));
                        }
                        acc0 += wgt * pr[0];
                        acc1 += wgt * pr[1];
                        acc2 += wgt * pr[2];
                        den += wgt;
                    }
            }
        }
    if (den > 0.0f) {
        const float inv = 1.0f / den;
        out[q] = acc0 * inv;
        out[n + q] = acc1 * inv;
        out[2 * n + q] = acc2 * inv;
    } else {
        out[q] = t[q];
        out[n + q] = t[n + q];
        out[2 * n + q] = t[2 * n + q];
    }
}

__global__ void __launch_bounds__(kBlk) k_scale(const float* src, int64_t n, const float* scales, int planes,
                                                float* dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * planes) return;
    dst[i] = src[i] * scales[i / n];
}

// PlaneOps.BoxSmooth (PlaneOps.cs:55-85): running row sums, then column sums
__global__ void __launch_bounds__(kBlk) k_box_rows(const float* src, float* tmp, int w, int h, int r) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= h) return;
    const int64_t row = (int64_t)blockIdx.y * w * h + (int64_t)y * w;
    float run = 0.0f;
    for (int x = 0; x < min(r + 1, w); x++) run += src[row + x];
    for (int x = 0; x < w; x++) {
        const int lo = max(0, x - r), hi = min(w - 1, x + r);
        tmp[row + x] = run / (float)(hi - lo + 1);
        if (x + r + 1 < w) run += src[row + x + r + 1];
        if (x - r >= 0) run -= src[row + x - r];
    }
}

__global__ void __launch_bounds__(kBlk) k_box_cols(const float* tmp, float* dst, int w, int h, int r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int64_t base = (int64_t)blockIdx.y * w * h;
    const int x = (int)(i % w), y = (int)(i / w);
    const int lo = max(0, y - r), hi = min(h - 1, y + r);
    const float inv = 1.0f / (hi - lo + 1);
    float sum = 0.0f;
    for (int yy = lo; yy <= hi; yy++) sum += tmp[base + (int64_t)yy * w + x];
    dst[base + i] = sum * inv;
}

// candidate MSE maps before smoothing (NforRegression.cs:95-133)
__global__ void __launch_bounds__(kBlk) k_mse_noisy(const float* hv, int64_t n, float* tmp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v0 = 0.5f * hv[i], v1 = 0.5f * hv[n + i], v2 = 0.5f * hv[2 * n + i];
    tmp[i] = (v0 + v1 + v2) * (1.0f / 3.0f);
}

__global__ void __launch_bounds__(kBlk) k_mse_filtered(const float* fa, const float* fb, const float* bA,
                                                       const float* bB, const float* hv, int64_t n, float margin,
                                                       float* tmp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float e = 0.0f, vSum = 0.0f;
    for (int c = 0; c < 3; c++) {
        const int64_t p = c * n + i;
        const float dA = fa[p] - bB[p], dB = fb[p] - bA[p];
        e += 0.5f * (dA * dA + dB * dB) - hv[p];
        vSum += 0.5f * hv[p];
    }
    tmp[i] = (e + margin * vSum) * (1.0f / 3.0f);
}

// indicator planes: 1 where candidate ci has the (first) smallest smoothed MSE
__global__ void __launch_bounds__(kBlk) k_indicator(const float* mse, int ncand, int64_t n, float* ind) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float best = mse[i];
    int bi = 0;
    for (int cj = 1; cj < ncand; cj++)
        if (mse[cj * n + i] < best) {
            best = mse[cj * n + i];
            bi = cj;
        }
    for (int ci = 0; ci < ncand; ci++) ind[ci * n + i] = bi == ci ? 1.0f : 0.0f;
}

// softened blend of the candidates, clamped at zero (NforRegression.cs:136-167)
__global__ void __launch_bounds__(kBlk) k_blend(const float* wts, const float* cand, int ncand, int64_t n, float* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float wSum = 0.0f;
    for (int c = 0; c < ncand; c++) wSum += wts[c * n + i];
    const float inv = wSum > 0.0f ? 1.0f / wSum : 0.0f;
    for (int ch = 0; ch < 3; ch++) {
        float acc = 0.0f;
        for (int c = 0; c < ncand; c++) acc += wts[c * n + i] * cand[(int64_t)c * 3 * n + ch * n + i];
        const float r = acc * inv;
        out[ch * n + i] = r < 0.0f ? 0.0f : r;
    }
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

constexpr int kMaxPlanes = 768;   // dh planes per k_nlm_rows / k_nlm_accum pair (jobs x offsets)
constexpr int kMaxJobs = 8;

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
    float* dh;        // kMaxPlanes planes
    float* num;       // kMaxJobs x 3 planes
    float* den;       // kMaxJobs planes
    int32_t* hist;    // kDigitBins bins
    float* bw;        // 8 inverse bandwidths
    NlmJob* jobs;     // kMaxJobs job records (device): slots 0-5 prefilter, 6-7 main filter
    NlmJob host_jobs[8];   // their host copies, alive until the call's final synchronisation
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

// NlMeansCore.Filter for a batch of jobs with one search radius, strength
// and feature guide set, offsets in chunks of at most kMaxPlanes / jobs; the
// batch's job records go to slots [slot0, slot0 + jobs) (stream-ordered copy)
int nlm(Pipe& Q, const std::vector<NlmJob>& jobs, int slot0, int R, float K, int nf, const float* feat) {
    rt_ctx* ctx = Q.ctx;
    const int nj = (int)jobs.size();
    if (slot0 + nj > kMaxJobs) return ctx_fail(ctx, RT_ERR_INTERNAL, "too many NL-means jobs");
    for (int j = 0; j < nj; j++) Q.host_jobs[slot0 + j] = jobs[j];
    DN_TRY(hipMemcpyAsync(Q.jobs + slot0, Q.host_jobs + slot0, sizeof(NlmJob) * nj, hipMemcpyHostToDevice, Q.st));
    const NlmJob* dj = Q.jobs + slot0;
    const int side = 2 * R + 1, offsets = side * side, chunk = kMaxPlanes / nj;
    const float invK2 = 1.0f / (K * K);
    for (int o0 = 0; o0 < offsets; o0 += chunk) {
        Geom G{Q.w, Q.h, R, side, o0, std::min(chunk, offsets - o0)};
        const int64_t rows = (int64_t)Q.h * G.no;
        hipLaunchKernelGGL(k_nlm_rows, dim3(blocks(rows), nj), dim3(kBlk), 0, Q.st, dj, G, Q.dh);
        hipLaunchKernelGGL(k_nlm_accum, dim3(blocks(Q.n), nj), dim3(kBlk), 0, Q.st, dj, G,
                           (const float*)Q.dh, invK2, nf, feat, (const float*)Q.bw, Q.num, Q.den, o0 == 0 ? 1 : 0,
                           o0 + chunk >= offsets ? 1 : 0);
    }
    return RT_OK;
}

// FeaturePrefilter.Prefilter of the albedo, normal and depth halves at once:
// per feature the mean, the half variance, then the six cross NL-means
// filterings (R 5, k 1) in one batch, then the sample-weighted combination.
struct FeatureHalves {
    const float* A;
    const float* B;
    int c;
    float *mean, *hv, *fa, *fb, *out;
};

int prefilter(Pipe& Q, FeatureHalves* F, int nfeat, int nA, int nB, float* s0, float* s1) {
    std::vector<NlmJob> jobs;
    for (int i = 0; i < nfeat; i++) {
        const int64_t len = (int64_t)F[i].c * Q.n;
        int rc = combine(Q, F[i].A, F[i].B, len, nA, nB, F[i].mean);
        if (!rc) rc = half_variance(Q, F[i].A, F[i].B, F[i].mean, F[i].c, s0, s1, F[i].hv);
        if (rc) return rc;
        jobs.push_back(NlmJob{F[i].B, F[i].hv, F[i].A, F[i].fa, F[i].c, F[i].c, 0, 0});
        jobs.push_back(NlmJob{F[i].A, F[i].hv, F[i].B, F[i].fb, F[i].c, F[i].c, 0, 0});
    }
    int rc = nlm(Q, jobs, 0, 5, 1.0f, 0, nullptr);
    for (int i = 0; i < nfeat && !rc; i++) rc = combine(Q, F[i].fa, F[i].fb, (int64_t)F[i].c * Q.n, nA, nB, F[i].out);
    return rc;
}

// PlaneOps.PercentileOfNonNegative (0.99): an exact radix select over the
// order keys of the non-negative values, digits of 11, 11 and 10 bits
int percentile99(Pipe& Q, const float* x, float* result) {
    rt_ctx* ctx = Q.ctx;
    std::vector<int32_t> hist(kDigitBins);
    const int shifts[3] = {21, 10, 0};
    uint32_t prefix = 0, mask = 0;
    int64_t before = 0, k = 0;
    const unsigned grid = (unsigned)std::min<int64_t>(1024, blocks(Q.n));
    for (int pass = 0; pass < 3; pass++) {
        const int nbins = pass == 2 ? 1024 : kDigitBins;
        DN_TRY(hipMemsetAsync(Q.hist, 0, kDigitBins * 4, Q.st));
        hipLaunchKernelGGL(k_pct_hist, dim3(grid), dim3(kBlk), 0, Q.st, x, Q.n, shifts[pass], (uint32_t)(nbins - 1),
                           mask, prefix, Q.hist);
        DN_TRY(hipMemcpyAsync(hist.data(), Q.hist, kDigitBins * 4, hipMemcpyDeviceToHost, Q.st));
        DN_TRY(hipStreamSynchronize(Q.st));
        if (pass == 0) {
            int64_t count = 0;
            for (int i = 0; i < kDigitBins; i++) count += hist[i];
            if (count == 0) {
                *result = 1.0f;
                return RT_OK;
            }
            k = (int64_t)(0.99f * (float)(count - 1));
            k = std::min<int64_t>(std::max<int64_t>(k, 0), count - 1);
        }
        int d = 0;
        for (; d < nbins; d++) {
            if (before + hist[d] > k) break;
            before += hist[d];
        }
        prefix |= (uint32_t)d << shifts[pass];
        mask |= (uint32_t)(nbins - 1) << shifts[pass];
    }
    memcpy(result, &prefix, 4);
    return RT_OK;
}

// PlaneOps.BoxSmooth of `planes` consecutive planes (tmp: as many scratch planes)
void box_smooth(Pipe& Q, const float* src, float* tmp, float* dst, int r, int planes) {
    hipLaunchKernelGGL(k_box_rows, dim3(blocks(Q.h), planes), dim3(kBlk), 0, Q.st, src, tmp, Q.w, Q.h, r);
    hipLaunchKernelGGL(k_box_cols, dim3(blocks(Q.n), planes), dim3(kBlk), 0, Q.st, tmp, dst, Q.w, Q.h, r);
}

// NforRegression.BuildScaledFeatures (NforRegression.cs:173-193): the global
// standard deviation of each feature plane from sequential double sums, in
// the reference's order (host threads, one per plane), then the scaling on
// the device.
int scale_features(Pipe& Q, const float* feat, float* scales_dev, float* out) {
    rt_ctx* ctx = Q.ctx;
    const int64_t n = Q.n;
    std::vector<float> host((size_t)kFeat * n);
    DN_TRY(hipMemcpyAsync(host.data(), feat, host.size() * 4, hipMemcpyDeviceToHost, Q.st));
    DN_TRY(hipStreamSynchronize(Q.st));
    float scales[kFeat];
    std::vector<std::thread> th;
    for (int pl = 0; pl < kFeat; pl++)
        th.emplace_back([&, pl] {
            const float* src = host.data() + (size_t)pl * n;
            double sum = 0, sumSq = 0;
            for (int64_t i = 0; i < n; i++) {
                sum += src[i];
                sumSq += (double)src[i] * src[i];
            }
            const double mean = sum / n;
            const float sd = (float)std::sqrt(std::max(sumSq / n - mean * mean, 0.0));
            scales[pl] = 1.0f / std::max(sd, 1e-4f);
        });
    for (auto& t : th) t.join();
    DN_TRY(hipMemcpy(scales_dev, scales, sizeof(scales), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_scale, dim3(blocks(kFeat * n)), dim3(kBlk), 0, Q.st, feat, n, (const float*)scales_dev, kFeat,
                       out);
    return RT_OK;
}

struct NforBufs {
    float *featS, *dist, *W, *beta, *reg, *cand, *mse, *wts, *tmp, *tmp2, *scales;
};

// NforRegression.Run (NforRegression.cs:54-170)
int nfor(Pipe& Q, const NforBufs& B, const float* bA, const float* bB, const float* bmean, const float* bhv,
         const float* feat, int quality, float margin, int nA, int nB, float* res) {
    rt_ctx* ctx = Q.ctx;
    const int64_t n = Q.n;
    int rc = scale_features(Q, feat, B.scales, B.featS);
    if (rc) return rc;
    const float ksHigh[2] = {0.5f, 1.0f}, ksFast[1] = {0.7f};
    const float* ks = quality == RT_DENOISE_HIGH ? ksHigh : ksFast;
    const int nk = quality == RT_DENOISE_HIGH ? 2 : 1;
    const int R = quality == RT_DENOISE_HIGH ? 9 : 7;
    NforGeom G{Q.w, Q.h, R, 2 * R + 1, (2 * R + 1) * (2 * R + 1), (Q.w + 1) / 2, (Q.h + 1) / 2};
    const int64_t nc = (int64_t)G.cxn * G.cyn;
    const int tiles = ((G.cxn + kTileCX - 1) / kTileCX) * ((G.cyn + kTileCY - 1) / kTileCY);
    // regressions: guide B -> target A (reg[0..nk)), guide A -> target B (reg[nk..2nk))
    for (int half = 0; half < 2; half++) {
        const float* g = half == 0 ? bB : bA;
        const float* t = half == 0 ? bA : bB;
        hipLaunchKernelGGL(k_nfor_dist, dim3(tiles), dim3(kBlk), 0, Q.st, g, bhv, G, B.dist);
        for (int ki = 0; ki < nk; ki++) {
            const float invK2 = 1.0f / (ks[ki] * ks[ki]);
            hipLaunchKernelGGL(k_nfor_solve, dim3(blocks(nc)), dim3(kBlk), 0, Q.st, (const float*)B.dist,
                               (const float*)B.featS, t, G, invK2, B.W, B.beta);
            hipLaunchKernelGGL(k_nfor_gather, dim3(blocks(n)), dim3(kBlk), 0, Q.st, (const float*)B.W,
                               (const float*)B.beta, (const float*)B.featS, t, G,
                               B.reg + (int64_t)(half * nk + ki) * 3 * n);
        }
    }
    // candidates: the unfiltered mean, then one per k
    const int ncand = nk + 1;
    DN_TRY(hipMemcpyAsync(B.cand, bmean, (size_t)n * 12, hipMemcpyDeviceToDevice, Q.st));
    for (int ki = 0; ki < nk; ki++) {
        rc = combine(Q, B.reg + (int64_t)ki * 3 * n, B.reg + (int64_t)(nk + ki) * 3 * n, 3 * n, nA, nB,
                     B.cand + (int64_t)(ki + 1) * 3 * n);
        if (rc) return rc;
    }
    // per-candidate MSE maps (tmp: ncand planes), smoothed together
    hipLaunchKernelGGL(k_mse_noisy, dim3(blocks(n)), dim3(kBlk), 0, Q.st, bhv, n, B.tmp);
    for (int ki = 0; ki < nk; ki++)
        hipLaunchKernelGGL(k_mse_filtered, dim3(blocks(n)), dim3(kBlk), 0, Q.st,
                           (const float*)(B.reg + (int64_t)ki * 3 * n), (const float*)(B.reg + (int64_t)(nk + ki) * 3 * n),
                           bA, bB, bhv, n, margin, B.tmp + (int64_t)(ki + 1) * n);
    box_smooth(Q, B.tmp, B.tmp2, B.mse, 4, ncand);
    hipLaunchKernelGGL(k_indicator, dim3(blocks(n)), dim3(kBlk), 0, Q.st, (const float*)B.mse, ncand, n, B.tmp);
    box_smooth(Q, B.tmp, B.tmp2, B.wts, 2, ncand);
    hipLaunchKernelGGL(k_blend, dim3(blocks(n)), dim3(kBlk), 0, Q.st, (const float*)B.wts, (const float*)B.cand, ncand,
                       n, res);
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
    const int nA = opts->samples_a, nB = opts->samples_b;
    if (nA <= 0 || nB < 0) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "samples_a must be > 0 and samples_b >= 0");
    if ((int64_t)width * height > (1ll << 28)) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "film too large");
    DN_TRY(hipSetDevice(ctx_device(ctx)));

    Pipe Q{ctx, ctx_stream(ctx), width, height, (int64_t)width * height};
    const int64_t n = Q.n;
    // device layout (floats): inputs 20 n, per-feature temporaries 28 n,
    // features 7 n, beauty temporaries, kMaxPlanes dh planes, NFOR buffers
    float *bA, *bB, *aA, *aB, *nmA, *nmB, *dA, *dB, *comb, *dNa, *dNb, *feat, *ftmp, *s0, *s1, *bmean, *bhv, *oa,
        *ob, *res;
    NforBufs NB{};
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
        ftmp = A.f(28 * n), s0 = A.f(n), s1 = A.f(n);
        bmean = A.f(3 * n), bhv = A.f(3 * n), oa = A.f(3 * n), ob = A.f(3 * n), res = A.f(3 * n);
        Q.dh = A.f((int64_t)kMaxPlanes * n);
        Q.num = A.f((int64_t)kMaxJobs * 3 * n);
        Q.den = A.f((int64_t)kMaxJobs * n);
        Q.hist = (int32_t*)A.f(kDigitBins);
        Q.bw = A.f(8);
        Q.jobs = (NlmJob*)A.f(kMaxJobs * (int64_t)sizeof(NlmJob) / 4);
        if (kind == RT_DENOISE_NFOR) {
            const int R = quality == RT_DENOISE_HIGH ? 9 : 7, nk = quality == RT_DENOISE_HIGH ? 2 : 1;
            const int64_t offs = (int64_t)(2 * R + 1) * (2 * R + 1);
            const int64_t nc = (int64_t)((width + 1) / 2) * ((height + 1) / 2);
            NB.featS = A.f(kFeat * n), NB.dist = A.f(offs * nc), NB.W = A.f(offs * nc), NB.beta = A.f(nc * 3 * kDim);
            NB.reg = A.f(2 * nk * 3 * n), NB.cand = A.f((nk + 1) * 3 * n), NB.mse = A.f((nk + 1) * n);
            NB.wts = A.f((nk + 1) * n), NB.tmp = A.f((nk + 1) * n), NB.tmp2 = A.f((nk + 1) * n), NB.scales = A.f(8);
        }
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
    static const float bw[8] = {1.0f / 2e-3f, 1.0f / 2e-3f, 1.0f / 2e-3f, 1.0f / 1e-2f,
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
    // 2. prefiltered features: albedo, normal, depth planes back to back in feat
    FeatureHalves F[3] = {{aA, aB, 3, ftmp, ftmp + 3 * n, ftmp + 6 * n, ftmp + 9 * n, feat},
                          {nmA, nmB, 3, ftmp + 12 * n, ftmp + 15 * n, ftmp + 18 * n, ftmp + 21 * n, feat + 3 * n},
                          {dNa, dNb, 1, ftmp + 24 * n, ftmp + 25 * n, ftmp + 26 * n, ftmp + 27 * n, feat + 6 * n}};
    int rc = prefilter(Q, F, 3, nA, nB, s0, s1);
    // 3. beauty variance
    if (!rc) rc = combine(Q, bA, bB, 3 * n, nA, nB, bmean);
    if (!rc) rc = half_variance(Q, bA, bB, bmean, 3, s0, s1, bhv);
    // 4. main filter: NFOR regression or joint NL-means (both halves in one batch)
    if (!rc && kind == RT_DENOISE_NFOR) {
        const float margin = opts->selection_margin >= 0.0f ? opts->selection_margin : 0.5f;
        rc = nfor(Q, NB, bA, bB, bmean, bhv, feat, quality, margin, nA, nB, res);
    } else if (!rc) {
        const float k = quality == RT_DENOISE_HIGH ? 1.0f : 0.7f;
        const int R = quality == RT_DENOISE_HIGH ? 9 : 7;
        std::vector<NlmJob> jobs = {NlmJob{bB, bhv, bA, oa, 3, 3, 0, 0}, NlmJob{bA, bhv, bB, ob, 3, 3, 0, 0}};
        rc = nlm(Q, jobs, 6, R, k, 7, feat);
        if (!rc) rc = combine(Q, oa, ob, 3 * n, nA, nB, res);
    }
    if (rc) return rc;
    DN_TRY(hipGetLastError());
    DN_TRY(hipMemcpyAsync(beauty_out, res, p3, hipMemcpyDeviceToHost, Q.st));
    DN_TRY(hipStreamSynchronize(Q.st));
    return RT_OK;
}
