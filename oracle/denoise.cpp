// CPU restatement of the reference's denoiser (src/RayTracer/Denoising/*.cs).
//
// TEST INFRASTRUCTURE ONLY: built into liboracle.so and loaded by tests/ and
// bench.py's cpu_baseline leg, never by the product.
//
//   orc_nlm_filter        NlMeansCore.Filter        NlMeansCore.cs:147-283
//                         (DistanceRow :103-139, FastExpNegScalar :90-94)
//   orc_binomial_smooth7  PlaneOps.BinomialSmooth7  PlaneOps.cs:18-48
//   orc_box_smooth        PlaneOps.BoxSmooth        PlaneOps.cs:55-85
//   orc_percentile_nonneg PlaneOps.PercentileOfNonNegative  PlaneOps.cs:90-99
//   orc_regression_solve  RegressionSolver.Solve    RegressionSolver.cs:19-66
//   orc_denoise           NforDenoiser.Denoise      NforDenoiser.cs:29-57
//                         (NormaliseDepth :92-112, JointNlm :66-84,
//                          FeaturePrefilter.Prefilter FeaturePrefilter.cs:19-28,
//                          VarianceEstimator.cs:24-70, PlaneOps.Combine PlaneOps.cs:104-118,
//                          NforRegression.Run :54-170, BuildScaledFeatures :173-193,
//                          RegressAllCandidates :227-313, ProcessTile :316-518)
//
// Loops follow the reference's order and fp32 operations one by one (the
// library is built with -ffp-contract=off).  Where the reference runs a
// System.Numerics.Vector<float> loop with a scalar tail that rounds
// differently (the vertical patch average multiplies by the reciprocal of
// the count in the vector body and divides in the tail), the vector width is
// 8 — Vector<float> on x64 — and the same split is made.  Where it uses
// Vector256 (the regression's right-hand sides and the splat's dot
// products, with Vector256 hardware acceleration), the association of those
// products is the vector one, and Vector256.Sum adds its 128-bit halves,
// each the sum of its 64-bit halves: ((e0+e1)+(e2+e3)) + ((e4+e5)+(e6+e7)).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kVecW = 8;
constexpr float kEps = 1e-10f;
constexpr float kCutoff = 10.0f;

inline float fast_exp_neg(float d) {
    // NlMeansCore.FastExpNegScalar: Schraudolph's exp(-d) bit trick
    int i = (int)(-12102203.16f * d + 1064866805.0f);
    float f;
    memcpy(&f, &i, 4);
    return f;
}

inline float fmaxz(float a) { return a > 0.0f ? a : 0.0f; }

void distance_row(const float* g, const float* v, int n, int channels, int p0, int q0, int count, float* dst,
                  int dst_base) {
    const float invC = 1.0f / channels;
    for (int i = 0; i < count; i++) {
        float d = 0.0f;
        for (int c = 0; c < channels; c++) {
            const int cb = c * n;
            const float du = g[cb + p0 + i] - g[cb + q0 + i];
            const float vp = v[cb + p0 + i], vq = v[cb + q0 + i];
            const float vmin = vp < vq ? vp : vq;
            d += (du * du - (vp + vmin)) / (kEps + vp + vq);
        }
        dst[dst_base + i] = d * invC;
    }
}

void nlm_filter(int w, int h, int gc, const float* g, const float* v, int tc, const float* t, int R, int P, float K,
                int fplanes, const float* fdata, const float* fbw, float* out) {
    const int n = w * h;
    const float invK2 = 1.0f / (K * K);
    std::vector<float> num((size_t)tc * n, 0.0f), den(n, 0.0f), dpt(n, 0.0f), dh(n, 0.0f);
    for (int dy = -R; dy <= R; dy++)
        for (int dx = -R; dx <= R; dx++) {
            const int x0 = std::max(0, -dx), x1 = w - 1 - std::max(0, dx);
            const int y0 = std::max(0, -dy), y1 = h - 1 - std::max(0, dy);
            if (x0 > x1 || y0 > y1) continue;
            const int qOff = dy * w + dx;
            for (int y = y0; y <= y1; y++) {
                const int row = y * w;
                distance_row(g, v, n, gc, row + x0, row + x0 + qOff, x1 - x0 + 1, dpt.data(), row + x0);
                float run = 0.0f;
                const int warm = std::min(x0 + P, x1);
                for (int xx = x0; xx <= warm; xx++) run += dpt[row + xx];
                for (int x = x0; x <= x1; x++) {
                    const int lo = std::max(x0, x - P), hi = std::min(x1, x + P);
                    dh[row + x] = run / (float)(hi - lo + 1);
                    if (x + P + 1 <= x1) run += dpt[row + x + P + 1];
                    if (x - P >= x0) run -= dpt[row + x - P];
                }
            }
            const int nvec = ((x1 - x0 + 1) / kVecW) * kVecW;
            for (int y = y0; y <= y1; y++) {
                const int row = y * w;
                const int vlo = std::max(y0, y - P), vhi = std::min(y1, y + P);
                const int cnt = vhi - vlo + 1;
                const float invCntY = 1.0f / cnt;
                for (int x = x0; x <= x1; x++) {
                    const int p = row + x;
                    float sum = 0.0f;
                    for (int yy = vlo; yy <= vhi; yy++) sum += dh[yy * w + x];
                    float d = (x - x0 < nvec) ? fmaxz(sum * invCntY) * invK2 : fmaxz(sum / (float)cnt) * invK2;
                    const int q = p + qOff;
                    for (int fp = 0; fp < fplanes; fp++) {
                        const float df = fdata[(size_t)fp * n + p] - fdata[(size_t)fp * n + q];
                        d += df * df * fbw[fp];
                    }
                    if (!(d < kCutoff)) continue;
                    const float wgt = fast_exp_neg(d);
                    den[p] += wgt;
                    for (int c = 0; c < tc; c++) num[(size_t)c * n + p] += wgt * t[(size_t)c * n + q];
                }
            }
        }
    for (int p = 0; p < n; p++) {
        const float invDen = den[p] > 0.0f ? 1.0f / den[p] : 0.0f;
        for (int c = 0; c < tc; c++) out[(size_t)c * n + p] = num[(size_t)c * n + p] * invDen;
    }
}

void separable_pass(const float* src, float* dst, int w, int h, const float* k, int klen, bool horizontal) {
    const int r = klen / 2;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float sum = 0.0f, wSum = 0.0f;
            for (int i = -r; i <= r; i++) {
                const int xx = horizontal ? x + i : x, yy = horizontal ? y : y + i;
                if (xx < 0 || xx >= w || yy < 0 || yy >= h) continue;
                const float kw = k[i + r];
                sum += kw * src[yy * w + xx];
                wSum += kw;
            }
            dst[y * w + x] = sum / wSum;
        }
}

void binomial_smooth7(const float* src, float* dst, int w, int h) {
    const float k[7] = {1.0f, 6.0f, 15.0f, 20.0f, 15.0f, 6.0f, 1.0f};
    std::vector<float> tmp((size_t)w * h);
    separable_pass(src, tmp.data(), w, h, k, 7, true);
    separable_pass(tmp.data(), dst, w, h, k, 7, false);
}

void box_smooth(const float* src, float* dst, int w, int h, int r) {
    std::vector<float> tmp((size_t)w * h);
    for (int y = 0; y < h; y++) {
        const int row = y * w;
        float run = 0.0f;
        for (int x = 0; x < std::min(r + 1, w); x++) run += src[row + x];
        for (int x = 0; x < w; x++) {
            const int lo = std::max(0, x - r), hi = std::min(w - 1, x + r);
            tmp[row + x] = run / (float)(hi - lo + 1);
            if (x + r + 1 < w) run += src[row + x + r + 1];
            if (x - r >= 0) run -= src[row + x - r];
        }
    }
    for (int y = 0; y < h; y++) {
        const int lo = std::max(0, y - r), hi = std::min(h - 1, y + r);
        const float inv = 1.0f / (hi - lo + 1);
        for (int x = 0; x < w; x++) {
            float sum = 0.0f;
            for (int yy = lo; yy <= hi; yy++) sum += tmp[yy * w + x];
            dst[y * w + x] = sum * inv;
        }
    }
}

float percentile_nonneg(const float* plane, int n, float pct, float fallback) {
    std::vector<float> f;
    f.reserve(n);
    for (int i = 0; i < n; i++)
        if (plane[i] >= 0.0f) f.push_back(plane[i]);
    if (f.empty()) return fallback;
    std::sort(f.begin(), f.end());
    int idx = (int)(pct * (float)(f.size() - 1));
    idx = std::min(std::max(idx, 0), (int)f.size() - 1);
    return f[idx];
}

bool regression_solve(double* gram, double* rhs, int dim, int rhsCount, double lambdaScale) {
    double trace = 0;
    for (int i = 0; i < dim; i++) trace += gram[i * dim + i];
    const double lambda = lambdaScale * trace / dim + 1e-12;
    for (int i = 0; i < dim; i++) gram[i * dim + i] += lambda;
    for (int i = 0; i < dim; i++)
        for (int j = 0; j <= i; j++) {
            double sum = gram[i * dim + j];
            for (int k = 0; k < j; k++) sum -= gram[i * dim + k] * gram[j * dim + k];
            if (i == j) {
                if (sum <= 0) return false;
                gram[i * dim + i] = std::sqrt(sum);
            } else {
                gram[i * dim + j] = sum / gram[j * dim + j];
            }
        }
    for (int r = 0; r < rhsCount; r++) {
        double* b = rhs + r * dim;
        for (int i = 0; i < dim; i++) {
            double sum = b[i];
            for (int k = 0; k < i; k++) sum -= gram[i * dim + k] * b[k];
            b[i] = sum / gram[i * dim + i];
        }
        for (int i = dim - 1; i >= 0; i--) {
            double sum = b[i];
            for (int k = i + 1; k < dim; k++) sum -= gram[k * dim + i] * b[k];
            b[i] = sum / gram[i * dim + i];
        }
    }
    return true;
}

// PlaneOps.Combine: wa = nA / (nA + nB), wb = 1 - wa; nB <= 0 copies a
void combine(const float* a, const float* b, size_t len, int nA, int nB, float* dst) {
    if (nB <= 0) {
        memcpy(dst, a, len * 4);
        return;
    }
    const float wa = (float)nA / (nA + nB), wb = 1.0f - wa;
    for (size_t i = 0; i < len; i++) dst[i] = wa * a[i] + wb * b[i];
}

// VarianceEstimator.FullVariance then HalfFromFull (x2)
void half_variance(const float* A, const float* B, const float* mean, int w, int h, int c, float* half) {
    const int n = w * h;
    std::vector<float> raw(n), sm(n);
    for (int ch = 0; ch < c; ch++) {
        for (int i = 0; i < n; i++) {
            const float d = 0.5f * (A[(size_t)ch * n + i] - B[(size_t)ch * n + i]);
            raw[i] = d * d;
        }
        binomial_smooth7(raw.data(), sm.data(), w, h);
        for (int i = 0; i < n; i++) {
            const float m = mean[(size_t)ch * n + i];
            const float full = std::max(sm[i], std::max(1e-5f * m * m, 1e-12f));
            half[(size_t)ch * n + i] = 2.0f * full;
        }
    }
}

// FeaturePrefilter.Prefilter: cross NL-means of the halves (R 5, P 3, k 1)
void prefilter(const float* A, const float* B, int w, int h, int c, int nA, int nB, float* out) {
    const size_t len = (size_t)c * w * h;
    std::vector<float> mean(len), hv(len), fa(len), fb(len);
    combine(A, B, len, nA, nB, mean.data());
    half_variance(A, B, mean.data(), w, h, c, hv.data());
    nlm_filter(w, h, c, B, hv.data(), c, A, 5, 3, 1.0f, 0, nullptr, nullptr, fa.data());
    nlm_filter(w, h, c, A, hv.data(), c, B, 5, 3, 1.0f, 0, nullptr, nullptr, fb.data());
    combine(fa.data(), fb.data(), len, nA, nB, out);
}

}  // namespace

extern "C" {

void orc_nlm_filter(int w, int h, int gc, const float* g, const float* v, int tc, const float* t, int R, int P, float K,
                    int fplanes, const float* fdata, const float* fbw, float* out) {
    nlm_filter(w, h, gc, g, v, tc, t, R, P, K, fplanes, fdata, fbw, out);
}

void orc_binomial_smooth7(int w, int h, const float* src, float* dst) { binomial_smooth7(src, dst, w, h); }

void orc_box_smooth(int w, int h, const float* src, float* dst, int r) { box_smooth(src, dst, w, h, r); }

float orc_percentile_nonneg(int n, const float* plane, float pct, float fallback) {
    return percentile_nonneg(plane, n, pct, fallback);
}

int orc_regression_solve(double* gram, double* rhs, int dim, int rhs_count, double lambda_scale) {
    return regression_solve(gram, rhs, dim, rhs_count, lambda_scale) ? 1 : 0;
}

// planes: beauty_a/b, albedo_a/b, normal_a/b (3 x w*h), depth_a/b (w*h);
// kind 1 = NL-means, 2 = NFOR; quality 0 = fast, 1 = high
int orc_denoise(int w, int h, const float* beauty_a, const float* beauty_b, const float* albedo_a,
                const float* albedo_b, const float* normal_a, const float* normal_b, const float* depth_a,
                const float* depth_b, int nA, int nB, int kind, int quality, float margin, float* out);

}  // extern "C"

namespace {

// NforRegression (NforRegression.cs:56-518), declared here and defined below
void nfor_run(int w, int h, const float* bA, const float* bB, int nA, int nB, const float* halfVar,
              const float* albedo, const float* normal, const float* depth, int quality, float margin, float* out);

}  // namespace

int orc_denoise(int w, int h, const float* beauty_a, const float* beauty_b, const float* albedo_a,
                const float* albedo_b, const float* normal_a, const float* normal_b, const float* depth_a,
                const float* depth_b, int nA, int nB, int kind, int quality, float margin, float* out) {
    const int n = w * h;
    // 1. depth -> [0, 1] (RenderBuffers.CombineDepthHalves, then the far plane)
    std::vector<float> comb(n);
    {
        const int tot = nA + nB;
        const float wa = tot > 0 ? (float)nA / tot : 1.0f, wb = 1.0f - wa;
        for (int i = 0; i < n; i++) {
            const float va = depth_a[i], vb = depth_b[i];
            const bool ma = va < 0.0f, mb = vb < 0.0f;
            comb[i] = (ma && mb) ? -1.0f : ma ? vb : mb ? va : (nB == 0 ? va : wa * va + wb * vb);
        }
    }
    const float p99 = percentile_nonneg(comb.data(), n, 0.99f, 1.0f);
    const float far = std::max(1.05f * p99, 1e-6f), invFar = 1.0f / far;
    std::vector<float> dA(n), dB(n);
    for (int i = 0; i < n; i++) {
        dA[i] = depth_a[i] < 0.0f ? 1.0f : std::min(depth_a[i] * invFar, 1.0f);
        dB[i] = depth_b[i] < 0.0f ? 1.0f : std::min(depth_b[i] * invFar, 1.0f);
    }
    // 2. prefiltered guide features
    std::vector<float> albedo(3 * (size_t)n), normal(3 * (size_t)n), depth(n);
    prefilter(albedo_a, albedo_b, w, h, 3, nA, nB, albedo.data());
    prefilter(normal_a, normal_b, w, h, 3, nA, nB, normal.data());
    prefilter(dA.data(), dB.data(), w, h, 1, nA, nB, depth.data());
    // 3. beauty variance
    std::vector<float> mean(3 * (size_t)n), hv(3 * (size_t)n);
    combine(beauty_a, beauty_b, 3 * (size_t)n, nA, nB, mean.data());
    half_variance(beauty_a, beauty_b, mean.data(), w, h, 3, hv.data());
    // 4. main filter
    if (kind == 1) {
        std::vector<float> feat(7 * (size_t)n);
        memcpy(feat.data(), albedo.data(), 12 * (size_t)n);
        memcpy(feat.data() + 3 * (size_t)n, normal.data(), 12 * (size_t)n);
        memcpy(feat.data() + 6 * (size_t)n, depth.data(), 4 * (size_t)n);
        const float bw[7] = {1.0f / 2e-3f, 1.0f / 2e-3f, 1.0f / 2e-3f, 1.0f / 1e-2f,
                             1.0f / 1e-2f, 1.0f / 1e-2f, 1.0f / 2e-3f};
        const float k = quality == 1 ? 1.0f : 0.7f;
        const int R = quality == 1 ? 9 : 7;
        std::vector<float> fa(3 * (size_t)n), fb(3 * (size_t)n);
        nlm_filter(w, h, 3, beauty_b, hv.data(), 3, beauty_a, R, 3, k, 7, feat.data(), bw, fa.data());
        nlm_filter(w, h, 3, beauty_a, hv.data(), 3, beauty_b, R, 3, k, 7, feat.data(), bw, fb.data());
        combine(fa.data(), fb.data(), 3 * (size_t)n, nA, nB, out);
        return 0;
    }
    if (kind == 2) {
        nfor_run(w, h, beauty_a, beauty_b, nA, nB, hv.data(), albedo.data(), normal.data(), depth.data(), quality,
                 margin, out);
        return 0;
    }
    return -1;
}

namespace {

constexpr int kTile = 64, kDim = 8, kFeat = 7, kStride = 2, kPatch = 3;

// BuildScaledFeatures (NforRegression.cs:174-199): each feature plane scaled
// to unit global standard deviation (sums in double)
std::vector<float> scaled_features(const float* albedo, const float* normal, const float* depth, int n) {
    std::vector<float> feat((size_t)kFeat * n);
    const float* src[kFeat] = {albedo, albedo + n, albedo + 2 * (size_t)n, normal, normal + n, normal + 2 * (size_t)n,
                               depth};
    for (int pl = 0; pl < kFeat; pl++) {
        double sum = 0, sumSq = 0;
        for (int i = 0; i < n; i++) {
            sum += src[pl][i];
            sumSq += (double)src[pl][i] * src[pl][i];
        }
        const double mean = sum / n;
        const float sd = (float)std::sqrt(std::max(sumSq / n - mean * mean, 0.0));
        const float scale = 1.0f / std::max(sd, 1e-4f);
        for (int i = 0; i < n; i++) feat[(size_t)pl * n + i] = src[pl][i] * scale;
    }
    return feat;
}

// RegressAllCandidates + ProcessTile (NforRegression.cs:224-518)
void regress(const float* g, const float* v, const float* t, const float* feat, int w, int h, const float* ks, int nk,
             int R, std::vector<std::vector<float>>& result) {
    const int n = w * h, side = 2 * R + 1, offsets = side * side, P = kPatch;
    std::vector<int> dxs(offsets), dys(offsets);
    for (int o = 0; o < offsets; o++) {
        dxs[o] = o % side - R;
        dys[o] = o / side - R;
    }
    std::vector<std::vector<float>> num(nk, std::vector<float>(3 * (size_t)n, 0.0f)),
        den(nk, std::vector<float>(n, 0.0f));
    const int tilesX = (w + kTile - 1) / kTile, tilesY = (h + kTile - 1) / kTile;
    const int bs = kTile + 2 * P;
    std::vector<float> dptb((size_t)bs * bs), wbuf(offsets);
    for (int pass = 0; pass < 4; pass++)
        for (int ty = 0; ty < tilesY; ty++) {
            if ((ty & 1) != (pass >> 1)) continue;
            for (int tx = (pass & 1); tx < tilesX; tx += 2) {
                const int tx0 = tx * kTile, ty0 = ty * kTile;
                const int tx1 = std::min(tx0 + kTile, w) - 1, ty1 = std::min(ty0 + kTile, h) - 1;
                std::vector<int> cx, cy;
                const int firstY = ty0 + ((kStride - ty0 % kStride) % kStride);
                const int firstX = tx0 + ((kStride - tx0 % kStride) % kStride);
                for (int gy = firstY; gy <= ty1; gy += kStride)
                    for (int gx = firstX; gx <= tx1; gx += kStride) {
                        cy.push_back(gy);
                        cx.push_back(gx);
                    }
                const int cc = (int)cx.size();
                if (cc == 0) continue;
                const int bx0 = tx0 - P, by0 = ty0 - P;
                std::vector<float> dist((size_t)cc * offsets);
                for (int o = 0; o < offsets; o++) {
                    const int dx = dxs[o], dy = dys[o], qOff = dy * w + dx;
                    const int pxLo = std::max(0, -dx), pxHi = w - 1 - std::max(0, dx);
                    const int pyLo = std::max(0, -dy), pyHi = h - 1 - std::max(0, dy);
                    const int rowLo = std::max(by0, pyLo), rowHi = std::min(by0 + bs - 1, pyHi);
                    const int colLo = std::max(bx0, pxLo), colHi = std::min(bx0 + bs - 1, pxHi);
                    for (int gy = rowLo; gy <= rowHi; gy++) {
                        const int bRow = (gy - by0) * bs - bx0;
                        const int pRow = gy * w;
                        if (colHi >= colLo)
                            distance_row(g, v, n, 3, pRow + colLo, pRow + colLo + qOff, colHi - colLo + 1, dptb.data(),
                                         bRow + colLo);
                    }
                    for (int c = 0; c < cc; c++) {
                        const int gx = cx[c], gy = cy[c];
                        if (gx < pxLo || gx > pxHi || gy < pyLo || gy > pyHi) {
                            dist[(size_t)c * offsets + o] = INFINITY;
                            continue;
                        }
                        const int xLo = std::max(gx - P, pxLo), xHi = std::min(gx + P, pxHi);
                        const int yLo = std::max(gy - P, pyLo), yHi = std::min(gy + P, pyHi);
                        float sum = 0.0f;
                        for (int yy = yLo; yy <= yHi; yy++) {
                            const int bRow = (yy - by0) * bs - bx0;
                            for (int xx = xLo; xx <= xHi; xx++) sum += dptb[bRow + xx];
                        }
                        dist[(size_t)c * offsets + o] = sum / (float)((xHi - xLo + 1) * (yHi - yLo + 1));
                    }
                }
                float f[kDim], fp[kFeat], gram[kDim * kDim], rhs[3 * kDim], beta[3 * kDim];
                double G[kDim * kDim], B[3 * kDim], mf[3];
                for (int c = 0; c < cc; c++) {
                    const int gx = cx[c], gy = cy[c], p = gy * w + gx;
                    for (int i = 0; i < kFeat; i++) fp[i] = feat[(size_t)i * n + p];
                    for (int ki = 0; ki < nk; ki++) {
                        const float invK2 = 1.0f / (ks[ki] * ks[ki]);
                        std::fill(gram, gram + kDim * kDim, 0.0f);
                        std::fill(rhs, rhs + 3 * kDim, 0.0f);
                        for (int o = 0; o < offsets; o++) {
                            const float dd = dist[(size_t)c * offsets + o];
                            if (std::isinf(dd) && dd > 0) {
                                wbuf[o] = 0.0f;
                                continue;
                            }
                            const float d = fmaxz(dd) * invK2;
                            if (d >= kCutoff) {
                                wbuf[o] = 0.0f;
                                continue;
                            }
                            const float wgt = fast_exp_neg(d);
                            wbuf[o] = wgt;
                            const int q = p + dys[o] * w + dxs[o];
                            f[0] = 1.0f;
                            for (int i = 0; i < kFeat; i++) f[i + 1] = feat[(size_t)i * n + q] - fp[i];
                            const float y0 = t[q], y1 = t[n + q], y2 = t[2 * (size_t)n + q];
                            for (int i = 0; i < kDim; i++) {
                                const float wf = wgt * f[i];
                                for (int j = 0; j < kDim; j++) gram[i * kDim + j] += wf * f[j];
                            }
                            const float wy0 = wgt * y0, wy1 = wgt * y1, wy2 = wgt * y2;
                            for (int j = 0; j < kDim; j++) {
                                rhs[j] += wy0 * f[j];
                                rhs[kDim + j] += wy1 * f[j];
                                rhs[2 * kDim + j] += wy2 * f[j];
                            }
                        }
                        const float sumW = gram[0];
                        if (sumW <= 0.0f) continue;
                        for (int i = 0; i < kDim * kDim; i++) G[i] = gram[i];
                        for (int i = 0; i < 3 * kDim; i++) B[i] = rhs[i];
                        for (int ch = 0; ch < 3; ch++) mf[ch] = B[ch * kDim] / sumW;
                        if (regression_solve(G, B, kDim, 3, 1e-3)) {
                            for (int i = 0; i < 3 * kDim; i++) beta[i] = (float)B[i];
                        } else {
                            std::fill(beta, beta + 3 * kDim, 0.0f);
                            for (int ch = 0; ch < 3; ch++) beta[ch * kDim] = (float)mf[ch];
                        }
                        float* numK = num[ki].data();
                        float* denK = den[ki].data();
                        for (int o = 0; o < offsets; o++) {
                            const float wgt = wbuf[o];
                            if (wgt <= 0.0f) continue;
                            const int q = p + dys[o] * w + dxs[o];
                            f[0] = 1.0f;
                            for (int i = 0; i < kFeat; i++) f[i + 1] = feat[(size_t)i * n + q] - fp[i];
                            for (int ch = 0; ch < 3; ch++) {
                                const float* bt = beta + ch * kDim;
                                float e[kDim];
                                for (int i = 0; i < kDim; i++) e[i] = bt[i] * f[i];
                                const float pred = ((e[0] + e[1]) + (e[2] + e[3])) + ((e[4] + e[5]) + (e[6] + e[7]));
                                numK[(size_t)ch * n + q] += wgt * pred;
                            }
                            denK[q] += wgt;
                        }
                    }
                }
            }
        }
    result.assign(nk, std::vector<float>(3 * (size_t)n));
    for (int ki = 0; ki < nk; ki++)
        for (int p = 0; p < n; p++) {
            if (den[ki][p] > 0.0f) {
                const float inv = 1.0f / den[ki][p];
                for (int ch = 0; ch < 3; ch++) result[ki][(size_t)ch * n + p] = num[ki][(size_t)ch * n + p] * inv;
            } else {
                for (int ch = 0; ch < 3; ch++) result[ki][(size_t)ch * n + p] = t[(size_t)ch * n + p];
            }
        }
}

void nfor_run(int w, int h, const float* bA, const float* bB, int nA, int nB, const float* halfVar,
              const float* albedo, const float* normal, const float* depth, int quality, float margin, float* out) {
    const int n = w * h;
    std::vector<float> feat = scaled_features(albedo, normal, depth, n);
    const float ksHigh[2] = {0.5f, 1.0f}, ksFast[1] = {0.7f};
    const float* ks = quality == 1 ? ksHigh : ksFast;
    const int nk = quality == 1 ? 2 : 1;
    const int R = quality == 1 ? 9 : 7;
    std::vector<std::vector<float>> regA, regB;
    regress(bB, halfVar, bA, feat.data(), w, h, ks, nk, R, regA);
    regress(bA, halfVar, bB, feat.data(), w, h, ks, nk, R, regB);
    const int nc = nk + 1;
    std::vector<std::vector<float>> cand(nc, std::vector<float>(3 * (size_t)n));
    combine(bA, bB, 3 * (size_t)n, nA, nB, cand[0].data());
    for (int ki = 0; ki < nk; ki++) combine(regA[ki].data(), regB[ki].data(), 3 * (size_t)n, nA, nB, cand[ki + 1].data());
    std::vector<std::vector<float>> mse(nc, std::vector<float>(n));
    std::vector<float> tmp(n);
    for (int i = 0; i < n; i++) {
        const float f0 = 0.5f * halfVar[i], f1 = 0.5f * halfVar[n + i], f2 = 0.5f * halfVar[2 * (size_t)n + i];
        tmp[i] = (f0 + f1 + f2) * (1.0f / 3.0f);
    }
    box_smooth(tmp.data(), mse[0].data(), w, h, 4);
    for (int ki = 0; ki < nk; ki++) {
        const float* fa = regA[ki].data();
        const float* fb = regB[ki].data();
        for (int i = 0; i < n; i++) {
            float e = 0.0f, vSum = 0.0f;
            for (int c = 0; c < 3; c++) {
                const size_t p = (size_t)c * n + i;
                const float dA = fa[p] - bB[p], dB = fb[p] - bA[p];
                e += 0.5f * (dA * dA + dB * dB) - halfVar[p];
                vSum += 0.5f * halfVar[p];
            }
            tmp[i] = (e + margin * vSum) * (1.0f / 3.0f);
        }
        box_smooth(tmp.data(), mse[ki + 1].data(), w, h, 4);
    }
    std::vector<std::vector<float>> wts(nc, std::vector<float>(n));
    std::vector<float> ind(n);
    for (int ci = 0; ci < nc; ci++) {
        for (int i = 0; i < n; i++) {
            float best = mse[0][i];
            int bi = 0;
            for (int cj = 1; cj < nc; cj++)
                if (mse[cj][i] < best) {
                    best = mse[cj][i];
                    bi = cj;
                }
            ind[i] = bi == ci ? 1.0f : 0.0f;
        }
        box_smooth(ind.data(), wts[ci].data(), w, h, 2);
    }
    for (int i = 0; i < n; i++) {
        float wSum = 0.0f;
        for (int ci = 0; ci < nc; ci++) wSum += wts[ci][i];
        const float inv = wSum > 0.0f ? 1.0f / wSum : 0.0f;
        for (int c = 0; c < 3; c++) {
            const size_t p = (size_t)c * n + i;
            float acc = 0.0f;
            for (int ci = 0; ci < nc; ci++) acc += wts[ci][i] * cand[ci][p];
            const float r = acc * inv;
            out[p] = r < 0.0f ? 0.0f : r;
        }
    }
}

}  // namespace
