// render.hip — wavefront path tracer for gfx950 and the C ABI (include/rt3d.h).
//
// One camera sample = one path slot for the lifetime of a chunk.  A chunk is
// K whole sample passes over the image (path p = s_local * W*H + pixel), so
// the per-pixel film sum is taken in sample order exactly like the
// reference's `cumulativeColor += sample` (Renderer.cs:668-672).  Each bounce
// runs as its own stage:
//   k_raygen   camera rays, Sobol dims 0-1 (+ lens)      Renderer.cs:630-653
//   k_trace    closest hit of the active queue          Renderer.cs:1609-1644
//   k_shade    medium free flight, emission with MIS, NEE light samples
//              (shadow rays queued), BSDF sample, Russian roulette, next
//              ray; survivors are compacted into the next queue with a
//              wave-aggregated atomic                    Renderer.cs:1008-1383, 1700-1782
//   k_shadow   transparent shadow walk of the queued shadow rays
//                                                        ShadowRay.cs:59-158
//   k_finalize per-pixel clamp + film accumulation       Renderer.cs:668-721
//
// The recursion's radiance sum L = Le0 + Ld0 + clampInd(att0 · L1) is
// carried forward with two accumulators per path (primary-vertex terms and
// the post-primary "indirect" sum that the depth-aware clamp acts on,
// Renderer.cs:1222-1223), a throughput for roulette/dead-path decisions
// (the reference's pathThroughput, which excludes Beer-Lambert) and a
// separate radiance weight (which includes it).
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "rt_scene.h"
#include "shade.h"
#include "adjoint.h"

namespace rt {

// ═════════════════════════════════════════════════════════════════════════
// Device-side scene view and render parameters
// ═════════════════════════════════════════════════════════════════════════
struct DevScene {
    const GNode* nodes;
    const GNode4* nodes4;       // any-hit BVH (null: any-hit queries use the binary tree)
    const GEntry* entries;
    const float4* aux;
    const GMat* mats;
    const GLight* lights;
    const float* lut_e;
    const float* light_cdf;
    const float* light_pdf;
    uint32_t top_start, top_count;
    int num_lights;
    int medium;
    int opaque_shadows;
    int env_light;              // first EnvironmentLight of the list, -1 without one
    int thin_film;              // some Disney material has a thin film
    f3 sky;                     // flat sky radiance (FlatSky.EvaluateRadiance)
    int caustics;               // photon map active: gather at diffuse/glossy hits, suppress L S+ re-hits
    int block_spec;             // ShadowRay.BlockSpecularTransmission
    f3 cam_o, cam_llc, cam_h, cam_v, cam_u, cam_vv;
    float lens_radius;
    f3 sig_t, sig_s;
    // texture set of textured Disney materials (kMatTextured)
    const rt_material* raw;     // material records as given (knobs the texture does not replace)
    const float4* tuv;          // 2 per primitive or null (HitRecord U/V)
    const float4* smooth;       // 4 per primitive or null: (N0,Tx) (N1,Ty) (N2,Tz) (B, smooth)
    const float* tex;           // planar [kTexChannels][tex_h][tex_w]
    int tex_w, tex_h;
    // image textures (kMatImage)
    const GImage* images;
    const GLevel* img_levels;
    const float4* img_texels;
    // object hierarchy: wrapper transforms (8 float4 each); null = the scene
    // has no wrapper entries
    const float4* xf;
};

struct RParams {
    int W, H, spp, max_depth, npix;
    float max_rad, ind_max_rad;
    int power_mis, picked;      // picked: one light per NEE event (power / uniform)
    int rr_min_b;
    float rr_min_s;
    uint32_t seed_xor;
    int aovs;
    int maxs;                   // shadow slots per path
    int base_sample;            // first sample pass of the window
    int soff;                   // rt_settings.sample_offset: global index of pass 0
    int spp_fp;                 // spp of the whole render (footprint shrink)
    int prng;                   // PRNG sampler: stratified camera jitter, hashed draws
    int sq;                     // strata per axis, ceil(sqrt(spp)) (PRNG sampler)
    float inv_sq;
    int footprints;             // camera-ray differentials feed image-texture footprints
    int diff_rays;              // camera rays as GetRayWithDifferentials builds them (filtering on)
};

// A window is a run of whole sample passes of one or more jobs; part k
// holds passes [pass0, pass0 + npass) of job `job` at window samples
// [off, off + npass * npix), pass-major.
struct PartDev {
    int job, pass0, off, npix, npass, pad[3];
};

// Batch job on the device: its scene view and render parameters.
struct JobDev {
    DevScene S;
    RParams R;
};

// Scene/parameter source of the kernels: the kernel arguments of a single
// render (uniform, scalar registers) or the path's job record of a batch.
// run(act, j, f) calls f(scene, params) for the lanes with act set; the
// batch form loads the job record with scalar loads, once per distinct job
// among those lanes (a single pass when they share a job, the common case:
// slots are refilled in sample order and jobs are contiguous sample ranges).
// It contains ballots, so f must not use wave-level operations.  run_lane
// is the same with per-lane record loads.
template <class T>
__device__ __forceinline__ T load_uniform(const T* p) {
    // the record is read through the constant address space: with a
    // wave-uniform address these become scalar loads into SGPRs
    typedef const __attribute__((address_space(4))) uint32_t CU;
    CU* src = (CU*)p;
    uint32_t buf[sizeof(T) / 4];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) buf[i] = src[i];
    T r;
    __builtin_memcpy(&r, buf, sizeof(T));
    return r;
}
struct OneJob {
    DevScene S;
    RParams R;
    __device__ __forceinline__ int job_of(const int32_t*, int) const { return 0; }
    __device__ __forceinline__ const DevScene& scene(int) const { return S; }
    __device__ __forceinline__ const RParams& params(int) const { return R; }
    template <class F>
    __device__ __forceinline__ void run(bool act, int, F&& f) const {
        if (act) f(S, R);
    }
    template <class F>
    __device__ __forceinline__ void run_lane(bool act, int, F&& f) const {
        if (act) f(S, R);
    }
};
struct ManyJobs {
    const JobDev* J;
    int thin_film;              // some job's scene has a thin-film material (host-side launch choice)
    int hier;                   // some job's scene has an object hierarchy (host-side launch choice)
    int all_lights;             // every job samples all lights (host-side launch choice)
    __device__ __forceinline__ int job_of(const int32_t* job, int p) const { return job[p]; }
    __device__ __forceinline__ const DevScene& scene(int j) const { return J[j].S; }
    __device__ __forceinline__ const RParams& params(int j) const { return J[j].R; }
    template <class F>
    __device__ __forceinline__ void run(bool act, int j, F&& f) const {
        bool done = !act;
        for (;;) {
            const unsigned long long pend = __ballot(!done);
            if (pend == 0ull) break;
            // the job of the lowest pending lane, as a wave-uniform value; the
            // pointer goes through an empty asm so the compiler cannot rewrite
            // it from the per-lane job inside the branch below
            const int ju = __builtin_amdgcn_readfirstlane(__shfl(j, __ffsll((long long)pend) - 1, 64));
            const JobDev* pj = J + ju;
            asm volatile("" : "+s"(pj));
            if (!done && j == ju) {
                const JobDev jd = load_uniform(pj);
                f(jd.S, jd.R);
                done = true;
            }
        }
    }
    // per-lane records (vector loads): for kernels whose register pressure
    // leaves no room for the scalar copy of the record (the shading kernel)
    template <class F>
    __device__ __forceinline__ void run_lane(bool act, int j, F&& f) const {
        if (act) f(J[j].S, J[j].R);
    }
};

// counters[] (device-resident loop state):
//   per iteration (cleared by k_iter_begin): paths active in the next
//   iteration, shadow-queue size (ungrouped form), retired paths, gather
//   requests, one size per light-sample segment of the grouped shadow queue;
//   per window: next window sample to start, window size, paths active in
//   the current iteration; per render: stack-overflow flag.
constexpr int kCntNext = 0, kCntShadow = 1, kCntRetire = 2, kCntGather = 3, kCntSeg = 4, kMaxSeg = 32,
              kCntClear = 40, kCntSid = 40, kCntWs = 41, kCntActive = 42, kCntSidSnap = 43, kCntWindow = 48,
              kCntPend = 52, kCntOverflow = 63;
// tcount[] (64-bit): traversal tests (trace node/tri/other, shadow node/tri/other), closest-hit segments, shadow rays
constexpr int kTcRays = 6, kTcShadowRays = 7;
// any-hit (4-wide BVH) traversal work of the shadow kernel: node fetches, triangle/quad/plane tests, sphere tests
constexpr int kTcAnyNode = 8, kTcAnyTri = 9, kTcAnyOther = 10;
// per-wave count arrays (wc): continuing paths, retired paths, gather
// requests, shadow rays (ungrouped form), then one per light-sample segment
constexpr int kWcNext = 0, kWcRetire = 1, kWcGather = 2, kWcShadow = 3, kWcSeg = 4;
constexpr int kCntSidBase = 44;   // first window sample of this iteration's restarts
// kCntPend..+2 (written by k_batch_end for the host): paths pending in the next
// iteration, the stack-overflow flag, and whether the window's samples are all issued
constexpr int kIterBatch = 4;   // bounce iterations enqueued between host checks
// kSlotBounce: shaded by the NEE phase of a split shade, waiting for its
// BSDF bounce (k_shade<..., kPhBounce>) in the same iteration
enum : int32_t { kSlotIdle = 0, kSlotActive = 1, kSlotRetired = 2, kSlotBounce = 3 };
// Shading phases: the whole vertex in one kernel (kept for the adjoint's
// replay pass), or split in two launches — medium, emission, NEE and the
// caustic gather request first, then the BSDF sample, roulette and next ray
// of surface vertices — so that neither kernel holds both the light-sample
// and the BSDF-sample state.
enum : int { kPhAll = 0, kPhNee = 1, kPhBounce = 2 };

// Shadow queue view: segment j holds the shadow rays of light sample j
// (in light-list order) at q_sh[j*stride ...]; pre[] is the prefix of the
// segment sizes, so launch index i maps to segment j with pre[j] <= i < pre[j+1].
struct ShQ {
    int nseg, stride;
    int pre[kMaxSeg + 1];
};

struct PathBufs {
    float4 *o, *d, *thr, *wt, *w0, *a0, *ai, *ab;
    float* ior;                 // 8 × N (slot k of path p at k*N + p)
    int32_t* hit_e;
    float4* hit_tuv;
    int4* hit_xf;               // wrapper chain of the hit (scenes with an object hierarchy)
    float4 *av0, *av1, *av2, *av3, *av4;   // AOV probe
    float4 *sh_o, *sh_d, *sh_c, *sh_r;     // shadow records, record k of slot p at k*N + p (N × maxs)
    int32_t* state;             // per slot: kSlotIdle, kSlotActive (to trace), kSlotRetired
    int32_t* gflag;             // per slot: a caustic gather request was queued this iteration
    int32_t* q_g;               // gather queue: slots with a request
    int32_t* wc;                // per-wave counts, kWc* arrays of ceil(N/64) entries, scanned in place
    int nw;                     // row stride of wc: ceil(N/64) rounded up to a multiple of 4
    int32_t* smask;             // per slot: light samples with a pending shadow ray (grouped form)
    int32_t* q_sh;              // shadow queue: shadow record indices
    int32_t* sid;               // sample of the window each path slot is tracing
    int32_t *job, *pix, *pass;  // its job (batch), pixel index and sample pass
    const struct PartDev* parts;   // window parts (job, first pass, window offset)
    int nparts;
    float4 *res_l, *res_a, *res_n;         // per-sample results of the window: (L, depth), albedo, normal
    float4 *gq0, *gq1, *gq2, *gq3;         // caustic gather requests: (o,t) (d,entry) (u,v,eta,path) (weight,target)
    int32_t* counters;          // kCnt* slots below
    unsigned long long* tcount; // traversal counters: trace node/tri/other, shadow node/tri/other
    int n;
    // path-replay adjoint: adj_pass 1 = primal (res_l[sample] = the sample's
    // radiance sum a0 + ai), 2 = replay (texture gradients); 0 = plain render
    int adj_pass;
    float adj_inv_spp;
    float4* fp;                 // per slot: image-texture footprint of the camera hit (dudx, dvdx, dudy, dvdy)
    float4* adj_q;              // 5 per slot: pending bounce gradient (15 coefficients), (u, v, flag, -)
    float4* adj_g;              // 2 per shadow record: NEE gradient (5), (u, v, flag)
    const float* adj_db;        // dLoss/dBeauty, planar [3][npix]
    float* adj_dt;              // texture gradient, planar [kTexChannels][tex_h][tex_w]
};

__device__ __forceinline__ void flush_counts(unsigned long long* dst, const TravCount& tc) {
    unsigned long long a = tc.node, b = tc.tri, c = tc.other;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        c += __shfl_down(c, off, 64);
    }
    if (__lane_id() == 0) {
        atomicAdd(dst, a);
        atomicAdd(dst + 1, b);
        atomicAdd(dst + 2, c);
    }
}

struct FilmBufs {
    float *sum, *sa, *sb, *alba, *albb, *nrma, *nrmb, *depa, *depb;
    int32_t *dca, *dcb;
    float *out_beauty, *out_a, *out_b, *out_disp, *out_alba, *out_albb, *out_nrma, *out_nrmb, *out_depa, *out_depb;
};

// path flag bits (stored as the int bits of d.w)
enum : int32_t {
    kPfPrevDelta = 1,
    kPfCatShift = 1,           // 2 bits: 0 camera, 1 diffuse, 2 glossy, 3 transmission
    kPfTgtInd = 8,             // contributions go to the indirect accumulator
    kPfPendShift = 8,          // 8 bits: pending shadow slots
    kPfPendInd = 1 << 16,      // pending shadows belong to the indirect accumulator
    kPfCstShift = 17,          // 2 bits: caustic-chain state (none, diffuse, suppress)
};
enum : int { kCstNone = 0, kCstDiffuse = 1, kCstSuppress = 2 };
// NextCausticState (Renderer.cs:1434-1438)
__device__ __forceinline__ int next_cst(int st, bool delta) {
    if (!delta) return kCstDiffuse;
    return st == kCstNone ? kCstNone : kCstSuppress;
}
// AOV probe flag bits (av0.w)
enum : int32_t { kAovPending = 1, kAovHaveDepth = 2 };

constexpr int kBlock = 64;

__device__ __forceinline__ int pf_cat(int32_t f) { return (f >> kPfCatShift) & 3; }
__device__ __forceinline__ int pf_pend(int32_t f) { return (f >> kPfPendShift) & 0xff; }

// XCD-aware block order: gfx950 dispatches consecutive workgroups round-robin
// over its 8 XCDs; renumber so each XCD walks one contiguous range of the
// queue (neighbouring rays share BVH nodes in that XCD's L2).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    uint32_t q = nb / 8u, r = nb % 8u, x = b % 8u;
    return x * q + min(x, r) + b / 8u;
}
__device__ __forceinline__ uint32_t xcd_block() { return xcd_remap(blockIdx.x, gridDim.x); }


// ── surface reconstruction from an entry (primitive Hit bodies) ─────────
__device__ __forceinline__ Surf make_surf(const DevScene& S, int32_t eidx, f3 o, f3 d, float t, float u, float v) {
    GEntry e = S.entries[eidx];
    int32_t tag = tag_of(e.a);
    int kind = tag & 0xff;
    Surf s;
    s.t = t;
    s.rel_eta = 0.0f;
    s.p = o + t * d;
    s.mat = __float_as_int(e.c.w);
    s.prim = __float_as_int(e.b.w);
    if (kind == kTri) {
        f3 e1 = f3of(e.b), e2 = f3of(e.c);
        s.n = face(d, f3of(e.d), s.front);
        s.tan = length_sq(e1) > 0.0f ? normalize(e1) : mk(1.0f, 0.0f, 0.0f);
        s.bitan = length_sq(e2) > 0.0f ? normalize(e2) : mk(0.0f, 0.0f, 1.0f);
        if (S.smooth) {
            // SmoothTriangle.Hit (SmoothTriangle.cs:166-193): interpolated
            // shading normal, oriented by the face normal's front test
            const float4* q = S.smooth + 4 * (size_t)s.prim;
            const float4 q3 = q[3];
            if (q3.w != 0.0f) {
                const float4 q0 = q[0], q1 = q[1], q2 = q[2];
                const float w0 = 1.0f - u - v;
                f3 sn = normalize((w0 * f3of(q0) + u * f3of(q1)) + v * f3of(q2));
                s.n = s.front ? sn : -sn;
                s.tan = mk(q0.w, q1.w, q2.w);
                s.bitan = f3of(q3);
            }
        }
    } else if (kind == kSphere) {
        f3 c = f3of(e.a);
        float r = e.b.x;
        f3 outward = (s.p - c) / r;
        s.n = face(d, outward, s.front);
        f3 td = cross(mk(0.0f, 1.0f, 0.0f), outward);
        if (length_sq(td) < 1e-4f) td = mk(1.0f, 0.0f, 0.0f);
        s.tan = normalize(td);
        s.bitan = normalize(cross(outward, s.tan));
    } else if (kind == kQuad) {
        s.n = face(d, f3of(e.b), s.front);
        int ai = __float_as_int(e.d.y);
        s.tan = f3of(S.aux[4 * ai + 2]);
        s.bitan = f3of(S.aux[4 * ai + 3]);
    } else {
        f3 n = f3of(e.b);
        s.n = face(d, n, s.front);
        f3 ua = fabsf(n.y) < 0.99f ? normalize(cross(n, mk(0.0f, 1.0f, 0.0f))) : mk(1.0f, 0.0f, 0.0f);
        s.tan = ua;
        s.bitan = cross(n, ua);
    }
    return s;
}

// Surface of a hit found under wrapper entries (rt_node): the primitive's
// Hit in the frame of the innermost transform, then MapHitToWorld through
// each transform outward (point by the matrix, normal by the transposed
// inverse, tangent frame by the matrix, each normalised; Transform.cs:162-207)
// and the outermost material override last (Instance.cs:71-82).  lo / ld
// receive the ray of the hit's frame (texture coordinates are read there).
__device__ __forceinline__ Surf make_surf_w(const DevScene& S, int32_t eidx, int4 chain, f3 o, f3 d, float t,
                                            float u, float v, f3& lo, f3& ld) {
    lo = o;
    ld = d;
    chain_ray(S.entries, S.xf, chain, lo, ld);
    Surf s = make_surf(S, eidx, lo, ld, t, u, v);
    const int32_t c[4] = {chain.x, chain.y, chain.z, chain.w};
#pragma unroll
    for (int k = kMaxWrap - 1; k >= 0; k--) {
        if (c[k] < 0) continue;
        const float4 ec = S.entries[c[k]].c;
        const int32_t xi = __float_as_int(ec.x), om = __float_as_int(ec.y);
        if (xi >= 0) {
            const float4* M = S.xf + 8 * xi;
            s.p = xf_point(s.p, M);
            s.n = normalize(xf_normal(s.n, M + 4));
            s.tan = normalize(xf_vec(s.tan, M));
            s.bitan = normalize(xf_vec(s.bitan, M));
        }
        if (om >= 0) s.mat = om;
    }
    return s;
}

// A hit's camera visibility: its primitive's flag or a camera-invisible
// wrapper above it.
__device__ __forceinline__ bool hit_cam_invisible(const DevScene& S, int32_t e, int4 chain) {
    if ((tag_of(S.entries[e].a) >> 8) & kEntryCamInvisible) return true;
    if (!S.xf) return false;
    const int32_t c[4] = {chain.x, chain.y, chain.z, chain.w};
#pragma unroll
    for (int k = 0; k < kMaxWrap; k++)
        if (c[k] >= 0 && ((tag_of(S.entries[c[k]].a) >> 8) & kEntryCamInvisible)) return true;
    return false;
}

// Texture coordinate of a hit: the primitive's interpolated vertex UVs when
// the scene has them (SmoothTriangle.cs:168,186), else HitRecord.U/V.
__device__ __forceinline__ float2 tex_coord(const DevScene& S, int32_t he, f3 o, f3 d, float4 htuv) {
    const float4 ea = S.entries[he].a;
    if (S.tuv && (tag_of(ea) & 0xff) == kTri) {
        const int prim = __float_as_int(S.entries[he].b.w);
        const float4 a = S.tuv[2 * prim], b = S.tuv[2 * prim + 1];
        const float w1 = htuv.y, w2 = htuv.z, w0 = 1.0f - w1 - w2;
        return make_float2((w0 * a.x + w1 * a.z) + w2 * b.x, (w0 * a.y + w1 * a.w) + w2 * b.y);
    }
    HitOut h{he, htuv.x, htuv.y, htuv.z};
    float u, v;
    report_uv(S.entries, h, o, d, u, v);
    return make_float2(u, v);
}

// ── ImageTexture (ImageTexture.cs:148-289) ─────────────────────────────
__device__ __forceinline__ int img_wrap(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// SampleBilinear: texel centres at (i + 0.5) / w, repeat addressing.
__device__ __forceinline__ f3 img_bilinear(const DevScene& S, const GLevel L, float u, float v) {
    const float px = u * (float)L.w - 0.5f, py = v * (float)L.h - 0.5f;
    const int x0 = (int)floorf(px), y0 = (int)floorf(py);
    const float fx = px - (float)x0, fy = py - (float)y0;
    const int x0w = img_wrap(x0, L.w), y0w = img_wrap(y0, L.h);
    const int x1w = img_wrap(x0 + 1, L.w), y1w = img_wrap(y0 + 1, L.h);
    const float4* T = S.img_texels + L.off;
    const f3 c00 = f3of(T[y0w * L.w + x0w]), c10 = f3of(T[y0w * L.w + x1w]);
    const f3 c01 = f3of(T[y1w * L.w + x0w]), c11 = f3of(T[y1w * L.w + x1w]);
    return lerp3(lerp3(c00, c10, fx), lerp3(c01, c11, fx), fy);
}

// SampleTrilinear
__device__ __forceinline__ f3 img_trilinear(const DevScene& S, const GImage& im, float u, float v, float lod) {
    const int lo = (int)floorf(lod);
    const int hi = min(lo + 1, im.mips - 1);
    const float f = lod - (float)lo;
    const f3 a = img_bilinear(S, S.img_levels[im.level0 + lo], u, v);
    if (f <= 0.0f) return a;
    return lerp3(a, img_bilinear(S, S.img_levels[im.level0 + hi], u, v), f);
}

// ImageTexture.Value: point-sampled level 0 without a footprint; with one
// (fp = dudx, dvdx, dudy, dvdy; all zero = none) ratio-trilinear filtering:
// LOD from the minor axis (anisotropy capped at 8), up to 8 taps along the
// major axis.
__device__ __forceinline__ f3 img_value(const DevScene& S, int img, float u, float v, float4 fp) {
    const GImage im = S.images[img];
    float dudx = fp.x * im.scale_u, dvdx = fp.y * im.scale_v;
    float dudy = fp.z * im.scale_u, dvdy = fp.w * im.scale_v;
    float lx = sqrtf(dudx * dudx + dvdx * dvdx), ly = sqrtf(dudy * dudy + dvdy * dvdy);
    float minor = lx <= ly ? lx : ly, major = lx >= ly ? lx : ly;
    if (im.mips <= 1 || !(major > 0.0f)) {
        float uu = u * im.scale_u, vv = v * im.scale_v;
        uu = uu - floorf(uu);
        vv = vv - floorf(vv);
        return img_bilinear(S, S.img_levels[im.level0], uu, 1.0f - vv);
    }
    if (minor * 8.0f < major) minor = major / 8.0f;
    const GLevel L0 = S.img_levels[im.level0];
    const float maxdim = fmaxf((float)L0.w, (float)L0.h);
    float lod = log2f(fmaxf(minor * maxdim, 1.0f));
    lod = fminf(fmaxf(lod, 0.0f), (float)(im.mips - 1));
    int n = (int)ceilf(major / minor);
    n = n < 1 ? 1 : (n > 8 ? 8 : n);
    const float mu = lx >= ly ? dudx : dudy, mv = lx >= ly ? dvdx : dvdy;
    const float uc = u * im.scale_u, vc = v * im.scale_v;
    const float inv = 1.0f / (float)n;
    f3 acc = splat(0.0f);
    for (int i = 0; i < n; i++) {
        const float t = ((float)i + 0.5f) * inv - 0.5f;
        float uu = uc + t * mu, vv = vc + t * mv;
        uu = uu - floorf(uu);
        vv = vv - floorf(vv);
        acc = acc + img_trilinear(S, im, uu, 1.0f - vv, lod);
    }
    return acc * inv;
}


// Material record of a textured hit, one disney_gmat site for both
// sources: the texture set (Disney base colour, roughness and metallic) or
// an ImageTexture (Lambertian albedo or Disney baseColor,
// SceneLoader.cs:1027-1031; fp = the hit's footprint).
__device__ __forceinline__ GMat textured_gmat(const DevScene& S, const GMat& g, int mat, float2 uv,
                                              float tp[kTexChannels], bool image, float4 fp) {
    const rt_material& m = S.raw[mat];
    f3 c;
    float metallic = m.metallic, roughness = m.roughness;
    if (image) {
        c = img_value(S, m.image, uv.x, uv.y, fp);
        if (g.m0.x == RT_MAT_LAMBERTIAN) {
            GMat t = g;
            t.m1 = make_float4(c.x, c.y, c.z, g.m1.w);
            return t;
        }
    } else {
        const TexFoot foot = tex_footprint(S.tex_w, S.tex_h, uv.x, uv.y);
#pragma unroll
        for (int k = 0; k < kTexChannels; k++) tp[k] = tex_value(S.tex, S.tex_w, S.tex_h, foot, k);
        c = mk(tp[0], tp[1], tp[2]);
        metallic = tp[4];
        roughness = tp[3];
    }
    GMat t = disney_gmat(m, c, metallic, roughness, S.lut_e + 32 * 32);
    t.m0.z = g.m0.z;
    return t;
}

// ── ray-differential footprint (FootprintMath.cs:38-119) ────────────────
__device__ __forceinline__ bool tangent_plane_hit(f3 o, f3 d, f3 P, f3 N, f3& hit) {
    const float den = dot(d, N);
    if (fabsf(den) < 1e-12f) return false;
    const float t = dot(P - o, N) / den;
    hit = o + t * d;
    return isfinite(t);
}

// UV partials (dudx, dvdx, dudy, dvdy) of the camera differentials
// (ox, dx), (oy, dy) at hit point P with normal N; zero when a ray is
// parallel to the tangent plane or the dpdu/dpdv basis is singular.
__device__ __forceinline__ float4 uv_footprint(f3 ox, f3 dx, f3 dy, f3 P, f3 N, f3 dpdu, f3 dpdv) {
    f3 px, py;
    if (!tangent_plane_hit(ox, dx, P, N, px) || !tangent_plane_hit(ox, dy, P, N, py)) return make_float4(0, 0, 0, 0);
    const f3 ddx = px - P, ddy = py - P;
    const float ax = fabsf(N.x), ay = fabsf(N.y), az = fabsf(N.z);
    int d0, d1;
    if (ax > ay && ax > az) { d0 = 1; d1 = 2; }
    else if (ay > az) { d0 = 0; d1 = 2; }
    else { d0 = 0; d1 = 1; }
    const float a00 = get(dpdu, d0), a01 = get(dpdv, d0), a10 = get(dpdu, d1), a11 = get(dpdv, d1);
    const float det = a00 * a11 - a01 * a10;
    if (fabsf(det) < 1e-20f) return make_float4(0, 0, 0, 0);
    const float inv = 1.0f / det;
    const float bx0 = get(ddx, d0), bx1 = get(ddx, d1), by0 = get(ddy, d0), by1 = get(ddy, d1);
    return make_float4((a11 * bx0 - a01 * bx1) * inv, (-a10 * bx0 + a00 * bx1) * inv,
                       (a11 * by0 - a01 * by1) * inv, (-a10 * by0 + a00 * by1) * inv);
}

// ── lights (SphereLight/AreaLight/DirectionalLight/GeometryLight) ───────
struct LSample {
    bool skip;      // in shadow / zero contribution / early return
    bool delta;
    f3 color;       // unoccluded light colour of the estimator
    f3 dir, so;     // direction to light, shadow-ray origin
    float tmax;     // shadow-ray tMax
    float dist;     // distance reported to the medium transmittance
};

__device__ __forceinline__ void grid_xi(int si, int sq, Rng& rng, float& xi1, float& xi2) {
    float inv = 1.0f / (float)sq;
    int su = si % sq, sv = si / sq;
    xi1 = ((float)su + rng.next()) * inv;
    xi2 = ((float)sv + rng.next()) * inv;
}

// Sphere.SolidAnglePdf (Sphere.cs:206-229) / SphereLight.PdfSolidAngle
__device__ __forceinline__ float sphere_cone_pdf(f3 c, float r, f3 hp, f3 wi) {
    f3 to_c = c - hp;
    float d2 = length_sq(to_c);
    float r2 = r * r;
    if (d2 <= r2) return 1.0f / (4.0f * kPi);
    float ctm = sqrtf(fmaxf(0.0f, 1.0f - r2 / d2));
    float wl = length(wi);
    if (wl < kEps) return 0.0f;
    f3 w = to_c / sqrtf(d2);
    float ct = dot(wi, w) / wl;
    if (ct < ctm) return 0.0f;
    float sa = 2.0f * kPi * (1.0f - ctm);
    return sa > kEps ? 1.0f / sa : 0.0f;
}

// FlatSky.Pdf: uniform over the sphere (FlatSky.cs:31-35)
constexpr float kUniformSpherePdf = 1.0f / (4.0f * kPi);

// ILight.PdfSolidAngle for every supported light
__device__ float light_pdf(const GLight& L, f3 hp, f3 wi) {
    int kind = L.l0.x;
    if (kind == RT_LIGHT_SPHERE) return sphere_cone_pdf(f3of(L.l2), L.l2.w, hp, wi);
    if (kind == RT_LIGHT_ENVIRONMENT) return L.l0.w ? kUniformSpherePdf : 0.0f;
    if (kind == RT_LIGHT_AREA) {
        f3 n = f3of(L.l5), corner = f3of(L.l2), u = f3of(L.l3), v = f3of(L.l4);
        float denom = dot(n, wi);
        if (fabsf(denom) < 1e-7f) return 0.0f;
        float t = dot(corner - hp, n) / denom;
        if (t <= kEps) return 0.0f;
        f3 lc = (hp + t * wi) - corner;
        float uu = dot(u, u), vv = dot(v, v), uv = dot(u, v);
        float det = uu * vv - uv * uv;
        if (det < 1e-12f) return 0.0f;
        float lu = dot(lc, u), lv = dot(lc, v);
        float a = (lu * vv - lv * uv) / det, b = (lv * uu - lu * uv) / det;
        if (a < 0.0f || a > 1.0f || b < 0.0f || b > 1.0f) return 0.0f;
        float cl = fabsf(denom);
        if (cl < 1e-6f) return 0.0f;
        return (t * t) / (L.l2.w * cl);
    }
    if (kind == RT_LIGHT_DIRECTIONAL) {
        if (L.l2.w <= 0.0f) return 0.0f;
        float solid = L.l3.y;
        if (solid <= 0.0f) return 0.0f;
        float wl = length(wi);
        if (wl < kEps) return 0.0f;
        float ct = dot(wi / wl, -f3of(L.l2));
        if (ct < L.l3.x) return 0.0f;
        return 1.0f / solid;
    }
    // geometry light
    int sub = L.l0.w;
    if (sub == kGeoSphere) return sphere_cone_pdf(f3of(L.l2), L.l2.w, hp, wi);
    float t, u, v;
    bool hit;
    f3 n = f3of(L.l5);
    if (sub == kGeoTri) hit = hit_triangle(hp, wi, f3of(L.l2), f3of(L.l3), f3of(L.l4), kEps, kInfinity, t, u, v);
    else {
        float denom = dot(n, wi);
        hit = false;
        if (fabsf(denom) >= 1e-8f) {
            f3 q = f3of(L.l2), U = f3of(L.l3), V = f3of(L.l4);
            float D = dot(n, q);
            t = (D - dot(n, hp)) / denom;
            if (!(t < kEps || t > kInfinity)) {
                f3 nn = cross(U, V);
                f3 w = nn / dot(nn, nn);
                f3 pv = (hp + t * wi) - q;
                float a = dot(w, cross(pv, V)), b = dot(w, cross(U, pv));
                hit = !(a < 0.0f || a > 1.0f || b < 0.0f || b > 1.0f);
            }
        }
    }
    if (!hit) return 0.0f;
    float cl = fabsf(dot(-wi, n));
    if (cl < 1e-6f) return 0.0f;
    float area = L.l5.w;
    if (area < 1e-12f) return 0.0f;
    return (t * t) / (area * cl);
}

// IlluminateAndTestStratified without the shadow test (the shadow ray is
// queued; its transmittance multiplies the colour later).
__device__ LSample light_sample(const GLight& L, f3 hp, f3 sn, int si, Rng& rng) {
    LSample o;
    o.skip = true;
    o.delta = false;
    o.color = splat(0.0f);
    o.dir = mk(0.0f, 1.0f, 0.0f);
    o.so = hp;
    o.tmax = 0.0f;
    o.dist = 0.0f;
    int kind = L.l0.x, ss = L.l0.y, sq = L.l0.z;
    if (kind == RT_LIGHT_ENVIRONMENT) {
        // EnvironmentLight.IlluminateAndTest (EnvironmentLight.cs:66-104) over a
        // flat sky: a uniform direction (FlatSky.ImportanceSample), below the
        // horizon of a surface point discarded, an unbounded shadow ray, the
        // colour L / (pdf N); a black sky cannot sample and draws nothing
        if (L.l0.w == 0) return o;
        const f3 dir = random_unit_vector(rng);
        if (length_sq(sn) > 0.0f && dot(sn, dir) <= 0.0f) return o;
        o.so = offset_origin(hp, sn);
        o.tmax = kInfinity;
        o.color = f3of(L.l1) / (kUniformSpherePdf * (float)ss);
        o.dir = dir;
        o.dist = kInfinity;
        o.skip = false;
        return o;
    }
    if (kind == RT_LIGHT_SPHERE) {
        f3 c = f3of(L.l2);
        float r = L.l2.w;
        f3 to_c = c - hp;
        float d2 = length_sq(to_c);
        if (d2 < kEps * kEps) return o;
        float dist = sqrtf(d2);
        f3 w = to_c / dist;
        float s2 = r * r / d2;
        float ctm = s2 >= 1.0f ? -1.0f : sqrtf(1.0f - s2);
        float xi1, xi2;
        grid_xi(si, sq, rng, xi1, xi2);
        float ct = 1.0f - xi1 * (1.0f - ctm);
        float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        float phi = 2.0f * kPi * xi2;
        f3 dw = local_to_world(mk(st * cosf(phi), st * sinf(phi), ct), w);
        f3 oc = hp - c;
        float b = dot(oc, dw);
        float cc = length_sq(oc) - r * r;
        float disc = b * b - cc;
        if (disc < 0.0f) return o;
        float sd = sqrtf(disc);
        float t = -b - sd;
        if (t < kEps) t = -b + sd;
        if (t < kEps) return o;
        f3 sp = hp + t * dw;
        f3 ln = (sp - c) / r;
        float cl = fmaxf(0.0f, dot(-dw, ln));
        if (cl <= 0.0f) return o;
        o.so = offset_origin(hp, sn);
        o.tmax = length(sp - o.so) - kEps;
        float sa = 2.0f * kPi * (1.0f - ctm);
        o.color = f3of(L.l1) * (L.l1.w * sa / (float)ss);
        o.dir = dw;
        o.dist = t;
        o.skip = false;
        return o;
    }
    if (kind == RT_LIGHT_AREA) {
        float ru, rv;
        grid_xi(si, sq, rng, ru, rv);
        f3 sp = (f3of(L.l2) + ru * f3of(L.l3)) + rv * f3of(L.l4);
        f3 tl = sp - hp;
        float d2 = length_sq(tl);
        if (d2 < kEps * kEps) return o;
        float dist = sqrtf(d2);
        f3 dl = tl / dist;
        float cl = fmaxf(0.0f, dot(-dl, f3of(L.l5)));
        if (cl <= 0.0f) return o;
        o.so = offset_origin(hp, sn);
        o.tmax = length(sp - o.so) - kEps;
        float ad2 = d2;
        float soft = L.l3.w;
        if (soft > 0.0f) ad2 = fmaxf(ad2, soft * soft);
        o.color = f3of(L.l1) * (L.l1.w * L.l2.w * cl / (ad2 * (float)ss));
        o.dir = dl;
        o.dist = dist;
        o.skip = false;
        return o;
    }
    if (kind == RT_LIGHT_DIRECTIONAL) {
        f3 dir = f3of(L.l2);
        f3 tl;
        if (L.l2.w > 0.0f) {
            float xi1, xi2;
            grid_xi(si, sq, rng, xi1, xi2);
            float ct = 1.0f - xi1 * (1.0f - L.l3.x);
            float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
            float phi = 2.0f * kPi * xi2;
            tl = local_to_world(mk(st * cosf(phi), st * sinf(phi), ct), -dir);
        } else {
            tl = -dir;
            o.delta = true;
        }
        o.so = offset_origin(hp, sn);
        o.tmax = kInfinity;
        o.color = f3of(L.l1) * (L.l1.w / (float)ss);
        o.dir = tl;
        o.dist = kInfinity;
        o.skip = false;
        return o;
    }
    // GeometryLight (GeometryLight.cs:133-279)
    int sub = L.l0.w;
    f3 em = f3of(L.l1);
    if (sub == kGeoSphere) {
        float xi1, xi2;
        grid_xi(si, sq, rng, xi1, xi2);
        f3 c = f3of(L.l2);
        float r = L.l2.w;
        f3 to_c = c - hp;
        float d2c = length_sq(to_c);
        float r2 = r * r;
        f3 sp, ln;
        float pdf;
        if (d2c <= r2) {
            f3 pu = random_unit_vector(rng);
            sp = c + pu * r;
            ln = pu;
            pdf = 1.0f / (4.0f * kPi);
        } else {
            float dd = sqrtf(d2c);
            f3 w = to_c / dd;
            float ctm = sqrtf(fmaxf(0.0f, 1.0f - r2 / d2c));
            float ct = 1.0f - xi1 * (1.0f - ctm);
            float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
            float phi = 2.0f * kPi * xi2;
            f3 dw = local_to_world(mk(st * cosf(phi), st * sinf(phi), ct), w);
            f3 oc = hp - c;
            float b = dot(oc, dw);
            float cc = length_sq(oc) - r2;
            float disc = b * b - cc;
            float t;
            if (disc <= 0.0f) t = -b;
            else {
                float sd = sqrtf(disc);
                t = -b - sd;
                if (t < kEps) t = -b + sd;
            }
            sp = hp + t * dw;
            ln = normalize(sp - c);
            float sa = 2.0f * kPi * (1.0f - ctm);
            pdf = sa > kEps ? 1.0f / sa : 0.0f;
        }
        if (pdf <= 0.0f) return o;
        f3 tl = sp - hp;
        float d2 = length_sq(tl);
        if (d2 < kEps * kEps) return o;
        float dist = sqrtf(d2);
        f3 dl = tl / dist;
        float cl = fmaxf(0.0f, dot(-dl, ln));
        if (cl <= 0.0f) return o;
        o.so = offset_origin(hp, sn);
        o.tmax = length(sp - o.so) - kEps;
        o.color = em * (1.0f / (pdf * (float)ss));
        o.dir = dl;
        o.dist = dist;
        o.skip = false;
        return o;
    }
    f3 sp, ln = f3of(L.l5);
    float area = L.l5.w;
    {
        float inv = 1.0f / (float)sq;
        int su = si % sq, sv = si / sq;
        float a = ((float)su + rng.next()) * inv;
        float b = ((float)sv + rng.next()) * inv;
        if (sub == kGeoTri) {
            float r1 = sqrtf(a);
            float uu = 1.0f - r1, vv = b * r1;
            sp = f3of(L.l2) + uu * f3of(L.l3) + vv * f3of(L.l4);
        } else {
            sp = (f3of(L.l2) + a * f3of(L.l3)) + b * f3of(L.l4);
        }
    }
    f3 tl = sp - hp;
    float d2 = length_sq(tl);
    if (d2 < kEps * kEps) return o;
    float dist = sqrtf(d2);
    f3 dl = tl / dist;
    float cl = fmaxf(0.0f, dot(-dl, ln));
    if (cl <= 0.0f) return o;
    o.so = offset_origin(hp, sn);
    o.tmax = length(sp - o.so) - kEps;
    float ad2 = d2;
    float soft = L.l3.w;
    if (soft > 0.0f) ad2 = fmaxf(ad2, soft * soft);
    o.color = em * (area * cl / (ad2 * (float)ss));
    o.dir = dl;
    o.dist = dist;
    o.skip = false;
    return o;
}

__device__ __forceinline__ float mis_w(float p, float q, int power) {
    if (power) {
        float p2 = p * p, q2 = q * q;
        return p2 / (p2 + q2 + 1e-30f);
    }
    return p / (p + q + 1e-30f);
}
__device__ __forceinline__ f3 sanitize(f3 c) {
    if (isnan(c.x) || isinf(c.x)) c.x = 0.0f;
    if (isnan(c.y) || isinf(c.y)) c.y = 0.0f;
    if (isnan(c.z) || isinf(c.z)) c.z = 0.0f;
    return mk(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f));
}
__device__ __forceinline__ f3 clamp_lum(f3 c, float m, bool require_pos) {
    c = sanitize(c);
    float l = luminance(c);
    if (l > m && (!require_pos || l > 0.0f)) c *= m / l;
    return c;
}
__device__ __forceinline__ f3 clamp01(f3 v) {
    return mk(clampf(v.x, 0.0f, 1.0f), clampf(v.y, 0.0f, 1.0f), clampf(v.z, 0.0f, 1.0f));
}
__device__ __forceinline__ f3 medium_tr(const DevScene& S, float tmax) {
    float d = fminf(fmaxf(tmax, 0.0f), 1e30f);
    return mk(expf(-S.sig_t.x * d), expf(-S.sig_t.y * d), expf(-S.sig_t.z * d));
}

// ═════════════════════════════════════════════════════════════════════════
// Kernels
// ═════════════════════════════════════════════════════════════════════════
// Sample `sid` of the window: pass-major (sid = pass_local * npix + pixel).
__device__ __forceinline__ void sample_of(const RParams& R, int sid, int& i, int& j, int& s) {
    int pix = sid % R.npix;
    s = R.base_sample + sid / R.npix;
    i = pix % R.W;
    j = pix / R.W;
}

// The sample's jitter inside its pixel: two Sobol draws, or with the PRNG
// sampler two draws inside stratum (sx, sy) = (s mod sq, s div sq) of the
// sq x sq grid (Renderer.cs:677-682).
__device__ __forceinline__ void pixel_jitter(const RParams& R, Rng& rng, float& ju, float& jv) {
    if (R.prng) {
        const int s = (int)rng.index;
        ju = ((float)(s % R.sq) + rng.next()) * R.inv_sq;
        jv = ((float)(s / R.sq) + rng.next()) * R.inv_sq;
    } else {
        ju = rng.next();
        jv = rng.next();
    }
}

// Camera ray (Renderer.cs:630-653, Camera.cs:127-153)
__device__ __forceinline__ void camera_ray(const DevScene& S, const RParams& R, int i, int j, Rng& rng, f3& o, f3& d) {
    float ju, jv;
    pixel_jitter(R, rng, ju, jv);
    float u = ((float)i + ju) / (float)R.W;
    float v = ((float)(R.H - j - 1) + jv) / (float)R.H;
    f3 target = (S.cam_llc + u * S.cam_h) + v * S.cam_v;
    if (S.lens_radius == 0.0f) {
        o = S.cam_o;
        d = target - S.cam_o;
        return;
    }
    float a = 2.0f * rng.next() - 1.0f;
    float b = 2.0f * rng.next() - 1.0f;
    f3 rd = splat(0.0f);
    if (!(a == 0.0f && b == 0.0f)) {
        float r, phi;
        if (a * a > b * b) { r = a; phi = (kPi / 4.0f) * (b / a); }
        else { r = b; phi = (kPi / 2.0f) - (kPi / 4.0f) * (a / b); }
        rd = mk(r * cosf(phi), r * sinf(phi), 0.0f);
    }
    rd = S.lens_radius * rd;
    f3 off = S.cam_u * rd.x + S.cam_vv * rd.y;
    o = S.cam_o + off;
    // GetRayWithDifferentials aims from the lens point; GetRay (texture
    // filtering off) subtracts the origin and the offset in turn
    d = R.diff_rays ? target - o : (target - S.cam_o) - off;
}

// Window sample `sid` → (part, job, pixel, pass) through the part table.
__device__ __forceinline__ void resolve_sample(const PathBufs& P, int sid, int& job, int& pix, int& pass) {
    int lo = 0, hi = P.nparts - 1;
    while (lo < hi) {   // last part with off <= sid
        int mid = (lo + hi + 1) / 2;
        if (P.parts[mid].off <= sid) lo = mid;
        else hi = mid - 1;
    }
    const PartDev pt = P.parts[lo];
    int local = sid - pt.off;
    job = pt.job;
    pix = local % pt.npix;
    pass = pt.pass0 + local / pt.npix;
}

// Starts the camera path of window sample `sid` (pixel `pix`, sample pass
// `pass` of job `job`) in slot p.
__device__ __forceinline__ void init_path(const DevScene& S, const RParams& R, PathBufs& P, int p, int sid, int job,
                                          int pix, int pass) {
    int i = pix % R.W, j = pix / R.W;
    pass += R.soff;   // the sample's global index (Sobol index, half parity)
    Rng rng{(uint32_t)pass, pixel_seed(i, j, R.seed_xor), 0u, R.prng != 0};
    f3 o, d;
    camera_ray(S, R, i, j, rng, o, d);
    P.o[p] = make_float4(o.x, o.y, o.z, 0.0f);
    P.d[p] = make_float4(d.x, d.y, d.z, __int_as_float(kPfPrevDelta));
    P.thr[p] = make_float4(1.0f, 1.0f, 1.0f, __int_as_float(R.max_depth));
    P.wt[p] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(rng.dim));
    P.w0[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    P.a0[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.ai[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    P.ab[p] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
    if (R.aovs) {
        P.av0[p] = make_float4(1.0f, 1.0f, 1.0f, __int_as_float(kAovPending));
        P.av1[p] = make_float4(1.0f, 1.0f, 1.0f, -1.0f);
        P.av2[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        P.av3[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        P.av4[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    P.sid[p] = sid;
    P.job[p] = job;
    P.pix[p] = pix;
    P.pass[p] = pass;
}

// Initial fill of the path pool: slot p < n traces window sample p; the
// rest of the pool stays idle.
template <class JOBS>
__global__ void __launch_bounds__(kBlock) k_raygen(JOBS jobs, PathBufs P, int n, int ws) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P.n) return;
    int st = kSlotIdle;
    if (p < n) {
        int job, pix, pass;
        resolve_sample(P, p, job, pix, pass);
        const RParams& R = jobs.params(job);
        init_path(jobs.scene(job), R, P, p, p, job, pix, pass);
        st = R.max_depth > 0 ? kSlotActive : kSlotRetired;
    }
    P.state[p] = st;
    if (p == 0) {
        P.counters[kCntNext] = n;
        P.counters[kCntSid] = n;
        P.counters[kCntSidSnap] = n;
        P.counters[kCntWs] = ws;
    }
}

// Start of one bounce iteration: count the paths pending in it (paths that
// continued from the previous iteration + slots the previous k_retire
// restarted, derived from the window's sample counter), add the previous
// iteration's traced and shadow segments to the statistics, clear the
// per-iteration counters.  One wave.
__global__ void __launch_bounds__(kBlock) k_iter_begin(PathBufs P, int traced) {
    if (threadIdx.x == 0) {
        const int sid = P.counters[kCntSid], snap = P.counters[kCntSidSnap], ws = P.counters[kCntWs];
        const int started = min(sid - snap, max(0, ws - snap));
        int active = P.counters[kCntNext] + started;
        P.counters[kCntSidSnap] = sid;
        long long sh = P.counters[kCntShadow];
        for (int j = 0; j < kMaxSeg; j++) sh += P.counters[kCntSeg + j];
        P.counters[kCntActive] = active;
        if (traced && active > 0) P.tcount[kTcRays] += (unsigned long long)active;
        if (sh > 0) P.tcount[kTcShadowRays] += (unsigned long long)sh;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kCntClear; k += kBlock) P.counters[k] = 0;
}

// End of a batch of iterations: the paths pending in the next iteration
// (k_iter_begin's count, without its side effects), the overflow flag and
// whether every sample of the window has been issued, for one host copy.
__global__ void __launch_bounds__(kBlock) k_batch_end(PathBufs P) {
    if (threadIdx.x == 0) {
        const int sid = P.counters[kCntSid], snap = P.counters[kCntSidSnap], ws = P.counters[kCntWs];
        const int started = min(sid - snap, max(0, ws - snap));
        P.counters[kCntPend] = P.counters[kCntNext] + started;
        P.counters[kCntPend + 1] = P.counters[kCntOverflow];
        P.counters[kCntPend + 2] = sid >= ws ? 1 : 0;
    }
}

// Traversal stack placement: LDS (lane-interleaved) or the wave's private
// scratch (PRIV); chosen per context.
// Single flat scenes build for 8 waves/SIMD (64 VGPRs, 2 spilled; 2.6 %
// faster than the 7-wave build), batches of flat scenes for 7; the
// wrapper-entry traversal takes the registers it needs (3-4 waves).
template <class JOBS, int SD, bool COUNT, bool PRIV, bool XF = false>
__global__ void __launch_bounds__(kBlock, XF ? 1 : (std::is_same<JOBS, OneJob>::value ? 8 : 7))
    k_trace(JOBS jobs, PathBufs P) {
    __shared__ uint2 lds_stk[PRIV ? 1 : SD * kBlock];
    uint2 priv_stk[PRIV ? SD : 1];
    uint2* stk = PRIV ? priv_stk : &lds_stk[threadIdx.x];
    constexpr int STRIDE = PRIV ? 1 : kBlock;
    TravCount tc{0u, 0u, 0u};
    const int p = xcd_block() * kBlock + threadIdx.x;
    if (p >= P.n || P.state[p] != kSlotActive) {
        if (COUNT) flush_counts(P.tcount, tc);
        return;
    }
    jobs.run_lane(true, jobs.job_of(P.job, p), [&](const DevScene& S, const RParams&) {
        float4 o4 = P.o[p], d4 = P.d[p];
        f3 o = f3of(o4), d = f3of(d4);
        bool camera = pf_cat(__float_as_int(d4.w)) == 0;
        HitOut h;
        bool moved = false;
        for (int skip = 0;;) {
            h = trace_closest<SD, STRIDE, COUNT, false, XF>(S.nodes, S.entries, S.aux, S.top_start, S.top_count, o, d,
                                                            kEps, kInfinity, stk, (uint32_t*)&P.counters[kCntOverflow],
                                                            tc, S.xf);
            if (h.entry < 0 || !camera) break;
            if (XF) {
                if (!hit_cam_invisible(S, h.entry, h.chain)) break;
                f3 lo, ld;
                o = offset_origin(make_surf_w(S, h.entry, h.chain, o, d, h.t, h.u, h.v, lo, ld).p, d);
            } else {
                int32_t fl = tag_of(S.entries[h.entry].a) >> 8;
                if (!(fl & kEntryCamInvisible)) break;
                o = offset_origin(o + h.t * d, d);
            }
            moved = true;
            if (++skip >= 8) break;
        }
        if (moved) P.o[p] = make_float4(o.x, o.y, o.z, o4.w);
        P.hit_e[p] = h.entry;
        P.hit_tuv[p] = make_float4(h.t, h.u, h.v, 0.0f);
        if (XF) P.hit_xf[p] = h.chain;
    });
    if (COUNT) flush_counts(P.tcount, tc);
}

// HitRecord.DpDu/DpDv an entry's hit writes (Triangle.cs:83-84, Quad.cs:79-80,
// Sphere.cs:65-70); false for SmoothTriangle and InfinitePlane.
__device__ __forceinline__ bool entry_dpdu(const DevScene& S, int32_t e, f3 o, f3 d, float t, f3& du, f3& dv) {
    const GEntry en = S.entries[e];
    const int kind = tag_of(en.a) & 0xff;
    if (kind == kTri) {
        if (S.smooth && S.smooth[4 * (size_t)__float_as_int(en.b.w) + 3].w != 0.0f) return false;
        du = f3of(en.b);
        dv = f3of(en.c);
        return true;
    }
    if (kind == kQuad) {
        const int ai = __float_as_int(en.d.y);
        du = f3of(S.aux[4 * ai]);
        dv = f3of(S.aux[4 * ai + 1]);
        return true;
    }
    if (kind == kSphere) {
        const Surf s = make_surf(S, e, o, d, t, 0.0f, 0.0f);
        const f3 outward = (s.p - f3of(en.a)) / en.b.x;
        const float st = sqrtf(fmaxf(0.0f, 1.0f - outward.y * outward.y));
        du = s.tan * ((2.0f * kPi) * en.b.x * st);
        dv = s.bitan * (kPi * en.b.x);
        return true;
    }
    return false;
}

// Footprint of every camera-ray hit on an image-textured material
// (Renderer.cs:1665-1677): the differentials of GetRayWithDifferentials
// (Camera.cs:127-153, pixel deltas shrunk by ceil(sqrt(spp)),
// Renderer.cs:616-624) projected on the hit's tangent plane and solved for
// the UV partials against DpDu/DpDv, or the hit's tangent frame when the
// record holds none.  A SmoothTriangle or plane hit keeps the DpDu/DpDv an
// earlier accepted hit of the same traversal left on the reference's
// record, so those hits re-trace the segment with trace_closest_dp.
template <class JOBS, int SD, bool COUNT, bool PRIV>
__global__ void __launch_bounds__(kBlock) k_fp(JOBS jobs, PathBufs P) {
    __shared__ uint2 lds_stk[PRIV ? 1 : SD * kBlock];
    uint2 priv_stk[PRIV ? SD : 1];
    uint2* stk = PRIV ? priv_stk : &lds_stk[threadIdx.x];
    constexpr int STRIDE = PRIV ? 1 : kBlock;
    const int p = xcd_block() * kBlock + threadIdx.x;
    if (p >= P.n || P.state[p] != kSlotActive) return;
    jobs.run(true, jobs.job_of(P.job, p), [&](const DevScene& S, const RParams& R) {
        float4 fp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const int he = P.hit_e[p];
        const float4 d4 = P.d[p];
        const bool camera = pf_cat(__float_as_int(d4.w)) == 0;
        if (R.footprints && he >= 0 && camera && __float_as_int(P.thr[p].w) == R.max_depth &&
            (S.mats[__float_as_int(S.entries[he].c.w)].m0.y & kMatImage)) {
            const f3 o = f3of(P.o[p]), d = f3of(d4);
            const float4 h = P.hit_tuv[p];
            const Surf s = make_surf(S, he, o, d, h.x, h.y, h.z);
            f3 du = splat(0.0f), dv = splat(0.0f);
            if (!entry_dpdu(S, he, o, d, h.x, du, dv)) {
                int32_t dp;
                float dpt;
                trace_closest_dp<SD, STRIDE>(S.nodes, S.entries, S.aux, S.smooth, S.top_start, S.top_count, o, d,
                                             kEps, kInfinity, stk, (uint32_t*)&P.counters[kCntOverflow], dp, dpt);
                if (dp < 0 || !entry_dpdu(S, dp, o, d, dpt, du, dv)) du = dv = splat(0.0f);
            }
            const f3 dpdu = dot(du, du) > 0.0f ? du : s.tan;
            const f3 dpdv = dot(dv, dv) > 0.0f ? dv : s.bitan;
            // GetRayWithDifferentials: the primary's jitter (and lens) sample
            const int pix = P.pix[p], i = pix % R.W, j = pix / R.W;
            Rng rng{(uint32_t)P.pass[p], pixel_seed(i, j, R.seed_xor), 0u, R.prng != 0};
            float ju, jv;
            pixel_jitter(R, rng, ju, jv);
            const float su = ((float)i + ju) / (float)R.W;
            const float tv = ((float)(R.H - j - 1) + jv) / (float)R.H;
            const int sq = (int)ceilf(sqrtf((float)R.spp_fp));
            const float shrink = 1.0f / (float)sq;
            const float dsdx = shrink / (float)R.W, dtdy = shrink / (float)R.H;
            f3 org = S.cam_o;
            if (S.lens_radius != 0.0f) {
                f3 lo, ld;
                Rng r2{(uint32_t)P.pass[p], pixel_seed(i, j, R.seed_xor), 0u, R.prng != 0};
                camera_ray(S, R, i, j, r2, lo, ld);
                org = lo;
            }
            const f3 tx = (S.cam_llc + (su + dsdx) * S.cam_h) + tv * S.cam_v;
            const f3 ty = (S.cam_llc + su * S.cam_h) + (tv - dtdy) * S.cam_v;
            fp = uv_footprint(org, tx - org, ty - org, s.p, s.n, dpdu, dpdv);
        }
        P.fp[p] = fp;
    });
}

__device__ __forceinline__ void add_target(PathBufs& P, int p, bool ind, f3 c) {
    float4* acc = ind ? P.ai : P.a0;
    float4 a = acc[p];
    acc[p] = make_float4(a.x + c.x, a.y + c.y, a.z + c.z, a.w);
}

// Adds the transmittance-weighted shadow contributions of the path's last
// vertex to its accumulator, in light/sample order.
__device__ __forceinline__ int32_t fold_shadows(PathBufs& P, const RParams& R, int p, int32_t flags) {
    int pend = pf_pend(flags);
    if (pend == 0) return flags;
    f3 s = splat(0.0f);
    if (R.maxs <= kMaxSeg) {
        // record of light sample j at j*N + p, summed in light-sample order
        // (the slot's mask of light samples with a shadow ray)
        for (uint32_t m = (uint32_t)P.smask[p]; m != 0u; m &= m - 1u)
            s += f3of(P.sh_r[(int64_t)(__ffs(m) - 1) * P.n + p]);
    } else {
        for (int k = 0; k < pend; k++) s += f3of(P.sh_r[(int64_t)k * P.n + p]);
    }
    add_target(P, p, (flags & kPfPendInd) != 0, s);
    return flags & ~(0xff << kPfPendShift) & ~kPfPendInd;
}

// ADJ: the replay pass of the path-replay adjoint (rt_adjoint) — the same
// shading plus, at textured Disney hits, the texture-gradient terms: NEE
// records carry dBrdf/dp contracted with the pixel's adjoint weight (added by
// k_adj_nee when the shadow ray is unoccluded), and a sampled bounce leaves
// dF/dp / F (per channel) pending until the next shade, where the radiance
// gathered after this vertex, D = (primal a0 + ai) - (a0 + ai so far), is
// known (path replay backpropagation, Vicini, Speierer and Jakob 2021).
//
// PH = kPhNee leaves surface vertices of scattering materials in state
// kSlotBounce with the vertex's throughput (medium weight applied), radiance
// weight (segment weight applied), sampler dimension and pending-shadow
// flags stored; PH = kPhBounce picks them up.  Sampler dimensions are drawn
// in the same order either way.
// RND: the PRNG sampler mode, fixed per kernel so the Sobol kernels carry no
// code of the other mode.  UL: every job of the launch samples all lights
// (no light pick), so the NEE light loop runs over a wave-uniform index.
template <class JOBS, int WPE, bool ADJ = false, int PH = kPhAll, bool TF = true, bool RND = false, bool UL = false>
__global__ void __launch_bounds__(kBlock, WPE) k_shade(JOBS jobs, PathBufs P, int maxs) {
    static_assert(!ADJ || PH == kPhAll, "the adjoint replay shades whole vertices");
    constexpr bool kFront = PH != kPhBounce;    // medium, emission, NEE, gather
    constexpr bool kBack = PH != kPhNee;        // BSDF bounce
    const int p = xcd_block() * kBlock + threadIdx.x;
    const int st_in = p < P.n ? P.state[p] : kSlotIdle;
    const bool active = st_in == (PH == kPhBounce ? kSlotBounce : kSlotActive);
    const int job = active ? jobs.job_of(P.job, p) : 0;
    int n_shadow = 0;
    uint32_t smask = 0;   // bit j: light sample j produced a shadow ray
    bool cont = false;
    bool bounce = false;  // kPhNee: the vertex waits for its BSDF bounce
    bool gather = false;  // a caustic gather request was queued
    // per-lane shadow records are written straight into the path's slots
    jobs.run_lane(active, job, [&](const DevScene& S, const RParams& R) {
        float4 o4 = P.o[p], d4 = P.d[p], wt4 = P.wt[p];
        // the bounce phase reads the throughput and the medium absorption where
        // it uses them (after its BSDF sample), not at entry
        float4 thr4;
        if constexpr (PH == kPhBounce) thr4 = make_float4(0.0f, 0.0f, 0.0f, reinterpret_cast<const float*>(P.thr + p)[3]);
        else thr4 = P.thr[p];
        int32_t flags = __float_as_int(d4.w);
        if constexpr (kFront) flags = fold_shadows(P, R, p, flags);
        bool adj_pend = false;   // this vertex leaves a bounce gradient pending
        float adj_qv[16];
        float2 adj_uv = make_float2(0.0f, 0.0f);
        if constexpr (ADJ) {
            const float4 qf = P.adj_q[5 * (int64_t)p + 4];
            if (qf.z != 0.0f) {
                // the previous vertex's bounce: everything gathered after it
                const float4 T = P.res_l[P.sid[p]];
                const float4 a0 = P.a0[p], ai = P.ai[p];
                const float D[3] = {T.x - (a0.x + ai.x), T.y - (a0.y + ai.y), T.z - (a0.z + ai.z)};
                float q[16];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    float4 v = P.adj_q[5 * (int64_t)p + k];
                    q[4 * k] = v.x;
                    q[4 * k + 1] = v.y;
                    q[4 * k + 2] = v.z;
                    q[4 * k + 3] = v.w;
                }
                float gsum[kNd];
#pragma unroll
                for (int j = 0; j < kNd; j++) gsum[j] = (D[0] * q[j] + D[1] * q[kNd + j]) + D[2] * q[2 * kNd + j];
                tex_scatter(P.adj_dt, S.tex_w, S.tex_h, qf.x, qf.y, gsum);
            }
        }
        int depth = __float_as_int(thr4.w);
        const int pix = P.pix[p];
        Rng rng{(uint32_t)P.pass[p], pixel_seed(pix % R.W, pix / R.W, R.seed_xor), __float_as_uint(wt4.w), RND};
        f3 o = f3of(o4), d = f3of(d4);
        float prev_pdf = o4.w;
        bool prev_delta = (flags & kPfPrevDelta) != 0;
        int cat = pf_cat(flags);
        bool tgt = (flags & kPfTgtInd) != 0;
        const int cst = (flags >> kPfCstShift) & 3;
        int next_cs = cst;
        f3 thr = f3of(thr4), wt = f3of(wt4);
        float4 ab4;
        if constexpr (PH == kPhBounce) ab4 = make_float4(0.0f, 0.0f, 0.0f, reinterpret_cast<const float*>(P.ab + p)[3]);
        else ab4 = P.ab[p];
        f3 sigma_abs = f3of(ab4);
        bool abs_set = PH != kPhBounce;   // sigma_abs holds the path's value
        int ior_n = __float_as_int(ab4.w);
        int32_t he = P.hit_e[p];
        float4 htuv = P.hit_tuv[p];
        bool hit = he >= 0;
        bool primary = depth == R.max_depth;
        int used = R.max_depth - depth;

        // ── medium free flight (HomogeneousMedium.Sample) ──────────────
        f3 mb = splat(1.0f);
        bool scat = false;
        float tmed = 0.0f;
        if (kFront && S.medium) {
            float tmax = hit ? htuv.x : 1e30f;
            int ch = (int)(rng.next() * 3.0f);
            if (ch > 2) ch = 2;
            float st = ch == 0 ? S.sig_t.x : (ch == 1 ? S.sig_t.y : S.sig_t.z);
            if (st <= 0.0f) tmed = tmax;
            else {
                float dist = -logf(1.0f - rng.next()) / st;
                scat = dist < tmax;
                tmed = scat ? dist : tmax;
                f3 tr = medium_tr(S, tmed);
                f3 dens = scat ? S.sig_t * tr : tr;
                float pdf = (dens.x + dens.y + dens.z) / 3.0f;
                if (pdf < 1e-20f) pdf = 1.0f;
                mb = scat ? (tr * S.sig_s / pdf) : (tr / pdf);
            }
        }
        // ── Beer-Lambert on the segment (ApplyBeerLambert) ──────────────
        f3 bl = splat(1.0f);
        if (kFront && (sigma_abs.x > 0.0f || sigma_abs.y > 0.0f || sigma_abs.z > 0.0f)) {
            if (!scat && !hit) bl = splat(0.0f);
            else {
                float tseg = scat ? tmed : htuv.x;
                if (tseg > 0.0f)
                    bl = mk(sigma_abs.x > 0.0f ? expf(-sigma_abs.x * tseg) : 1.0f,
                            sigma_abs.y > 0.0f ? expf(-sigma_abs.y * tseg) : 1.0f,
                            sigma_abs.z > 0.0f ? expf(-sigma_abs.z * tseg) : 1.0f);
            }
        }
        f3 seg = bl * mb;
        if constexpr (kFront) {
            if (primary) P.w0[p] = make_float4(seg.x, seg.y, seg.z, 0.0f);
            else wt = wt * seg;
        }

        int aovf = 0;
        float4 av0, av1;
        bool aov_dirty = false;   // kPhBounce: the AOV state was fetched (and is stored back)
        // the bounce phase fetches the AOV state where it uses it (after the
        // BSDF sample), so it is not held in registers through the light
        // samples and the sampling; the NEE phase already set the depth
        if (PH != kPhBounce && R.aovs) {
            av0 = P.av0[p];
            av1 = P.av1[p];
            aovf = __float_as_int(av0.w);
            if (hit && cat == 0 && (aovf & kAovPending) && !(aovf & kAovHaveDepth)) {
                av1.w = htuv.x * length(d);
                aovf |= kAovHaveDepth;
            }
        }
        auto aov_fetch = [&]() {
            if constexpr (PH == kPhBounce) {
                av0 = P.av0[p];
                av1 = P.av1[p];
                aovf = __float_as_int(av0.w);
                aov_dirty = true;
            }
        };
        auto aov_commit = [&](f3 alb, f3 nrm) {
            f3 aw = f3of(av0);
            f3 a = aw * alb;
            P.av3[p] = make_float4(a.x, a.y, a.z, 0.0f);
            P.av4[p] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
            aovf &= ~kAovPending;
        };

        const bool grouped = maxs <= kMaxSeg;
        f3 next_o = o, next_d = d;
        float next_pdf = 0.0f;
        bool next_delta = false;
        int next_cat = 0;
        f3 thr_next = thr, wt_next = wt;
        bool next_tgt = tgt;
        if constexpr (PH == kPhBounce) {
            // the light samples the NEE phase left for the BSDF evaluation
            const int32_t m = P.smask[p];
            smask = grouped ? (uint32_t)m : 0u;
            n_shadow = grouped ? __popc((uint32_t)m) : m;
        }

        // ── NEE light sampling (ComputeDirectLighting / ComputeDirectLightingMedium,
        // Renderer.cs:1894-2036, 2099-2206): the light pick (one dimension when a
        // strategy picks a light), then every shadow sample of the chosen
        // lights in list order with each light's own draws.  A medium vertex
        // weighs each sample with the phase function at once; a surface vertex
        // (records = true) leaves a candidate record per sample for the BSDF
        // evaluation: direction and light-pdf term (-1: delta light), light
        // colour and pick probability (-1: all lights), and the medium
        // transmittance in sh_r.
        auto sample_lights = [&](f3 pt, f3 nrm, bool medium_vtx, f3 w_vtx, bool records) {
            int first_l = 0, last_l = S.num_lights;
            float ppick = -1.0f;
            if (R.picked && S.num_lights > 0) {
                float xi = rng.next();
                int lo = 0, hi = S.num_lights - 1;
                if (S.num_lights > 1)
                    while (lo < hi) {
                        int mid = (lo + hi) / 2;
                        if (S.light_cdf[mid] < xi) lo = mid + 1;
                        else hi = mid;
                    }
                first_l = lo;
                last_l = lo + 1;
                ppick = S.light_pdf[lo];
            }
            int jn = 0;
            // UL (Strategy All): every lane walks the whole list, so the
            // light index is wave-uniform and the light record is read with
            // scalar loads into scalar registers; otherwise each lane walks
            // its own range (its picked light)
            const int l_begin = UL ? 0 : first_l, l_end = UL ? S.num_lights : last_l;
            for (int li = l_begin; li < l_end; li++) {
                GLight L = S.lights[li];
                for (int si = 0; si < L.l0.y; si++, jn++) {
                    LSample ls = light_sample(L, pt, nrm, si, rng);
                    if (ls.skip || !records) continue;
                    const f3 tr = S.medium ? medium_tr(S, ls.dist) : splat(1.0f);
                    float pn = -1.0f;
                    if (!ls.delta) {
                        const float pl = light_pdf(L, pt, ls.dir);
                        pn = ppick > 0.0f ? ppick * pl : pl;
                    }
                    const int64_t sl = (int64_t)(grouped ? jn : n_shadow) * P.n + p;
                    if (medium_vtx) {
                        float wn = 1.0f;
                        if (!ls.delta && pn + kInv4Pi > 0.0f) wn = mis_w(pn, kInv4Pi, R.power_mis);
                        f3 c = wn * ls.color * kInv4Pi * tr;
                        if (ppick > 0.0f) c = c / ppick;
                        c = w_vtx * c;
                        if (is_zero(c)) continue;
                        P.sh_d[sl] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, 0.0f);
                        P.sh_c[sl] = make_float4(c.x, c.y, c.z, 0.0f);
                    } else {
                        P.sh_d[sl] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, pn);
                        P.sh_c[sl] = make_float4(ls.color.x, ls.color.y, ls.color.z, ppick);
                        if (S.medium) P.sh_r[sl] = make_float4(tr.x, tr.y, tr.z, 0.0f);
                    }
                    P.sh_o[sl] = make_float4(ls.so.x, ls.so.y, ls.so.z, ls.tmax);
                    if (jn < kMaxSeg) smask |= 1u << jn;
                    n_shadow++;
                }
            }
        };

        // ── the vertex before its light samples ─────────────────────────
        f3 ps3 = o;                 // medium scattering point
        Surf s{};
        GMat g{};
        float tp[kTexChannels];
        float2 tuv = make_float2(0.0f, 0.0f);
        bool textured = false, image = false, scatters = false, has_brdf = false;
        int kind = -1, mflags = 0;
        // roulette throughput with the medium weight (the bounce phase fetches it)
        auto thr_at = [&]() -> f3 {
            if constexpr (PH == kPhBounce) return f3of(P.thr[p]);
            else return thr * mb;
        };
        const f3 vw = primary ? splat(1.0f) : wt;   // weight of this vertex's own terms
        f3 view = splat(0.0f);
        bool nee = false;           // the vertex samples the lights
        if (kFront && scat) {
            // ── medium scattering event (Renderer.cs:1703-1770) ─────────
            ps3 = o + d * tmed;
            if (R.aovs && (aovf & kAovPending)) {
                if (!(aovf & kAovHaveDepth)) { av1.w = tmed * length(d); aovf |= kAovHaveDepth; }
                aov_commit(splat(1.0f), splat(0.0f));
            }
            nee = true;
        } else if (kFront && !hit) {
            // ── miss: the flat sky (SampleSky, Renderer.cs:1820-1861), weighed
            // against the environment light's NEE by MIS after a non-delta
            // bounce when the light list holds one
            const f3 sky = S.sky;
            if (R.aovs && (aovf & kAovPending)) aov_commit(clamp01(sky), splat(0.0f));
            if (sky.x > 0.0f || sky.y > 0.0f || sky.z > 0.0f) {
                f3 sw = sky;
                if (!prev_delta && S.env_light >= 0 && S.lights[S.env_light].l0.w) {
                    const float pl = kUniformSpherePdf;
                    sw = prev_pdf + pl <= 1e-20f ? splat(0.0f) : sky * mis_w(prev_pdf, pl, R.power_mis);
                }
                add_target(P, p, tgt, sanitize((primary ? splat(1.0f) : wt) * sw));
            }
        } else {
            // ── surface (ShadeSurface, Renderer.cs:1008-1228) ──────────
            f3 lo = o, ld = d;      // the ray of the hit's frame
            if (S.xf) s = make_surf_w(S, he, P.hit_xf[p], o, d, htuv.x, htuv.y, htuv.z, lo, ld);
            else s = make_surf(S, he, o, d, htuv.x, htuv.y, htuv.z);
            g = S.mats[s.mat];
            textured = (g.m0.y & kMatTextured) != 0;
            image = (g.m0.y & kMatImage) != 0;
            // textures and images feed the BSDF and the albedo AOV of a
            // scattering vertex: the bounce phase's work
            if (kBack && (textured || image)) {
                tuv = tex_coord(S, he, lo, ld, htuv);
                // camera hits carry the footprint k_fp left; later hits are point-sampled
                const float4 fpv = image && P.fp ? P.fp[p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                g = textured_gmat(S, g, s.mat, tuv, tp, image, fpv);
            }
            kind = g.m0.x;
            mflags = g.m0.y;
            scatters = kind == RT_MAT_LAMBERTIAN || kind == RT_MAT_DISNEY || kind == RT_MAT_DIELECTRIC;
            // tentative AOVs (AovProbe, Renderer.cs:777-815): in the phase
            // that holds the vertex's textured material
            const bool aov_here = PH == kPhAll || (PH == kPhNee ? !scatters : true);
            if constexpr (PH == kPhBounce) {
                if (R.aovs && (__float_as_int(P.av0[p].w) & kAovPending)) {
                    P.av2[p] = make_float4(s.n.x, s.n.y, s.n.z, 0.0f);
                    const f3 ta = clamp01(f3of(g.m1));
                    float* a1 = reinterpret_cast<float*>(P.av1 + p);
                    a1[0] = ta.x;
                    a1[1] = ta.y;
                    a1[2] = ta.z;
                }
            } else if (aov_here && R.aovs && (aovf & kAovPending)) {
                P.av2[p] = make_float4(s.n.x, s.n.y, s.n.z, 0.0f);
                f3 ta = clamp01(f3of(g.m1));
                av1 = make_float4(ta.x, ta.y, ta.z, av1.w);
            }
            // emission with MIS (WeightEmission, Renderer.cs:1549-1565)
            if (kFront && kind == RT_MAT_EMISSIVE && s.front && !(S.caustics && cst == kCstSuppress)) {
                f3 raw = f3of(g.m1);
                if (raw.x > 0.0f || raw.y > 0.0f || raw.z > 0.0f) {
                    f3 em = raw;
                    int lix = g.m0.z;
                    if (!prev_delta && lix >= 0) {
                        float pl = light_pdf(S.lights[lix], o, d);
                        if (prev_pdf + pl <= 1e-20f) em = splat(0.0f);
                        else em = raw * mis_w(prev_pdf, pl, R.power_mis);
                    }
                    add_target(P, p, tgt, sanitize(vw * em));
                }
            }
            view = normalize(-d);
            if ((kBack || S.caustics) && (mflags & kMatDielectricIor)) {
                float top = ior_n > 0 ? P.ior[(int64_t)(ior_n - 1) * P.n + p] : 1.0f;
                float enc = ior_n >= 2 ? P.ior[(int64_t)(ior_n - 2) * P.n + p] : 1.0f;
                float mi = g.m3.w;
                if (mi > 0.0f) s.rel_eta = s.front ? (top / mi) : (mi / enc);
            }
            has_brdf = kind == RT_MAT_LAMBERTIAN || kind == RT_MAT_DISNEY;
            nee = kFront && kind != RT_MAT_EMISSIVE;
        }
        if constexpr (PH == kPhNee) {
            // what the light draws leave unchanged is stored before them, so
            // it is not held in registers across them: the AOV state, the
            // caustic gather request (Renderer.cs:1108-1109) and the bounce
            // phase's throughput and weight (the sampler dimension follows
            // the draws)
            if (R.aovs) {
                P.av0[p] = make_float4(av0.x, av0.y, av0.z, __int_as_float(aovf));
                P.av1[p] = av1;
            }
            if (!scat && hit) {
                if (S.caustics && has_brdf) {
                    gather = true;
                    P.gq0[p] = make_float4(o.x, o.y, o.z, htuv.x);
                    P.gq1[p] = make_float4(d.x, d.y, d.z, __int_as_float(he));
                    P.gq2[p] = make_float4(htuv.y, htuv.z, s.rel_eta, __int_as_float(p));
                    P.gq3[p] = make_float4(vw.x, vw.y, vw.z, __int_as_float(tgt ? 1 : 0));
                }
                // the NEE phase hands scattering vertices to the bounce phase
                bounce = scatters;
                if (bounce) {
                    const f3 ts = thr * mb;
                    P.thr[p] = make_float4(ts.x, ts.y, ts.z, __int_as_float(depth));
                    P.wt[p] = make_float4(wt.x, wt.y, wt.z, 0.0f);
                }
            }
        }
        // NEE light samples (one site for medium and surface vertices)
        if (kFront && nee) sample_lights(scat ? ps3 : s.p, scat ? splat(0.0f) : s.n, scat, vw, scat || has_brdf);
        if (kFront && scat) {
            // ── roulette and phase sampling of a medium vertex
            f3 mt = thr * mb;
            float iscale = 1.0f;
            bool kill = false;
            if (used >= R.rr_min_b) {
                float sp = fmax3(mt);
                sp = fmaxf(sp, R.rr_min_s);
                sp = fminf(sp, 0.95f);
                if (rng.next() > sp) kill = true;
                else iscale = 1.0f / sp;
            }
            if (!kill) {
                float u1 = rng.next(), u2 = rng.next();
                float z = 1.0f - 2.0f * u1;
                float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
                float phi = 2.0f * kPi * u2;
                f3 wi = mk(rr * cosf(phi), rr * sinf(phi), z);
                float pw = kInv4Pi / kInv4Pi;
                thr_next = mt * (pw * iscale);
                wt_next = (primary ? splat(1.0f) : wt) * (pw * iscale);
                next_o = ps3;
                next_d = wi;
                next_pdf = kInv4Pi;
                next_delta = false;
                next_cat = 0;
                next_cs = kCstDiffuse;
                cont = true;
            }
        } else if (hit) {
            // ── the surface vertex after its light samples
            if constexpr (PH == kPhAll) {
                // caustic gather request (Renderer.cs:1108-1109)
                if (S.caustics && has_brdf) {
                    gather = true;
                    P.gq0[p] = make_float4(o.x, o.y, o.z, htuv.x);
                    P.gq1[p] = make_float4(d.x, d.y, d.z, __int_as_float(he));
                    P.gq2[p] = make_float4(htuv.y, htuv.z, s.rel_eta, __int_as_float(p));
                    P.gq3[p] = make_float4(vw.x, vw.y, vw.z, __int_as_float(tgt ? 1 : 0));
                }
            }
            if constexpr (kBack) {
                DisneyP DP;
                if (kind == RT_MAT_DISNEY) {
                    DP = load_disney(g);
                    // kernels for scenes without thin films: no film (the
                    // iridescent Fresnel compiles out)
                    if constexpr (!TF) DP.tf_thick = 0.0f;
                }
                // adjoint weight of this vertex's terms: dLoss/dPixel * w0 / spp
                DisneyD DD;
                f3 adj_w = splat(0.0f);
                const bool adj_on = ADJ && P.adj_pass == 2 && textured && kind == RT_MAT_DISNEY;
                if constexpr (ADJ) {
                    if (adj_on) {
                        DD = disney_d(S.raw[s.mat], tp, S.lut_e + 32 * 32);
                        const int pix = P.pix[p];
                        const f3 w0v = primary ? seg : f3of(P.w0[p]);
                        adj_w = mk(P.adj_db[pix], P.adj_db[R.npix + pix], P.adj_db[2 * R.npix + pix]) * w0v * P.adj_inv_spp;
                        adj_uv = tuv;
                    }
                }
                // ── BSDF evaluation of the light samples (EvaluateDirect, Pdf,
                // MisWeight): c = w_nee Le brdf Tr / pPick times the vertex weight;
                // a zero contribution drops the sample's shadow ray (the
                // ungrouped queue keeps it with c = 0)
                if (has_brdf && n_shadow > 0) {
                    const float view_term = shadow_terminator(s.n, s.n, view);
                    uint32_t left = smask;
                    for (int k = 0;; k++) {
                        int j;
                        if (grouped) {
                            if (left == 0u) break;
                            j = __ffs(left) - 1;
                            left &= left - 1u;
                        } else {
                            if (k >= n_shadow) break;
                            j = k;
                        }
                        const int64_t sl = (int64_t)j * P.n + p;
                        const float4 rd = P.sh_d[sl], rc = P.sh_c[sl];
                        const f3 dir = f3of(rd), col = f3of(rc);
                        const float pn = rd.w, ppick = rc.w;
                        f3 brdf;
                        if (kind == RT_MAT_LAMBERTIAN) {
                            float ndl = fmaxf(dot(s.n, dir), 0.0f);
                            brdf = ndl <= 0.0f ? splat(0.0f) : f3of(g.m1) * (ndl / kPi);
                        } else {
                            brdf = disney_eval_direct(DP, S.lut_e, dir, view, s);
                        }
                        brdf *= view_term * shadow_terminator(s.n, s.n, dir);
                        f3 c = splat(0.0f), tr = splat(1.0f);
                        float wn = 1.0f;
                        if (!is_zero(brdf)) {
                            if (S.medium) tr = f3of(P.sh_r[sl]);
                            if (pn >= 0.0f) {   // not a delta light
                                float pb;
                                if (kind == RT_MAT_LAMBERTIAN) {
                                    float cs = dot(s.n, dir);
                                    pb = cs > 0.0f ? cs * (1.0f / kPi) : 0.0f;
                                } else pb = disney_pdf(DP, view, dir, s);
                                if (pb > 0.0f) wn = (pn + pb > 0.0f) ? mis_w(pn, pb, R.power_mis) : 1.0f;
                            }
                            c = wn * col * brdf * tr;
                            if (ppick > 0.0f) c = c / ppick;
                            c = vw * c;
                        }
                        const bool drop = is_zero(c);
                        if (drop && grouped) {
                            smask &= ~(1u << j);
                            n_shadow--;
                            continue;
                        }
                        P.sh_c[sl] = make_float4(c.x, c.y, c.z, 0.0f);
                        if constexpr (ADJ) {
                            float4 g0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), g1 = g0;
                            if (adj_on && !drop) {
                                // c = vw * (wn Le tr / ppick) * st * brdf(p): dc/dp = k * dbrdf/dp
                                const D3 db = disney_eval_direct_d(DP, DD, S.lut_e, dir, view, s);
                                const float st = view_term * shadow_terminator(s.n, s.n, dir);
                                f3 k = wn * col * tr;
                                if (ppick > 0.0f) k = k / ppick;
                                k = vw * k * adj_w * st;
                                float gj[kNd];
    #pragma unroll
                                for (int q = 0; q < kNd; q++) gj[q] = (k.x * db.x.d[q] + k.y * db.y.d[q]) + k.z * db.z.d[q];
                                g0 = make_float4(gj[0], gj[1], gj[2], gj[3]);
                                g1 = make_float4(gj[4], tuv.x, tuv.y, 1.0f);
                            }
                            P.adj_g[2 * sl] = g0;
                            P.adj_g[2 * sl + 1] = g1;
                        }
                    }
                }
                // ── indirect bounce ─────────────────────────────────────────
                bool have = false, legacy = false, is_delta = false;
                int legacy_lobe = -1;   // Disney lobe of a fallback (Scatter) bounce
                f3 att = splat(0.0f), wo = splat(0.0f), origin = s.p;
                float spdf = 0.0f;
                BsdfS bs;
                bs.has_next_abs = false;
                bs.transition = 0;
                if (kind == RT_MAT_LAMBERTIAN) {
                    f3 alb = f3of(g.m1);
                    f3 dir = s.n + random_unit_vector(rng);
                    if (near_zero(dir)) dir = s.n;
                    f3 w = normalize(dir);
                    float ndw = dot(s.n, w);
                    if (ndw > 0.0f) {
                        bs.wo = w;
                        bs.f = alb * (1.0f / kPi);
                        bs.pdf = ndw * (1.0f / kPi);
                        bs.delta = false;
                        have = true;
                    } else {
                        f3 d2 = s.n + random_unit_vector(rng);
                        if (near_zero(d2)) d2 = s.n;
                        wo = d2;
                        att = alb;
                        legacy = true;
                    }
                } else if (kind == RT_MAT_DISNEY) {
                    if (disney_sample(DP, S.lut_e, view, s, rng, bs)) have = true;
                    else {
                        if (disney_scatter(DP, S.lut_e, s, d, rng, att, wo, legacy_lobe)) legacy = true;
                    }
                } else if (kind == RT_MAT_DIELECTRIC) {
                    dielectric_scatter(g, s, d, rng, att, wo, origin);
                    legacy = true;
                    is_delta = true;
                }
                if (have) {
                    // ShadeSampleBounce (Renderer.cs:1235-1383)
                    bool ok = true;
                    if (bs.delta) att = bs.f;
                    else {
                        float a = fabsf(dot(s.n, bs.wo));
                        if (a <= 0.0f || bs.pdf <= 0.0f) ok = false;
                        else att = bs.f * a / bs.pdf;
                    }
                    if (ok) {
                        if (R.aovs) aov_fetch();
                        if (R.aovs && (aovf & kAovPending)) {
                            if (bs.delta) {
                                f3 aw = f3of(av0) * clamp01(bs.f);
                                av0 = make_float4(aw.x, aw.y, aw.z, av0.w);
                            } else aov_commit(f3of(av1), s.n);
                        }
                        f3 nt = thr_at() * att;
                        bool killed = false;
                        if (used >= R.rr_min_b) {
                            float sp = fmax3(nt);
                            sp = fmaxf(sp, R.rr_min_s);
                            sp = fminf(sp, 0.95f);
                            if (rng.next() > sp) killed = true;
                            else {
                                float inv = 1.0f / sp;
                                att *= inv;
                                nt *= inv;
                            }
                        }
                        if (!killed) {
                            f3 od = dot(bs.wo, s.n) >= 0.0f ? s.n : -s.n;
                            next_o = offset_origin(s.p, od);
                            next_d = bs.wo;
                            next_pdf = bs.delta ? 0.0f : bs.pdf;
                            next_delta = bs.delta;
                            next_cat = dot(s.n, bs.wo) < 0.0f ? 3 : (bs.delta ? 2 : 1);
                            if (bs.has_next_abs) {
                                sigma_abs = bs.next_abs;
                                abs_set = true;
                            }
                            if (bs.transition == 1) {
                                float v = (mflags & kMatDielectricIor) ? g.m3.w
                                                                       : (ior_n > 0 ? P.ior[(int64_t)(ior_n - 1) * P.n + p] : 1.0f);
                                if (ior_n >= 8) {
                                    for (int k = 0; k < 7; k++) P.ior[(int64_t)k * P.n + p] = P.ior[(int64_t)(k + 1) * P.n + p];
                                    P.ior[(int64_t)7 * P.n + p] = v;
                                } else {
                                    P.ior[(int64_t)ior_n * P.n + p] = v;
                                    ior_n++;
                                }
                            } else if (bs.transition == 2 && ior_n > 0) ior_n--;
                            thr_next = nt;
                            wt_next = primary ? att : wt * att;
                            if (primary) next_tgt = true;
                            next_cs = next_cst(cst, bs.delta);
                            cont = true;
                            if constexpr (ADJ) {
                                if (adj_on) {
                                    // non-delta: att = F(wo) |cos| / pdf / q, so d(att)/dp / att =
                                    // dF/dp / F; delta: the lobe weight over its fixed probability
                                    const D3 fw = bs.delta ? lobe_att_d(bs.lobe, DP, DD, S.lut_e, view, bs.wo, s)
                                                           : disney_eval_d(DP, DD, S.lut_e, view, bs.wo, s);
                                    log_deriv(fw, adj_w, adj_qv);
                                    adj_pend = true;
                                }
                            }
                        }
                    }
                } else if (legacy) {
                    // legacy Scatter path (Renderer.cs:1127-1224)
                    if (R.aovs) aov_fetch();
                    if (R.aovs && (aovf & kAovPending)) {
                        if (is_delta) {
                            f3 aw = f3of(av0) * clamp01(att);
                            av0 = make_float4(aw.x, aw.y, aw.z, av0.w);
                        } else aov_commit(f3of(av1), s.n);
                    }
                    f3 nt = thr_at() * att;
                    bool killed = false;
                    if (used >= R.rr_min_b) {
                        float sp = fmax3(nt);
                        sp = fmaxf(sp, R.rr_min_s);
                        sp = fminf(sp, 0.95f);
                        if (rng.next() > sp) killed = true;
                        else {
                            float inv = 1.0f / sp;
                            att *= inv;
                            nt *= inv;
                        }
                    }
                    if (!killed) {
                        next_o = origin;
                        next_d = wo;
                        next_pdf = 0.0f;
                        next_delta = is_delta;
                        next_cat = dot(s.n, wo) < 0.0f ? 3 : (is_delta ? 2 : 1);
                        if ((mflags & kMatDielectricIor) && dot(wo, s.n) < 0.0f) {
                            if (s.front) {
                                float v = g.m3.w;
                                if (ior_n >= 8) {
                                    for (int k = 0; k < 7; k++) P.ior[(int64_t)k * P.n + p] = P.ior[(int64_t)(k + 1) * P.n + p];
                                    P.ior[(int64_t)7 * P.n + p] = v;
                                } else {
                                    P.ior[(int64_t)ior_n * P.n + p] = v;
                                    ior_n++;
                                }
                            } else if (ior_n > 0) ior_n--;
                        }
                        thr_next = nt;
                        wt_next = primary ? att : wt * att;
                        if (primary) next_tgt = true;
                        next_cs = next_cst(cst, is_delta);
                        cont = true;
                        if constexpr (ADJ) {
                            if (adj_on && legacy_lobe >= 0) {
                                const f3 V = normalize(-d);
                                log_deriv(lobe_att_d(legacy_lobe, DP, DD, S.lut_e, V, wo, s), adj_w, adj_qv);
                                adj_pend = true;
                            }
                        }
                    }
                }
            }
        }
        // pending shadows of this vertex belong to this vertex's target
        flags = flags & ~(0xff << kPfPendShift) & ~kPfPendInd;
        flags |= (n_shadow << kPfPendShift) | (tgt ? kPfPendInd : 0);
        (void)grouped;
        // TraceRay entry checks of the next segment (Renderer.cs:1589-1603)
        if (cont) {
            if (depth - 1 <= 0 || fmax3(thr_next) < 1e-4f) cont = false;
        }
        if (bounce) {
            // the bounce phase continues from here: roulette throughput with
            // the medium weight and radiance weight with the segment weight
            // (stored before the light draws), the path flags (direction
            // unchanged) and the next sampler dimension
            reinterpret_cast<float*>(P.d + p)[3] = __int_as_float(flags);
            reinterpret_cast<float*>(P.wt + p)[3] = __uint_as_float(rng.dim);
        } else if (cont) {
            flags &= ~(kPfPrevDelta | (3 << kPfCatShift) | kPfTgtInd | (3 << kPfCstShift));
            flags |= (next_delta ? kPfPrevDelta : 0) | (next_cat << kPfCatShift) | (next_tgt ? kPfTgtInd : 0) |
                     (next_cs << kPfCstShift);
            P.o[p] = make_float4(next_o.x, next_o.y, next_o.z, next_pdf);
            P.d[p] = make_float4(next_d.x, next_d.y, next_d.z, __int_as_float(flags));
            P.thr[p] = make_float4(thr_next.x, thr_next.y, thr_next.z, __int_as_float(depth - 1));
            P.wt[p] = make_float4(wt_next.x, wt_next.y, wt_next.z, __uint_as_float(rng.dim));
            const f3 sab = abs_set ? sigma_abs : f3of(P.ab[p]);
            P.ab[p] = make_float4(sab.x, sab.y, sab.z, __int_as_float(ior_n));
        } else {
            P.d[p] = make_float4(d.x, d.y, d.z, __int_as_float(flags));
        }
        if (PH == kPhAll ? R.aovs != 0 : (PH == kPhBounce && aov_dirty)) {
            P.av0[p] = make_float4(av0.x, av0.y, av0.z, __int_as_float(aovf));
            P.av1[p] = av1;
        }
        if constexpr (ADJ) {
            if (adj_pend && cont) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    P.adj_q[5 * (int64_t)p + k] =
                        make_float4(adj_qv[4 * k], adj_qv[4 * k + 1], adj_qv[4 * k + 2], adj_qv[4 * k + 3]);
                P.adj_q[5 * (int64_t)p + 4] = make_float4(adj_uv.x, adj_uv.y, 1.0f, 0.0f);
            } else {
                P.adj_q[5 * (int64_t)p + 4] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
        }
    });
    // slot state, per-slot outputs and this wave's counts (no atomics: the
    // queues are laid out by k_wave_scan + k_compact from the counts)
    const bool retire = active && !cont && !bounce;
    if (p < P.n) {
        if (PH == kPhBounce) {
            if (active) {
                P.state[p] = cont ? kSlotActive : kSlotRetired;
                P.smask[p] = maxs <= kMaxSeg ? (int32_t)smask : n_shadow;
            }
        } else {
            if (retire) P.state[p] = kSlotRetired;
            else if (bounce) P.state[p] = kSlotBounce;
            P.smask[p] = maxs <= kMaxSeg ? (int32_t)smask : n_shadow;
            P.gflag[p] = gather ? 1 : 0;
        }
    }
    const int w = p >> 6;
    if (PH != kPhNee) {
        // continuing and retired paths: retired this iteration, plus slots
        // restarted as already-retired (samples of depth-0 jobs), which
        // k_retire finishes as well; the bounce phase also counts the
        // vertices the NEE phase ended or continued (medium scattering)
        const bool is_cont = PH == kPhBounce ? (active ? cont : st_in == kSlotActive) : cont;
        const bool is_ret = PH == kPhBounce ? (active ? !cont : st_in == kSlotRetired) : (retire || st_in == kSlotRetired);
        const int n_cont = __popcll(__ballot(is_cont)), n_ret = __popcll(__ballot(is_ret));
        // shadow rays of every slot: the bounce phase's own vertices after the
        // BSDF evaluation, the others as the NEE phase left them
        uint32_t sm = smask;
        int nsh = n_shadow;
        if (PH == kPhBounce && !active) {
            const int32_t m = p < P.n ? P.smask[p] : 0;
            sm = maxs <= kMaxSeg ? (uint32_t)m : 0u;
            nsh = maxs <= kMaxSeg ? 0 : m;
        }
        for (int off = 32; off > 0; off >>= 1) nsh += __shfl_xor(nsh, off, 64);
        if (threadIdx.x == 0) {
            P.wc[kWcNext * P.nw + w] = n_cont;
            P.wc[kWcRetire * P.nw + w] = n_ret;
            P.wc[kWcShadow * P.nw + w] = maxs <= kMaxSeg ? 0 : nsh;
        }
        if (maxs <= kMaxSeg)
            for (int j = 0; j < maxs; j++) {
                int c = __popcll(__ballot((sm >> j) & 1u));
                if (threadIdx.x == 0) P.wc[(kWcSeg + j) * P.nw + w] = c;
            }
    }
    if (PH != kPhBounce) {
        const int n_g = __popcll(__ballot(gather));
        if (threadIdx.x == 0) P.wc[kWcGather * P.nw + w] = n_g;
    }
}

// Per-wave counts when no path is traced (depth-0 renders): every pending
// slot is retired; the other count arrays are zero.
__global__ void __launch_bounds__(kBlock) k_count_retired(PathBufs P, int narr) {
    const int p = xcd_block() * kBlock + threadIdx.x;
    const bool ret = p < P.n && P.state[p] == kSlotRetired;
    const int n = __popcll(__ballot(ret));
    if (threadIdx.x == 0)
        for (int a = 0; a < narr; a++) P.wc[a * P.nw + (p >> 6)] = a == kWcRetire ? n : 0;
}

// Exclusive scan of one per-wave count array (block a scans array a) and
// its total into the counters; the retire array also advances the window's
// sample counter (the restarted slots take samples [sid, sid + total)).
// Each thread owns kScanPer consecutive counts, read and written as int4
// (a wave covers 8 KB of whole cache lines per round); a round scans
// 1024 * kScanPer counts and the rounds carry the running total.
constexpr int kScanPer = 32;
__global__ void __launch_bounds__(1024) k_wave_scan(PathBufs P) {
    __shared__ int32_t part[1024 / 64];
    const int a = blockIdx.x;
    int32_t* x = P.wc + (int64_t)a * P.nw;          // row stride P.nw: a multiple of 4
    const int nwr = (P.n + kBlock - 1) / kBlock;    // counts written this iteration
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int carry = 0;
    for (int r0 = 0; r0 < nwr; r0 += 1024 * kScanPer) {
        const int b0 = r0 + (int)threadIdx.x * kScanPer;
        int v[kScanPer];
#pragma unroll
        for (int k = 0; k < kScanPer; k += 4) {
            const int i = b0 + k;
            int4 q = make_int4(0, 0, 0, 0);
            if (i + 3 < nwr) {
                q = *(const int4*)(x + i);
            } else if (i < nwr) {
                q.x = x[i];
                if (i + 1 < nwr) q.y = x[i + 1];
                if (i + 2 < nwr) q.z = x[i + 2];
            }
            v[k] = q.x;
            v[k + 1] = q.y;
            v[k + 2] = q.z;
            v[k + 3] = q.w;
        }
        int t = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; k++) t += v[k];
        int incl = t;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (lane == 63) part[wid] = incl;
        __syncthreads();
        int wpre = 0, round = 0;
        for (int k = 0; k < 1024 / 64; k++) {
            const int c = part[k];
            wpre += k < wid ? c : 0;
            round += c;
        }
        int run = carry + wpre + incl - t;
#pragma unroll
        for (int k = 0; k < kScanPer; k += 4) {
            const int i = b0 + k;
            int4 q;
            q.x = run;
            q.y = q.x + v[k];
            q.z = q.y + v[k + 1];
            q.w = q.z + v[k + 2];
            run = q.w + v[k + 3];
            if (i + 3 < nwr) {
                *(int4*)(x + i) = q;
            } else if (i < nwr) {
                x[i] = q.x;
                if (i + 1 < nwr) x[i + 1] = q.y;
                if (i + 2 < nwr) x[i + 2] = q.z;
            }
        }
        carry += round;
        __syncthreads();   // part[] is rewritten by the next round
    }
    if (threadIdx.x == 0) {
        const int total = carry;
        const int slot = a == kWcNext ? kCntNext : a == kWcRetire ? kCntRetire : a == kWcGather ? kCntGather
                       : a == kWcShadow ? kCntShadow : kCntSeg + (a - kWcSeg);
        P.counters[slot] = total;
        if (a == kWcRetire) {
            P.counters[kCntSidBase] = P.counters[kCntSid];
            P.counters[kCntSid] += total;
        }
    }
}

// Lays out the shadow queue (grouped by light sample, slot order within a
// segment) and the gather queue from the scanned per-wave counts.
__global__ void __launch_bounds__(kBlock) k_compact(PathBufs P, int maxs) {
    const int p = xcd_block() * kBlock + threadIdx.x;
    const bool ok = p < P.n;
    const int w = p >> 6;
    const unsigned long long lt = (1ull << (threadIdx.x & 63)) - 1ull;
    const int st = ok ? P.state[p] : kSlotIdle;
    const bool shaded = st != kSlotIdle;
    const int m = shaded ? P.smask[p] : 0;
    if (maxs <= kMaxSeg) {
        for (int j = 0; j < maxs; j++) {
            bool has = (m >> j) & 1;
            unsigned long long b = __ballot(has);
            if (has)
                P.q_sh[(int64_t)j * P.n + P.wc[(kWcSeg + j) * P.nw + w] + __popcll(b & lt)] =
                    (int32_t)((int64_t)j * P.n + p);
        }
    } else {
        int incl = m;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(incl, off, 64);
            if ((int)(threadIdx.x & 63) >= off) incl += y;
        }
        int base = P.wc[kWcShadow * P.nw + w] + incl - m;
        for (int k = 0; k < m; k++) P.q_sh[base + k] = (int32_t)((int64_t)k * P.n + p);
    }
    const bool g = shaded && P.gflag[p];
    unsigned long long gb = __ballot(g);
    if (g) P.q_g[P.wc[kWcGather * P.nw + w] + __popcll(gb & lt)] = p;
}

}  // namespace rt
#include "photon.h"
namespace rt {

// Caustic radiance at the diffuse/glossy vertices shaded this iteration
// (Renderer.GatherCaustics, Renderer.cs:1449-1508; CausticOccupancyWeight
// :1520-1526): k nearest photons, cone filter k_f = 1 (normalisation 1/3),
// adaptive radius, occupancy fade for sparse gathers.  Added to the target
// accumulator of the path with the vertex weight.
__global__ void __launch_bounds__(kBlock) k_gather(DevScene S, PathBufs P, DevPhotons M) {
    const int count = P.counters[kCntGather];
    const uint32_t nbk = (uint32_t)((count + kBlock - 1) / kBlock);
    if (blockIdx.x >= nbk) return;
    int i = xcd_remap(blockIdx.x, nbk) * kBlock + threadIdx.x;
    if (i >= count) return;
    const int slot = P.q_g[i];
    float4 a = P.gq0[slot], b = P.gq1[slot], c = P.gq2[slot], w4 = P.gq3[slot];
    f3 o = f3of(a), d = f3of(b);
    Surf s = make_surf(S, __float_as_int(b.w), o, d, a.w, c.x, c.y);
    s.rel_eta = c.z;
    const int p = __float_as_int(c.w);
    // the lane's heap of the k nearest photons, lane-interleaved in LDS
    __shared__ float2 lds_h[kGatherK * kBlock];
    float2* h = &lds_h[threadIdx.x];
    float radius;
    int n = photon_knn<kBlock>(M, s.p, h, radius);
    if (n == 0) return;
    bool under = n < kGatherK;
    if (under) radius = M.max_r;
    if (radius <= 0.0f) return;
    const GMat g = S.mats[s.mat];
    const bool disney = g.m0.x == RT_MAT_DISNEY;
    DisneyP DP;
    if (disney) DP = load_disney(g);
    f3 view = normalize(-d);
    float inv_r = 1.0f / radius;
    f3 sum = splat(0.0f);
    // heap order, as PhotonMap hands the set back
    for (int k = 0; k < n; k++) {
        const int kk = __float_as_int(h[k * kBlock].y);
        f3 inc = f3of(M.sinc[kk]);
        if (dot(inc, s.n) <= 0.0f) continue;
        float wgt = 1.0f - length(f3of(M.spos[kk]) - s.p) * inv_r / 1.0f;
        if (wgt <= 0.0f) continue;
        f3 f;
        if (disney) f = disney_eval(DP, S.lut_e, view, inc, s);
        else f = dot(s.n, inc) <= 0.0f ? splat(0.0f) : f3of(g.m1) * (1.0f / kPi);
        sum += f * (f3of(M.spw[kk]) * wgt);
    }
    const float norm = 1.0f - 2.0f / (3.0f * 1.0f);
    f3 est = sum * (1.0f / (norm * kPi * radius * radius));
    if (under) {
        float t = (float)n / (float)kGatherK;
        est *= t * t * (3.0f - 2.0f * t);
    }
    add_target(P, p, __float_as_int(w4.w) != 0, sanitize(f3of(w4) * est));
}

// Transparent shadow walk (ShadowRay.Transmittance, ShadowRay.cs:59-158)
template <class JOBS, int SD, bool COUNT, bool PRIV, bool XF = false>
__global__ void __launch_bounds__(kBlock, XF ? 4 : 7) k_shadow(JOBS jobs, PathBufs P, int maxs) {
    __shared__ uint2 lds_stk[PRIV ? 1 : SD * kBlock];
    uint2 priv_stk[PRIV ? SD : 1];
    uint2* stk = PRIV ? priv_stk : &lds_stk[threadIdx.x];
    constexpr int STRIDE = PRIV ? 1 : kBlock;
    TravCount tc{0u, 0u, 0u};
    // shadow queue view from the segment sizes (uniform over the grid)
    const bool grouped = maxs <= kMaxSeg;
    int total = grouped ? 0 : P.counters[kCntShadow];
    if (grouped)
        for (int j = 0; j < maxs; j++) total += P.counters[kCntSeg + j];
    const uint32_t nb = (uint32_t)((total + kBlock - 1) / kBlock);
    if (blockIdx.x >= nb) {
        if (COUNT) flush_counts(P.tcount + 3, tc);
        return;
    }
    int i = xcd_remap(blockIdx.x, nb) * kBlock + threadIdx.x;
    if (i >= total) {
        if (COUNT) flush_counts(P.tcount + 3, tc);
        return;
    }
    int64_t qi = i;
    int seg = 0;
    if (grouped) {
        int pre = 0;
        for (;;) {
            int c = P.counters[kCntSeg + seg];
            if (i < pre + c || seg + 1 >= maxs) break;
            pre += c;
            seg++;
        }
        qi = (int64_t)seg * P.n + (i - pre);
    }
    int64_t sl = P.q_sh[qi];
    // slot of the record: segment j holds light sample j's records at j*N + slot
    const int slot = grouped ? (int)(sl - (int64_t)seg * P.n) : (int)(sl % P.n);
    jobs.run_lane(true, jobs.job_of(P.job, slot), [&](const DevScene& S, const RParams&) {
        float4 o4 = P.sh_o[sl], d4 = P.sh_d[sl];
        f3 org = f3of(o4), dir = f3of(d4);
        float rem = o4.w;
        if (S.opaque_shadows) {
            // Every material is opaque to shadow rays (no Dielectric, no Disney
            // spec_trans > 0), so ShadowRay.Transmittance is 0 at the first
            // interface and 1 without one: an any-hit query.
            bool blocked;
            if (S.nodes4) {
                TravCount t4{0u, 0u, 0u};
                blocked = trace_any4<SD, STRIDE, COUNT>(S.nodes4, S.entries, S.aux, S.top_start, S.top_count, org, dir,
                                                        kEps, rem, stk, (uint32_t*)&P.counters[kCntOverflow], t4);
                if (COUNT) {
                    // the reference traversal's work on the binary tree, for
                    // the algorithmic-bytes figures (same answer)
                    trace_closest<SD, STRIDE, COUNT, true>(S.nodes, S.entries, S.aux, S.top_start, S.top_count, org,
                                                           dir, kEps, rem, stk, (uint32_t*)&P.counters[kCntOverflow], tc);
                    atomicAdd(&P.tcount[kTcAnyNode], (unsigned long long)t4.node);
                    atomicAdd(&P.tcount[kTcAnyTri], (unsigned long long)t4.tri);
                    atomicAdd(&P.tcount[kTcAnyOther], (unsigned long long)t4.other);
                }
            } else
                blocked = trace_closest<SD, STRIDE, COUNT, true, XF>(S.nodes, S.entries, S.aux, S.top_start,
                                                                     S.top_count, org, dir, kEps, rem, stk,
                                                                     (uint32_t*)&P.counters[kCntOverflow], tc, S.xf)
                              .entry >= 0;
            f3 c = blocked ? splat(0.0f) : f3of(P.sh_c[sl]);
            P.sh_r[sl] = make_float4(c.x, c.y, c.z, 0.0f);
            return;
        }
        f3 thr = splat(1.0f), sigma = splat(0.0f), entry = splat(0.0f);
        bool in_med = false;
        bool done = false;
        for (int it = 0; it < 8 && !done; it++) {
            HitOut h = trace_closest<SD, STRIDE, COUNT, false, XF>(S.nodes, S.entries, S.aux, S.top_start,
                                                                   S.top_count, org, dir, kEps, rem, stk,
                                                                   (uint32_t*)&P.counters[kCntOverflow], tc, S.xf);
            if (h.entry < 0) { done = true; break; }
            f3 lo, ld;
            Surf s = XF ? make_surf_w(S, h.entry, h.chain, org, dir, h.t, h.u, h.v, lo, ld)
                        : make_surf(S, h.entry, org, dir, h.t, h.u, h.v);
            const GMat g = S.mats[s.mat];
            if (S.block_spec && (g.m0.y & kMatSpecTransmissive)) { thr = splat(0.0f); done = true; break; }
            if (in_med) {
                float dd = length(s.p - entry);
                thr = mk(thr.x * expf(-sigma.x * dd), thr.y * expf(-sigma.y * dd), thr.z * expf(-sigma.z * dd));
                in_med = false;
                if (near_zero(thr)) { thr = splat(0.0f); done = true; break; }
            }
            f3 t = mat_shadow_tr(g, dir, s.n);
            if (near_zero(t)) { thr = splat(0.0f); done = true; break; }
            thr *= t;
            if (near_zero(thr)) { thr = splat(0.0f); done = true; break; }
            if (s.front) {
                f3 sg = mat_shadow_abs(g);
                if (!near_zero(sg)) { in_med = true; entry = s.p; sigma = sg; }
            }
            f3 on = dot(dir, s.n) >= 0.0f ? s.n : -s.n;
            org = offset_origin(s.p, on);
            rem -= h.t + 2.0f * kEps;
            if (rem <= kEps) { done = true; break; }
        }
        if (!done || near_zero(thr)) thr = splat(0.0f);
        f3 c = f3of(P.sh_c[sl]) * thr;
        P.sh_r[sl] = make_float4(c.x, c.y, c.z, 0.0f);
    });
    if (COUNT) flush_counts(P.tcount + 3, tc);
}

// Replay pass of the adjoint: the NEE gradient of every unoccluded shadow
// record of a textured hit goes to the texture (opaque shadows only, so a
// record's result is its full contribution or zero).  Same queue view as
// k_shadow.
__global__ void __launch_bounds__(kBlock) k_adj_nee(PathBufs P, int maxs, int tex_w, int tex_h) {
    const bool grouped = maxs <= kMaxSeg;
    int total = grouped ? 0 : P.counters[kCntShadow];
    if (grouped)
        for (int j = 0; j < maxs; j++) total += P.counters[kCntSeg + j];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= total) return;
    int64_t qi = i;
    if (grouped) {
        int j = 0, pre = 0;
        for (;;) {
            int c = P.counters[kCntSeg + j];
            if (i < pre + c || j + 1 >= maxs) break;
            pre += c;
            j++;
        }
        qi = (int64_t)j * P.n + (i - pre);
    }
    const int64_t sl = P.q_sh[qi];
    const float4 g1 = P.adj_g[2 * sl + 1];
    if (g1.w == 0.0f) return;
    const float4 r = P.sh_r[sl];
    if (r.x == 0.0f && r.y == 0.0f && r.z == 0.0f) return;
    const float4 g0 = P.adj_g[2 * sl];
    const float g[kNd] = {g0.x, g0.y, g0.z, g0.w, g1.x};
    tex_scatter(P.adj_dt, tex_w, tex_h, g1.y, g1.z, g);
}

// Ends the path in slot p: folds its last shadow results and stores the
// camera sample's radiance L = w0 * (Le0 + Ld0 + clampInd(Lind)), clamped
// (Renderer.cs:658-666, 1222-1223), and its AOV probe, under its sample id.
__device__ __forceinline__ void finish_path(const RParams& R, PathBufs& P, int p) {
    int32_t flags = __float_as_int(P.d[p].w);
    fold_shadows(P, R, p, flags);
    f3 a0 = f3of(P.a0[p]), ai = f3of(P.ai[p]), w0 = f3of(P.w0[p]);
    if (P.adj_pass) {
        // adjoint primal: the unweighted radiance sum the replay subtracts from
        if (P.adj_pass == 1) {
            f3 t = a0 + ai;
            P.res_l[P.sid[p]] = make_float4(t.x, t.y, t.z, 0.0f);
        }
        return;
    }
    f3 L = w0 * (sanitize(a0) + clamp_lum(ai, R.ind_max_rad, true));
    L = clamp_lum(L, R.max_rad, false);
    int sid = P.sid[p];
    float depth = -1.0f;
    if (R.aovs) {
        float4 av0 = P.av0[p], av1 = P.av1[p];
        int aovf = __float_as_int(av0.w);
        f3 alb, nrm;
        if (aovf & kAovPending) {
            alb = f3of(av0) * f3of(av1);
            nrm = f3of(P.av2[p]);
        } else {
            alb = f3of(P.av3[p]);
            nrm = f3of(P.av4[p]);
        }
        if (aovf & kAovHaveDepth) depth = av1.w;
        P.res_a[sid] = make_float4(alb.x, alb.y, alb.z, 0.0f);
        P.res_n[sid] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
    }
    P.res_l[sid] = make_float4(L.x, L.y, L.z, depth);
}

// Retires the paths that ended in this iteration and restarts their slots
// on the next window samples (allocated from the window's sample counter);
// once the window is exhausted the slots go idle.  Which slot traces which
// sample does not change any result.
template <class JOBS>
__global__ void __launch_bounds__(kBlock) k_retire(JOBS jobs, PathBufs P) {
    const int p = xcd_block() * kBlock + threadIdx.x;
    const bool ret = p < P.n && P.state[p] == kSlotRetired;
    if (ret) finish_path(jobs.params(jobs.job_of(P.job, p)), P, p);
    const int ws = P.counters[kCntWs];
    const unsigned long long lt = (1ull << (threadIdx.x & 63)) - 1ull;
    const unsigned long long rb = __ballot(ret);
    const int sid = P.counters[kCntSidBase] + P.wc[kWcRetire * P.nw + (p >> 6)] + __popcll(rb & lt);
    bool start = ret && sid < ws;
    if (ret) {
        int st = kSlotIdle;
        if (start) {
            int job, pix, pass;
            resolve_sample(P, sid, job, pix, pass);
            const RParams& R = jobs.params(job);
            init_path(jobs.scene(job), R, P, p, sid, job, pix, pass);
            st = R.max_depth > 0 ? kSlotActive : kSlotRetired;
        }
        P.state[p] = st;
    }
}

// Per-pixel film accumulation in sample order (Renderer.cs:668-672,
// CaptureAccumulator.Add, Renderer.cs:829-856): the window's samples of a
// pixel are summed pass by pass, even passes into half A, odd into half B.
__global__ void __launch_bounds__(kBlock) k_accum(RParams R, PathBufs P, FilmBufs F, int pass0, int passes, int off) {
    int pix = blockIdx.x * kBlock + threadIdx.x;
    if (pix >= R.npix) return;
    const int np = R.npix;
    f3 sum = mk(F.sum[pix], F.sum[np + pix], F.sum[2 * np + pix]);
    f3 sa = mk(F.sa[pix], F.sa[np + pix], F.sa[2 * np + pix]);
    f3 sb = mk(F.sb[pix], F.sb[np + pix], F.sb[2 * np + pix]);
    for (int sl = 0; sl < passes; sl++) {
        int s = R.soff + pass0 + sl;
        int64_t k = (int64_t)off + (int64_t)sl * np + pix;
        float4 l = P.res_l[k];
        f3 L = f3of(l);
        sum += L;
        bool isA = (s & 1) == 0;
        if (isA) sa += L;
        else sb += L;
        if (R.aovs) {
            f3 alb = f3of(P.res_a[k]), nrm = f3of(P.res_n[k]);
            float* A = isA ? F.alba : F.albb;
            float* N = isA ? F.nrma : F.nrmb;
            A[pix] += alb.x; A[np + pix] += alb.y; A[2 * np + pix] += alb.z;
            N[pix] += nrm.x; N[np + pix] += nrm.y; N[2 * np + pix] += nrm.z;
            if (l.w >= 0.0f) {
                if (isA) { F.depa[pix] += l.w; F.dca[pix]++; }
                else { F.depb[pix] += l.w; F.dcb[pix]++; }
            }
        }
    }
    F.sum[pix] = sum.x; F.sum[np + pix] = sum.y; F.sum[2 * np + pix] = sum.z;
    F.sa[pix] = sa.x; F.sa[np + pix] = sa.y; F.sa[2 * np + pix] = sa.z;
    F.sb[pix] = sb.x; F.sb[np + pix] = sb.y; F.sb[2 * np + pix] = sb.z;
}

__global__ void k_resolve(RParams R, FilmBufs F, float exposure) {
    int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= R.npix) return;
    const int np = R.npix;
    // even global indices form half A (RenderBuffers.cs:43-46)
    const int nA = (R.soff & 1) ? R.spp / 2 : (R.spp + 1) / 2, nB = R.spp - nA;
    for (int c = 0; c < 3; c++) {
        int k = c * np + pix;
        float lin = F.sum[k] / (float)R.spp;
        F.out_beauty[k] = lin;
        // a half without samples repeats the other (one sample per pixel)
        float mA = nA > 0 ? F.sa[k] / (float)nA : F.sb[k] / (float)nB;
        F.out_a[k] = mA;
        F.out_b[k] = nB > 0 ? F.sb[k] / (float)nB : mA;
        float x = fmaxf(lin * exposure, 0.0f);
        float a = x * (2.51f * x + 0.03f);
        float b = x * (2.43f * x + 0.59f) + 0.14f;
        F.out_disp[k] = powf(clampf(a / b, 0.0f, 1.0f), 1.0f / 2.2f);
        if (R.aovs) {
            float aa = nA > 0 ? F.alba[k] / (float)nA : F.albb[k] / (float)nB;
            float na = nA > 0 ? F.nrma[k] / (float)nA : F.nrmb[k] / (float)nB;
            F.out_alba[k] = aa;
            F.out_albb[k] = nB > 0 ? F.albb[k] / (float)nB : aa;
            F.out_nrma[k] = na;
            F.out_nrmb[k] = nB > 0 ? F.nrmb[k] / (float)nB : na;
        }
    }
    if (R.aovs) {
        float dA = F.dca[pix] > 0 ? F.depa[pix] / (float)F.dca[pix] : -1.0f;
        float dB = nB > 0 ? (F.dcb[pix] > 0 ? F.depb[pix] / (float)F.dcb[pix] : -1.0f) : dA;
        if (nA == 0) dA = dB;
        F.out_depa[pix] = dA;
        F.out_depb[pix] = dB;
    }
}

// rt_trace_primary / rt_trace_rays
template <int SD>
__global__ void __launch_bounds__(kBlock) k_trace_primary(DevScene S, RParams R, int n, int32_t* prim, float* to,
                                                          float* uo, float* vo) {
    __shared__ uint2 stk[SD * kBlock];
    int p = xcd_block() * kBlock + threadIdx.x;
    if (p >= n) return;
    int i, j, s;
    sample_of(R, p, i, j, s);
    Rng rng{(uint32_t)s, pixel_seed(i, j, R.seed_xor), 0u, false};
    f3 o, d;
    camera_ray(S, R, i, j, rng, o, d);
    HitOut h;
    TravCount tc{0u, 0u, 0u};
    for (int skip = 0;;) {
        if (S.xf)
            h = trace_closest<SD, kBlock, false, false, true>(S.nodes, S.entries, S.aux, S.top_start, S.top_count, o,
                                                              d, kEps, kInfinity, &stk[threadIdx.x],
                                                              (uint32_t*)(prim + n), tc, S.xf);
        else
            h = trace_closest<SD, kBlock, false>(S.nodes, S.entries, S.aux, S.top_start, S.top_count, o, d, kEps,
                                                 kInfinity, &stk[threadIdx.x], (uint32_t*)(prim + n), tc);
        if (h.entry < 0) break;
        if (!hit_cam_invisible(S, h.entry, h.chain)) break;
        if (S.xf) {
            f3 lo, ld;
            o = offset_origin(make_surf_w(S, h.entry, h.chain, o, d, h.t, h.u, h.v, lo, ld).p, d);
        } else o = offset_origin(o + h.t * d, d);
        if (++skip >= 8) break;
    }
    float hu, hv;
    if (S.xf && h.entry >= 0) chain_ray(S.entries, S.xf, h.chain, o, d);
    report_uv(S.entries, h, o, d, hu, hv);
    prim[p] = h.entry < 0 ? -1 : __float_as_int(S.entries[h.entry].b.w);
    to[p] = h.entry < 0 ? 0.0f : h.t;
    uo[p] = h.entry < 0 ? 0.0f : hu;
    vo[p] = h.entry < 0 ? 0.0f : hv;
}

template <int SD>
__global__ void __launch_bounds__(kBlock) k_trace_rays(DevScene S, const float* __restrict__ rays, int n, int32_t* prim,
                                                       float* to, float* uo, float* vo) {
    __shared__ uint2 stk[SD * kBlock];
    int p = xcd_block() * kBlock + threadIdx.x;
    if (p >= n) return;
    const float* r = rays + 6 * (int64_t)p;
    TravCount tc{0u, 0u, 0u};
    f3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
    HitOut h;
    if (S.xf) {
        h = trace_closest<SD, kBlock, false, false, true>(S.nodes, S.entries, S.aux, S.top_start, S.top_count, o, d,
                                                          kEps, kInfinity, &stk[threadIdx.x], (uint32_t*)(prim + n), tc,
                                                          S.xf);
        if (h.entry >= 0) chain_ray(S.entries, S.xf, h.chain, o, d);
    } else
        h = trace_closest<SD, kBlock, false>(S.nodes, S.entries, S.aux, S.top_start, S.top_count, o, d, kEps,
                                             kInfinity, &stk[threadIdx.x], (uint32_t*)(prim + n), tc);
    float hu, hv;
    report_uv(S.entries, h, o, d, hu, hv);
    prim[p] = h.entry < 0 ? -1 : __float_as_int(S.entries[h.entry].b.w);
    to[p] = h.entry < 0 ? 0.0f : h.t;
    uo[p] = h.entry < 0 ? 0.0f : hu;
    vo[p] = h.entry < 0 ? 0.0f : hv;
}

// Occlusion queries (rt_occluded): the any-hit traversal the shadow kernel
// runs for opaque scenes, on caller-given segments.
template <int SD>
__global__ void __launch_bounds__(kBlock) k_occluded(DevScene S, const float* __restrict__ rays, int n, int32_t* occ) {
    __shared__ uint2 stk[SD * kBlock];
    int p = xcd_block() * kBlock + threadIdx.x;
    if (p >= n) return;
    const float* r = rays + 8 * (int64_t)p;
    f3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
    bool b;
    if (S.nodes4) {
        TravCount tc{0u, 0u, 0u};
        b = trace_any4<SD, kBlock>(S.nodes4, S.entries, S.aux, S.top_start, S.top_count, o, d, r[6], r[7],
                                   &stk[threadIdx.x], (uint32_t*)(occ + n), tc);
    } else if (S.xf) {
        TravCount tc{0u, 0u, 0u};
        b = trace_closest<SD, kBlock, false, true, true>(S.nodes, S.entries, S.aux, S.top_start, S.top_count, o, d,
                                                         r[6], r[7], &stk[threadIdx.x], (uint32_t*)(occ + n), tc, S.xf)
                .entry >= 0;
    } else {
        TravCount tc{0u, 0u, 0u};
        b = trace_closest<SD, kBlock, false, true>(S.nodes, S.entries, S.aux, S.top_start, S.top_count, o, d, r[6], r[7],
                                                   &stk[threadIdx.x], (uint32_t*)(occ + n), tc)
                .entry >= 0;
    }
    occ[p] = b ? 1 : 0;
}

// Staging of rt_occluded_noalloc: the segments (32 B each) at offset 0, the
// results (4 B each) at kNoAllocMaxRays * 32, then the traversal-overflow
// flag after the last result.  One copy per GPU, alive until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns for the device (rt_ctx::scratch).
constexpr int64_t kNoAllocMaxRays = 1 << 20;
constexpr size_t kNoAllocResults = (size_t)kNoAllocMaxRays * 32;
constexpr size_t kNoAllocBytes = kNoAllocResults + 4 * ((size_t)kNoAllocMaxRays + 64);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

__global__ void __launch_bounds__(kBlock) k_zero_i32(int32_t* x, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = 0;
}

}  // namespace rt

// ═════════════════════════════════════════════════════════════════════════
// Host side: context, scene upload, render loop, C ABI
// ═════════════════════════════════════════════════════════════════════════
using namespace rt;

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // cached path/film buffers (grow-only)
    void* path_mem = nullptr;
    size_t path_bytes = 0;
    void* film_mem = nullptr;
    size_t film_bytes = 0;
    hipEvent_t ev[8] = {};
    bool priv_stack = true;    // traversal stack in private scratch (false: LDS)
    int shade_wpe = 3;         // minimum waves per SIMD the NEE shading kernel is compiled for
    int bounce_wpe = 3;        // ... and the bounce shading kernel
    void* ph_mem = nullptr;    // photon pass buffers (grow-only)
    size_t ph_bytes = 0;
    hipEvent_t iev[2][kIterBatch][5] = {};   // per-iteration kernel timing, two batches
    hipEvent_t chk[2] = {};                   // batch completion
    int32_t* pinned = nullptr;                // host-visible pending counts (hipHostMalloc)
    void* parts_mem = nullptr;                // window part tables (grow-only)
    size_t parts_bytes = 0;
    void* jobs_mem = nullptr;                 // batch job table (grow-only)
    size_t jobs_bytes = 0;
    bool use_graph = false;                   // replay captured bounce iterations (hipGraph)
    void* adj_mem = nullptr;                  // adjoint buffers (grow-only)
    size_t adj_bytes = 0;
    void* io_mem = nullptr;                   // film-file encoder buffers (grow-only)
    size_t io_bytes = 0;
    void* dn_mem = nullptr;                   // denoiser buffers (grow-only)
    size_t dn_bytes = 0;
    char* scratch = nullptr;                  // this device's copy of rt::scratch (hipGetSymbolAddress)
};

struct rt_scene {
    rt_ctx* ctx = nullptr;
    HostScene hs;
    GNode* d_nodes = nullptr;
    GNode4* d_nodes4 = nullptr;
    GEntry* d_entries = nullptr;
    float4* d_aux = nullptr;
    GMat* d_mats = nullptr;
    GLight* d_lights = nullptr;
    float* d_lut = nullptr;
    float* d_ldist = nullptr;   // cdf[n], pdf[n]
    rt_material* d_raw = nullptr;
    float4* d_tuv = nullptr;
    float4* d_smooth = nullptr;
    float* d_tex = nullptr;
    GImage* d_images = nullptr;
    GLevel* d_levels = nullptr;
    float4* d_imtex = nullptr;
    float4* d_xf = nullptr;
    int stack_depth = 0;
    int64_t device_bytes = 0;
};

#define HIP_TRY(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return e_ == hipErrorOutOfMemory ? RT_ERR_OUT_OF_MEMORY : RT_ERR_DEVICE; \
        }                                                                          \
    } while (0)

static int fail(rt_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

// Context services for the film-file encoders (film_io.hip).
namespace rt {
hipStream_t ctx_stream(rt_ctx* ctx) { return ctx->stream; }
int ctx_device(rt_ctx* ctx) { return ctx->device; }
int ctx_fail(rt_ctx* ctx, int code, const std::string& msg) { return fail(ctx, code, msg); }
// Grow-only device buffer of at least `bytes` (contents not kept across growth).
static int grow_buffer(rt_ctx* ctx, void** mem, size_t* have, size_t bytes, void** out) {
    if (bytes > *have) {
        if (*mem) {
            hipStreamSynchronize(ctx->stream);
            hipFree(*mem);
        }
        *mem = nullptr;
        *have = 0;
        hipError_t e = hipMalloc(mem, bytes);
        if (e != hipSuccess) return fail(ctx, RT_ERR_OUT_OF_MEMORY, hipGetErrorString(e));
        *have = bytes;
    }
    *out = *mem;
    return RT_OK;
}
int ctx_io_buffer(rt_ctx* ctx, size_t bytes, void** out) { return grow_buffer(ctx, &ctx->io_mem, &ctx->io_bytes, bytes, out); }
int ctx_dn_buffer(rt_ctx* ctx, size_t bytes, void** out) { return grow_buffer(ctx, &ctx->dn_mem, &ctx->dn_bytes, bytes, out); }
}  // namespace rt

template <typename T>
static int upload(rt_ctx* ctx, T** dst, const T* src, size_t n, int64_t& bytes) {
    size_t b = std::max<size_t>(n, 1) * sizeof(T);
    HIP_TRY(hipMalloc((void**)dst, b));
    if (n) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    bytes += (int64_t)b;
    return RT_OK;
}

static DevScene dev_view(const rt_scene* sc) {
    const HostScene& h = sc->hs;
    DevScene S;
    S.nodes = sc->d_nodes;
    S.nodes4 = h.nodes4.empty() ? nullptr : sc->d_nodes4;
    S.entries = sc->d_entries;
    S.aux = sc->d_aux;
    S.mats = sc->d_mats;
    S.lights = sc->d_lights;
    S.lut_e = sc->d_lut;
    S.light_cdf = sc->d_ldist;
    S.light_pdf = sc->d_ldist + std::max<size_t>(h.lights.size(), 1);
    S.top_start = h.top_start;
    S.top_count = h.top_count;
    S.num_lights = (int)h.lights.size();
    S.medium = h.medium ? 1 : 0;
    S.env_light = h.env_light;
    S.thin_film = h.thin_film ? 1 : 0;
    S.sky = h.sky;
    S.opaque_shadows = h.opaque_shadows ? 1 : 0;
    S.cam_o = h.cam_origin;
    S.cam_llc = h.cam_llc;
    S.cam_h = h.cam_horiz;
    S.cam_v = h.cam_vert;
    S.cam_u = h.cam_u;
    S.cam_vv = h.cam_v;
    S.lens_radius = h.lens_radius;
    S.sig_t = h.sig_t;
    S.sig_s = h.sig_s;
    S.caustics = 0;
    S.block_spec = 0;
    S.raw = sc->d_raw;
    S.tuv = h.tex_uv.empty() ? nullptr : sc->d_tuv;
    S.smooth = h.smooth.empty() ? nullptr : sc->d_smooth;
    S.images = h.images.empty() ? nullptr : sc->d_images;
    S.img_levels = sc->d_levels;
    S.img_texels = sc->d_imtex;
    S.tex = h.textured ? sc->d_tex : nullptr;
    S.tex_w = h.tex_w;
    S.tex_h = h.tex_h;
    // non-null for every scene with wrapper entries (the table is allocated
    // even when no wrapper transforms: Instance / camera-invisible only)
    S.xf = h.hier ? sc->d_xf : nullptr;
    return S;
}

// Launches a traversal kernel with the LDS stack depth the scene needs.
// Traversal kernels: job source J, stack depth, traversal counting, stack
// placement (batches always use the private stack).
#define DISPATCH_SD_C(J, sd, cnt, priv, KERNEL, grid, ...)                                            \
    do {                                                                                               \
        if constexpr (std::is_same<J, OneJob>::value) {                                                \
            if (cnt) {                                                                                 \
                if (priv) DISPATCH_SD_T(J, sd, KERNEL, true, true, grid, __VA_ARGS__);                 \
                else DISPATCH_SD_T(J, sd, KERNEL, true, false, grid, __VA_ARGS__);                     \
            } else {                                                                                   \
                if (priv) DISPATCH_SD_T(J, sd, KERNEL, false, true, grid, __VA_ARGS__);                \
                else DISPATCH_SD_T(J, sd, KERNEL, false, false, grid, __VA_ARGS__);                    \
            }                                                                                          \
        } else {                                                                                       \
            if (cnt) DISPATCH_SD_T(J, sd, KERNEL, true, true, grid, __VA_ARGS__);                      \
            else DISPATCH_SD_T(J, sd, KERNEL, false, true, grid, __VA_ARGS__);                         \
        }                                                                                              \
    } while (0)
#define DISPATCH_SD_T(J, sd, KERNEL, C_, P_, grid, ...)                                                        \
    do {                                                                                                       \
        if ((sd) <= 24) hipLaunchKernelGGL((KERNEL<J, 24, C_, P_>), grid, dim3(kBlock), 0, __VA_ARGS__);       \
        else if ((sd) <= 32) hipLaunchKernelGGL((KERNEL<J, 32, C_, P_>), grid, dim3(kBlock), 0, __VA_ARGS__);  \
        else hipLaunchKernelGGL((KERNEL<J, 64, C_, P_>), grid, dim3(kBlock), 0, __VA_ARGS__);                  \
    } while (0)
// scenes with wrapper entries: the XF traversal, private-scratch stacks
#define DISPATCH_SD_X(J, sd, cnt, KERNEL, grid, ...)                                   \
    do {                                                                                \
        if (cnt) DISPATCH_SD_TX(J, sd, KERNEL, true, grid, __VA_ARGS__);                \
        else DISPATCH_SD_TX(J, sd, KERNEL, false, grid, __VA_ARGS__);                   \
    } while (0)
#define DISPATCH_SD_TX(J, sd, KERNEL, C_, grid, ...)                                                                 \
    do {                                                                                                             \
        if ((sd) <= 24) hipLaunchKernelGGL((KERNEL<J, 24, C_, true, true>), grid, dim3(kBlock), 0, __VA_ARGS__);      \
        else if ((sd) <= 32) hipLaunchKernelGGL((KERNEL<J, 32, C_, true, true>), grid, dim3(kBlock), 0, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERNEL<J, 64, C_, true, true>), grid, dim3(kBlock), 0, __VA_ARGS__);                 \
    } while (0)
#define DISPATCH_SD(sd, KERNEL, grid, ...)                                                       \
    do {                                                                                         \
        if ((sd) <= 24) hipLaunchKernelGGL(KERNEL<24>, grid, dim3(kBlock), 0, __VA_ARGS__);      \
        else if ((sd) <= 32) hipLaunchKernelGGL(KERNEL<32>, grid, dim3(kBlock), 0, __VA_ARGS__); \
        else if ((sd) <= 48) hipLaunchKernelGGL(KERNEL<48>, grid, dim3(kBlock), 0, __VA_ARGS__); \
        else hipLaunchKernelGGL(KERNEL<64>, grid, dim3(kBlock), 0, __VA_ARGS__);                 \
    } while (0)

// ── bounce loop over sample windows (single render and batch) ─────────────
struct WindowPlan {
    std::vector<PartDev> parts;            // parts of all windows, in window order
    std::vector<int> first, count, ws;     // per window: first part, part count, samples
    void add_window(const std::vector<PartDev>& w, int samples) {
        first.push_back((int)parts.size());
        count.push_back((int)w.size());
        ws.push_back(samples);
        parts.insert(parts.end(), w.begin(), w.end());
    }
};

struct LoopTiming {
    double trace = 0, shade = 0, shadow = 0;
    int64_t launches = 0;   // bounce iterations enqueued (one k_trace and one k_shadow each)
};

// Runs every window through the device-driven bounce loop.  Iterations are
// enqueued in batches of kIterBatch without host waits (every grid is sized
// to the pool, every count lives on the device); after each batch the
// pending-path count goes to pinned memory, the host waits for the previous
// batch's count (one batch stays queued) and ends the window once it is
// zero.  accum(part) enqueues the film accumulation of one window part.
template <class JOBS, class ACCUM>
static int run_windows(rt_ctx* ctx, const JOBS& jobs, PathBufs P, int maxs, bool traced, bool counting,
                       int stack_depth, const DevScene* gather_scene, const DevPhotons* M, const WindowPlan& plan,
                       ACCUM accum, LoopTiming& tm) {
    const int64_t N = P.n;
    bool thin_film, hier, rnd = false, all_lights;
    if constexpr (std::is_same<JOBS, OneJob>::value) {
        rnd = jobs.R.prng != 0;
        all_lights = jobs.R.picked == 0;
        thin_film = jobs.S.thin_film != 0;
        hier = jobs.S.xf != nullptr;
    } else {
        thin_film = jobs.thin_film != 0;
        hier = jobs.hier != 0;
        all_lights = jobs.all_lights != 0;
    }
    const int pool_blocks = (int)((N + kBlock - 1) / kBlock);
    const int shadow_blocks = (int)(((int64_t)N * maxs + kBlock - 1) / kBlock);
    size_t pbytes = sizeof(PartDev) * std::max<size_t>(plan.parts.size(), 1);
    if (pbytes > ctx->parts_bytes) {
        if (ctx->parts_mem) hipFree(ctx->parts_mem);
        ctx->parts_mem = nullptr;
        ctx->parts_bytes = 0;
        HIP_TRY(hipMalloc(&ctx->parts_mem, pbytes));
        ctx->parts_bytes = pbytes;
    }
    PartDev* d_parts = (PartDev*)ctx->parts_mem;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(d_parts, plan.parts.data(), sizeof(PartDev) * plan.parts.size(), hipMemcpyHostToDevice));
    auto enqueue_iteration = [&](const PathBufs& Pw, hipEvent_t* e) -> int {
        hipLaunchKernelGGL(k_iter_begin, dim3(1), dim3(kBlock), 0, ctx->stream, Pw, traced ? 1 : 0);
        if (traced) {
            dim3 grid(pool_blocks);
            if (e) HIP_TRY(hipEventRecord(e[0], ctx->stream));
            if (hier) DISPATCH_SD_X(JOBS, stack_depth, counting, k_trace, grid, ctx->stream, jobs, Pw);
            else DISPATCH_SD_C(JOBS, stack_depth, counting, ctx->priv_stack, k_trace, grid, ctx->stream, jobs, Pw);
            if (e) HIP_TRY(hipEventRecord(e[1], ctx->stream));
            if (Pw.fp) DISPATCH_SD_C(JOBS, stack_depth, false, ctx->priv_stack, k_fp, grid, ctx->stream, jobs, Pw);
            bool adj_replay = false;
            if constexpr (std::is_same<JOBS, OneJob>::value) adj_replay = Pw.adj_pass == 2;
            if (adj_replay) {
                if constexpr (std::is_same<JOBS, OneJob>::value)
                    hipLaunchKernelGGL((k_shade<JOBS, 1, true>), grid, dim3(kBlock), 0, ctx->stream, jobs, Pw, maxs);
            } else {
                // split shade: NEE phase, then the bounce phase (with the
                // thin-film Fresnel only when a scene of the launch has one);
                // single renders with the PRNG sampler use their own pair
                if (rnd) {
                    if constexpr (std::is_same<JOBS, OneJob>::value) {
                        if (all_lights)
                            hipLaunchKernelGGL((k_shade<JOBS, 2, false, kPhNee, true, true, true>), grid, dim3(kBlock), 0,
                                               ctx->stream, jobs, Pw, maxs);
                        else
                            hipLaunchKernelGGL((k_shade<JOBS, 2, false, kPhNee, true, true, false>), grid, dim3(kBlock),
                                               0, ctx->stream, jobs, Pw, maxs);
                        if (thin_film)
                            hipLaunchKernelGGL((k_shade<JOBS, 2, false, kPhBounce, true, true>), grid, dim3(kBlock), 0,
                                               ctx->stream, jobs, Pw, maxs);
                        else
                            hipLaunchKernelGGL((k_shade<JOBS, 2, false, kPhBounce, false, true>), grid, dim3(kBlock),
                                               0, ctx->stream, jobs, Pw, maxs);
                    }
                } else {
                    if (ctx->shade_wpe <= 2) {
                        if (all_lights)
                            hipLaunchKernelGGL((k_shade<JOBS, 2, false, kPhNee, true, false, true>), grid, dim3(kBlock),
                                               0, ctx->stream, jobs, Pw, maxs);
                        else
                            hipLaunchKernelGGL((k_shade<JOBS, 2, false, kPhNee>), grid, dim3(kBlock), 0, ctx->stream,
                                               jobs, Pw, maxs);
                    } else {
                        if (all_lights)
                            hipLaunchKernelGGL((k_shade<JOBS, 3, false, kPhNee, true, false, true>), grid, dim3(kBlock),
                                               0, ctx->stream, jobs, Pw, maxs);
                        else
                            hipLaunchKernelGGL((k_shade<JOBS, 3, false, kPhNee>), grid, dim3(kBlock), 0, ctx->stream,
                                               jobs, Pw, maxs);
                    }
                    if (thin_film) {
                        if (ctx->bounce_wpe <= 2)
                            hipLaunchKernelGGL((k_shade<JOBS, 2, false, kPhBounce, true>), grid, dim3(kBlock), 0,
                                               ctx->stream, jobs, Pw, maxs);
                        else
                            hipLaunchKernelGGL((k_shade<JOBS, 3, false, kPhBounce, true>), grid, dim3(kBlock), 0,
                                               ctx->stream, jobs, Pw, maxs);
                    } else {
                        if (ctx->bounce_wpe <= 2)
                            hipLaunchKernelGGL((k_shade<JOBS, 2, false, kPhBounce, false>), grid, dim3(kBlock), 0,
                                               ctx->stream, jobs, Pw, maxs);
                        else
                            hipLaunchKernelGGL((k_shade<JOBS, 3, false, kPhBounce, false>), grid, dim3(kBlock), 0,
                                               ctx->stream, jobs, Pw, maxs);
                    }
                }
            }
            if (e) HIP_TRY(hipEventRecord(e[2], ctx->stream));
            hipLaunchKernelGGL(k_wave_scan, dim3(kWcSeg + (maxs <= kMaxSeg ? maxs : 0)), dim3(1024), 0, ctx->stream, Pw);
            hipLaunchKernelGGL(k_compact, grid, dim3(kBlock), 0, ctx->stream, Pw, maxs);
            if (gather_scene) hipLaunchKernelGGL(k_gather, grid, dim3(kBlock), 0, ctx->stream, *gather_scene, Pw, *M);
            if (e) HIP_TRY(hipEventRecord(e[3], ctx->stream));
            if (hier)
                DISPATCH_SD_X(JOBS, stack_depth, counting, k_shadow, dim3(shadow_blocks), ctx->stream, jobs, Pw, maxs);
            else
                DISPATCH_SD_C(JOBS, stack_depth, counting, ctx->priv_stack, k_shadow, dim3(shadow_blocks), ctx->stream,
                              jobs, Pw, maxs);
            if (e) HIP_TRY(hipEventRecord(e[4], ctx->stream));
            if constexpr (std::is_same<JOBS, OneJob>::value)
                if (adj_replay)
                    hipLaunchKernelGGL(k_adj_nee, dim3(shadow_blocks), dim3(kBlock), 0, ctx->stream, Pw, maxs, jobs.S.tex_w,
                                       jobs.S.tex_h);
        } else {
            const int narr = kWcSeg + (maxs <= kMaxSeg ? maxs : 0);
            hipLaunchKernelGGL(k_count_retired, dim3(pool_blocks), dim3(kBlock), 0, ctx->stream, Pw, narr);
            hipLaunchKernelGGL(k_wave_scan, dim3(narr), dim3(1024), 0, ctx->stream, Pw);
        }
        hipLaunchKernelGGL((k_retire<JOBS>), dim3(pool_blocks), dim3(kBlock), 0, ctx->stream, jobs, Pw);
        HIP_TRY(hipGetLastError());
        return RT_OK;
    };
    auto collect = [&](int set, int n) {
        if (!traced) return;
        for (int k = 0; k < n; k++) {
            hipEvent_t* e = ctx->iev[set][k];
            float ms = 0;
            hipEventElapsedTime(&ms, e[0], e[1]);
            tm.trace += ms;
            hipEventElapsedTime(&ms, e[1], e[2]);
            tm.shade += ms;
            hipEventElapsedTime(&ms, e[3], e[4]);
            tm.shadow += ms;
        }
    };
    for (size_t w = 0; w < plan.ws.size(); w++) {
        PathBufs Pw = P;
        Pw.parts = d_parts + plan.first[w];
        Pw.nparts = plan.count[w];
        const int ws = plan.ws[w];
        const int n0 = (int)std::min<int64_t>(N, ws);
        HIP_TRY(hipMemsetAsync(P.counters, 0, 4 * kCntWindow, ctx->stream));
        hipLaunchKernelGGL((k_raygen<JOBS>), dim3(pool_blocks), dim3(kBlock), 0, ctx->stream, jobs, Pw, n0, ws);
        // graph mode: the kIterBatch iterations of a batch are captured once
        // per window (no memset nodes: k_iter_begin clears the counters) and
        // replayed; per-kernel timings are not recorded
        hipGraphExec_t gexec = nullptr;
        if (ctx->use_graph) {
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            const bool saved = traced;
            int rc = RT_OK;
            for (int k = 0; k < kIterBatch && rc == RT_OK; k++) rc = enqueue_iteration(Pw, nullptr);
            hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
            (void)saved;
            if (rc) return rc;
            if (ce != hipSuccess) return fail(ctx, RT_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
            ce = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
            if (ce != hipSuccess) return fail(ctx, RT_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ce));
        }
        struct GraphGuard {
            hipGraphExec_t g;
            ~GraphGuard() { if (g) hipGraphExecDestroy(g); }
        } guard{gexec};
        // Once every sample of the window has been issued (known one batch
        // late), batches shrink to one iteration, so the window ends at most
        // one empty iteration after its last path instead of a whole batch.
        int nb[2] = {0, 0};   // iterations with timing events in each in-flight batch
        bool drain = false;
        for (int b = 0;; b++) {
            const int set = b & 1;
            const int iters = drain ? 1 : kIterBatch;
            if (traced) tm.launches += iters;
            if (gexec && !drain) {
                HIP_TRY(hipGraphLaunch(gexec, ctx->stream));
                nb[set] = 0;
            } else {
                for (int k = 0; k < iters; k++) {
                    int rc = enqueue_iteration(Pw, gexec ? nullptr : ctx->iev[set][k]);
                    if (rc) return rc;
                }
                nb[set] = gexec ? 0 : iters;
            }
            hipLaunchKernelGGL(k_batch_end, dim3(1), dim3(kBlock), 0, ctx->stream, Pw);
            HIP_TRY(hipMemcpyAsync(&ctx->pinned[4 * set], &P.counters[kCntPend], 12, hipMemcpyDeviceToHost,
                                   ctx->stream));
            HIP_TRY(hipEventRecord(ctx->chk[set], ctx->stream));
            if (b == 0) continue;
            const int prev = set ^ 1;
            HIP_TRY(hipEventSynchronize(ctx->chk[prev]));
            collect(prev, nb[prev]);
            if (ctx->pinned[4 * prev + 1]) return fail(ctx, RT_ERR_INTERNAL, "traversal stack overflow");
            if (ctx->pinned[4 * prev] == 0) {
                HIP_TRY(hipEventSynchronize(ctx->chk[set]));
                collect(set, nb[set]);
                break;
            }
            drain = ctx->pinned[4 * prev + 2] != 0;
        }
        for (int k = 0; k < plan.count[w]; k++) accum(plan.parts[plan.first[w] + k]);
        HIP_TRY(hipGetLastError());
    }
    return RT_OK;
}

// ── caustic photon pass (host orchestration) ───────────────────────────────
struct PhotonPass {
    DevPhotons M{};
    int64_t emitted = 0;
    int n = 0, nb = 0;
    const uint32_t* key0 = nullptr;   // bucket of each photon, photon order
};

// Exclusive scan of n ints (out may alias nothing); bsum needs
// ceil(n / kScanTile) + 1 ints; the total lands in bsum[tiles].
static void dev_scan(hipStream_t st, const int32_t* in, int32_t* out, int n, int32_t* bsum) {
    int tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 0) return;
    hipLaunchKernelGGL(k_scan_tile, dim3(tiles), dim3(kScanBlock), 0, st, in, out, n, bsum);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(1024), 0, st, bsum, tiles);
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kScanBlock), 0, st, out, n, bsum);
}

// CausticPhotonTracer.Build + PhotonMap construction on the device
// (Renderer.cs:385-406).  Per-light budgets are set on the host exactly as
// the reference does (luminous power shares, banker's rounding, >= 1 each).
static int photon_pass(rt_ctx* ctx, const rt_scene* sc, const DevScene& S, int budget, float cell_size,
                       PhotonPass& out) {
    const HostScene& h = sc->hs;
    std::vector<PhotonSrc> src;
    std::vector<float> lum;
    float total = 0.0f;
    for (size_t i = 0; i < h.lights.size(); i++) {
        // CausticPhotonTracer.CanEmit: an environment light emits no photons
        if (h.lights[i].l0.x == RT_LIGHT_ENVIRONMENT) continue;
        float p = std::max(0.0f, h.light_power[i]);
        if (p <= 0.0f) continue;
        src.push_back(PhotonSrc{(int)i, 0, 0, 0});
        lum.push_back(p);
        total += p;
    }
    out = PhotonPass{};
    if (budget <= 0 || src.empty() || total <= 0.0f) return RT_OK;
    int64_t ntot64 = 0;
    for (size_t i = 0; i < src.size(); i++) {
        int n = (int)nearbyintf((float)budget * (lum[i] / total));
        src[i].count = std::max(1, n);
        src[i].offset = (int)ntot64;
        ntot64 += src[i].count;
    }
    if (ntot64 > (1ll << 30)) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "photon budget too large");
    const int ntot = (int)ntot64;
    out.emitted = ntot;
    int nbmax = 1;
    while (nbmax < ntot) nbmax <<= 1;
    const int tiles_max = (std::max(ntot, nbmax) + 1023) / 1024;
    const size_t hist_n = (size_t)256 * ((ntot + 1023) / 1024);
    size_t need = 0;
    auto lay = [&](size_t bytes) { size_t o = need; need += (bytes + 255) & ~(size_t)255; return o; };
    size_t o_src = lay(sizeof(PhotonSrc) * src.size());
    size_t o_sp = lay(16ull * ntot), o_si = lay(16ull * ntot), o_sw = lay(16ull * ntot);
    size_t o_flag = lay(4ull * ntot), o_posof = lay(4ull * ntot);
    size_t o_cp = lay(16ull * ntot), o_ci = lay(16ull * ntot), o_cw = lay(16ull * ntot);
    size_t o_k0 = lay(4ull * ntot), o_ka = lay(4ull * ntot), o_kb = lay(4ull * ntot);
    size_t o_v0 = lay(4ull * ntot), o_va = lay(4ull * ntot), o_vb = lay(4ull * ntot);
    size_t o_cnt = lay(4ull * nbmax), o_start = lay(4ull * (nbmax + 1));
    size_t o_hist = lay(4 * hist_n), o_hscan = lay(4 * hist_n);
    size_t o_bsum = lay(4ull * ((std::max(hist_n, (size_t)std::max(ntot, nbmax)) + kScanTile - 1) / kScanTile + 2));
    size_t o_part = lay(16 * 256), o_org = lay(16), o_ovf = lay(16);
    (void)tiles_max;
    if (need > ctx->ph_bytes) {
        if (ctx->ph_mem) hipFree(ctx->ph_mem);
        ctx->ph_mem = nullptr;
        ctx->ph_bytes = 0;
        HIP_TRY(hipMalloc(&ctx->ph_mem, need));
        ctx->ph_bytes = need;
    }
    char* m = (char*)ctx->ph_mem;
    auto at = [&](size_t o) { return (void*)(m + o); };
    PhotonSrc* d_src = (PhotonSrc*)at(o_src);
    float4 *sp = (float4*)at(o_sp), *si = (float4*)at(o_si), *sw = (float4*)at(o_sw);
    int32_t *flag = (int32_t*)at(o_flag), *posof = (int32_t*)at(o_posof);
    float4 *cp = (float4*)at(o_cp), *ci = (float4*)at(o_ci), *cw = (float4*)at(o_cw);
    uint32_t *k0 = (uint32_t*)at(o_k0), *ka = (uint32_t*)at(o_ka), *kb = (uint32_t*)at(o_kb);
    int32_t *v0 = (int32_t*)at(o_v0), *va = (int32_t*)at(o_va), *vb = (int32_t*)at(o_vb);
    int32_t *cnt = (int32_t*)at(o_cnt), *start = (int32_t*)at(o_start);
    int32_t *hist = (int32_t*)at(o_hist), *hscan = (int32_t*)at(o_hscan), *bsum = (int32_t*)at(o_bsum);
    float4 *part = (float4*)at(o_part), *org = (float4*)at(o_org);
    int32_t* ovf = (int32_t*)at(o_ovf);
    hipStream_t st = ctx->stream;

    HIP_TRY(hipMemcpyAsync(d_src, src.data(), sizeof(PhotonSrc) * src.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(ovf, 0, 4, st));
    dim3 g((ntot + kBlock - 1) / kBlock);
    DISPATCH_SD(sc->stack_depth, k_photon_trace, g, st, S, d_src, (int)src.size(), ntot, h.bounds.mn, h.bounds.mx, sp,
                si, sw, flag, (uint32_t*)ovf);
    HIP_TRY(hipGetLastError());
    dev_scan(st, flag, posof, ntot, bsum);
    int32_t hv[2] = {0, 0};
    const int tiles = (ntot + kScanTile - 1) / kScanTile;
    HIP_TRY(hipMemcpyAsync(&hv[0], bsum + tiles, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&hv[1], ovf, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (hv[1]) return fail(ctx, RT_ERR_INTERNAL, "traversal stack overflow in the photon pass");
    const int n = hv[0];
    if (n == 0) return RT_OK;
    hipLaunchKernelGGL(k_photon_compact, g, dim3(kBlock), 0, st, ntot, flag, posof, sp, si, sw, cp, ci, cw);
    // grid: origin = min position, cell, power-of-two bucket table
    int nb = 1;
    while (nb < n) nb <<= 1;
    const float cell = std::max(cell_size, 1e-6f);
    GridParams G{cell, 1.0f / cell, (uint32_t)(nb - 1), n};
    int pb = std::min(256, (n + 255) / 256);
    hipLaunchKernelGGL(k_min3, dim3(pb), dim3(256), 0, st, cp, n, part);
    hipLaunchKernelGGL(k_min3, dim3(1), dim3(256), 0, st, part, pb, org);
    HIP_TRY(hipMemsetAsync(cnt, 0, 4ull * nb, st));
    hipLaunchKernelGGL(k_photon_keys, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, cp, org, G, k0, v0, cnt);
    // stable LSD radix sort of (bucket, photon index)
    int bits = 0;
    while ((1 << bits) < nb) bits++;
    const uint32_t* kin = k0;
    const int32_t* vin = v0;
    const int rtiles = (n + 1023) / 1024;
    for (int shift = 0, pass = 0; shift < bits; shift += 8, pass++) {
        uint32_t* kout = (pass & 1) ? kb : ka;
        int32_t* vout = (pass & 1) ? vb : va;
        hipLaunchKernelGGL(k_rs_hist, dim3(rtiles), dim3(256), 0, st, kin, n, shift, hist);
        dev_scan(st, hist, hscan, 256 * rtiles, bsum);
        hipLaunchKernelGGL(k_rs_scatter, dim3(rtiles), dim3(256), 0, st, kin, vin, kout, vout, n, shift, hscan);
        kin = kout;
        vin = vout;
    }
    // photon records in bucket order, into the (now free) per-slot planes
    hipLaunchKernelGGL(k_photon_order, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, n, vin, cp, ci, cw, sp, si,
                       sw);
    // CSR bucket starts
    dev_scan(st, cnt, start, nb, bsum);
    HIP_TRY(hipMemcpyAsync(start + nb, &hv[0], 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    out.n = n;
    out.nb = nb;
    out.key0 = k0;
    DevPhotons& M = out.M;
    M.pos = cp;
    M.inc = ci;
    M.pw = cw;
    M.spos = sp;
    M.sinc = si;
    M.spw = sw;
    M.start = start;
    M.idx = vin;
    M.origin = org;
    M.cell = G.cell;
    M.inv_cell = G.inv_cell;
    M.mask = G.mask;
    M.n = n;
    return RT_OK;
}

static bool chunked(const rt_settings* s) { return s->sample_offset != 0 || s->sample_total != 0; }

// q with q * q == n, or 0 when n is not a positive square.
static int exact_isqrt(int n) {
    if (n <= 0) return 0;
    int q = (int)std::lround(std::sqrt((double)n));
    return (int64_t)q * q == n ? q : 0;
}

// Camera samples per pixel the render takes: spp (Sobol) or the
// ceil(sqrt(spp))^2 strata of the PRNG sampler (Renderer.cs:556-558).  A
// chunk of a PRNG render takes exactly its spp samples of the whole
// render's sample_total = q^2 strata.
static int render_spp(const rt_settings* s) {
    if (s->sampler != RT_SAMPLER_PRNG || chunked(s)) return s->samples_per_pixel;
    const int q = (int)ceilf(sqrtf((float)s->samples_per_pixel));
    return q * q;
}

// Strata per axis of the PRNG sampler's grid.
static int strata_per_axis(const rt_settings* s) {
    if (s->sampler == RT_SAMPLER_PRNG && chunked(s)) return exact_isqrt(s->sample_total);
    return (int)ceilf(sqrtf((float)s->samples_per_pixel));
}

static int check_settings(rt_ctx* ctx, const rt_settings* s) {
    if (!s) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "settings is null");
    if (s->width <= 0 || s->height <= 0 || s->samples_per_pixel <= 0 || s->max_depth < 0)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "width, height and samples_per_pixel must be positive, max_depth >= 0");
    if ((int64_t)s->width * s->height > (1ll << 30)) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "image too large");
    if (s->sampler != RT_SAMPLER_SOBOL && s->sampler != RT_SAMPLER_PRNG)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "bad sampler");
    if (s->sampler == RT_SAMPLER_PRNG && chunked(s) && exact_isqrt(s->sample_total) == 0)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT,
                    "a chunk of a PRNG render needs sample_total = the whole render's q x q strata");
    if (s->caustics && (s->caustic_photons < 0 || s->caustic_gather_radius < 0.0f))
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "negative photon budget or gather radius");
    if (s->light_sampling < RT_LIGHTS_ALL || s->light_sampling > RT_LIGHTS_UNIFORM)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "bad light_sampling");
    if (s->texture_filtering < 0 || s->texture_filtering > 2)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "bad texture_filtering");
    if (s->sample_offset < 0 || (int64_t)s->sample_offset + s->samples_per_pixel > INT32_MAX)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "sample_offset must be >= 0 with sample_offset + spp < 2^31");
    if (s->sample_total < 0 || (s->sample_total > 0 && s->sample_total < s->sample_offset + s->samples_per_pixel))
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "sample_total must be 0 or cover sample_offset + spp");
    return RT_OK;
}

extern "C" {

int rt_get_version(void) { return RT3D_ABI_VERSION; }

void rt_material_defaults(rt_material* m, int32_t kind) {
    if (!m) return;
    std::memset(m, 0, sizeof(*m));
    m->kind = kind;
    m->color[0] = m->color[1] = m->color[2] = 0.8f;
    m->intensity = 1.0f;
    m->roughness = 0.5f;
    m->specular = 0.5f;
    m->sheen_tint = 0.5f;
    m->sheen_roughness = 0.3f;
    m->clearcoat_gloss = 1.0f;
    m->ior = 1.5f;
    m->coat_ior = 1.5f;
    m->thin_film_ior = 1.5f;
    m->transmission_color[0] = m->transmission_color[1] = m->transmission_color[2] = 1.0f;
}

void rt_settings_defaults(rt_settings* s) {
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->width = 960;
    s->height = 540;
    s->samples_per_pixel = 16;
    s->max_depth = 8;
    s->sampler = RT_SAMPLER_SOBOL;
    s->mis = RT_MIS_BALANCE;
    s->light_sampling = RT_LIGHTS_ALL;
    s->max_sample_radiance = 10.0f;
    s->indirect_clamp_factor = 0.25f;
    s->exposure_ev = 0.0f;
    s->rr_min_bounces = -1;
    s->rr_min_survival = -1.0f;
}

int rt_create(int32_t device, rt_ctx** out) {
    if (!out) return RT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RT_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return RT_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return RT_ERR_DEVICE;
    rt_ctx* ctx = new rt_ctx();
    ctx->device = device;
    const char* stk = getenv("RT3D_STACK");
    ctx->priv_stack = !(stk && std::string(stk) == "lds");
    const char* gr = getenv("RT3D_GRAPH");
    ctx->use_graph = gr && std::string(gr) == "1";
    const char* wpe = getenv("RT3D_SHADE_WPE");
    if (wpe) ctx->shade_wpe = std::max(1, std::min(4, atoi(wpe)));
    const char* bwpe = getenv("RT3D_BOUNCE_WPE");
    if (bwpe) ctx->bounce_wpe = std::max(1, std::min(4, atoi(bwpe)));
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return RT_ERR_DEVICE;
    }
    // the device address of this GPU's static staging array (the array's
    // name in host code is not a device address)
    void* scr = nullptr;
    if (hipGetSymbolAddress(&scr, HIP_SYMBOL(rt::scratch)) != hipSuccess || !scr) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return RT_ERR_DEVICE;
    }
    ctx->scratch = (char*)scr;
    bool ok = true;
    for (auto& e : ctx->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    for (auto& set : ctx->iev)
        for (auto& it : set)
            for (auto& e : it) ok = ok && hipEventCreate(&e) == hipSuccess;
    for (auto& e : ctx->chk) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&ctx->pinned, 64, hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        rt_destroy(ctx);
        return RT_ERR_DEVICE;
    }
    *out = ctx;
    return RT_OK;
}

void rt_destroy(rt_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->path_mem) hipFree(ctx->path_mem);
    if (ctx->film_mem) hipFree(ctx->film_mem);
    if (ctx->ph_mem) hipFree(ctx->ph_mem);
    for (auto& e : ctx->ev)
        if (e) hipEventDestroy(e);
    for (auto& set : ctx->iev)
        for (auto& it : set)
            for (auto& e : it)
                if (e) hipEventDestroy(e);
    for (auto& e : ctx->chk)
        if (e) hipEventDestroy(e);
    if (ctx->pinned) hipHostFree(ctx->pinned);
    if (ctx->parts_mem) hipFree(ctx->parts_mem);
    if (ctx->jobs_mem) hipFree(ctx->jobs_mem);
    if (ctx->adj_mem) hipFree(ctx->adj_mem);
    if (ctx->io_mem) hipFree(ctx->io_mem);
    if (ctx->dn_mem) hipFree(ctx->dn_mem);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* rt_last_error(const rt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rt_scene_create(rt_ctx* ctx, const rt_scene_desc* desc, rt_scene** out) {
    if (!ctx || !out) return RT_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    hipSetDevice(ctx->device);
    rt_scene* sc = new rt_scene();
    sc->ctx = ctx;
    std::string err;
    int rc = build_host_scene(desc, sc->hs, err);
    if (rc != RT_OK) {
        delete sc;
        return fail(ctx, rc, err);
    }
    HostScene& h = sc->hs;
    // Stack bound: deepest world path + a leaf continuation + the deepest
    // nested mesh path (build_host_scene records the maximum over all BVHs).
    sc->stack_depth = h.world_depth + (h.mesh_depth >= 0 ? 1 + h.mesh_depth : 0) + 1;
    // an object hierarchy: nested BVHs, leaf continuations and one frame
    // marker per wrapper along the deepest path
    if (h.hier) sc->stack_depth = std::max(sc->stack_depth, h.hier_stack + 1);
    // the any-hit BVH needs up to 3 pushes per level; use it only when its
    // bound fits the stacks the binary traversal's depth class provides
    if (h.any_stack > 64) h.nodes4.clear();
    else sc->stack_depth = std::max(sc->stack_depth, h.any_stack);
    if (sc->stack_depth > 64) {
        delete sc;
        return fail(ctx, RT_ERR_UNSUPPORTED, "BVH too deep for the 64-entry traversal stack");
    }
    int64_t& b = sc->device_bytes;
    auto cleanup = [&](int code) {
        rt_scene_destroy(sc);
        return code;
    };
    if ((rc = upload(ctx, &sc->d_nodes, h.nodes.data(), h.nodes.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_nodes4, h.nodes4.data(), h.nodes4.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_entries, h.entries.data(), h.entries.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_aux, h.quad_aux.data(), h.quad_aux.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_mats, h.mats.data(), h.mats.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_lights, h.lights.data(), h.lights.size(), b))) return cleanup(rc);
    std::vector<float> lut(h.lut_e, h.lut_e + 32 * 32);
    lut.insert(lut.end(), h.lut_eavg, h.lut_eavg + 32);   // E then Eavg
    if ((rc = upload(ctx, &sc->d_lut, lut.data(), lut.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_raw, h.raw_mats.data(), h.raw_mats.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_tuv, h.tex_uv.data(), h.tex_uv.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_smooth, h.smooth.data(), h.smooth.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_images, h.images.data(), h.images.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_levels, h.img_levels.data(), h.img_levels.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_imtex, h.img_texels.data(), h.img_texels.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_tex, h.texels.data(), h.texels.size(), b))) return cleanup(rc);
    if ((rc = upload(ctx, &sc->d_xf, h.xf.data(), h.xf.size(), b))) return cleanup(rc);
    std::vector<float> ld(2 * std::max<size_t>(h.lights.size(), 1), 0.0f);
    if ((rc = upload(ctx, &sc->d_ldist, ld.data(), ld.size(), b))) return cleanup(rc);
    *out = sc;
    return RT_OK;
}

void rt_scene_destroy(rt_scene* sc) {
    if (!sc) return;
    if (sc->ctx) hipSetDevice(sc->ctx->device);
    hipFree(sc->d_nodes);
    hipFree(sc->d_nodes4);
    hipFree(sc->d_entries);
    hipFree(sc->d_aux);
    hipFree(sc->d_mats);
    hipFree(sc->d_lights);
    hipFree(sc->d_lut);
    hipFree(sc->d_ldist);
    hipFree(sc->d_raw);
    hipFree(sc->d_tuv);
    hipFree(sc->d_smooth);
    hipFree(sc->d_tex);
    hipFree(sc->d_images);
    hipFree(sc->d_levels);
    hipFree(sc->d_imtex);
    hipFree(sc->d_xf);
    delete sc;
}

int rt_scene_set_textures(rt_ctx* ctx, rt_scene* sc, int32_t width, int32_t height, const float* texels) {
    if (!ctx || !sc || !texels || width <= 0 || height <= 0 || (int64_t)width * height > (1ll << 28))
        return RT_ERR_INVALID_ARGUMENT;
    if (!sc->hs.textured) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "scene has no textured material");
    hipSetDevice(ctx->device);
    HostScene& h = sc->hs;
    const size_t n = (size_t)kTexChannels * width * height;
    if (width != h.tex_w || height != h.tex_h) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        float* d = nullptr;
        HIP_TRY(hipMalloc(&d, n * sizeof(float)));
        hipFree(sc->d_tex);
        sc->device_bytes += (int64_t)(n - h.texels.size()) * (int64_t)sizeof(float);
        sc->d_tex = d;
        h.tex_w = width;
        h.tex_h = height;
    }
    h.texels.assign(texels, texels + n);
    HIP_TRY(hipMemcpyAsync(sc->d_tex, h.texels.data(), n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return RT_OK;
}

int rt_scene_get_info(const rt_scene* sc, rt_scene_info* info) {
    if (!sc || !info) return RT_ERR_INVALID_ARGUMENT;
    std::memset(info, 0, sizeof(*info));
    const HostScene& h = sc->hs;
    info->num_primitives = h.num_prims;
    info->num_objects = h.num_objects;
    info->num_nodes = h.num_internal;
    info->num_leaves = h.num_leaves;
    info->max_depth = h.max_depth;
    info->has_bvh = h.has_bvh;
    info->rr_min_bounces = h.rr_min_bounces;
    info->rr_min_survival = h.rr_min_survival;
    info->scene_radius = h.scene_radius;
    info->mean_irradiance = h.mean_irradiance;
    info->device_bytes = sc->device_bytes;
    return RT_OK;
}

int rt_trace_primary(rt_ctx* ctx, const rt_scene* sc, const rt_settings* s, int32_t sample_begin, int32_t sample_count,
                     int32_t* prim_id, float* t, float* u, float* v) {
    if (!ctx || !sc) return RT_ERR_INVALID_ARGUMENT;
    int rc = check_settings(ctx, s);
    if (rc) return rc;
    if (sample_begin < 0 || sample_count <= 0) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "bad sample range");
    if (s->sampler != RT_SAMPLER_SOBOL) return fail(ctx, RT_ERR_UNSUPPORTED, "primary hits are traced with the Sobol sampler");
    hipSetDevice(ctx->device);
    int64_t npix = (int64_t)s->width * s->height;
    int64_t n = npix * sample_count;
    if (n > (1ll << 31) - 2) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "too many rays");
    RParams R{};
    R.W = s->width;
    R.H = s->height;
    R.npix = (int)npix;
    R.base_sample = sample_begin;
    R.seed_xor = s->seed_xor;
    R.diff_rays = s->texture_filtering != 2 ? 1 : 0;
    DevScene S = dev_view(sc);
    char* mem = nullptr;
    size_t bytes = (size_t)n * 16 + 64;
    HIP_TRY(hipMalloc((void**)&mem, bytes));
    int32_t* d_prim = (int32_t*)mem;
    float* d_t = (float*)(mem + 4 * (size_t)n + 16);
    float* d_u = d_t + n;
    float* d_v = d_u + n;
    hipMemsetAsync(mem, 0, bytes, ctx->stream);
    dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
    DISPATCH_SD(sc->stack_depth, k_trace_primary, grid, ctx->stream, S, R, (int)n, d_prim, d_t, d_u, d_v);
    hipError_t e = hipGetLastError();
    int32_t ovf = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&ovf, d_prim + n, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && prim_id) e = hipMemcpyAsync(prim_id, d_prim, 4 * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && t) e = hipMemcpyAsync(t, d_t, 4 * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && u) e = hipMemcpyAsync(u, d_u, 4 * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && v) e = hipMemcpyAsync(v, d_v, 4 * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(mem);
    if (e != hipSuccess) return fail(ctx, RT_ERR_DEVICE, std::string("trace_primary: ") + hipGetErrorString(e));
    if (ovf) return fail(ctx, RT_ERR_INTERNAL, "traversal stack overflow");
    return RT_OK;
}

int rt_occluded(rt_ctx* ctx, const rt_scene* sc, const float* rays, int64_t num_rays, int32_t* occluded) {
    if (!ctx || !sc || num_rays < 0 || (num_rays > 0 && (!rays || !occluded))) return RT_ERR_INVALID_ARGUMENT;
    if (num_rays == 0) return RT_OK;
    if (num_rays > (1ll << 31) - 2) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "too many rays");
    hipSetDevice(ctx->device);
    int n = (int)num_rays;
    DevScene S = dev_view(sc);
    char* mem = nullptr;
    HIP_TRY(hipMalloc((void**)&mem, (size_t)n * 32 + 4 * (size_t)(n + 4)));
    float* d_rays = (float*)mem;
    int32_t* d_occ = (int32_t*)(mem + (size_t)n * 32);
    hipError_t e = hipMemcpyAsync(d_rays, rays, (size_t)n * 32, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_occ, 0, (size_t)(n + 4) * 4, ctx->stream);
    if (e == hipSuccess) {
        dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
        DISPATCH_SD(sc->stack_depth, k_occluded, grid, ctx->stream, S, d_rays, n, d_occ);
        e = hipGetLastError();
    }
    int32_t ovf = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&ovf, d_occ + n, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(occluded, d_occ, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(mem);
    if (e != hipSuccess) return fail(ctx, RT_ERR_DEVICE, std::string("occluded: ") + hipGetErrorString(e));
    if (ovf) return fail(ctx, RT_ERR_INTERNAL, "traversal stack overflow");
    return RT_OK;
}

int rt_occluded_noalloc(rt_ctx* ctx, const rt_scene* sc, const float* rays, int64_t num_rays, int32_t* occluded) {
    if (!ctx || !sc || num_rays < 0 || (num_rays > 0 && (!rays || !occluded))) return RT_ERR_INVALID_ARGUMENT;
    if (num_rays == 0) return RT_OK;
    if (num_rays > kNoAllocMaxRays)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT,
                    "rt_occluded_noalloc: at most " + std::to_string(kNoAllocMaxRays) + " segments per call");
    if (!ctx->scratch) return fail(ctx, RT_ERR_DEVICE, "rt_occluded_noalloc: no staging on this device");
    hipSetDevice(ctx->device);
    const int n = (int)num_rays;
    DevScene S = dev_view(sc);
    float* d_rays = (float*)ctx->scratch;
    int32_t* d_occ = (int32_t*)(ctx->scratch + kNoAllocResults);
    // the overflow flag after the results starts at zero (a fill kernel: the
    // call may sit in captured work); every result is written by the kernel
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(kBlock), 0, ctx->stream, d_occ + n, 4);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(d_rays, rays, (size_t)n * 32, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
        DISPATCH_SD(sc->stack_depth, k_occluded, grid, ctx->stream, S, d_rays, n, d_occ);
        e = hipGetLastError();
    }
    int32_t ovf = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&ovf, d_occ + n, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(occluded, d_occ, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(ctx, RT_ERR_DEVICE, std::string("occluded_noalloc: ") + hipGetErrorString(e));
    if (ovf) return fail(ctx, RT_ERR_INTERNAL, "traversal stack overflow");
    return RT_OK;
}

int rt_trace_rays(rt_ctx* ctx, const rt_scene* sc, const float* rays, int64_t num_rays, int32_t* prim_id, float* t,
                  float* u, float* v) {
    if (!ctx || !sc || (num_rays > 0 && !rays) || num_rays < 0) return RT_ERR_INVALID_ARGUMENT;
    if (num_rays == 0) return RT_OK;
    if (num_rays > (1ll << 31) - 2) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "too many rays");
    hipSetDevice(ctx->device);
    int n = (int)num_rays;
    DevScene S = dev_view(sc);
    char* mem = nullptr;
    size_t bytes = (size_t)n * (24 + 16) + 64;
    HIP_TRY(hipMalloc((void**)&mem, bytes));
    float* d_rays = (float*)mem;
    int32_t* d_prim = (int32_t*)(mem + (size_t)n * 24);
    float* d_t = (float*)(d_prim + n + 4);
    float* d_u = d_t + n;
    float* d_v = d_u + n;
    hipError_t e = hipMemcpyAsync(d_rays, rays, (size_t)n * 24, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_prim, 0, (size_t)(n + 4) * 4, ctx->stream);
    if (e == hipSuccess) {
        dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
        DISPATCH_SD(sc->stack_depth, k_trace_rays, grid, ctx->stream, S, d_rays, n, d_prim, d_t, d_u, d_v);
        e = hipGetLastError();
    }
    int32_t ovf = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&ovf, d_prim + n, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && prim_id) e = hipMemcpyAsync(prim_id, d_prim, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && t) e = hipMemcpyAsync(t, d_t, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && u) e = hipMemcpyAsync(u, d_u, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && v) e = hipMemcpyAsync(v, d_v, 4 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(mem);
    if (e != hipSuccess) return fail(ctx, RT_ERR_DEVICE, std::string("trace_rays: ") + hipGetErrorString(e));
    if (ovf) return fail(ctx, RT_ERR_INTERNAL, "traversal stack overflow");
    return RT_OK;
}

int rt_photon_map(rt_ctx* ctx, const rt_scene* sc, const rt_settings* s, int64_t cap_photons, int64_t cap_buckets,
                  int64_t* n_photons, int64_t* n_emitted, int64_t* n_buckets, float* pos, float* inc, float* power,
                  int32_t* key, int32_t* idx, int32_t* start, float* grid) {
    if (!ctx || !sc || !s) return RT_ERR_INVALID_ARGUMENT;
    if (s->caustic_photons < 0 || s->caustic_gather_radius < 0.0f)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "negative photon budget or gather radius");
    hipSetDevice(ctx->device);
    const HostScene& h = sc->hs;
    if (h.hier) return fail(ctx, RT_ERR_UNSUPPORTED, "photon maps of a scene with an object hierarchy are not supported");
    float gr = s->caustic_gather_radius > 0.0f ? s->caustic_gather_radius : std::max(h.scene_radius * 0.06f, 1e-4f);
    DevScene S = dev_view(sc);
    PhotonPass PP;
    int rc = photon_pass(ctx, sc, S, s->caustic_photons, gr * 0.5f, PP);
    if (rc) return rc;
    if (n_photons) *n_photons = PP.n;
    if (n_emitted) *n_emitted = PP.emitted;
    if (n_buckets) *n_buckets = PP.n > 0 ? PP.nb : 0;
    if (PP.n == 0) return RT_OK;
    const int n = PP.n;
    if (cap_photons >= n) {
        std::vector<float4> tmp(n);
        const float4* src3[3] = {PP.M.pos, PP.M.inc, PP.M.pw};
        float* dst3[3] = {pos, inc, power};
        for (int a = 0; a < 3; a++) {
            if (!dst3[a]) continue;
            HIP_TRY(hipMemcpy(tmp.data(), src3[a], 16ull * n, hipMemcpyDeviceToHost));
            for (int i = 0; i < n; i++) {
                dst3[a][3 * i] = tmp[i].x;
                dst3[a][3 * i + 1] = tmp[i].y;
                dst3[a][3 * i + 2] = tmp[i].z;
            }
        }
        if (key) HIP_TRY(hipMemcpy(key, PP.key0, 4ull * n, hipMemcpyDeviceToHost));
        if (idx) HIP_TRY(hipMemcpy(idx, PP.M.idx, 4ull * n, hipMemcpyDeviceToHost));
    }
    if (start && cap_buckets >= PP.nb + 1) HIP_TRY(hipMemcpy(start, PP.M.start, 4ull * (PP.nb + 1), hipMemcpyDeviceToHost));
    if (grid) {
        float4 o;
        HIP_TRY(hipMemcpy(&o, PP.M.origin, 16, hipMemcpyDeviceToHost));
        grid[0] = o.x;
        grid[1] = o.y;
        grid[2] = o.z;
        grid[3] = PP.M.cell;
    }
    return RT_OK;
}

int rt_photon_gather(rt_ctx* ctx, const rt_scene* sc, const rt_settings* s, const float* points, int64_t n_points,
                     int32_t* count, float* radius, int32_t* photon_idx) {
    if (!ctx || !sc || !s) return RT_ERR_INVALID_ARGUMENT;
    if (n_points < 0 || (n_points > 0 && (!points || !count || !radius || !photon_idx)))
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "points and all three outputs are required");
    if (n_points > (1ll << 26)) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "more than 2^26 query points");
    if (s->caustic_photons < 0 || s->caustic_gather_radius < 0.0f)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "negative photon budget or gather radius");
    hipSetDevice(ctx->device);
    const HostScene& h = sc->hs;
    if (h.hier) return fail(ctx, RT_ERR_UNSUPPORTED, "photon maps of a scene with an object hierarchy are not supported");
    const float gr = s->caustic_gather_radius > 0.0f ? s->caustic_gather_radius
                                                     : std::max(h.scene_radius * 0.06f, 1e-4f);
    DevScene S = dev_view(sc);
    PhotonPass PP;
    int rc = photon_pass(ctx, sc, S, s->caustic_photons, gr * 0.5f, PP);
    if (rc) return rc;
    const int n = (int)n_points;
    if (n == 0) return RT_OK;
    if (PP.n == 0) {
        for (int i = 0; i < n; i++) {
            count[i] = 0;
            radius[i] = 0.0f;
            for (int j = 0; j < kGatherK; j++) photon_idx[(int64_t)i * kGatherK + j] = -1;
        }
        return RT_OK;
    }
    PP.M.max_r = gr;
    const size_t b_pts = 12ull * n, b_cnt = 4ull * n, b_idx = 4ull * kGatherK * n;
    char* mem = nullptr;
    HIP_TRY(hipMalloc((void**)&mem, b_pts + 2 * b_cnt + b_idx + 64));
    float* d_pts = (float*)mem;
    int32_t* d_cnt = (int32_t*)(mem + ((b_pts + 15) & ~15ull));
    float* d_rad = (float*)((char*)d_cnt + b_cnt);
    int32_t* d_idx = (int32_t*)((char*)d_rad + b_cnt);
    hipError_t e = hipMemcpyAsync(d_pts, points, b_pts, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_photon_knn_query, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, PP.M,
                           d_pts, n, d_cnt, d_rad, d_idx);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(count, d_cnt, b_cnt, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(radius, d_rad, b_cnt, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(photon_idx, d_idx, b_idx, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(mem);
    if (e != hipSuccess) return fail(ctx, RT_ERR_DEVICE, std::string("photon_gather: ") + hipGetErrorString(e));
    return RT_OK;
}

}  // extern "C"

// ── render set-up helpers shared by rt_render and rt_render_batch ─────────
// Path pool of N slots (and per-sample results of a window of WS samples)
// in the context's grow-only buffer.
static int alloc_paths(rt_ctx* ctx, int64_t N, int64_t WS, int maxs, bool aovs, bool gather, bool fp, bool xf,
                       PathBufs& P) {
    size_t per_path = 8 * 16 + 8 * 4 + 4 + 16 + (aovs ? 5 * 16 : 0) + 4 * 16 * (size_t)maxs + 8 * 4 + 4 * (size_t)maxs +
                      (gather ? 64 : 0) + (fp ? 16 : 0) + (xf ? 16 : 0);
    size_t per_sample = aovs ? 48 : 16;
    size_t need = (size_t)N * per_path + (size_t)WS * per_sample + 40 * 256 + 16384 + 8 * N +
                  4 * (size_t)((N + kBlock - 1) / kBlock + 3) * (kWcSeg + kMaxSeg);
    if (need > ctx->path_bytes) {
        if (ctx->path_mem) hipFree(ctx->path_mem);
        ctx->path_mem = nullptr;
        ctx->path_bytes = 0;
        HIP_TRY(hipMalloc(&ctx->path_mem, need));
        ctx->path_bytes = need;
    }
    char* m = (char*)ctx->path_mem;
    auto take = [&](size_t bytes) {
        char* r = m;
        m += (bytes + 255) & ~(size_t)255;
        return (void*)r;
    };
    P = PathBufs{};
    P.n = (int)N;
    P.counters = (int32_t*)take(256);
    P.tcount = (unsigned long long*)take(128);
    P.o = (float4*)take(16 * N);
    P.d = (float4*)take(16 * N);
    P.thr = (float4*)take(16 * N);
    P.wt = (float4*)take(16 * N);
    P.w0 = (float4*)take(16 * N);
    P.a0 = (float4*)take(16 * N);
    P.ai = (float4*)take(16 * N);
    P.ab = (float4*)take(16 * N);
    P.ior = (float*)take(4 * 8 * N);
    P.hit_e = (int32_t*)take(4 * N);
    P.hit_tuv = (float4*)take(16 * N);
    if (xf) P.hit_xf = (int4*)take(16 * N);
    if (aovs) {
        P.av0 = (float4*)take(16 * N);
        P.av1 = (float4*)take(16 * N);
        P.av2 = (float4*)take(16 * N);
        P.av3 = (float4*)take(16 * N);
        P.av4 = (float4*)take(16 * N);
    }
    P.sh_o = (float4*)take(16 * N * maxs);
    P.sh_d = (float4*)take(16 * N * maxs);
    P.sh_c = (float4*)take(16 * N * maxs);
    P.sh_r = (float4*)take(16 * N * maxs);
    P.q_sh = (int32_t*)take(4 * N * maxs);
    P.state = (int32_t*)take(4 * N);
    P.smask = (int32_t*)take(4 * N);
    P.gflag = (int32_t*)take(4 * N);
    P.q_g = (int32_t*)take(4 * N);
    P.nw = (int)(((N + kBlock - 1) / kBlock + 3) & ~(int64_t)3);   // row stride of the count arrays (int4 rows)
    P.wc = (int32_t*)take(4 * (size_t)P.nw * (kWcSeg + std::min(maxs, kMaxSeg)));
    P.job = (int32_t*)take(4 * N);
    P.pix = (int32_t*)take(4 * N);
    P.pass = (int32_t*)take(4 * N);
    P.sid = (int32_t*)take(4 * N);
    P.res_l = (float4*)take(16 * WS);
    if (gather) {
        P.gq0 = (float4*)take(16 * N);
        P.gq1 = (float4*)take(16 * N);
        P.gq2 = (float4*)take(16 * N);
        P.gq3 = (float4*)take(16 * N);
    }
    if (aovs) {
        P.res_a = (float4*)take(16 * WS);
        P.res_n = (float4*)take(16 * WS);
    }
    if (fp) P.fp = (float4*)take(16 * N);
    if ((size_t)(m - (char*)ctx->path_mem) > ctx->path_bytes) return fail(ctx, RT_ERR_INTERNAL, "path buffer layout");
    return RT_OK;
}

// Film accumulators + output planes of an npix image at `base` (null: size
// only).  Returns the bytes used; *acc_bytes = the accumulator part.
static size_t film_layout(char* base, int64_t npix, FilmBufs& F, size_t* acc_bytes) {
    size_t off = 0;
    auto ftake = [&](size_t bytes) {
        char* r = base ? base + off : nullptr;
        off += (bytes + 255) & ~(size_t)255;
        return (void*)r;
    };
    F = FilmBufs{};
    F.sum = (float*)ftake(12 * (size_t)npix);
    F.sa = (float*)ftake(12 * (size_t)npix);
    F.sb = (float*)ftake(12 * (size_t)npix);
    F.alba = (float*)ftake(12 * (size_t)npix);
    F.albb = (float*)ftake(12 * (size_t)npix);
    F.nrma = (float*)ftake(12 * (size_t)npix);
    F.nrmb = (float*)ftake(12 * (size_t)npix);
    F.depa = (float*)ftake(4 * (size_t)npix);
    F.depb = (float*)ftake(4 * (size_t)npix);
    F.dca = (int32_t*)ftake(4 * (size_t)npix);
    F.dcb = (int32_t*)ftake(4 * (size_t)npix);
    if (acc_bytes) *acc_bytes = off;
    F.out_beauty = (float*)ftake(12 * (size_t)npix);
    F.out_a = (float*)ftake(12 * (size_t)npix);
    F.out_b = (float*)ftake(12 * (size_t)npix);
    F.out_disp = (float*)ftake(12 * (size_t)npix);
    F.out_alba = (float*)ftake(12 * (size_t)npix);
    F.out_albb = (float*)ftake(12 * (size_t)npix);
    F.out_nrma = (float*)ftake(12 * (size_t)npix);
    F.out_nrmb = (float*)ftake(12 * (size_t)npix);
    F.out_depa = (float*)ftake(4 * (size_t)npix);
    F.out_depb = (float*)ftake(4 * (size_t)npix);
    return off;
}

static int grow(rt_ctx* ctx, void** mem, size_t* bytes, size_t need) {
    if (need <= *bytes) return RT_OK;
    if (*mem) hipFree(*mem);
    *mem = nullptr;
    *bytes = 0;
    HIP_TRY(hipMalloc(mem, need));
    *bytes = need;
    return RT_OK;
}

// LightDistribution (LightDistribution.cs:41-99) for the power / uniform
// strategies, uploaded into the scene's table.
static int upload_light_dist(rt_ctx* ctx, const rt_scene* sc, int strategy) {
    const HostScene& h = sc->hs;
    int nl = (int)h.lights.size();
    if (strategy == RT_LIGHTS_ALL || nl == 0) return RT_OK;
    std::vector<float> cdf(nl), pdf(nl), pw(nl, 0.0f);
    float sum = 0.0f;
    if (strategy == RT_LIGHTS_POWER)
        for (int i = 0; i < nl; i++) { pw[i] = std::max(0.0f, h.light_power[i]); sum += pw[i]; }
    float cum = 0.0f;
    if (sum > 0.0f) {
        float inv = 1.0f / sum;
        for (int i = 0; i < nl; i++) { pdf[i] = pw[i] * inv; cum += pdf[i]; cdf[i] = cum; }
    } else {
        float un = 1.0f / (float)nl;
        for (int i = 0; i < nl; i++) { pdf[i] = un; cum += un; cdf[i] = cum; }
    }
    cdf[nl - 1] = 1.0f;
    rt_scene* msc = const_cast<rt_scene*>(sc);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(msc->d_ldist, cdf.data(), 4 * nl, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(msc->d_ldist + nl, pdf.data(), 4 * nl, hipMemcpyHostToDevice));
    return RT_OK;
}

static int shadow_slots(const rt_settings* s, const HostScene& h) {
    return std::max(1, s->light_sampling != RT_LIGHTS_ALL ? h.max_light_samples : h.max_shadow_per_vertex);
}

// Renderer constructor arguments (Renderer.cs:300-378) as device parameters.
// Default path pool: 32 M slots (~19 GB of path state with 5 shadow records
// per path, 6.5 % of HBM).  Larger pools mean fewer bounce iterations per
// frame (each with fixed scan/compaction/launch costs) and more rays per
// launch over the same tree; the 1080p x 64 spp frame of the 1.31 M-triangle
// scene renders at 397 / 397 / 412 / 381 / 388 Msamples/s with 16 / 28 / 32
// / 36 / 40 M slots (the 32 M figure repeated on two boxes within 0.4 %).
constexpr int64_t kDefaultPool = 1ll << 25;
// rt_render_batch's default pool: 16 M slots.  Jobs of at least two pools of
// samples run alone; with the 32 M pool more of a batch's jobs fall below
// that and share windows, where their waves mix scenes (16-scene check:
// batched/plain 0.978 at 16 M, 0.909 at 32 M).
constexpr int64_t kDefaultBatchPool = 1ll << 24;

// Samples per window for a pool of `pool` slots.  A window's per-sample
// results stay resident until it completes (16 B per sample, 48 B with
// AOVs: 4 GiB / 12 GiB at the cap), and every window ends with a drain of
// the pool, so windows are made as large as HBM comfortably allows.
static int64_t window_cap(int64_t pool) { return std::max<int64_t>(pool, std::min<int64_t>(1ll << 28, 128 * pool)); }

static RParams make_params(const rt_settings* s, const HostScene& h, int maxs) {
    RParams R{};
    R.W = s->width;
    R.H = s->height;
    R.spp = render_spp(s);
    R.prng = s->sampler == RT_SAMPLER_PRNG ? 1 : 0;
    R.sq = strata_per_axis(s);
    R.inv_sq = 1.0f / (float)R.sq;
    R.max_depth = s->max_depth;
    R.npix = s->width * s->height;
    R.max_rad = s->max_sample_radiance;
    R.ind_max_rad = s->max_sample_radiance * std::max(0.0f, s->indirect_clamp_factor);
    R.power_mis = s->mis == RT_MIS_POWER;
    R.picked = (s->light_sampling != RT_LIGHTS_ALL && !h.lights.empty()) ? 1 : 0;
    R.rr_min_b = s->rr_min_bounces >= 0 ? s->rr_min_bounces : h.rr_min_bounces;
    R.rr_min_s = s->rr_min_survival >= 0.0f ? s->rr_min_survival : h.rr_min_survival;
    R.seed_xor = s->seed_xor;
    R.soff = s->sample_offset;
    R.spp_fp = s->sample_total > 0 ? s->sample_total : R.spp;
    R.aovs = s->capture_aovs ? 1 : 0;
    R.maxs = maxs;
    R.diff_rays = s->texture_filtering != 2 ? 1 : 0;
    R.footprints = R.diff_rays && !h.images.empty() ? 1 : 0;
    return R;
}

// Resolve (mean, halves, display) and copy the requested planes out.
static int finish_film(rt_ctx* ctx, const RParams& R, const FilmBufs& F, const rt_settings* s, const rt_film* film) {
    const int64_t npix = (int64_t)R.npix;
    float exposure = powf(2.0f, s->exposure_ev);
    hipLaunchKernelGGL(k_resolve, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, ctx->stream, R, F, exposure);
    HIP_TRY(hipGetLastError());
    if (!film) return RT_OK;
    const bool aovs = R.aovs != 0;
    size_t c3 = 12 * (size_t)npix, c1 = 4 * (size_t)npix;
    struct { void* dst; const void* src; size_t n; } copies[] = {
        {film->beauty, F.out_beauty, c3}, {film->beauty_a, F.out_a, c3}, {film->beauty_b, F.out_b, c3},
        {film->display, F.out_disp, c3},
        {aovs ? film->albedo_a : nullptr, F.out_alba, c3}, {aovs ? film->albedo_b : nullptr, F.out_albb, c3},
        {aovs ? film->normal_a : nullptr, F.out_nrma, c3}, {aovs ? film->normal_b : nullptr, F.out_nrmb, c3},
        {aovs ? film->depth_a : nullptr, F.out_depa, c1}, {aovs ? film->depth_b : nullptr, F.out_depb, c1},
    };
    for (auto& c : copies)
        if (c.dst) HIP_TRY(hipMemcpyAsync(c.dst, c.src, c.n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (film->sample_count)
        for (int64_t k = 0; k < npix; k++) film->sample_count[k] = R.spp;
    return RT_OK;
}

static void fill_counts(const unsigned long long* tcv, bool counting, int64_t launches, rt_render_stats* st) {
    st->rays_closest = (int64_t)tcv[kTcRays];
    st->rays_shadow = (int64_t)tcv[kTcShadowRays];
    st->trace_launches = launches;
    st->shadow_launches = launches;
    if (counting) {
        st->trace_node_tests = (int64_t)tcv[0];
        st->trace_tri_tests = (int64_t)tcv[1];
        st->trace_other_tests = (int64_t)tcv[2];
        st->shadow_node_tests = (int64_t)tcv[3];
        st->shadow_tri_tests = (int64_t)tcv[4];
        st->shadow_other_tests = (int64_t)tcv[5];
        st->any_node_fetches = (int64_t)tcv[kTcAnyNode];
        st->any_tri_tests = (int64_t)tcv[kTcAnyTri];
        st->any_other_tests = (int64_t)tcv[kTcAnyOther];
    }
}

extern "C" {

int rt_render(rt_ctx* ctx, const rt_scene* sc, const rt_settings* s, const rt_film* film, rt_render_stats* stats) {
    auto t_start = std::chrono::steady_clock::now();
    if (!ctx || !sc) return RT_ERR_INVALID_ARGUMENT;
    int rc = check_settings(ctx, s);
    if (rc) return rc;
    hipSetDevice(ctx->device);
    const HostScene& h = sc->hs;
    const int spp = render_spp(s);
    const int npix = s->width * s->height;
    // Path pool of N slots; samples are taken in windows of K whole passes
    // over the image whose per-sample results are summed into the film in
    // sample order when the window completes.  A slot whose path ends is
    // restarted on the window's next sample, so the pool stays full.
    int64_t target = s->paths_in_flight > 0 ? s->paths_in_flight : kDefaultPool;
    const int64_t kWindowCap = window_cap(target);
    int K = (int)std::max<int64_t>(1, std::min<int64_t>(spp, kWindowCap / npix));
    const int64_t WS = (int64_t)K * npix;
    const int64_t N = std::min<int64_t>(target, WS);
    if (WS > (1ll << 31) - 1 || N > (1ll << 30)) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "image too large");
    const int maxs = shadow_slots(s, h);
    if ((int64_t)maxs * N > (1ll << 31) - 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "too many shadow slots");
    if (maxs > 255) return fail(ctx, RT_ERR_UNSUPPORTED, "more than 255 shadow samples per shading point");
    const bool want_caustics = s->caustics && s->caustic_photons > 0 && !h.lights.empty();
    if (want_caustics && (h.textured || !h.images.empty()))
        return fail(ctx, RT_ERR_UNSUPPORTED, "photon caustics with textured materials are not supported");
    if (want_caustics && h.hier)
        return fail(ctx, RT_ERR_UNSUPPORTED, "photon caustics in a scene with an object hierarchy are not supported");

    PathBufs P;
    rc = alloc_paths(ctx, N, WS, maxs, s->capture_aovs != 0, want_caustics,
                     !h.images.empty() && s->texture_filtering != 2, h.hier, P);
    if (rc) return rc;
    FilmBufs F;
    size_t acc_bytes = 0;
    rc = grow(ctx, &ctx->film_mem, &ctx->film_bytes, film_layout(nullptr, npix, F, nullptr) + 256);
    if (rc) return rc;
    film_layout((char*)ctx->film_mem, npix, F, &acc_bytes);
    HIP_TRY(hipMemsetAsync(ctx->film_mem, 0, acc_bytes, ctx->stream));
    rc = upload_light_dist(ctx, sc, s->light_sampling);
    if (rc) return rc;
    DevScene S = dev_view(sc);
    RParams R = make_params(s, h, maxs);

    // ── caustic photon pre-pass (Renderer.cs:385-406) ──────────────────────
    PhotonPass PP;
    if (want_caustics) {
        float gr = s->caustic_gather_radius > 0.0f ? s->caustic_gather_radius : std::max(h.scene_radius * 0.06f, 1e-4f);
        rc = photon_pass(ctx, sc, S, s->caustic_photons, gr * 0.5f, PP);
        if (rc) return rc;
        PP.M.max_r = gr;
        if (PP.n > 0) {
            S.caustics = 1;
            S.block_spec = 1;
        }
    }

    const bool counting = s->count_traversal != 0;
    HIP_TRY(hipMemsetAsync(P.tcount, 0, 128, ctx->stream));
    HIP_TRY(hipMemsetAsync(P.counters, 0, 256, ctx->stream));
    hipEvent_t* ev = ctx->ev;
    HIP_TRY(hipEventRecord(ev[6], ctx->stream));
    WindowPlan plan;
    for (int base = 0; base < spp; base += K) {
        const int passes = std::min(K, spp - base);
        plan.add_window({PartDev{0, base, 0, npix, passes, {0, 0, 0}}}, passes * npix);
    }
    LoopTiming tm;
    OneJob one{S, R};
    rc = run_windows(ctx, one, P, maxs, s->max_depth > 0, counting, sc->stack_depth, S.caustics ? &S : nullptr, &PP.M,
                     plan,
                     [&](const PartDev& pt) {
                         hipLaunchKernelGGL(k_accum, dim3((npix + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream,
                                            R, P, F, pt.pass0, pt.npass, pt.off);
                     },
                     tm);
    if (rc) return rc;
    const double t_trace = tm.trace, t_shade = tm.shade, t_shadow = tm.shadow;
    rc = finish_film(ctx, R, F, s, film);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ev[7], ctx->stream));
    HIP_TRY(hipEventSynchronize(ev[7]));
    float dev_ms = 0;
    hipEventElapsedTime(&dev_ms, ev[6], ev[7]);
    int32_t ovf = 0;
    HIP_TRY(hipMemcpy(&ovf, P.counters + kCntOverflow, 4, hipMemcpyDeviceToHost));
    if (ovf) return fail(ctx, RT_ERR_INTERNAL, "traversal stack overflow");
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        stats->device_ms = dev_ms;
        stats->trace_ms = t_trace;
        stats->shade_ms = t_shade;
        stats->shadow_ms = t_shadow;
        stats->camera_samples = (int64_t)npix * spp;
        unsigned long long tcv[16];
        HIP_TRY(hipMemcpy(tcv, P.tcount, sizeof(tcv), hipMemcpyDeviceToHost));
        fill_counts(tcv, counting, tm.launches, stats);
    }
    return RT_OK;
}


int rt_adjoint(rt_ctx* ctx, const rt_scene* sc, const rt_settings* s, const float* d_beauty, float* d_texels,
               rt_render_stats* stats) {
    auto t_start = std::chrono::steady_clock::now();
    if (!ctx || !sc || !d_beauty || !d_texels) return RT_ERR_INVALID_ARGUMENT;
    int rc = check_settings(ctx, s);
    if (rc) return rc;
    if (s->sampler != RT_SAMPLER_SOBOL) return fail(ctx, RT_ERR_UNSUPPORTED, "adjoint: the Sobol sampler only");
    hipSetDevice(ctx->device);
    const HostScene& h = sc->hs;
    if (!h.textured) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "scene has no textured material");
    if (h.medium) return fail(ctx, RT_ERR_UNSUPPORTED, "adjoint: participating media are not supported");
    if (!h.opaque_shadows) return fail(ctx, RT_ERR_UNSUPPORTED, "adjoint: transmissive materials are not supported");
    if (s->caustics) return fail(ctx, RT_ERR_UNSUPPORTED, "adjoint: photon caustics are not supported");
    if (h.hier) return fail(ctx, RT_ERR_UNSUPPORTED, "adjoint: scenes with an object hierarchy are not supported");
    if ((double)s->max_sample_radiance * (double)s->indirect_clamp_factor < 1e30)
        return fail(ctx, RT_ERR_UNSUPPORTED, "adjoint: radiance clamps must be inactive "
                                             "(max_sample_radiance * indirect_clamp_factor >= 1e30)");
    const int spp = s->samples_per_pixel;
    const int npix = s->width * s->height;
    const int64_t plane = (int64_t)h.tex_w * h.tex_h;
    int64_t target = s->paths_in_flight > 0 ? s->paths_in_flight : kDefaultPool;
    const int64_t kWindowCap = window_cap(target);
    int K = (int)std::max<int64_t>(1, std::min<int64_t>(spp, kWindowCap / npix));
    const int64_t WS = (int64_t)K * npix;
    const int64_t N = std::min<int64_t>(target, WS);
    if (WS > (1ll << 31) - 1 || N > (1ll << 30)) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "image too large");
    const int maxs = shadow_slots(s, h);
    if ((int64_t)maxs * N > (1ll << 31) - 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "too many shadow slots");
    if (maxs > 255) return fail(ctx, RT_ERR_UNSUPPORTED, "more than 255 shadow samples per shading point");

    PathBufs P;
    rc = alloc_paths(ctx, N, WS, maxs, false, false, !h.images.empty() && s->texture_filtering != 2, false, P);
    if (rc) return rc;
    // adjoint buffers: pending bounce terms, NEE terms, dLoss/dBeauty, gradient
    const size_t b_q = 16 * 5 * (size_t)N, b_g = 16 * 2 * (size_t)N * std::max(maxs, 1);
    const size_t b_db = 4 * 3 * (size_t)npix, b_dt = 4 * (size_t)kTexChannels * plane;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    rc = grow(ctx, &ctx->adj_mem, &ctx->adj_bytes, al(b_q) + al(b_g) + al(b_db) + al(b_dt));
    if (rc) return rc;
    char* am = (char*)ctx->adj_mem;
    P.adj_q = (float4*)am;
    P.adj_g = (float4*)(am + al(b_q));
    float* d_db = (float*)(am + al(b_q) + al(b_g));
    P.adj_db = d_db;
    P.adj_dt = (float*)(am + al(b_q) + al(b_g) + al(b_db));
    P.adj_inv_spp = 1.0f / (float)spp;
    HIP_TRY(hipMemcpyAsync(d_db, d_beauty, b_db, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(P.adj_dt, 0, b_dt, ctx->stream));
    rc = upload_light_dist(ctx, sc, s->light_sampling);
    if (rc) return rc;
    DevScene S = dev_view(sc);
    RParams R = make_params(s, h, maxs);
    HIP_TRY(hipMemsetAsync(P.tcount, 0, 128, ctx->stream));
    HIP_TRY(hipMemsetAsync(P.counters, 0, 256, ctx->stream));
    hipEvent_t* ev = ctx->ev;
    HIP_TRY(hipEventRecord(ev[6], ctx->stream));
    LoopTiming tm_primal, tm_replay;
    OneJob one{S, R};
    auto no_accum = [](const PartDev&) {};
    if (s->max_depth > 0) {
        // per window: the primal pass stores every sample's radiance sum, the
        // replay pass retraces the same samples (same streams) with gradients
        for (int base = 0; base < spp; base += K) {
            const int passes = std::min(K, spp - base);
            WindowPlan plan;
            plan.add_window({PartDev{0, base, 0, npix, passes, {0, 0, 0}}}, passes * npix);
            P.adj_pass = 1;
            rc = run_windows(ctx, one, P, maxs, true, false, sc->stack_depth, nullptr, nullptr, plan, no_accum,
                             tm_primal);
            if (rc) return rc;
            HIP_TRY(hipMemsetAsync(P.adj_q, 0, b_q, ctx->stream));
            P.adj_pass = 2;
            rc = run_windows(ctx, one, P, maxs, true, false, sc->stack_depth, nullptr, nullptr, plan, no_accum,
                             tm_replay);
            if (rc) return rc;
        }
    }
    HIP_TRY(hipEventRecord(ev[7], ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_texels, P.adj_dt, b_dt, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    float dev_ms = 0;
    hipEventElapsedTime(&dev_ms, ev[6], ev[7]);
    int32_t ovf = 0;
    HIP_TRY(hipMemcpy(&ovf, P.counters + kCntOverflow, 4, hipMemcpyDeviceToHost));
    if (ovf) return fail(ctx, RT_ERR_INTERNAL, "traversal stack overflow");
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        stats->device_ms = dev_ms;
        stats->trace_ms = tm_primal.trace + tm_replay.trace;
        stats->shade_ms = tm_primal.shade + tm_replay.shade;
        stats->shadow_ms = tm_primal.shadow + tm_replay.shadow;
        stats->camera_samples = (int64_t)npix * spp;
        unsigned long long tcv[16];
        HIP_TRY(hipMemcpy(tcv, P.tcount, sizeof(tcv), hipMemcpyDeviceToHost));
        fill_counts(tcv, false, tm_primal.launches + tm_replay.launches, stats);
    }
    return RT_OK;
}

int rt_render_batch(rt_ctx* ctx, const rt_batch_job* jobs, int32_t n_jobs, int32_t pool_paths,
                    rt_render_stats* total) {
    auto t_start = std::chrono::steady_clock::now();
    if (!ctx || !jobs || n_jobs <= 0) return RT_ERR_INVALID_ARGUMENT;
    hipSetDevice(ctx->device);
    int rc;
    // validate; shadow-slot count, traversal stack depth and AOV buffers are
    // shared by the pool, so they take the maximum over the jobs
    int maxs = 1, stack_depth = 0;
    bool aovs = false, counting = false, traced = false;
    std::vector<std::pair<const rt_scene*, int>> strategy;
    for (int j = 0; j < n_jobs; j++) {
        const rt_batch_job& b = jobs[j];
        if (!b.scene || !b.settings) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "batch job without scene or settings");
        rc = check_settings(ctx, b.settings);
        if (rc) return rc;
        if (b.settings->caustics) return fail(ctx, RT_ERR_UNSUPPORTED, "photon caustics are not supported in a batch");
        if (b.settings->sampler != RT_SAMPLER_SOBOL)
            return fail(ctx, RT_ERR_UNSUPPORTED, "a batch renders with the Sobol sampler");
        for (auto& p : strategy)
            if (p.first == b.scene && p.second != b.settings->light_sampling)
                return fail(ctx, RT_ERR_UNSUPPORTED, "one scene with two light-sampling strategies in a batch");
        strategy.push_back({b.scene, b.settings->light_sampling});
        maxs = std::max(maxs, shadow_slots(b.settings, b.scene->hs));
        stack_depth = std::max(stack_depth, b.scene->stack_depth);
        aovs = aovs || b.settings->capture_aovs;
        counting = counting || b.settings->count_traversal;
        traced = traced || b.settings->max_depth > 0;
    }
    if (maxs > 255) return fail(ctx, RT_ERR_UNSUPPORTED, "more than 255 shadow samples per shading point");
    // Segments of the batch, in job order: a job of at least two pools of
    // samples runs alone (single-job kernels, its windows like rt_render's);
    // smaller jobs are packed job-major into shared windows (batch kernels),
    // cut at pass boundaries, so they fill the pool together.
    const int64_t N0 = pool_paths > 0 ? pool_paths : kDefaultBatchPool;
    const int64_t cap = window_cap(N0);
    struct Segment {
        int job;   // >= 0: that job alone; -1: packed jobs
        WindowPlan plan;
    };
    std::vector<Segment> segs;
    {
        std::vector<PartDev> cur;
        int64_t used = 0;
        WindowPlan packed;
        auto flush_packed = [&]() {
            if (!cur.empty()) packed.add_window(cur, (int)used);
            cur.clear();
            used = 0;
            if (!packed.ws.empty()) segs.push_back(Segment{-1, packed});
            packed = WindowPlan();
        };
        for (int j = 0; j < n_jobs; j++) {
            const rt_settings* st = jobs[j].settings;
            const int64_t npix = (int64_t)st->width * st->height;
            const int spp = render_spp(st);
            if (npix * spp >= 2 * N0) {
                flush_packed();
                Segment sg{j, WindowPlan()};
                const int K = (int)std::max<int64_t>(1, std::min<int64_t>(spp, cap / npix));
                if ((int64_t)K * npix > (1ll << 31) - 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "image too large");
                for (int base = 0; base < spp; base += K) {
                    const int passes = std::min(K, spp - base);
                    sg.plan.add_window({PartDev{j, base, 0, (int)npix, passes, {0, 0, 0}}}, passes * (int)npix);
                }
                segs.push_back(sg);
                continue;
            }
            int pass = 0;
            while (pass < spp) {
                int64_t room = (cap - used) / npix;
                if (room <= 0 && !cur.empty()) {
                    packed.add_window(cur, (int)used);
                    cur.clear();
                    used = 0;
                    continue;
                }
                int np = (int)std::max<int64_t>(1, std::min<int64_t>(room, spp - pass));
                if (used + np * npix > (1ll << 31) - 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "image too large");
                cur.push_back(PartDev{j, pass, (int)used, (int)npix, np, {0, 0, 0}});
                used += np * npix;
                pass += np;
            }
        }
        flush_packed();
    }
    int64_t WS = 1;
    for (const Segment& sg : segs)
        for (int w : sg.plan.ws) WS = std::max<int64_t>(WS, w);
    const int64_t N = std::min<int64_t>(N0, WS);
    if (N > (1ll << 30) || (int64_t)maxs * N > (1ll << 31) - 1)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "path pool too large");
    bool any_fp = false, any_hier = false;
    for (int j = 0; j < n_jobs; j++) {
        any_fp |= !jobs[j].scene->hs.images.empty() && jobs[j].settings->texture_filtering != 2;
        any_hier |= jobs[j].scene->hs.hier;
    }
    PathBufs P;
    rc = alloc_paths(ctx, N, WS, maxs, aovs, false, any_fp, any_hier, P);
    if (rc) return rc;
    // per-job films, one buffer
    std::vector<FilmBufs> F(n_jobs);
    std::vector<size_t> foff(n_jobs + 1, 0);
    for (int j = 0; j < n_jobs; j++) {
        const int64_t npix = (int64_t)jobs[j].settings->width * jobs[j].settings->height;
        foff[j + 1] = foff[j] + ((film_layout(nullptr, npix, F[j], nullptr) + 255) & ~(size_t)255);
    }
    rc = grow(ctx, &ctx->film_mem, &ctx->film_bytes, foff[n_jobs] + 256);
    if (rc) return rc;
    for (int j = 0; j < n_jobs; j++) {
        size_t acc = 0;
        const int64_t npix = (int64_t)jobs[j].settings->width * jobs[j].settings->height;
        film_layout((char*)ctx->film_mem + foff[j], npix, F[j], &acc);
        HIP_TRY(hipMemsetAsync((char*)ctx->film_mem + foff[j], 0, acc, ctx->stream));
    }
    // job table
    std::vector<JobDev> jd(n_jobs);
    std::vector<RParams> R(n_jobs);
    for (int j = 0; j < n_jobs; j++) {
        rc = upload_light_dist(ctx, jobs[j].scene, jobs[j].settings->light_sampling);
        if (rc) return rc;
        jd[j].S = dev_view(jobs[j].scene);
        R[j] = make_params(jobs[j].settings, jobs[j].scene->hs, maxs);
        jd[j].R = R[j];
    }
    rc = grow(ctx, &ctx->jobs_mem, &ctx->jobs_bytes, sizeof(JobDev) * n_jobs);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(ctx->jobs_mem, jd.data(), sizeof(JobDev) * n_jobs, hipMemcpyHostToDevice));
    HIP_TRY(hipMemsetAsync(P.tcount, 0, 128, ctx->stream));
    HIP_TRY(hipMemsetAsync(P.counters, 0, 256, ctx->stream));
    hipEvent_t* ev = ctx->ev;
    HIP_TRY(hipEventRecord(ev[6], ctx->stream));
    LoopTiming tm;
    ManyJobs many{(const JobDev*)ctx->jobs_mem, 0, 0, 1};
    for (int j = 0; j < n_jobs; j++) {
        many.thin_film |= jd[j].S.thin_film;
        many.hier |= jd[j].S.xf != nullptr ? 1 : 0;
        if (R[j].picked) many.all_lights = 0;
    }
    auto accum = [&](const PartDev& pt) {
        const int64_t npix = pt.npix;
        hipLaunchKernelGGL(k_accum, dim3((unsigned)((npix + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                           R[pt.job], P, F[pt.job], pt.pass0, pt.npass, pt.off);
    };
    for (const Segment& sg : segs) {
        if (sg.job >= 0) {
            OneJob one{jd[sg.job].S, R[sg.job]};
            rc = run_windows(ctx, one, P, maxs, jobs[sg.job].settings->max_depth > 0, counting, stack_depth, nullptr,
                             nullptr, sg.plan, accum, tm);
        } else {
            rc = run_windows(ctx, many, P, maxs, traced, counting, stack_depth, nullptr, nullptr, sg.plan, accum, tm);
        }
        if (rc) return rc;
    }
    for (int j = 0; j < n_jobs; j++) {
        rc = finish_film(ctx, R[j], F[j], jobs[j].settings, jobs[j].film);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(ev[7], ctx->stream));
    HIP_TRY(hipEventSynchronize(ev[7]));
    float dev_ms = 0;
    hipEventElapsedTime(&dev_ms, ev[6], ev[7]);
    int32_t ovf = 0;
    HIP_TRY(hipMemcpy(&ovf, P.counters + kCntOverflow, 4, hipMemcpyDeviceToHost));
    if (ovf) return fail(ctx, RT_ERR_INTERNAL, "traversal stack overflow");
    int64_t samples = 0;
    for (int j = 0; j < n_jobs; j++) {
        int64_t sj = (int64_t)jobs[j].settings->width * jobs[j].settings->height * render_spp(jobs[j].settings);
        samples += sj;
        if (jobs[j].stats) {
            std::memset(jobs[j].stats, 0, sizeof(rt_render_stats));
            jobs[j].stats->camera_samples = sj;
        }
    }
    if (total) {
        std::memset(total, 0, sizeof(*total));
        total->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        total->device_ms = dev_ms;
        total->trace_ms = tm.trace;
        total->shade_ms = tm.shade;
        total->shadow_ms = tm.shadow;
        total->camera_samples = samples;
        unsigned long long tcv[16];
        HIP_TRY(hipMemcpy(tcv, P.tcount, sizeof(tcv), hipMemcpyDeviceToHost));
        fill_counts(tcv, counting, tm.launches, total);
    }
    return RT_OK;
}


int rt_context_tune(rt_ctx* ctx, int32_t shade_waves_per_simd, int32_t traversal_stack, int32_t bounce_graph) {
    if (!ctx) return RT_ERR_INVALID_ARGUMENT;
    if (shade_waves_per_simd < 0 || shade_waves_per_simd > 4 || traversal_stack < 0 || traversal_stack > 2 ||
        bounce_graph < 0 || bounce_graph > 2)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "shade waves per SIMD in [0, 4], traversal stack and graph in [0, 2]");
    if (shade_waves_per_simd > 0) ctx->shade_wpe = shade_waves_per_simd;
    if (traversal_stack > 0) ctx->priv_stack = traversal_stack == 1;
    if (bounce_graph > 0) ctx->use_graph = bounce_graph == 2;
    return RT_OK;
}

}  // extern "C"
