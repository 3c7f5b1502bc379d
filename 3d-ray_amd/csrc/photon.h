// photon.h — caustic photon pre-pass and photon-map gather on gfx950.
//
//   k_photon_trace   one lane per emitted photon: light emission
//                    (CausticPhotonTracer.TryEmit) and the specular walk
//                    (CausticPhotonTracer.Trace) to the first diffuse/glossy
//                    hit; the deposited photon goes to slot g (global photon
//                    index), so the photon order is the emission order
//   compaction       stable (exclusive scan of the deposit flags)
//   grid             origin = component-wise min, Teschner hash key per
//                    photon, stable LSD radix sort of (key, photon index),
//                    CSR bucket starts — the same bucket order as the host's
//                    counting sort (PhotonMap.cs:39-77)
//   k_gather         k-nearest expanding-shell search (PhotonMap.cs:209-317)
//                    and the cone-filtered density estimate
//                    (Renderer.GatherCaustics, Renderer.cs:1449-1526)
#pragma once

namespace rt {

constexpr int kGatherK = 40;          // Renderer.CausticGatherK
constexpr int kPhotonMaxBounces = 24; // CausticPhotonTracer.MaxPhotonBounces
constexpr int kPhotonRrStart = 4;     // CausticPhotonTracer.RrStartBounce

struct PhotonSrc {
    int light, count, offset, pad;
};

// ── emission (CausticPhotonTracer.cs:279-345) ──────────────────────────────
__device__ __forceinline__ f3 photon_face_toward(f3 n, f3 from, f3 bc) { return dot(n, bc - from) < 0.0f ? -n : n; }

__device__ __forceinline__ f3 photon_cosine_dir(f3 n, Rng& rng) {
    f3 d = n + random_unit_vector(rng);
    if (length_sq(d) < 1e-8f) d = n;
    return normalize(d);
}

__device__ bool photon_emit(const GLight& L, f3 bmn, f3 bmx, Rng& rng, f3& o, f3& d, f3& flux) {
    const int kind = L.l0.x;
    const f3 col = f3of(L.l1);
    const float I = L.l1.w;
    f3 bc = 0.5f * (bmn + bmx);
    if (kind == RT_LIGHT_SPHERE) {
        f3 dir = random_unit_vector(rng);
        f3 p = f3of(L.l2) + dir * L.l2.w;
        o = offset_origin(p, dir);
        d = photon_cosine_dir(dir, rng);
        flux = col * I * (4.0f * kPi);
        return true;
    }
    if (kind == RT_LIGHT_AREA) {
        f3 q = f3of(L.l2), u = f3of(L.l3), v = f3of(L.l4);
        float r1 = rng.next();
        float r2 = rng.next();
        f3 p = (q + r1 * u) + r2 * v;
        f3 n = photon_face_toward(f3of(L.l5), (q + 0.5f * u) + 0.5f * v, bc);
        o = offset_origin(p, n);
        d = photon_cosine_dir(n, rng);
        flux = col * I * L.l2.w * kPi;
        return true;
    }
    if (kind == RT_LIGHT_DIRECTIONAL) {
        float r = fmaxf(0.5f * length(bmx - bmn), 1e-3f);
        f3 w = f3of(L.l2);
        f3 a = fabsf(w.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
        f3 t = normalize(cross(a, w));
        f3 b = cross(w, t);
        float rr = sqrtf(rng.next());
        float ang = 2.0f * kPi * rng.next();
        float dx = rr * cosf(ang), dy = rr * sinf(ang);
        o = (bc - w * (r * 1.5f)) + (t * dx + b * dy) * r;
        d = w;
        flux = col * I * (kPi * r * r);
        return true;
    }
    // geometry light: uniform point on the emitter (Triangle/Quad/Sphere.Sample)
    f3 p, n;
    float ar = L.l5.w;
    int sub = L.l0.w;
    if (sub == kGeoSphere) {
        f3 dir = random_unit_vector(rng);
        p = f3of(L.l2) + dir * L.l2.w;
        n = dir;
    } else if (sub == kGeoTri) {
        float xi1 = rng.next(), xi2 = rng.next();
        float r1 = sqrtf(xi1);
        float uu = 1.0f - r1, vv = xi2 * r1;
        p = f3of(L.l2) + uu * f3of(L.l3) + vv * f3of(L.l4);
        n = f3of(L.l5);
    } else {
        float uu = rng.next(), vv = rng.next();
        p = (f3of(L.l2) + uu * f3of(L.l3)) + vv * f3of(L.l4);
        n = f3of(L.l5);
    }
    if (col.x <= 0.0f && col.y <= 0.0f && col.z <= 0.0f) return false;
    n = photon_face_toward(n, p, bc);
    o = offset_origin(p, n);
    d = photon_cosine_dir(n, rng);
    flux = col * ar * kPi;
    return true;
}

// ── specular walk (CausticPhotonTracer.cs:134-231) ─────────────────────────
template <int SD>
__global__ void __launch_bounds__(kBlock) k_photon_trace(DevScene S, const PhotonSrc* __restrict__ src, int nsrc,
                                                         int ntot, f3 bmn, f3 bmx, float4* __restrict__ out_pos,
                                                         float4* __restrict__ out_inc, float4* __restrict__ out_pow,
                                                         int32_t* __restrict__ flag, uint32_t* overflow) {
    uint2 stk[SD];
    TravCount tc{0u, 0u, 0u};
    int g = xcd_block() * kBlock + threadIdx.x;
    if (g >= ntot) return;
    int lo = 0, hi = nsrc - 1;
    while (lo < hi) {   // last source with offset <= g
        int mid = (lo + hi + 1) / 2;
        if (src[mid].offset <= g) lo = mid;
        else hi = mid - 1;
    }
    const PhotonSrc ps = src[lo];
    Rng rng{(uint32_t)(g - ps.offset), ((uint32_t)lo * 0x9E3779B1u) ^ 0xC2B2AE35u, 0u, false};
    const GLight L = S.lights[ps.light];
    f3 o, d, power;
    bool dep = false;
    f3 dpos = splat(0.0f), dinc = splat(0.0f);
    if (photon_emit(L, bmn, bmx, rng, o, d, power)) {
        power = power * (splat(1.0f) / (float)ps.count);
        float emax = fmaxf(power.x, fmaxf(power.y, power.z));
        if (emax > 0.0f) {
            float inv_emax = 1.0f / emax;
            int spec = 0;
            bool in_med = false;
            f3 sig = splat(0.0f), entry = splat(0.0f);
            for (int bounce = 0; bounce < kPhotonMaxBounces; bounce++) {
                HitOut h = trace_closest<SD, 1, false>(S.nodes, S.entries, S.aux, S.top_start, S.top_count, o, d, kEps,
                                                       kInfinity, stk, overflow, tc);
                if (h.entry < 0) break;
                Surf s = make_surf(S, h.entry, o, d, h.t, h.u, h.v);
                if (in_med) {
                    float dd = length(s.p - entry);
                    power = mk(power.x * expf(-sig.x * dd), power.y * expf(-sig.y * dd), power.z * expf(-sig.z * dd));
                    in_med = false;
                }
                const GMat gm = S.mats[s.mat];
                const int kind = gm.m0.x;
                f3 view = normalize(-d);
                bool follow = false, deposit = false;
                f3 att = splat(0.0f), wo = splat(0.0f), org = splat(0.0f);
                if (kind == RT_MAT_LAMBERTIAN) {
                    // Lambertian.Sample is non-delta; its Scatter fallback too
                    deposit = true;
                } else if (kind == RT_MAT_DISNEY) {
                    DisneyP DP = load_disney(gm);
                    BsdfS bs;
                    if (disney_sample(DP, S.lut_e, view, s, rng, bs)) {
                        if (bs.delta && bs.caster) {
                            follow = true;
                            att = bs.f;
                            wo = bs.wo;
                            f3 od = dot(wo, s.n) >= 0.0f ? s.n : -s.n;
                            org = offset_origin(s.p, od);
                        } else deposit = true;
                    } else {
                        int lobe;
                        f3 a2, w2;
                        // legacy Scatter: DisneyBsdf is not a delta scatterer
                        if (disney_scatter(DP, S.lut_e, s, d, rng, a2, w2, lobe)) deposit = true;
                    }
                } else if (kind == RT_MAT_DIELECTRIC) {
                    dielectric_scatter(gm, s, d, rng, att, wo, org);
                    follow = true;
                }
                if (deposit) {
                    if (spec >= 1 && (power.x > 0.0f || power.y > 0.0f || power.z > 0.0f)) {
                        dep = true;
                        dpos = s.p;
                        dinc = normalize(-d);
                    }
                    break;
                }
                if (!follow) break;   // absorbed (emitters, failed samples)
                power *= att;
                // EnterMediumIfRefracting
                if (dot(wo, s.n) < 0.0f && s.front) {
                    f3 sa = mat_shadow_abs(gm);
                    if (sa.x > 0.0f || sa.y > 0.0f || sa.z > 0.0f) {
                        in_med = true;
                        sig = sa;
                        entry = s.p;
                    }
                }
                o = org;
                d = wo;
                spec++;
                if (spec >= kPhotonRrStart) {
                    float survive = fminf(1.0f, fmaxf(power.x, fmaxf(power.y, power.z)) * inv_emax);
                    if (survive <= 1e-6f || rng.next() >= survive) break;
                    power = power / survive;
                }
            }
        }
    }
    flag[g] = dep ? 1 : 0;
    if (dep) {
        out_pos[g] = make_float4(dpos.x, dpos.y, dpos.z, 0.0f);
        out_inc[g] = make_float4(dinc.x, dinc.y, dinc.z, 0.0f);
        out_pow[g] = make_float4(power.x, power.y, power.z, 0.0f);
    }
}

// ── scans, compaction, grid build ──────────────────────────────────────────
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

// Exclusive scan of one tile per block; writes the tile total to bsum.
__global__ void __launch_bounds__(kScanBlock) k_scan_tile(const int32_t* __restrict__ in, int32_t* __restrict__ out,
                                                          int n, int32_t* __restrict__ bsum) {
    __shared__ int32_t wsum[kScanBlock / 64];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int t = 0;
    for (int k = 0; k < kScanItems; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        t += v[k];
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = t;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    int wpre = 0;
    for (int w = 0; w < wid; w++) wpre += wsum[w];
    int run = wpre + incl - t;
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = wpre + incl;
}

// Exclusive scan of the tile totals in one block (any count), in place;
// total -> bsum[nb].
__global__ void __launch_bounds__(1024) k_scan_totals(int32_t* bsum, int nb) {
    __shared__ int32_t part[1024 / 64];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int b0 = 0; b0 < nb; b0 += 1024) {
        int i = b0 + threadIdx.x;
        int x = i < nb ? bsum[i] : 0;
        int incl = x;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(incl, off, 64);
            if (lane >= off) incl += y;
        }
        if (lane == 63) part[wid] = incl;
        __syncthreads();
        int wpre = 0;
        for (int w = 0; w < wid; w++) wpre += part[w];
        int c = carry;
        if (i < nb) bsum[i] = c + wpre + incl - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry = c + wpre + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_add(int32_t* __restrict__ out, int n,
                                                         const int32_t* __restrict__ bsum) {
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const int add = bsum[blockIdx.x];
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) out[base + k] += add;
}

// Stable compaction of the deposited photons (slot order = photon order).
__global__ void __launch_bounds__(kBlock) k_photon_compact(int ntot, const int32_t* __restrict__ flag,
                                                           const int32_t* __restrict__ pos_of,
                                                           const float4* __restrict__ ip, const float4* __restrict__ ii,
                                                           const float4* __restrict__ iw, float4* __restrict__ op,
                                                           float4* __restrict__ oi, float4* __restrict__ ow) {
    int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= ntot || !flag[g]) return;
    int k = pos_of[g];
    op[k] = ip[g];
    oi[k] = ii[g];
    ow[k] = iw[g];
}

// Component-wise min of the photon positions (exact in any order):
// per-block partials, then one block folds the partials.
__global__ void __launch_bounds__(256) k_min3(const float4* __restrict__ p, int n, float4* __restrict__ part) {
    __shared__ float sm[3][256];
    float mx = __int_as_float(0x7f800000), my = mx, mz = mx;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float4 q = p[i];
        mx = fminf(mx, q.x);
        my = fminf(my, q.y);
        mz = fminf(mz, q.z);
    }
    sm[0][threadIdx.x] = mx;
    sm[1][threadIdx.x] = my;
    sm[2][threadIdx.x] = mz;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 3; c++) sm[c][threadIdx.x] = fminf(sm[c][threadIdx.x], sm[c][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = make_float4(sm[0][0], sm[1][0], sm[2][0], 0.0f);
}

struct GridParams {
    float cell, inv_cell;
    uint32_t mask;
    int n;
};

// PhotonMap.CellOf / Bucket (PhotonMap.cs:79-97)
__device__ __forceinline__ void photon_cell(f3 p, f3 org, float inv, int& x, int& y, int& z) {
    x = (int)floorf((p.x - org.x) * inv);
    y = (int)floorf((p.y - org.y) * inv);
    z = (int)floorf((p.z - org.z) * inv);
}
__device__ __forceinline__ uint32_t photon_bucket(int x, int y, int z, uint32_t mask) {
    uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u);
    return h & mask;
}

__global__ void __launch_bounds__(kBlock) k_photon_keys(const float4* __restrict__ p, const float4* __restrict__ org4,
                                                        GridParams G, uint32_t* __restrict__ key,
                                                        int32_t* __restrict__ val, int32_t* __restrict__ count) {
    int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= G.n) return;
    f3 org = f3of(org4[0]);
    int x, y, z;
    photon_cell(f3of(p[i]), org, G.inv_cell, x, y, z);
    uint32_t k = photon_bucket(x, y, z, G.mask);
    key[i] = k;
    val[i] = i;
    atomicAdd(&count[k], 1);
}

// ── stable LSD radix sort, 8-bit digits, 1024-key tiles ────────────────────
// Histogram of one digit per tile: hist[digit * ntiles + tile].
__global__ void __launch_bounds__(256) k_rs_hist(const uint32_t* __restrict__ key, int n, int shift,
                                                 int32_t* __restrict__ hist) {
    __shared__ int32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * 1024;
    for (int r = 0; r < 4; r++) {
        int i = base + r * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Scatter with the stable rank of each key inside its tile: the tile is
// walked in 4 rounds of 256 keys in index order; within a round a key's rank
// among equal digits is (equal digits in earlier waves) + (equal digits in
// lower lanes of its wave), found with 8 ballots.
__global__ void __launch_bounds__(256) k_rs_scatter(const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin,
                                                    uint32_t* __restrict__ kout, int32_t* __restrict__ vout, int n,
                                                    int shift, const int32_t* __restrict__ offs) {
    __shared__ int32_t run[256];
    __shared__ int32_t wcnt[4][256];
    run[threadIdx.x] = offs[threadIdx.x * gridDim.x + blockIdx.x];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int base = blockIdx.x * 1024;
    for (int r = 0; r < 4; r++) {
        for (int w = 0; w < 4; w++) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        int i = base + r * 256 + threadIdx.x;
        bool valid = i < n;
        uint32_t k = valid ? kin[i] : 0u;
        uint32_t dg = (k >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            bool bit = (dg >> b) & 1u;
            unsigned long long m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wid][dg] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pre = 0;
            for (int w = 0; w < wid; w++) pre += wcnt[w][dg];
            int pos = run[dg] + pre + rank;
            kout[pos] = k;
            vout[pos] = vin[i];
        }
        __syncthreads();
        run[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] + wcnt[2][threadIdx.x] + wcnt[3][threadIdx.x];
        __syncthreads();
    }
}

// ── gather (PhotonMap.GatherKNearest + Renderer.GatherCaustics) ────────────
struct DevPhotons {
    const float4* pos;      // photon order (the deposited set as traced)
    const float4* inc;
    const float4* pw;
    const float4* spos;     // the same records in bucket order: entry kk is
    const float4* sinc;     // photon idx[kk], so a cell's photons are one
    const float4* spw;      // contiguous run of start[b] .. start[b+1]
    const int32_t* start;
    const int32_t* idx;
    const float4* origin;   // device scalar (x, y, z)
    float cell, inv_cell, max_r;
    uint32_t mask;
    int n;
};

// Photon records in bucket order (a gather of the three planes by idx).
__global__ void __launch_bounds__(kBlock) k_photon_order(int n, const int32_t* __restrict__ idx,
                                                         const float4* __restrict__ ip, const float4* __restrict__ ii,
                                                         const float4* __restrict__ iw, float4* __restrict__ op,
                                                         float4* __restrict__ oi, float4* __restrict__ ow) {
    int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    int pi = idx[k];
    op[k] = ip[pi];
    oi[k] = ii[pi];
    ow[k] = iw[pi];
}

// Max-heap of the k nearest photons with PhotonMap's insertion order and
// sift rules (PhotonMap.cs:296-317).  Entry j is (squared distance, record
// index) at h[j * S]: the gather keeps one heap per lane in LDS, lane-
// interleaved (S = the block size), one 8-byte access per entry.  The sifts
// move a hole instead of swapping pairs; every comparison is the one the
// swapping form makes (the moving entry keeps its distance), so the layout
// after each insertion is the same.
template <int S>
__device__ __forceinline__ void heap_insert(float2* h, int& count, int id, float d2) {
    const float2 e = make_float2(d2, __int_as_float(id));
    if (count < kGatherK) {
        int i = count++;
        while (i > 0) {
            const int parent = (i - 1) >> 1;
            const float2 hp = h[parent * S];
            if (hp.x >= d2) break;
            h[i * S] = hp;
            i = parent;
        }
        h[i * S] = e;
    } else if (d2 < h[0].x) {
        int i = 0;
        for (;;) {
            const int l = 2 * i + 1, r = 2 * i + 2;
            int lg = i;
            float hlg = d2;
            float2 elg = e;
            if (l < count) {
                const float2 hl = h[l * S];
                if (hl.x > hlg) { lg = l; hlg = hl.x; elg = hl; }
            }
            if (r < count) {
                const float2 hr = h[r * S];
                if (hr.x > hlg) { lg = r; hlg = hr.x; elg = hr; }
            }
            if (lg == i) break;
            h[i * S] = elg;
            i = lg;
        }
        h[i * S] = e;
    }
}

// k nearest photons within max_r of p (PhotonMap.GatherKNearest,
// PhotonMap.cs:209-294): rings of cells by Chebyshev distance, cells beyond
// the radius cap or the current k-th distance skipped, stop once the k-th
// distance is inside the scanned reach.  Returns the count and the kernel
// radius (distance of the farthest gathered photon); the heap holds
// bucket-order record indices.
template <int S>
__device__ int photon_knn(const DevPhotons& M, f3 p, float2* h, float& radius) {
    radius = 0.0f;
    if (M.n <= 0) return 0;
    f3 org = f3of(M.origin[0]);
    int cx, cy, cz;
    photon_cell(p, org, M.inv_cell, cx, cy, cz);
    const float max_r2 = M.max_r * M.max_r;
    int count = 0;
    const int max_rings = (int)ceilf(M.max_r * M.inv_cell) + 1;
    for (int ring = 0; ring <= max_rings; ring++) {
        for (int z = cz - ring; z <= cz + ring; z++)
            for (int y = cy - ring; y <= cy + ring; y++) {
                // rows inside the shell hold only its two end cells
                const bool face = ring == 0 || abs(z - cz) == ring || abs(y - cy) == ring;
                const int xstep = face ? 1 : 2 * ring;
                for (int x = cx - ring; x <= cx + ring; x += xstep) {
                    float bx = org.x + (float)x * M.cell, by = org.y + (float)y * M.cell, bz = org.z + (float)z * M.cell;
                    float dx = fmaxf(fmaxf(bx - p.x, p.x - (bx + M.cell)), 0.0f);
                    float dy = fmaxf(fmaxf(by - p.y, p.y - (by + M.cell)), 0.0f);
                    float dz = fmaxf(fmaxf(bz - p.z, p.z - (bz + M.cell)), 0.0f);
                    float md2 = (dx * dx + dy * dy) + dz * dz;
                    if (md2 > max_r2) continue;
                    if (count >= kGatherK && md2 >= h[0].x) continue;
                    uint32_t b = photon_bucket(x, y, z, M.mask);
                    const int k0 = M.start[b], k1 = M.start[b + 1];
                    if (k0 >= k1) continue;
                    // the next record is in flight while this one is tested
                    float4 nxt = M.spos[k0];
                    for (int kk = k0; kk < k1; kk++) {
                        const f3 pos = f3of(nxt);
                        if (kk + 1 < k1) nxt = M.spos[kk + 1];
                        int px, py, pz;
                        photon_cell(pos, org, M.inv_cell, px, py, pz);
                        if (px != x || py != y || pz != z) continue;
                        float d2 = length_sq(pos - p);
                        if (d2 > max_r2) continue;
                        heap_insert<S>(h, count, kk, d2);
                    }
                }
            }
        float reach = (float)ring * M.cell;
        if (count >= kGatherK && h[0].x <= reach * reach) break;
        if (reach > M.max_r) break;
    }
    radius = count > 0 ? sqrtf(h[0].x) : 0.0f;
    return count;
}

// One k-nearest search per query point, results in heap order as photon-
// order indices (the record's idx entry), -1 past the count.
__global__ void __launch_bounds__(kBlock) k_photon_knn_query(DevPhotons M, const float* __restrict__ pts, int n,
                                                             int32_t* __restrict__ cnt, float* __restrict__ rad,
                                                             int32_t* __restrict__ out) {
    __shared__ float2 lds_h[kGatherK * kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    f3 p;
    p.x = pts[3 * i];
    p.y = pts[3 * i + 1];
    p.z = pts[3 * i + 2];
    float2* h = &lds_h[threadIdx.x];
    float r;
    const int c = photon_knn<kBlock>(M, p, h, r);
    cnt[i] = c;
    rad[i] = r;
    for (int j = 0; j < kGatherK; j++)
        out[(int64_t)i * kGatherK + j] = j < c ? M.idx[__float_as_int(h[j * kBlock].y)] : -1;
}

}  // namespace rt
