// film_io.hip — film file encoders for gfx950 and their C ABI (include/rt3d.h):
//
//   rt_exr_encode   ExrImage.Write  (src/RayTracer/Rendering/ExrImage.cs:52-137):
//                   single-part scanline OpenEXR v2, ZIP (16-line blocks, zlib
//                   framing, raw fallback), channels in ordinal name order,
//                   half (clamped to +-65504, NaN kept) or float samples
//   rt_pfm_encode   PfmImage.Write  (src/RayTracer/Rendering/PfmImage.cs:29-65):
//                   "PF"/"Pf" header, little-endian, rows bottom-to-top
//   rt_png_encode   Program.SaveImage (src/RayTracer/Program.cs:723-756) for
//                   .png: 8-bit RGBA, byte = clamp((int)(clamp(v,0,1)*256),0,255)
//
// The device does the per-byte work: it reads the plane-major film planes
// (row 0 = top) and writes the file's pixel payload directly — for EXR the
// interleaved scanline bytes of each 16-line block already put through the
// ZIP pre-transform (even/odd byte split, then the delta predictor), for PFM
// the interleaved bottom-to-top rows, for PNG the filtered (filter type 0)
// RGBA rows.  The host runs zlib on the blocks in parallel and assembles the
// header, offset table and chunks.
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rt3d.h"

namespace rt {

// Context services defined next to rt_ctx (render.hip).
hipStream_t ctx_stream(rt_ctx* ctx);
int ctx_device(rt_ctx* ctx);
int ctx_fail(rt_ctx* ctx, int code, const std::string& msg);
int ctx_io_buffer(rt_ctx* ctx, size_t bytes, void** out);

constexpr int kExrLines = 16;        // ZIP block height (ExrImage.cs:41)
constexpr float kHalfMax = 65504.0f;  // largest finite half (ExrImage.cs:42)

// fp32 -> fp16 bits, IEEE round to nearest even with subnormals, after the
// writer's clamp to +-65504 (ExrImage.cs:234-236: Math.Clamp maps +-Inf to
// +-65504 and passes NaN through; a NaN stays a quiet NaN of the same sign).
__host__ __device__ inline uint32_t half_bits(float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return sign | 0x7e00u | ((a >> 13) & 0x3ffu);
    if (a > 0x477fe000u) a = 0x477fe000u;   // |v| > 65504 (incl. Inf) -> 65504
    if (a >= 0x38800000u) {                  // normal half: |v| >= 2^-14
        const uint32_t m = a & 0x7fffffu;
        uint32_t h = ((((a >> 23) - 112u)) << 10) | (m >> 13);
        const uint32_t rem = m & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
        return sign | h;
    }
    if (a >= 0x33000000u) {                  // subnormal half: 2^-25 <= |v| < 2^-14
        const uint32_t M = (a & 0x7fffffu) | 0x800000u;
        const int shift = 126 - (int)(a >> 23);   // 14 .. 24
        uint32_t h = M >> shift;
        const uint32_t rem = M & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (h & 1u))) h++;
        return sign | h;
    }
    return sign;
}

// Channel table of one EXR file in the device io buffer: scanline byte
// offset of each channel (ordinal name order, off[nch] = scanline bytes),
// bytes per sample (2 half, 4 float) and its plane.
struct ExrChan {
    int64_t off;
    int64_t bps;
    const float* plane;
    int64_t pad;
};

// Byte r of the raw (untransformed) payload of the block starting at line y0.
__device__ inline uint32_t exr_raw_byte(const ExrChan* ch, int nch, int64_t sb, int width, int y0, int64_t r) {
    const int64_t ly = r / sb;
    const int64_t pos = r - ly * sb;
    int lo = 0, hi = nch - 1;   // last channel with off <= pos
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ch[mid].off <= pos) lo = mid;
        else hi = mid - 1;
    }
    const ExrChan c = ch[lo];
    const int64_t q = pos - c.off;
    const int64_t x = c.bps == 2 ? (q >> 1) : (q >> 2);
    const int k = (int)(q - x * c.bps);
    const float v = c.plane[(int64_t)(y0 + ly) * width + x];
    uint32_t bits;
    if (c.bps == 2) bits = half_bits(v);
    else memcpy(&bits, &v, 4);
    return (bits >> (8 * k)) & 0xffu;
}

// ZIP-transformed payload of every block, blocks back to back (block b at
// b * 16 * sb): byte i of a block of L bytes is t[i] - t[i-1] + 128 (mod
// 256) over the split order t (even raw bytes, then odd), t[0] unchanged
// (ExrImage.cs:248-266).  Four bytes per thread.
__global__ void __launch_bounds__(256) k_exr_blocks(const ExrChan* ch, int nch, int64_t sb, int width, int height,
                                                    uint8_t* out, int64_t total) {
    const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (g0 >= total) return;
    const int64_t bbytes = sb * kExrLines;
    uint32_t word = 0;
    const int n = (int)min<int64_t>(4, total - g0);
    for (int j = 0; j < n; j++) {
        const int64_t g = g0 + j;
        const int64_t b = g / bbytes;
        const int64_t i = g - b * bbytes;
        const int y0 = (int)(b * kExrLines);
        const int64_t L = (int64_t)min(kExrLines, height - y0) * sb;
        const int64_t half = (L + 1) / 2;
        auto src = [&](int64_t t) { return t < half ? 2 * t : 2 * (t - half) + 1; };
        const uint32_t cur = exr_raw_byte(ch, nch, sb, width, y0, src(i));
        uint32_t o = cur;
        if (i > 0) o = (cur - exr_raw_byte(ch, nch, sb, width, y0, src(i - 1)) + 128u) & 0xffu;
        word |= o << (8 * j);
    }
    if (n == 4) {
        *(uint32_t*)(out + g0) = word;
    } else {
        for (int j = 0; j < n; j++) out[g0 + j] = (uint8_t)(word >> (8 * j));
    }
}

// PFM payload: out[((H-1-y) * W + x) * C + c] = planes[c*W*H + y*W + x].
__global__ void __launch_bounds__(256) k_pfm_rows(const float* planes, int width, int height, int nc, float* out) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = (int64_t)width * nc;
    if (o >= row * height) return;
    const int64_t rf = o / row;
    const int64_t rem = o - rf * row;
    const int64_t x = rem / nc;
    const int c = (int)(rem - x * nc);
    const int64_t y = height - 1 - rf;
    out[o] = planes[(int64_t)c * width * height + y * width + x];
}

__device__ inline uint32_t png_byte(float v) {
    // Program.cs:733-735; NaN (a NaN survives Math.Clamp) maps to 0
    if (!(v == v)) return 0u;
    const float c = fminf(fmaxf(v, 0.0f), 1.0f);
    const int q = (int)(c * 256.0f);
    return (uint32_t)min(max(q, 0), 255);
}

// PNG scanlines: per row one filter byte (0) then W RGBA pixels, alpha 255;
// display = 3 planes of W*H floats.  One pixel (or a row's filter byte) per
// thread: thread t of a row handles byte 1 + 4 (t - 1) .. for t >= 1.
__global__ void __launch_bounds__(256) k_png_rows(const float* display, int width, int height, uint8_t* out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per_row = (int64_t)width + 1;
    if (t >= per_row * height) return;
    const int64_t y = t / per_row;
    const int64_t k = t - y * per_row;
    uint8_t* row = out + y * (1 + 4 * (int64_t)width);
    if (k == 0) {
        row[0] = 0;
        return;
    }
    const int64_t x = k - 1;
    const int64_t n = (int64_t)width * height;
    const int64_t p = y * width + x;
    uint8_t* px = row + 1 + 4 * x;
    px[0] = (uint8_t)png_byte(display[p]);
    px[1] = (uint8_t)png_byte(display[n + p]);
    px[2] = (uint8_t)png_byte(display[2 * n + p]);
    px[3] = 255;
}

// ── host side ──────────────────────────────────────────────────────────────

struct Bytes {
    std::vector<uint8_t> v;
    void raw(const void* p, size_t n) {
        const uint8_t* b = (const uint8_t*)p;
        v.insert(v.end(), b, b + n);
    }
    void i32(int32_t x) { raw(&x, 4); }
    void u32(uint32_t x) { raw(&x, 4); }
    void i64(int64_t x) { raw(&x, 8); }
    void f32(float x) { raw(&x, 4); }
    void u8(uint8_t x) { v.push_back(x); }
    void str0(const char* s) { raw(s, strlen(s) + 1); }
    void be32(uint32_t x) {
        u8((uint8_t)(x >> 24));
        u8((uint8_t)(x >> 16));
        u8((uint8_t)(x >> 8));
        u8((uint8_t)x);
    }
};

static int worker_count(int64_t items) {
    unsigned hc = std::thread::hardware_concurrency();
    int n = (int)std::min<int64_t>(items, std::max(1u, std::min(hc, 16u)));
    return std::max(1, n);
}

// Runs fn(i) for i in [0, n) on a few threads (ExrImage.cs:85 uses Parallel.For).
template <class F>
static void parallel_for(int64_t n, F fn) {
    const int nt = worker_count(n);
    if (nt <= 1) {
        for (int64_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int64_t i = t; i < n; i += nt) fn(i);
        });
    for (auto& x : th) x.join();
}

// Inverse of the ZIP pre-transform of one block (ExrImage.cs:272-287): the
// raw-block fallback stores the untransformed bytes.
static void zip_reconstruct(const uint8_t* enc, int64_t n, uint8_t* raw) {
    std::vector<uint8_t> t(enc, enc + n);
    for (int64_t i = 1; i < n; i++) t[i] = (uint8_t)(t[i - 1] + t[i] - 128);
    const int64_t half = (n + 1) / 2;
    int64_t i1 = 0, i2 = half;
    for (int64_t s = 0; s < n;) {
        raw[s++] = t[i1++];
        if (s < n) raw[s++] = t[i2++];
    }
}

struct SortedChan {
    std::string name;
    int type;
    const float* plane;
};

// Header of ExrImage.BuildHeader (ExrImage.cs:158-212): the required
// attributes in alphabetical order plus Rec.709 chromaticities.
static std::vector<uint8_t> exr_header(int width, int height, const std::vector<SortedChan>& ch) {
    Bytes chl;
    for (const auto& c : ch) {
        chl.str0(c.name.c_str());
        chl.i32(c.type);
        chl.u8(0);   // pLinear
        chl.u8(0);
        chl.u8(0);
        chl.u8(0);
        chl.i32(1);   // xSampling
        chl.i32(1);   // ySampling
    }
    chl.u8(0);
    Bytes cr;
    for (float f : {0.6400f, 0.3300f, 0.3000f, 0.6000f, 0.1500f, 0.0600f, 0.3127f, 0.3290f}) cr.f32(f);
    Bytes box;
    box.i32(0);
    box.i32(0);
    box.i32(width - 1);
    box.i32(height - 1);
    Bytes h;
    auto attr = [&](const char* name, const char* type, const std::vector<uint8_t>& payload) {
        h.str0(name);
        h.str0(type);
        h.i32((int32_t)payload.size());
        h.raw(payload.data(), payload.size());
    };
    Bytes one;
    one.f32(1.0f);
    attr("channels", "chlist", chl.v);
    attr("chromaticities", "chromaticities", cr.v);
    attr("compression", "compression", {3});   // ZIP
    attr("dataWindow", "box2i", box.v);
    attr("displayWindow", "box2i", box.v);
    attr("lineOrder", "lineOrder", {0});       // INCREASING_Y
    attr("pixelAspectRatio", "float", one.v);
    attr("screenWindowCenter", "v2f", std::vector<uint8_t>(8, 0));
    attr("screenWindowWidth", "float", one.v);
    h.u8(0);
    return h.v;
}

#define IO_TRY(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) return ctx_fail(ctx, RT_ERR_DEVICE, hipGetErrorString(e_)); \
    } while (0)

}  // namespace rt

using namespace rt;

extern "C" {

int rt_exr_encode(rt_ctx* ctx, int32_t width, int32_t height, const rt_exr_channel* channels, int32_t n_channels,
                  uint8_t* out, int64_t capacity, int64_t* size) {
    if (!ctx || !size) return RT_ERR_INVALID_ARGUMENT;
    if (width <= 0 || height <= 0)
        return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "invalid EXR dimensions");
    if (n_channels <= 0 || !channels) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "EXR requires at least one channel");
    std::vector<SortedChan> ch;
    for (int i = 0; i < n_channels; i++) {
        const char* nm = channels[i].name;
        const size_t len = nm ? strlen(nm) : 0;
        bool ok = len > 0 && len <= 31;
        for (size_t k = 0; ok && k < len; k++) ok = nm[k] > ' ' && nm[k] <= '~';
        if (!ok) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, std::string("invalid EXR channel name '") + (nm ? nm : "") + "'");
        if (channels[i].type != RT_EXR_HALF && channels[i].type != RT_EXR_FLOAT)
            return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "EXR pixel type must be half (1) or float (2)");
        if (!channels[i].plane) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "EXR channel plane is null");
        ch.push_back(SortedChan{std::string(nm), channels[i].type, channels[i].plane});
    }
    // ordinal name order (ExrImage.cs:77)
    std::stable_sort(ch.begin(), ch.end(), [](const SortedChan& a, const SortedChan& b) {
        return strcmp(a.name.c_str(), b.name.c_str()) < 0;
    });
    for (size_t i = 1; i < ch.size(); i++)
        if (ch[i].name == ch[i - 1].name)
            return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "duplicate EXR channel name '" + ch[i].name + "'");

    const int nch = (int)ch.size();
    int64_t sb = 0;
    std::vector<ExrChan> tab(nch);
    for (int i = 0; i < nch; i++) {
        tab[i].off = sb;
        tab[i].bps = ch[i].type == RT_EXR_HALF ? 2 : 4;
        sb += tab[i].bps * (int64_t)width;
    }
    const int64_t nblocks = (height + kExrLines - 1) / kExrLines;
    const int64_t total = sb * height;
    std::vector<uint8_t> header = exr_header(width, height, ch);
    const int64_t bound = 8 + (int64_t)header.size() + 8 * nblocks + 8 * nblocks + total;
    if (!out) {
        *size = bound;
        return RT_OK;
    }
    if (capacity < bound) {
        *size = bound;
        return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "EXR output buffer smaller than the bound (query with out = NULL)");
    }

    IO_TRY(hipSetDevice(ctx_device(ctx)));
    hipStream_t st = ctx_stream(ctx);
    const int64_t npix = (int64_t)width * height;
    const size_t tab_bytes = (sizeof(ExrChan) * nch + 255) & ~(size_t)255;
    const size_t plane_bytes = (size_t)npix * 4;
    const size_t planes_off = tab_bytes;
    const size_t out_off = planes_off + (((size_t)nch * plane_bytes + 255) & ~(size_t)255);
    void* mem = nullptr;
    int rc = ctx_io_buffer(ctx, out_off + (size_t)total + 256, &mem);
    if (rc) return rc;
    char* base = (char*)mem;
    for (int i = 0; i < nch; i++) {
        float* d = (float*)(base + planes_off + (size_t)i * plane_bytes);
        IO_TRY(hipMemcpyAsync(d, ch[i].plane, plane_bytes, hipMemcpyHostToDevice, st));
        tab[i].plane = d;
    }
    IO_TRY(hipMemcpyAsync(base, tab.data(), sizeof(ExrChan) * nch, hipMemcpyHostToDevice, st));
    uint8_t* d_out = (uint8_t*)(base + out_off);
    const int64_t threads = (total + 3) / 4;
    hipLaunchKernelGGL(k_exr_blocks, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, (const ExrChan*)base,
                       nch, sb, width, height, d_out, total);
    IO_TRY(hipGetLastError());
    std::vector<uint8_t> enc((size_t)total);
    IO_TRY(hipMemcpyAsync(enc.data(), d_out, (size_t)total, hipMemcpyDeviceToHost, st));
    IO_TRY(hipStreamSynchronize(st));

    // zlib per block (ZLibStream framing, CompressionLevel.Optimal); a block
    // that does not shrink is stored raw (ExrImage.cs:111-113)
    std::vector<std::vector<uint8_t>> payload((size_t)nblocks);
    std::vector<int> zerr((size_t)nblocks, Z_OK);
    parallel_for(nblocks, [&](int64_t b) {
        const int64_t off = b * kExrLines * sb;
        const int64_t L = std::min<int64_t>(kExrLines, height - b * kExrLines) * sb;
        uLongf zn = compressBound((uLong)L);
        std::vector<uint8_t> z(zn);
        int e = compress2(z.data(), &zn, enc.data() + off, (uLong)L, Z_DEFAULT_COMPRESSION);
        if (e != Z_OK) {
            zerr[b] = e;
            return;
        }
        if ((int64_t)zn < L) {
            z.resize(zn);
            payload[b] = std::move(z);
        } else {
            payload[b].resize((size_t)L);
            zip_reconstruct(enc.data() + off, L, payload[b].data());
        }
    });
    for (int64_t b = 0; b < nblocks; b++)
        if (zerr[b] != Z_OK) return ctx_fail(ctx, RT_ERR_INTERNAL, "zlib compress2 failed");

    Bytes f;
    f.v.reserve((size_t)bound);
    f.u32(0x01312F76u);   // magic
    f.i32(2);             // version 2, no flags
    f.raw(header.data(), header.size());
    int64_t offset = 8 + (int64_t)header.size() + nblocks * 8;
    for (int64_t b = 0; b < nblocks; b++) {
        f.i64(offset);
        offset += 8 + (int64_t)payload[b].size();
    }
    for (int64_t b = 0; b < nblocks; b++) {
        f.i32((int32_t)(b * kExrLines));
        f.i32((int32_t)payload[b].size());
        f.raw(payload[b].data(), payload[b].size());
    }
    memcpy(out, f.v.data(), f.v.size());
    *size = (int64_t)f.v.size();
    return RT_OK;
}

int rt_pfm_encode(rt_ctx* ctx, const float* planes, int32_t width, int32_t height, int32_t channels, uint8_t* out,
                  int64_t capacity, int64_t* size) {
    if (!ctx || !size) return RT_ERR_INVALID_ARGUMENT;
    if (channels != 1 && channels != 3)
        return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "PFM supports 1 or 3 channels, got " + std::to_string(channels));
    if (width <= 0 || height <= 0 || !planes) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "invalid PFM dimensions or planes");
    const std::string hdr = std::string(channels == 3 ? "PF" : "Pf") + "\n" + std::to_string(width) + " " +
                            std::to_string(height) + "\n-1.0\n";
    const int64_t nval = (int64_t)width * height * channels;
    const int64_t need = (int64_t)hdr.size() + 4 * nval;
    *size = need;
    if (!out) return RT_OK;
    if (capacity < need) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "PFM output buffer too small (query with out = NULL)");
    IO_TRY(hipSetDevice(ctx_device(ctx)));
    hipStream_t st = ctx_stream(ctx);
    const size_t bytes = (size_t)nval * 4;
    const size_t out_off = (bytes + 255) & ~(size_t)255;
    void* mem = nullptr;
    int rc = ctx_io_buffer(ctx, out_off + bytes + 256, &mem);
    if (rc) return rc;
    float* d_in = (float*)mem;
    float* d_out = (float*)((char*)mem + out_off);
    IO_TRY(hipMemcpyAsync(d_in, planes, bytes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_pfm_rows, dim3((unsigned)((nval + 255) / 256)), dim3(256), 0, st, d_in, width, height, channels,
                       d_out);
    IO_TRY(hipGetLastError());
    memcpy(out, hdr.data(), hdr.size());
    IO_TRY(hipMemcpyAsync(out + hdr.size(), d_out, bytes, hipMemcpyDeviceToHost, st));
    IO_TRY(hipStreamSynchronize(st));
    return RT_OK;
}

int rt_png_encode(rt_ctx* ctx, const float* display, int32_t width, int32_t height, uint8_t* out, int64_t capacity,
                  int64_t* size) {
    if (!ctx || !size) return RT_ERR_INVALID_ARGUMENT;
    if (width <= 0 || height <= 0 || !display) return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "invalid PNG dimensions or planes");
    const int64_t row = 1 + 4 * (int64_t)width;
    const int64_t raw_n = row * height;
    // zlib stream of raw_n bytes cut into segments deflated independently
    constexpr int64_t kSeg = 1 << 20;
    const int64_t nseg = (raw_n + kSeg - 1) / kSeg;
    const int64_t zbound = 2 + 4 + nseg * ((int64_t)compressBound((uLong)kSeg) + 16);
    // the zlib stream goes out in IDAT chunks of at most kIdat bytes (a PNG
    // chunk length is below 2^31)
    constexpr int64_t kIdat = 1 << 30;
    const int64_t nidat = (zbound + kIdat - 1) / kIdat;
    const int64_t bound = 8 + 25 + 12 * nidat + zbound + 12;
    if (!out) {
        *size = bound;
        return RT_OK;
    }
    if (capacity < bound) {
        *size = bound;
        return ctx_fail(ctx, RT_ERR_INVALID_ARGUMENT, "PNG output buffer smaller than the bound (query with out = NULL)");
    }
    IO_TRY(hipSetDevice(ctx_device(ctx)));
    hipStream_t st = ctx_stream(ctx);
    const int64_t npix = (int64_t)width * height;
    const size_t in_bytes = (size_t)npix * 12;
    const size_t out_off = (in_bytes + 255) & ~(size_t)255;
    void* mem = nullptr;
    int rc = ctx_io_buffer(ctx, out_off + (size_t)raw_n + 256, &mem);
    if (rc) return rc;
    float* d_in = (float*)mem;
    uint8_t* d_out = (uint8_t*)mem + out_off;
    IO_TRY(hipMemcpyAsync(d_in, display, in_bytes, hipMemcpyHostToDevice, st));
    const int64_t threads = (width + 1) * (int64_t)height;
    hipLaunchKernelGGL(k_png_rows, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, d_in, width, height, d_out);
    IO_TRY(hipGetLastError());
    std::vector<uint8_t> raw((size_t)raw_n);
    IO_TRY(hipMemcpyAsync(raw.data(), d_out, (size_t)raw_n, hipMemcpyDeviceToHost, st));
    IO_TRY(hipStreamSynchronize(st));

    // raw deflate per segment: all but the last end on a sync flush (byte
    // aligned, not final), the last finishes the stream
    std::vector<std::vector<uint8_t>> seg((size_t)nseg);
    std::vector<int> zerr((size_t)nseg, Z_OK);
    parallel_for(nseg, [&](int64_t s) {
        const int64_t off = s * kSeg, n = std::min<int64_t>(kSeg, raw_n - off);
        z_stream z{};
        if (deflateInit2(&z, Z_DEFAULT_COMPRESSION, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) {
            zerr[s] = Z_STREAM_ERROR;
            return;
        }
        std::vector<uint8_t> buf(compressBound((uLong)n) + 16);
        z.next_in = raw.data() + off;
        z.avail_in = (uInt)n;
        z.next_out = buf.data();
        z.avail_out = (uInt)buf.size();
        const int e = deflate(&z, s + 1 == nseg ? Z_FINISH : Z_SYNC_FLUSH);
        const bool ok = s + 1 == nseg ? e == Z_STREAM_END : (e == Z_OK && z.avail_in == 0);
        buf.resize(buf.size() - z.avail_out);
        deflateEnd(&z);
        if (!ok) {
            zerr[s] = e;
            return;
        }
        seg[s] = std::move(buf);
    });
    for (int64_t s = 0; s < nseg; s++)
        if (zerr[s] != Z_OK) return ctx_fail(ctx, RT_ERR_INTERNAL, "zlib deflate failed");
    Bytes idat;
    idat.u8(0x78);   // zlib header: deflate, 32K window, default level
    idat.u8(0x9C);
    for (auto& s : seg) idat.raw(s.data(), s.size());
    idat.be32((uint32_t)adler32_z(adler32(0L, Z_NULL, 0), raw.data(), (z_size_t)raw_n));

    Bytes f;
    const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    f.raw(sig, 8);
    auto chunk = [&](const char* type, const uint8_t* data, size_t n) {
        f.be32((uint32_t)n);
        const size_t at = f.v.size();
        f.raw(type, 4);
        if (n) f.raw(data, n);
        uLong crc = crc32(0L, Z_NULL, 0);
        crc = crc32_z(crc, f.v.data() + at, (z_size_t)(4 + n));
        f.be32((uint32_t)crc);
    };
    Bytes ihdr;
    ihdr.be32((uint32_t)width);
    ihdr.be32((uint32_t)height);
    ihdr.u8(8);   // bit depth
    ihdr.u8(6);   // RGBA
    ihdr.u8(0);
    ihdr.u8(0);
    ihdr.u8(0);
    chunk("IHDR", ihdr.v.data(), ihdr.v.size());
    for (size_t at = 0; at < idat.v.size(); at += (size_t)kIdat)
        chunk("IDAT", idat.v.data() + at, std::min<size_t>((size_t)kIdat, idat.v.size() - at));
    chunk("IEND", nullptr, 0);
    if ((int64_t)f.v.size() > capacity) return ctx_fail(ctx, RT_ERR_INTERNAL, "PNG larger than its bound");
    memcpy(out, f.v.data(), f.v.size());
    *size = (int64_t)f.v.size();
    return RT_OK;
}

}  // extern "C"
