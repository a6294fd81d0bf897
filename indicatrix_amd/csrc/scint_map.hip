// scint_map.hip -- the per-pixel scintillation maps of a finished metric
// image (IX/color/metrics/image_scint.rs:150-259).
//
//   spatial[i]  scintillation_spatial: the squashed coefficient of variation
//               of the brilliance fraction over the 5x5 window of hit map
//               pixels around i (clipped at the frame), 0 where hits == 0;
//   index[i]    scintillation_index: 0.6 spatial[i] + 0.4 of the pixel's
//               five-pose Bernoulli variance over its 0.24 ceiling; NaN
//               where the temporal term was not computed (None).
//
// One lane per map pixel in 16 x 16 tiles; each tile stages its 20 x 20 halo
// of brilliance fractions in LDS once (the 56-byte pixel records are read
// once per tile instead of 25 times), then every lane folds its window in
// the reference's row-major order with the reference's rounding (fused
// sum of squares, correctly rounded division and square root), so the maps
// equal the CPU pass bit for bit.  HBM-bound: 56 B read and 8 B written per
// pixel.
#include <cmath>

#include "device_common.h"

extern "C" int ixg_internal_lost(ixg_ctx* ctx);
extern "C" int ixg_internal_device(ixg_ctx* ctx);
extern "C" hipStream_t ixg_internal_stream(ixg_ctx* ctx);
extern "C" ixg_status ixg_internal_fail(ixg_ctx* ctx, ixg_status st, const char* what);
extern "C" ixg_status ixg_internal_hip_fail(ixg_ctx* ctx, hipError_t e, const char* where);

namespace ixg {

constexpr int kScintTile = 16;
constexpr int kScintRadius = 2;  // SCINTILLATION_WINDOW_RADIUS (image_scint.rs:27)
constexpr int kScintSpan = kScintTile + 2 * kScintRadius;

// spatial_scintillation_pct (scintillation.rs:146-171)
__host__ __device__ __forceinline__ float scint_spatial_pct(float sum, float sum_sq, uint32_t count) {
    if (count == 0) return 0.0f;
    const float mean = sum / (float)count;
    if (mean <= 1e-4f) return 0.0f;
#ifdef __HIP_DEVICE_COMPILE__
    const float var = fmaxf(fma_(mean, -mean, sum_sq / (float)count), 0.0f);
    const float cv = sqrt_rn(var) / mean;
#else
    const float var = std::fmax(std::fma(mean, -mean, sum_sq / (float)count), 0.0f);
    const float cv = std::sqrt(var) / mean;
#endif
    const float pct = cv / (1.0f + cv) * 100.0f;
    return pct < 0.0f ? 0.0f : (pct > 100.0f ? 100.0f : pct);
}

// OpticalImage::temporal_variance (image_types.rs:105-112) folded through
// temporal_scintillation_pct (scintillation.rs:173-180, one cell) and
// combine_scintillation_pct (scintillation.rs:194-198).
__host__ __device__ __forceinline__ float scint_index_of(float spatial, uint8_t returned, uint8_t poses) {
    const float fr = (float)returned / (float)poses;
#ifdef __HIP_DEVICE_COMPILE__
    const float variance = fma_(fr, -fr, fr);
#else
    const float variance = std::fma(fr, -fr, fr);
#endif
    float temporal = ((variance / 1.0f) / 0.24f) * 100.0f;
    temporal = temporal < 0.0f ? 0.0f : (temporal > 100.0f ? 100.0f : temporal);
#ifdef __HIP_DEVICE_COMPILE__
    const float c = fma_(0.4f, temporal, 0.6f * spatial);
#else
    const float c = std::fma(0.4f, temporal, 0.6f * spatial);
#endif
    return c < 0.0f ? 0.0f : (c > 100.0f ? 100.0f : c);
}

// fractions()[2] (image_types.rs:79-81) of a hit pixel, -1 for none
__host__ __device__ __forceinline__ float scint_fraction(const ixg_optical_pixel& q) {
    return q.hits > 0 ? (float)q.classes[2] / (float)(q.hits > 1 ? q.hits : 1) : -1.0f;
}

// window_spatial (image_scint.rs:150-173) over a row-major frame of
// fractions (-1 = no hits), centre (x, y), frame w x h.
__host__ __device__ __forceinline__ float scint_window(const float* frac, int stride, int w, int h, int x, int y) {
    float sum = 0.0f, sum_sq = 0.0f;
    uint32_t count = 0;
    const int y0 = y - kScintRadius > 0 ? y - kScintRadius : 0;
    const int y1 = y + kScintRadius + 1 < h ? y + kScintRadius + 1 : h;
    const int x0 = x - kScintRadius > 0 ? x - kScintRadius : 0;
    const int x1 = x + kScintRadius + 1 < w ? x + kScintRadius + 1 : w;
    for (int wy = y0; wy < y1; ++wy)
        for (int wx = x0; wx < x1; ++wx) {
            const float f = frac[wy * stride + wx];
            if (f >= 0.0f) {
                sum += f;
#ifdef __HIP_DEVICE_COMPILE__
                sum_sq = fma_(f, f, sum_sq);
#else
                sum_sq = std::fma(f, f, sum_sq);
#endif
                ++count;
            }
        }
    return scint_spatial_pct(sum, sum_sq, count);
}

// spatial_valid = 0 (a map whose size is not its pixel count): the launch
// covers the pixels as one row and every spatial term reads 0.
__global__ __launch_bounds__(kScintTile* kScintTile) void scint_map_kernel(const ixg_optical_pixel* __restrict__ px,
                                                                           uint32_t w, uint32_t h, int spatial_valid,
                                                                           float* __restrict__ spatial,
                                                                           float* __restrict__ index) {
    __shared__ float frac[kScintSpan * kScintSpan];
    const int gx0 = (int)(blockIdx.x * kScintTile) - kScintRadius;
    const int gy0 = (int)(blockIdx.y * kScintTile) - kScintRadius;
    for (int k = threadIdx.x; k < kScintSpan * kScintSpan; k += kScintTile * kScintTile) {
        const int gx = gx0 + k % kScintSpan, gy = gy0 + k / kScintSpan;
        float f = -1.0f;
        if (gx >= 0 && gy >= 0 && gx < (int)w && gy < (int)h) f = scint_fraction(px[(size_t)gy * w + gx]);
        frac[k] = f;
    }
    __syncthreads();
    const int tx = threadIdx.x % kScintTile, ty = threadIdx.x / kScintTile;
    const uint32_t x = blockIdx.x * kScintTile + tx, y = blockIdx.y * kScintTile + ty;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    float sp = 0.0f;
    if (spatial_valid && frac[(ty + kScintRadius) * kScintSpan + tx + kScintRadius] >= 0.0f) {
        // the window in frame coordinates, read through the tile's halo:
        // out-of-frame cells hold -1 and are skipped, as the clipped loop does
        float sum = 0.0f, sum_sq = 0.0f;
        uint32_t count = 0;
#pragma unroll
        for (int dy = 0; dy < 2 * kScintRadius + 1; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2 * kScintRadius + 1; ++dx) {
                const float f = frac[(ty + dy) * kScintSpan + tx + dx];
                if (f >= 0.0f) {
                    sum += f;
                    sum_sq = fma_(f, f, sum_sq);
                    ++count;
                }
            }
        sp = scint_spatial_pct(sum, sum_sq, count);
    }
    if (spatial) spatial[i] = sp;
    if (index) {
        const ixg_optical_pixel& q = px[i];
        index[i] = q.scint_poses ? scint_index_of(sp, q.scint_returned, q.scint_poses) : __builtin_nanf("");
    }
}

}  // namespace ixg

using namespace ixg;

namespace {

ixg_status maps_on_stream(ixg_ctx* ctx, const ixg_optical_pixel* d_px, uint32_t n, uint32_t width,
                          uint32_t height, float* d_spatial, float* d_index, hipStream_t s) {
    const bool valid = (uint64_t)width * height == n;
    const uint32_t w = valid ? width : n, h = valid ? height : 1;
    const dim3 grid((w + kScintTile - 1) / kScintTile, (h + kScintTile - 1) / kScintTile);
    hipLaunchKernelGGL(scint_map_kernel, grid, dim3(kScintTile * kScintTile), 0, s, d_px, w, h, valid ? 1 : 0,
                       d_spatial, d_index);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? IXG_OK : ixg_internal_hip_fail(ctx, e, "scintillation maps");
}

ixg_status maps_begin(ixg_ctx* ctx, uint32_t n, uint32_t width, uint32_t height, hipStream_t* s) {
    if (!ctx) return IXG_INVALID;
    if (ixg_internal_lost(ctx)) return ixg_internal_fail(ctx, IXG_DEVICE_LOST, "context is lost");
    if (n > (1u << 30)) return ixg_internal_fail(ctx, IXG_RESOURCE_LIMIT, "at most 2^30 map pixels");
    // a size that matches the count must also fit the launch's 2-D grid
    if ((uint64_t)width * height == n && height > 65535u * kScintTile)
        return ixg_internal_fail(ctx, IXG_RESOURCE_LIMIT, "map too tall");
    const hipError_t e = hipSetDevice(ixg_internal_device(ctx));
    if (e != hipSuccess) return ixg_internal_hip_fail(ctx, e, "hipSetDevice");
    *s = ixg_internal_stream(ctx);
    return IXG_OK;
}

}  // namespace

extern "C" {

ixg_status ixg_scintillation_maps_device(ixg_ctx* ctx, const ixg_optical_pixel* d_pixels, uint32_t n_pixels,
                                         uint32_t width, uint32_t height, float* d_spatial, float* d_index) {
    hipStream_t s;
    ixg_status st = maps_begin(ctx, n_pixels, width, height, &s);
    if (st != IXG_OK || n_pixels == 0) return st;
    if (!d_pixels) return ixg_internal_fail(ctx, IXG_INVALID, "null pixels");
    if (!d_spatial && !d_index) return IXG_OK;
    st = maps_on_stream(ctx, d_pixels, n_pixels, width, height, d_spatial, d_index, s);
    if (st != IXG_OK) return st;
    const hipError_t e = hipStreamSynchronize(s);
    return e == hipSuccess ? IXG_OK : ixg_internal_hip_fail(ctx, e, "scintillation maps");
}

ixg_status ixg_scintillation_maps(ixg_ctx* ctx, const ixg_optical_pixel* pixels, uint32_t n_pixels, uint32_t width,
                                  uint32_t height, float* spatial, float* index) {
    hipStream_t s;
    ixg_status st = maps_begin(ctx, n_pixels, width, height, &s);
    if (st != IXG_OK || n_pixels == 0) return st;
    if (!pixels) return ixg_internal_fail(ctx, IXG_INVALID, "null pixels");
    if (!spatial && !index) return IXG_OK;
    ixg_optical_pixel* d_px = nullptr;
    float* d_out = nullptr;
    hipError_t e = hipMalloc(&d_px, (size_t)n_pixels * sizeof(ixg_optical_pixel));
    if (e == hipSuccess) e = hipMalloc(&d_out, (size_t)n_pixels * 2 * sizeof(float));
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_px, pixels, (size_t)n_pixels * sizeof(ixg_optical_pixel), hipMemcpyHostToDevice, s);
    float* d_sp = d_out;
    float* d_ix = d_out ? d_out + n_pixels : nullptr;
    if (e == hipSuccess) {
        st = maps_on_stream(ctx, d_px, n_pixels, width, height, spatial ? d_sp : nullptr, index ? d_ix : nullptr, s);
        if (st == IXG_OK && spatial)
            e = hipMemcpyAsync(spatial, d_sp, (size_t)n_pixels * sizeof(float), hipMemcpyDeviceToHost, s);
        if (st == IXG_OK && e == hipSuccess && index)
            e = hipMemcpyAsync(index, d_ix, (size_t)n_pixels * sizeof(float), hipMemcpyDeviceToHost, s);
        if (st == IXG_OK && e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (d_px) (void)hipFree(d_px);
    if (d_out) (void)hipFree(d_out);
    if (st != IXG_OK) return st;
    return e == hipSuccess ? IXG_OK : ixg_internal_hip_fail(ctx, e, "scintillation maps");
}

float ixg_scintillation_spatial_at(const ixg_optical_pixel* pixels, uint32_t n_pixels, uint32_t width,
                                   uint32_t height, uint32_t i) {
    // scintillation_spatial_at (image_scint.rs:201-232): the window as its own
    // small frame, read by the same window function as the full pass
    if (!pixels || (uint64_t)width * height != n_pixels || i >= n_pixels) return 0.0f;
    const int col = (int)(i % width), row = (int)(i / width);
    const int x0 = col - kScintRadius > 0 ? col - kScintRadius : 0;
    const int x1 = col + kScintRadius + 1 < (int)width ? col + kScintRadius + 1 : (int)width;
    const int y0 = row - kScintRadius > 0 ? row - kScintRadius : 0;
    const int y1 = row + kScintRadius + 1 < (int)height ? row + kScintRadius + 1 : (int)height;
    float local[(2 * kScintRadius + 1) * (2 * kScintRadius + 1)];
    const int fw = x1 - x0, fh = y1 - y0;
    for (int wy = y0; wy < y1; ++wy)
        for (int wx = x0; wx < x1; ++wx) local[(wy - y0) * fw + (wx - x0)] = scint_fraction(pixels[(size_t)wy * width + wx]);
    if (local[(row - y0) * fw + (col - x0)] < 0.0f) return 0.0f;
    return scint_window(local, fw, fw, fh, col - x0, row - y0);
}

int ixg_scintillation_index_at(const ixg_optical_pixel* pixels, uint32_t n_pixels, uint32_t width, uint32_t height,
                               uint32_t i, float* out) {
    // scintillation_index_at (image_scint.rs:249-259)
    if (!pixels || !out || i >= n_pixels || pixels[i].scint_poses == 0) return 0;
    *out = scint_index_of(ixg_scintillation_spatial_at(pixels, n_pixels, width, height, i), pixels[i].scint_returned,
                          pixels[i].scint_poses);
    return 1;
}

}  // extern "C"
