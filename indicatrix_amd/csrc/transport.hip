// transport.hip -- the film path's per-pixel reduction and launch dispatch.
//
// The film kernels themselves (film_paths.h) are instantiated in the
// film_*.hip translation units, one (uniaxial, extra-features) class each;
// this file routes a launch to the class the scene needs.
#include "film_paths.h"

namespace ixg {

__global__ void zero_counter_kernel(uint32_t* c) { *c = 0u; }

// Per-pixel FP64 sums in sample order; a sample with any non-finite
// component is dropped but still counted (sampling.rs:237, reduce_xyz.wgsl:55).
__global__ __launch_bounds__(kBlock) void film_reduce_kernel(const float* __restrict__ xyz,
                                                             uint32_t spp, uint32_t n_pixels,
                                                             uint32_t first_pixel,
                                                             double* __restrict__ accum) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_pixels) return;
    const float* s = xyz + (size_t)i * spp * 3;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (uint32_t k = 0; k < spp; ++k) {
        float x = s[3 * k], y = s[3 * k + 1], z = s[3 * k + 2];
        if (__builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z)) {
            a0 += (double)x;
            a1 += (double)y;
            a2 += (double)z;
        }
    }
    double* dst = accum + 3 * (size_t)(first_pixel + i);
    dst[0] += a0;
    dst[1] += a1;
    dst[2] += a2;
}

}  // namespace ixg

// the per-class launchers (film_lean_iso.hip ... film_diag_aniso.hip)
#define IXG_PERSISTENT_ARGS                                                                                  \
    const ixg::FilmParams *P, const float4 *d_planes, float *d_xyz, uint32_t *d_counter, uint32_t grid_blocks, \
        hipStream_t stream, int iso
extern "C" hipError_t ixg_launch_persistent_lean_iso(IXG_PERSISTENT_ARGS);
extern "C" hipError_t ixg_launch_persistent_lean_aniso(IXG_PERSISTENT_ARGS);
extern "C" hipError_t ixg_launch_persistent_extra_iso(IXG_PERSISTENT_ARGS);
extern "C" hipError_t ixg_launch_persistent_extra_aniso(IXG_PERSISTENT_ARGS);
extern "C" hipError_t ixg_launch_persistent_lean_biax(IXG_PERSISTENT_ARGS);
extern "C" hipError_t ixg_launch_persistent_extra_biax(IXG_PERSISTENT_ARGS);
extern "C" hipError_t ixg_launch_events_biax(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                             ixg_path_events* d_ev, unsigned long long* d_totals, hipStream_t stream);
extern "C" hipError_t ixg_launch_rays_biax(const ixg::FilmParams* P, const float4* d_planes, const float* d_rays,
                                           uint32_t n_rays, float* d_xyz, ixg_path_events* d_ev, hipStream_t stream);
extern "C" hipError_t ixg_launch_events_iso(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                            ixg_path_events* d_ev, unsigned long long* d_totals, hipStream_t stream);
extern "C" hipError_t ixg_launch_events_aniso(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                              ixg_path_events* d_ev, unsigned long long* d_totals,
                                              hipStream_t stream);
extern "C" hipError_t ixg_launch_rays_iso(const ixg::FilmParams* P, const float4* d_planes, const float* d_rays,
                                          uint32_t n_rays, float* d_xyz, ixg_path_events* d_ev, hipStream_t stream);
extern "C" hipError_t ixg_launch_rays_aniso(const ixg::FilmParams* P, const float4* d_planes, const float* d_rays,
                                            uint32_t n_rays, float* d_xyz, ixg_path_events* d_ev,
                                            hipStream_t stream);

extern "C" hipError_t ixg_launch_film_events(const ixg::FilmParams* P, const float4* d_planes,
                                             float* d_xyz, ixg_path_events* d_ev,
                                             unsigned long long* d_totals, hipStream_t stream) {
    if (P->an.biax) return ixg_launch_events_biax(P, d_planes, d_xyz, d_ev, d_totals, stream);
    return P->an.on ? ixg_launch_events_aniso(P, d_planes, d_xyz, d_ev, d_totals, stream)
                    : ixg_launch_events_iso(P, d_planes, d_xyz, d_ev, d_totals, stream);
}

extern "C" hipError_t ixg_launch_reduce(const float* d_xyz, uint32_t spp, uint32_t n_pixels,
                                        uint32_t first_pixel, double* d_accum, hipStream_t stream) {
    uint32_t blocks = (n_pixels + ixg::kBlock - 1) / ixg::kBlock;
    hipLaunchKernelGGL(ixg::film_reduce_kernel, dim3(blocks), dim3(ixg::kBlock), 0, stream, d_xyz,
                       spp, n_pixels, first_pixel, d_accum);
    return hipGetLastError();
}

extern "C" hipError_t ixg_launch_rays(const ixg::FilmParams* P, const float4* d_planes, const float* d_rays,
                                      uint32_t n_rays, float* d_xyz, ixg_path_events* d_ev,
                                      hipStream_t stream) {
    if (P->an.biax) return ixg_launch_rays_biax(P, d_planes, d_rays, n_rays, d_xyz, d_ev, stream);
    return P->an.on ? ixg_launch_rays_aniso(P, d_planes, d_rays, n_rays, d_xyz, d_ev, stream)
                    : ixg_launch_rays_iso(P, d_planes, d_rays, n_rays, d_xyz, d_ev, stream);
}

extern "C" hipError_t ixg_launch_film_persistent(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                                 uint32_t* d_counter, uint32_t grid_blocks, hipStream_t stream) {
    // a fill kernel rather than a memset node, so the launch pair can be captured
    hipLaunchKernelGGL(ixg::zero_counter_kernel, dim3(1), dim3(1), 0, stream, d_counter);
    // the lit models: one scaled factor per direction (ASET is per wavelength)
    const int iso = P->env.model != ixg::ENV_STUDIO && P->env.model != ixg::ENV_ASET && P->env.model != ixg::ENV_HDR;
    // the rarely used features, the HDR panorama among them, run the EXTRA kernels
    const bool extra = P->frosted || P->sigma_s > 0.0f || P->edge_radius > 0.0f || P->env.model == ixg::ENV_HDR;
    if (P->an.biax)
        return extra ? ixg_launch_persistent_extra_biax(P, d_planes, d_xyz, d_counter, grid_blocks, stream, iso)
                     : ixg_launch_persistent_lean_biax(P, d_planes, d_xyz, d_counter, grid_blocks, stream, iso);
    if (extra)
        return P->an.on ? ixg_launch_persistent_extra_aniso(P, d_planes, d_xyz, d_counter, grid_blocks, stream, iso)
                        : ixg_launch_persistent_extra_iso(P, d_planes, d_xyz, d_counter, grid_blocks, stream, iso);
    return P->an.on ? ixg_launch_persistent_lean_aniso(P, d_planes, d_xyz, d_counter, grid_blocks, stream, iso)
                    : ixg_launch_persistent_lean_iso(P, d_planes, d_xyz, d_counter, grid_blocks, stream, iso);
}
