// film_diag_biax.hip -- the diagnostic film kernels (per-path events, explicit rays) for biaxial stones.
#include "film_paths.h"

extern "C" hipError_t ixg_launch_events_biax(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                            ixg_path_events* d_ev, unsigned long long* d_totals, hipStream_t stream) {
    return ixg::launch_events_subset<2>(P, d_planes, d_xyz, d_ev, d_totals, stream);
}

extern "C" hipError_t ixg_launch_rays_biax(const ixg::FilmParams* P, const float4* d_planes, const float* d_rays,
                                          uint32_t n_rays, float* d_xyz, ixg_path_events* d_ev, hipStream_t stream) {
    return ixg::launch_rays_subset<2>(P, d_planes, d_rays, n_rays, d_xyz, d_ev, stream);
}
