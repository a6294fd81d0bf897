// film_extra_biax.hip -- the persistent film kernel for biaxial stones with edge-rounded normals or an HDR panorama.
#include "film_paths.h"

extern "C" hipError_t ixg_launch_persistent_extra_biax(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                                     uint32_t* d_counter, uint32_t grid_blocks, hipStream_t stream,
                                                     int iso) {
    return ixg::launch_persistent_subset<2, true>(P, d_planes, d_xyz, d_counter, grid_blocks, stream, iso != 0);
}
