// film_lean_iso.hip -- the persistent film kernel for isotropic stones without edge rounding, frosted facets or volume scattering (the C1 film).
#include "film_paths.h"

extern "C" hipError_t ixg_launch_persistent_lean_iso(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                                    uint32_t* d_counter, uint32_t grid_blocks, hipStream_t stream,
                                                    int iso) {
    return ixg::launch_persistent_subset<false, false>(P, d_planes, d_xyz, d_counter, grid_blocks, stream, iso != 0);
}

// Resident blocks of the persistent film kernel: per CU at its register
// budget, times the CUs.
extern "C" hipError_t ixg_persistent_grid(int device, uint32_t* blocks) {
    int per_cu = 0, cus = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, ixg::film_persistent_kernel<true, false, false, false>, ixg::kBlock, 0);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return e;
    *blocks = (uint32_t)((per_cu > 0 ? per_cu : 1) * cus);
    return hipSuccess;
}
