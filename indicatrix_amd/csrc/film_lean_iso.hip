// film_lean_iso.hip -- the persistent film kernel for isotropic stones without edge rounding, frosted facets or volume scattering (the C1 film).
#include "film_paths.h"

extern "C" hipError_t ixg_launch_persistent_lean_iso(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                                    uint32_t* d_counter, uint32_t grid_blocks, hipStream_t stream,
                                                    int iso) {
    return ixg::launch_persistent_subset<false, false>(P, d_planes, d_xyz, d_counter, grid_blocks, stream, iso != 0);
}

extern "C" int ixg_lean_blocks_per_cu(void) {
    static const int per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ixg::film_persistent_kernel<true, false, false, false>,
                                                         ixg::kBlock, 0) != hipSuccess)
            n = 0;
        return n;
    }();
    return per_cu;
}

// Resident blocks of the lean isotropic persistent film kernel: per CU at
// its register and LDS budget, times the CUs (the other instantiations scale
// this to their own residency at launch).
extern "C" hipError_t ixg_persistent_grid(int device, uint32_t* blocks) {
    int cus = 0;
    const int per_cu = ixg_lean_blocks_per_cu();
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return e;
    *blocks = (uint32_t)((per_cu > 0 ? per_cu : 1) * cus);
    return hipSuccess;
}
