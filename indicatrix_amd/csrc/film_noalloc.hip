// film_noalloc.hip -- ixg_accumulate_device_noalloc: the film path on a
// fixed scratch area in device memory, for callers that cannot allocate
// (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).
//
// The scratch is one file-scope device array per GPU the library runs on,
// held until the process ends:
//   [0, 256)                     the persistent kernel's path counter
//   [256, 256 + 4128)            the facet set + frosted bitmask (IXG_MAX_PLANES + 2 float4)
//   then kScratchPaths x 12 B    the chunk's per-path XYZ records
//   then kScratchHdrTexels x 12 B the HDR panorama slot
// Its device address comes from hipGetSymbolAddress, once per device, with
// that device current; the array's name is never used as a device address.
// The counter is re-zeroed by a fill kernel before every launch and every
// path record is written before the reduction reads it, so nothing carries
// over from the previous call.  Calls that overlap on one GPU share this
// memory and race: serialise them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/ixg.h"

namespace {

constexpr uint64_t kScratchPaths = IXG_NOALLOC_MAX_PATHS;
constexpr uint64_t kScratchHdrTexels = IXG_NOALLOC_MAX_HDR_TEXELS;
constexpr size_t kCounterBytes = 256;
constexpr size_t kPlaneBytes = (size_t)(IXG_MAX_PLANES + 2) * 16;
constexpr size_t kPlaneOffset = kCounterBytes;
constexpr size_t kXyzOffset = (kPlaneOffset + kPlaneBytes + 255) / 256 * 256;
constexpr size_t kHdrOffset = kXyzOffset + kScratchPaths * 12;
constexpr size_t kScratchBytes = kHdrOffset + kScratchHdrTexels * 12;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_ptr[kMaxDevices] = {};

}  // namespace

extern "C" int ixg_internal_device(ixg_ctx* ctx);
extern "C" ixg_status ixg_internal_fail(ixg_ctx* ctx, ixg_status st, const char* what);
extern "C" ixg_status ixg_internal_hip_fail(ixg_ctx* ctx, hipError_t e, const char* where);
extern "C" ixg_status ixg_internal_accumulate_on(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset, uint32_t spp,
                                      double* d_accum_xyz, float4* d_planes, float* d_xyz, uint64_t xyz_paths,
                                      uint32_t* d_counter, float* d_hdr, size_t hdr_cap);

extern "C" ixg_status ixg_accumulate_device_noalloc(ixg_ctx* ctx, const ixg_scene* scene, uint32_t sample_offset,
                                                    uint32_t spp, double* d_accum_xyz) {
    if (!ctx) return IXG_INVALID;
    const int dev = ixg_internal_device(ctx);
    if (dev < 0 || dev >= kMaxDevices) return ixg_internal_fail(ctx, IXG_RESOURCE_LIMIT, "device ordinal too large");
    if (spp > kScratchPaths)
        return ixg_internal_fail(ctx, IXG_RESOURCE_LIMIT, "spp above IXG_NOALLOC_MAX_PATHS");
    if (scene && scene->hdr_map && (uint64_t)scene->hdr_map->width * scene->hdr_map->height > kScratchHdrTexels)
        return ixg_internal_fail(ctx, IXG_RESOURCE_LIMIT, "panorama above IXG_NOALLOC_MAX_HDR_TEXELS");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> guard(g_lock);
        if (!g_ptr[dev]) {
            // the first call on this device: resolve this GPU's copy of the array
            hipError_t e = hipSetDevice(dev);
            if (e == hipSuccess) e = hipGetSymbolAddress(&g_ptr[dev], HIP_SYMBOL(scratch));
            if (e != hipSuccess || !g_ptr[dev]) {
                g_ptr[dev] = nullptr;
                return ixg_internal_hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol,
                                             "hipGetSymbolAddress(scratch)");
            }
        }
        base = g_ptr[dev];
    }
    unsigned char* p = static_cast<unsigned char*>(base);
    return ixg_internal_accumulate_on(ctx, scene, sample_offset, spp, d_accum_xyz,
                                      reinterpret_cast<float4*>(p + kPlaneOffset), reinterpret_cast<float*>(p + kXyzOffset),
                                      kScratchPaths, reinterpret_cast<uint32_t*>(p), reinterpret_cast<float*>(p + kHdrOffset),
                                      (size_t)kScratchHdrTexels * 3);
}
