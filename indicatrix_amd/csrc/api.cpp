// api.cpp -- libixg's C ABI: context lifecycle, routing, chunked film launches.
//
// Mirrors the reference's GpuBackend contract (IX/renderer/gpu_backend/):
//  * results are ADDED to the caller's accumulator and only on IXG_OK;
//  * a HIP error or an injected fault marks the context lost (DeviceLost)
//    and leaves the caller's buffer untouched; ixg_try_recover clears it;
//  * the cancel flag is polled between chunks (GPU/frame/mod.rs:43-49).
// One context is used by one host thread at a time (the reference serialises
// GpuFrameRenderer behind a Mutex; callers here do the same).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ixg.h"
#include "host_scene.h"
#include "transport_params.h"

extern "C" hipError_t ixg_launch_film_persistent(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                                 uint32_t* d_counter, uint32_t grid_blocks, hipStream_t stream);
extern "C" hipError_t ixg_persistent_grid(int device, uint32_t* blocks);
extern "C" hipError_t ixg_launch_film_events(const ixg::FilmParams* P, const float4* d_planes,
                                             float* d_xyz, ixg_path_events* d_ev,
                                             unsigned long long* d_totals, hipStream_t stream);
extern "C" hipError_t ixg_launch_rays(const ixg::FilmParams* P, const float4* d_planes, const float* d_rays,
                                      uint32_t n_rays, float* d_xyz, ixg_path_events* d_ev,
                                      hipStream_t stream);
extern "C" hipError_t ixg_launch_probe(const ixg::FilmParams* P, const float4* d_planes, int32_t fn,
                                       const float* d_in, uint32_t n, float* d_out, hipStream_t stream);
extern "C" hipError_t ixg_launch_reduce(const float* d_xyz, uint32_t spp, uint32_t n_pixels,
                                        uint32_t first_pixel, double* d_accum, hipStream_t stream);

struct ixg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool lost = false;
    int32_t fail_countdown = -1;  // chunks until an injected DeviceLost (-1 = none)
    std::string last_error;
    std::string label;
    uint64_t chunk_paths = 1ull << 23;
    // device buffers, grown on demand
    float4* d_planes = nullptr;
    float* d_xyz = nullptr;
    size_t xyz_cap = 0;  // floats
    float* d_hdr = nullptr;  // the scene's HDR panorama texels
    size_t hdr_cap = 0;      // floats
    double* d_accum = nullptr;
    size_t accum_cap = 0;  // doubles
    uint32_t* d_counter = nullptr;  // persistent kernel's path counter
    uint32_t persistent_blocks = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_kernel_ms_total = 0.0;
    uint32_t last_kernel_launches = 0;
    // the metric sweep's persistent buffers, captured graph and tuning cache
    void* metric_plan = nullptr;
    void (*metric_plan_free)(void*) = nullptr;
};

namespace {

ixg_status fail(ixg_ctx* ctx, ixg_status st, const char* what) {
    if (ctx) ctx->last_error = what;
    return st;
}

ixg_status hip_fail(ixg_ctx* ctx, hipError_t e, const char* where) {
    ctx->lost = true;
    ctx->last_error = std::string(where) + ": " + hipGetErrorString(e);
    return IXG_DEVICE_LOST;
}

#define HIP_TRY(ctx, expr)                                   \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
    } while (0)

ixg_status ensure_xyz(ixg_ctx* ctx, size_t floats) {
    if (floats <= ctx->xyz_cap) return IXG_OK;
    if (ctx->d_xyz) (void)hipFree(ctx->d_xyz);
    ctx->d_xyz = nullptr;
    HIP_TRY(ctx, hipMalloc(&ctx->d_xyz, floats * sizeof(float)));
    ctx->xyz_cap = floats;
    return IXG_OK;
}

ixg_status ensure_accum(ixg_ctx* ctx, size_t doubles) {
    if (doubles <= ctx->accum_cap) return IXG_OK;
    if (ctx->d_accum) (void)hipFree(ctx->d_accum);
    ctx->d_accum = nullptr;
    HIP_TRY(ctx, hipMalloc(&ctx->d_accum, doubles * sizeof(double)));
    ctx->accum_cap = doubles;
    return IXG_OK;
}

bool any_frosted(const ixg_scene* s) {
    if (!s->facet_finishes) return false;
    for (uint32_t i = 0; i < s->n_planes; ++i)
        if (s->facet_finishes[i]) return true;
    return false;
}

bool material_routable(const ixg_material& m) {
    // The device path covers isotropic, uniaxial and biaxial
    // (RayMaterialContext::is_anisotropic, transport/inner.rs:181-198),
    // unzoned materials, with or without absorption bands and volume
    // scattering.  A biaxial stone is anisotropic (a "biaxial" set on a stone
    // that is not declines).
    if (!(m.edge_rounding_radius < 1e30f)) return false;
    if (m.biaxial) {
        if (m.cubic || !(std::fabs(m.birefringence_delta) > 1e-4f) || !std::isfinite(m.biaxial_delta_beta_alpha))
            return false;
        if (m.has_beta_bands && (m.n_bands_beta < 0 || m.n_bands_beta > IXG_MAX_BANDS)) return false;
    }
    // Henyey-Greenstein volume scattering: finite parameters (isotropic,
    // uniaxial and biaxial stones)
    if (!(m.scattering_sigma_s <= 0.0f) &&
        (!(m.scattering_sigma_s < 1e30f) || !(std::fabs(m.scattering_g) < 1.0f) || !(m.absorption_path_scale > 0.0f)))
        return false;
    if (m.has_e_dispersion && (m.e_dispersion.kind < IXG_DISP_SELLMEIER1 || m.e_dispersion.kind > IXG_DISP_CAUCHY))
        return false;
    if (m.n_bands_o < 0 || m.n_bands_o > IXG_MAX_BANDS || m.n_bands_e < 0 || m.n_bands_e > IXG_MAX_BANDS)
        return false;
    return m.dispersion.kind >= IXG_DISP_SELLMEIER1 && m.dispersion.kind <= IXG_DISP_CAUCHY;
}

ixg_status validate(ixg_ctx* ctx, const ixg_scene* s) {
    if (!s || !s->planes || !s->material) return fail(ctx, IXG_INVALID, "null scene pointer");
    if (s->width == 0 || s->height == 0) return fail(ctx, IXG_INVALID, "empty frame");
    if ((uint64_t)s->width * s->height > (1ull << 31)) return fail(ctx, IXG_RESOURCE_LIMIT, "frame too large");
    return IXG_OK;
}

// glam Vec3 helpers on the host (compiled without contraction, so these
// round like the reference).
struct H3 {
    float x, y, z;
};
float hdot(H3 a, H3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
H3 hscale(H3 a, float s) { return H3{a.x * s, a.y * s, a.z * s}; }
H3 hnormalize_or_zero(H3 a) {
    float r = 1.0f / std::sqrt(hdot(a, a));
    return (std::isfinite(r) && r > 0.0f) ? hscale(a, r) : H3{0.0f, 0.0f, 0.0f};
}
H3 hcross(H3 a, H3 b) { return H3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }

// AbsorptionTensor3::uniaxial's axes: stable_orthonormal_basis of the unit
// optic axis (birefringence/absorption_tensor.rs, birefringence/helpers.rs).
void fill_absorption(const ixg_material& m, ixg::AbsorbParams* A) {
    A->n_o = m.n_bands_o;
    A->n_e = m.n_bands_e;
    A->path_scale = m.absorption_path_scale;
    H3 c{m.c_axis[0], m.c_axis[1], m.c_axis[2]};
    H3 n = hnormalize_or_zero(c);
    H3 a = std::fabs(n.x) > 0.9f ? H3{0.0f, 1.0f, 0.0f} : H3{1.0f, 0.0f, 0.0f};
    float na = hdot(n, a);
    H3 t = hnormalize_or_zero(H3{a.x - n.x * na, a.y - n.y * na, a.z - n.z * na});
    H3 b = hcross(n, t);
    const float* src[4] = {&c.x, &t.x, &b.x, &n.x};
    float* dst[4] = {A->c_raw, A->a1, A->a2, A->cn};
    for (int i = 0; i < 4; ++i) std::memcpy(dst[i], src[i], 3 * sizeof(float));
    for (int i = 0; i < m.n_bands_o; ++i)
        A->o[i] = ixg::AbsBand{m.bands_o[i].center_nm, m.bands_o[i].width, m.bands_o[i].peak, m.bands_o[i].shape};
    for (int i = 0; i < m.n_bands_e; ++i)
        A->e[i] = ixg::AbsBand{m.bands_e[i].center_nm, m.bands_e[i].width, m.bands_e[i].peak, m.bands_e[i].shape};
    A->n_b = (m.biaxial && m.has_beta_bands) ? m.n_bands_beta : 0;
    for (int i = 0; i < A->n_b; ++i)
        A->b[i] = ixg::AbsBand{m.bands_beta[i].center_nm, m.bands_beta[i].width, m.bands_beta[i].peak,
                               m.bands_beta[i].shape};
}

// hdr_dst / hdr_cap (floats): where the panorama is uploaded; null = the
// context's own buffer, grown on demand.
ixg_status build_params(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset, uint32_t spp,
                        ixg::FilmParams* P, float* hdr_dst = nullptr, size_t hdr_cap = 0) {
    std::memset(P, 0, sizeof *P);
    std::memcpy(P->cam_origin, s->camera.origin, sizeof P->cam_origin);
    P->fov_tan = s->camera.fov_tan;
    std::memcpy(P->forward, s->camera.forward, sizeof P->forward);
    std::memcpy(P->right, s->camera.right, sizeof P->right);
    std::memcpy(P->up, s->camera.up, sizeof P->up);
    P->width = s->width;
    P->height = s->height;
    P->width_f = (float)s->width;
    P->height_f = (float)s->height;
    P->spp = spp;
    P->sample_offset = sample_offset;
    P->n_planes = s->n_planes;
    P->max_bounces = s->max_bounces;
    P->disp_kind = s->material->dispersion.kind;
    P->absorb = (s->material->n_bands_o > 0 || s->material->n_bands_e > 0 ||
                 (s->material->biaxial && s->material->has_beta_bands && s->material->n_bands_beta > 0))
                    ? 1
                    : 0;
    if (P->absorb) fill_absorption(*s->material, &P->abs);
    P->abs.path_scale = s->material->absorption_path_scale;  // the scattering step reads it too
    const ixg_material& m = *s->material;
    P->an.on = (!m.cubic && std::fabs(m.birefringence_delta) > 1e-4f) ? 1 : 0;
    P->an.biref = m.birefringence_delta;
    P->an.has_e = m.has_e_dispersion ? 1 : 0;
    P->an.e_kind = m.e_dispersion.kind;
    std::memcpy(P->an.e_disp, m.e_dispersion.coeff, sizeof P->an.e_disp);
    std::memcpy(P->an.c_axis, m.c_axis, sizeof P->an.c_axis);
    if (m.biaxial) {
        // BiaxialIndicatrix::from_gamma_axis (biaxial.rs:61-65): the frame the
        // absorption tensor uses too
        P->an.biax = 1;
        P->an.trichroic = m.has_beta_bands ? 1 : 0;
        P->an.dba = m.biaxial_delta_beta_alpha;
        ixg::AbsorbParams F;
        fill_absorption(m, &F);
        std::memcpy(P->an.bx, F.a1, sizeof P->an.bx);
        std::memcpy(P->an.by, F.a2, sizeof P->an.by);
        std::memcpy(P->an.bz, F.cn, sizeof P->an.bz);
    }
    std::memcpy(P->disp, s->material->dispersion.coeff, sizeof P->disp);
    P->edge_radius = s->material->edge_rounding_radius > 0.0f ? s->material->edge_rounding_radius : 0.0f;
    P->frosted = any_frosted(s) ? 1 : 0;
    P->sigma_s = s->material->scattering_sigma_s > 0.0f ? s->material->scattering_sigma_s : 0.0f;
    P->scat_g = s->material->scattering_g;
    float g = s->environment.surface_glare;
    P->surface_glare = (g >= 1.0f || std::isnan(g)) ? 1.0f : (g > 0.0f ? g : 0.0f);
    if (s->hdr_map) {  // EnvironmentSource::HdrMap: no exposure, observer, tray or white balance
        // the panorama first, copied in stream order, so the pointer below is
        // the buffer this call's kernels read
        const size_t n = (size_t)s->hdr_map->width * s->hdr_map->height * 3;
        if (hdr_dst) {
            if (n > hdr_cap) return fail(ctx, IXG_RESOURCE_LIMIT, "panorama larger than the scratch panorama slot");
        } else {
            if (n > ctx->hdr_cap) {
                if (ctx->d_hdr) (void)hipFree(ctx->d_hdr);
                ctx->d_hdr = nullptr;
                ctx->hdr_cap = 0;
                HIP_TRY(ctx, hipMalloc(&ctx->d_hdr, n * sizeof(float)));
                ctx->hdr_cap = n;
            }
            if (!ctx->d_hdr) return fail(ctx, IXG_INVALID, "no device buffer for the panorama");
            hdr_dst = ctx->d_hdr;
        }
        HIP_TRY(ctx, hipMemcpyAsync(hdr_dst, s->hdr_map->rgb, n * sizeof(float), hipMemcpyHostToDevice,
                                    ctx->stream));
        std::memset(&P->env, 0, sizeof P->env);
        P->env.model = ixg::ENV_HDR;
        P->env.exposure = 1.0f;
        P->env.wb[0] = P->env.wb[1] = P->env.wb[2] = 1.0f;
        P->env.hdr_w = s->hdr_map->width;
        P->env.hdr_h = s->hdr_map->height;
        P->env.hdr_rgb = hdr_dst;
        return IXG_OK;
    }
    if (!ixg::resolve_environment(s->environment, &P->env))
        return fail(ctx, IXG_DECLINED, "environment not supported on the device path");
    return IXG_OK;
}

// The facet set (+ the frosted bitmask) into d_planes[IXG_MAX_PLANES + 2]
// (the context's table unless another is given); ixg_route has bounded
// n_planes by IXG_MAX_PLANES.  The staging copy lives on the stack: a copy
// from pageable memory has read its source when hipMemcpyAsync returns.
ixg_status upload_planes(ixg_ctx* ctx, const ixg_scene* s, float4* d_planes = nullptr) {
    if (!d_planes) d_planes = ctx->d_planes;
    float4 pl[IXG_MAX_PLANES + 2];
    for (uint32_t i = 0; i < s->n_planes; ++i)
        pl[i] = make_float4(s->planes[i].normal[0], s->planes[i].normal[1], s->planes[i].normal[2],
                            s->planes[i].d);
    if (s->n_planes)
        HIP_TRY(ctx, hipMemcpyAsync(d_planes, pl, s->n_planes * sizeof(float4), hipMemcpyHostToDevice,
                                    ctx->stream));
    if (any_frosted(s)) {  // the FacetFinish bitmask after the facet set
        uint32_t mask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t i = 0; i < s->n_planes; ++i)
            if (s->facet_finishes[i]) mask[i >> 5] |= 1u << (i & 31);
        std::memcpy(&pl[IXG_MAX_PLANES], mask, sizeof mask);
        HIP_TRY(ctx, hipMemcpyAsync(d_planes + IXG_MAX_PLANES, &pl[IXG_MAX_PLANES], sizeof mask,
                                    hipMemcpyHostToDevice, ctx->stream));
    }
    return IXG_OK;
}

ixg_status begin_call(ixg_ctx* ctx, const ixg_scene* s) {
    if (!ctx) return IXG_INVALID;
    if (ctx->lost) return fail(ctx, IXG_DEVICE_LOST, "context is lost; call ixg_try_recover");
    ixg_status st = validate(ctx, s);
    if (st != IXG_OK) return st;
    ixg_route_reason r = ixg_route(s);
    if (r != IXG_ROUTE_OK) return fail(ctx, IXG_DECLINED, "scene declined by ixg_route");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return IXG_OK;
}

bool take_injected_fault(ixg_ctx* ctx) {
    if (ctx->fail_countdown < 0) return false;
    if (ctx->fail_countdown == 0) {
        ctx->fail_countdown = -1;
        ctx->lost = true;
        ctx->last_error = "injected device loss";
        return true;
    }
    --ctx->fail_countdown;
    return false;
}

// Runs transport + reduction for every chunk of whole pixels, adding into d_accum.
// The device buffers one film call runs on: the facet table, the chunk's
// path records (room for xyz_paths paths), the work counter and the
// panorama slot (null = the context's own, grown on demand).
struct FilmBuffers {
    float4* planes;
    float* xyz;
    uint64_t xyz_paths;
    uint32_t* counter;
    float* hdr;
    size_t hdr_cap;  // floats
};

ixg_status run_film_on(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset, uint32_t spp, double* d_accum,
                       const volatile uint8_t* cancel, bool timed, const FilmBuffers& B) {
    ixg::FilmParams P;
    ixg_status st = build_params(ctx, s, sample_offset, spp, &P, B.hdr, B.hdr_cap);
    if (st != IXG_OK) return st;
    st = upload_planes(ctx, s, B.planes);
    if (st != IXG_OK) return st;
    const uint32_t n_pixels = s->width * s->height;
    const uint64_t pix_per_chunk = std::min<uint64_t>(std::max<uint64_t>(1, B.xyz_paths / spp), n_pixels);
    if (pix_per_chunk * spp > B.xyz_paths) return fail(ctx, IXG_RESOURCE_LIMIT, "spp beyond the path buffer");
    ctx->last_kernel_ms_total = 0.0;
    ctx->last_kernel_launches = 0;
    for (uint64_t first = 0; first < n_pixels; first += pix_per_chunk) {
        if (cancel && *cancel) return fail(ctx, IXG_CANCELLED, "cancelled");
        if (take_injected_fault(ctx)) return IXG_DEVICE_LOST;
        uint32_t npx = (uint32_t)std::min<uint64_t>(pix_per_chunk, n_pixels - first);
        P.first_path = first * spp;
        P.n_paths = (uint64_t)npx * spp;
        if (timed) HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIP_TRY(ctx, ixg_launch_film_persistent(&P, B.planes, B.xyz, B.counter, ctx->persistent_blocks,
                                                ctx->stream));
        if (timed) {
            HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
            HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
            float ms = 0.0f;
            HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            ctx->last_kernel_ms_total += ms;
            ctx->last_kernel_launches += 1;
        }
        HIP_TRY(ctx, ixg_launch_reduce(B.xyz, spp, npx, (uint32_t)first, d_accum, ctx->stream));
    }
    return IXG_OK;
}

// run_film_on with the context's own buffers, the path buffer grown to one chunk.
ixg_status run_film(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset, uint32_t spp,
                    double* d_accum, const volatile uint8_t* cancel, bool timed) {
    const uint32_t n_pixels = s->width * s->height;
    const uint64_t pix_per_chunk = std::min<uint64_t>(std::max<uint64_t>(1, ctx->chunk_paths / spp), n_pixels);
    ixg_status st = ensure_xyz(ctx, (size_t)pix_per_chunk * spp * 3);
    if (st != IXG_OK) return st;
    const FilmBuffers B{ctx->d_planes, ctx->d_xyz, pix_per_chunk * spp, ctx->d_counter, nullptr, 0};
    return run_film_on(ctx, s, sample_offset, spp, d_accum, cancel, timed, B);
}

}  // namespace

extern "C" {

ixg_status ixg_create(const ixg_opts* opts, ixg_ctx** out) {
    if (!out) return IXG_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return IXG_DECLINED;
    int dev = opts ? opts->device : 0;
    if (dev < 0 || dev >= n) return IXG_INVALID;
    ixg_ctx* ctx = new ixg_ctx();
    ctx->device = dev;
    if (opts && opts->chunk_paths) ctx->chunk_paths = opts->chunk_paths;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
        hipMalloc(&ctx->d_planes, (IXG_MAX_PLANES + 2) * sizeof(float4)) != hipSuccess ||
        hipMalloc(&ctx->d_counter, sizeof(uint32_t)) != hipSuccess ||
        ixg_persistent_grid(dev, &ctx->persistent_blocks) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        ixg_destroy(ctx);
        return IXG_DEVICE_LOST;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess)
        ctx->label = std::string(prop.name) + " (" + prop.gcnArchName + ")";
    *out = ctx;
    return IXG_OK;
}

void ixg_destroy(ixg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->metric_plan && ctx->metric_plan_free) ctx->metric_plan_free(ctx->metric_plan);
    if (ctx->d_planes) (void)hipFree(ctx->d_planes);
    if (ctx->d_hdr) (void)hipFree(ctx->d_hdr);
    if (ctx->d_xyz) (void)hipFree(ctx->d_xyz);
    if (ctx->d_counter) (void)hipFree(ctx->d_counter);
    if (ctx->d_accum) (void)hipFree(ctx->d_accum);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int ixg_is_lost(const ixg_ctx* ctx) { return ctx ? (ctx->lost ? 1 : 0) : 1; }

ixg_status ixg_adapter_label(const ixg_ctx* ctx, char* buf, size_t n) {
    if (!ctx || !buf || n == 0) return IXG_INVALID;
    std::snprintf(buf, n, "%s", ctx->label.c_str());
    return IXG_OK;
}

ixg_status ixg_last_error(const ixg_ctx* ctx, char* buf, size_t n) {
    if (!ctx || !buf || n == 0) return IXG_INVALID;
    std::snprintf(buf, n, "%s", ctx->last_error.c_str());
    return IXG_OK;
}

ixg_status ixg_try_recover(ixg_ctx* ctx) {
    if (!ctx) return IXG_INVALID;
    if (!ctx->lost) return IXG_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return IXG_DEVICE_LOST;
    (void)hipGetLastError();
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return IXG_DEVICE_LOST;
    ctx->lost = false;
    ctx->last_error.clear();
    return IXG_OK;
}

ixg_status ixg_inject_fault(ixg_ctx* ctx, int32_t fail_on_chunk) {
    if (!ctx) return IXG_INVALID;
    ctx->fail_countdown = fail_on_chunk;
    return IXG_OK;
}

ixg_route_reason ixg_route(const ixg_scene* s) {
    if (!s || !s->material || !s->planes) return IXG_ROUTE_UNSUPPORTED_MATERIAL;
    if (s->n_planes == 0 || s->n_planes > IXG_MAX_PLANES) return IXG_ROUTE_DEVICE_LIMITS;
    if (!material_routable(*s->material)) return IXG_ROUTE_UNSUPPORTED_MATERIAL;
    ixg::PresetInfo pi;
    if (s->hdr_map) {
        // a panorama: frosted facets and scatters would light-sample it by its
        // importance distribution, which is not carried
        const ixg_hdr_map& h = *s->hdr_map;
        if (!h.rgb || h.width == 0 || h.height == 0 || (uint64_t)h.width * h.height > IXG_MAX_HDR_TEXELS)
            return IXG_ROUTE_UNSUPPORTED_ENVIRONMENT;
        if (any_frosted(s) || s->material->scattering_sigma_s > 0.0f) return IXG_ROUTE_UNSUPPORTED_ENVIRONMENT;
        return IXG_ROUTE_OK;
    }
    if (!ixg::preset_info(s->environment.preset, &pi)) return IXG_ROUTE_UNSUPPORTED_ENVIRONMENT;
    if (!std::isfinite(s->environment.backdrop)) return IXG_ROUTE_UNSUPPORTED_ENVIRONMENT;
    // DaylightSun: the sun disc is light-sampled at exterior frosted events
    // (built); scatters light-sample only a panorama (hg_nee.rs:160-162)
    return IXG_ROUTE_OK;
}

ixg_status ixg_accumulate(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset, uint32_t spp,
                          double* accum_xyz, const volatile uint8_t* cancel) {
    ixg_status st = begin_call(ctx, s);
    if (st != IXG_OK) return st;
    if (!accum_xyz) return fail(ctx, IXG_INVALID, "null accumulator");
    if (spp == 0) return IXG_OK;
    const size_t n = (size_t)s->width * s->height * 3;
    st = ensure_accum(ctx, n);
    if (st != IXG_OK) return st;
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_accum, 0, n * sizeof(double), ctx->stream));
    st = run_film(ctx, s, sample_offset, spp, ctx->d_accum, cancel, false);
    if (st != IXG_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        return st;
    }
    std::vector<double> host(n);
    HIP_TRY(ctx, hipMemcpyAsync(host.data(), ctx->d_accum, n * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < n; ++i) accum_xyz[i] += host[i];
    return IXG_OK;
}

ixg_status ixg_accumulate_batch(ixg_ctx* ctx, const ixg_batch_item* items, uint32_t n,
                                const volatile uint8_t* cancel, ixg_batch_done_fn on_done, void* user) {
    if (!ctx || (n && !items) || !on_done) return IXG_INVALID;
    uint32_t reported = 0;  // items [0, reported) have had their on_done
    auto report = [&](uint32_t i, ixg_status outcome) {
        on_done(user, i, outcome);
        reported = i + 1;
    };
    auto report_rest = [&](ixg_status outcome) {
        while (reported < n) report(reported, outcome);
    };
    if (n == 0) return IXG_OK;
    if (ctx->lost || hipSetDevice(ctx->device) != hipSuccess) {
        if (!ctx->lost) ctx->last_error = "hipSetDevice failed";
        report_rest(IXG_DECLINED);
        return IXG_OK;
    }
    // The output ring: two device slots and two pinned host slots, sized once
    // for the largest picture, so picture i+1 is queued while picture i is
    // read back.
    size_t cap = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const ixg_scene* s = items[i].scene;
        if (s && s->width && s->height && (uint64_t)s->width * s->height <= (1ull << 31))
            cap = std::max(cap, (size_t)s->width * s->height * 3);
    }
    double* d_slot[2] = {nullptr, nullptr};
    double* h_slot[2] = {nullptr, nullptr};
    hipEvent_t ready[2] = {nullptr, nullptr};
    auto release = [&]() {
        (void)hipStreamSynchronize(ctx->stream);
        for (int k = 0; k < 2; ++k) {
            if (d_slot[k]) (void)hipFree(d_slot[k]);
            if (h_slot[k]) (void)hipHostFree(h_slot[k]);
            if (ready[k]) (void)hipEventDestroy(ready[k]);
        }
    };
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess && cap; ++k) {
        e = hipMalloc(&d_slot[k], cap * sizeof(double));
        if (e == hipSuccess) e = hipHostMalloc((void**)&h_slot[k], cap * sizeof(double), hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ready[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        hip_fail(ctx, e, "batch output ring");
        release();
        report_rest(IXG_DECLINED);
        return IXG_OK;
    }
    int pending = -1, pending_slot = 0;
    size_t pending_n = 0;
    // Waits for the queued picture's readback and reports it Done.
    auto finish_pending = [&]() -> bool {
        if (pending < 0) return true;
        hipError_t w = hipEventSynchronize(ready[pending_slot]);
        if (w != hipSuccess) {
            hip_fail(ctx, w, "batch readback");
            return false;
        }
        double* out = items[pending].out;
        const double* src = h_slot[pending_slot];
        for (size_t k = 0; k < pending_n; ++k) out[k] += src[k];
        report((uint32_t)pending, IXG_OK);
        pending = -1;
        return true;
    };
    bool stopped = false;
    for (uint32_t i = 0; i < n && !stopped; ++i) {
        const ixg_batch_item& it = items[i];
        ixg_status st = it.out ? begin_call(ctx, it.scene) : fail(ctx, IXG_INVALID, "null batch output");
        if (st == IXG_DEVICE_LOST) break;
        if (st == IXG_OK && cancel && *cancel) st = IXG_CANCELLED;
        if (st == IXG_OK && it.samples == 0) {  // nothing to trace: Done, nothing added
            if (!finish_pending()) break;
            report(i, IXG_OK);
            continue;
        }
        const int slot = (int)(i & 1u);
        const size_t count = st == IXG_OK ? (size_t)it.scene->width * it.scene->height * 3 : 0;
        if (st == IXG_OK) {
            hipError_t z = hipMemsetAsync(d_slot[slot], 0, count * sizeof(double), ctx->stream);
            st = z == hipSuccess ? run_film(ctx, it.scene, it.first_sample, it.samples, d_slot[slot], cancel, false)
                                 : hip_fail(ctx, z, "batch slot clear");
        }
        if (st == IXG_DEVICE_LOST) break;  // declines the queued picture too
        if (st == IXG_CANCELLED) {
            // the queued picture was fully enqueued before the flag: it finishes
            if (finish_pending()) report_rest(IXG_CANCELLED);
            stopped = true;
            break;
        }
        if (st != IXG_OK) {  // declined or malformed: this picture only
            if (!finish_pending()) break;
            report(i, st);
            continue;
        }
        e = hipMemcpyAsync(h_slot[slot], d_slot[slot], count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(ready[slot], ctx->stream);
        if (e != hipSuccess) {
            hip_fail(ctx, e, "batch readback enqueue");
            break;
        }
        // picture i is queued behind picture i-1: now read i-1 back
        if (!finish_pending()) break;
        pending = (int)i;
        pending_slot = slot;
        pending_n = count;
    }
    if (!stopped && reported < n && !ctx->lost) finish_pending();
    release();
    report_rest(IXG_DECLINED);  // anything left after a device loss
    return IXG_OK;
}

ixg_status ixg_accumulate_device(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset,
                                 uint32_t spp, double* d_accum_xyz) {
    ixg_status st = begin_call(ctx, s);
    if (st != IXG_OK) return st;
    if (!d_accum_xyz) return fail(ctx, IXG_INVALID, "null accumulator");
    if (spp == 0) return IXG_OK;
    st = run_film(ctx, s, sample_offset, spp, d_accum_xyz, nullptr, true);
    if (st != IXG_OK) return st;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return IXG_OK;
}

ixg_status ixg_trace_paths(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset, uint32_t spp,
                           float* out_xyz, ixg_path_events* events) {
    ixg_status st = begin_call(ctx, s);
    if (st != IXG_OK) return st;
    const uint64_t paths = (uint64_t)s->width * s->height * spp;
    if (paths == 0) return IXG_OK;
    if (paths > (1ull << 26)) return fail(ctx, IXG_RESOURCE_LIMIT, "diagnostic trace limited to 2^26 paths");
    ixg::FilmParams P;
    st = build_params(ctx, s, sample_offset, spp, &P);
    if (st != IXG_OK) return st;
    st = upload_planes(ctx, s);
    if (st != IXG_OK) return st;
    float* d_xyz = nullptr;
    ixg_path_events* d_ev = nullptr;
    HIP_TRY(ctx, hipMalloc(&d_xyz, paths * 3 * sizeof(float)));
    if (hipMalloc(&d_ev, paths * sizeof(ixg_path_events)) != hipSuccess) {
        (void)hipFree(d_xyz);
        return hip_fail(ctx, hipErrorOutOfMemory, "hipMalloc(events)");
    }
    P.first_path = 0;
    P.n_paths = paths;
    hipError_t e = ixg_launch_film_events(&P, ctx->d_planes, d_xyz, d_ev, nullptr, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && out_xyz) e = hipMemcpy(out_xyz, d_xyz, paths * 3 * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && events) e = hipMemcpy(events, d_ev, paths * sizeof(ixg_path_events), hipMemcpyDeviceToHost);
    (void)hipFree(d_xyz);
    (void)hipFree(d_ev);
    if (e != hipSuccess) return hip_fail(ctx, e, "trace_paths");
    return IXG_OK;
}

ixg_status ixg_film_samples(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset, uint32_t spp,
                            float* out_xyz) {
    ixg_status st = begin_call(ctx, s);
    if (st != IXG_OK) return st;
    if (!out_xyz) return fail(ctx, IXG_INVALID, "null out_xyz");
    const uint64_t paths = (uint64_t)s->width * s->height * spp;
    if (paths == 0) return IXG_OK;
    if (paths > (1ull << 26)) return fail(ctx, IXG_RESOURCE_LIMIT, "diagnostic samples limited to 2^26 paths");
    ixg::FilmParams P;
    st = build_params(ctx, s, sample_offset, spp, &P);
    if (st != IXG_OK) return st;
    st = upload_planes(ctx, s);
    if (st != IXG_OK) return st;
    // the chunking and launch of run_film, with each chunk's path records
    // copied out instead of reduced
    const uint32_t n_pixels = s->width * s->height;
    uint64_t pix_per_chunk = std::max<uint64_t>(1, ctx->chunk_paths / spp);
    pix_per_chunk = std::min<uint64_t>(pix_per_chunk, n_pixels);
    st = ensure_xyz(ctx, (size_t)pix_per_chunk * spp * 3);
    if (st != IXG_OK) return st;
    for (uint64_t first = 0; first < n_pixels; first += pix_per_chunk) {
        const uint32_t npx = (uint32_t)std::min<uint64_t>(pix_per_chunk, n_pixels - first);
        P.first_path = first * spp;
        P.n_paths = (uint64_t)npx * spp;
        HIP_TRY(ctx, ixg_launch_film_persistent(&P, ctx->d_planes, ctx->d_xyz, ctx->d_counter,
                                                ctx->persistent_blocks, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(out_xyz + 3 * P.first_path, ctx->d_xyz, P.n_paths * 3 * sizeof(float),
                                    hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return IXG_OK;
}

ixg_status ixg_trace_rays(ixg_ctx* ctx, const ixg_scene* s, const float* rays, uint32_t n,
                          float* out_xyz, ixg_path_events* events) {
    ixg_status st = begin_call(ctx, s);
    if (st != IXG_OK) return st;
    if (!rays || !out_xyz) return fail(ctx, IXG_INVALID, "null rays/out_xyz");
    if (n == 0) return IXG_OK;
    if (n > (1u << 24)) return fail(ctx, IXG_RESOURCE_LIMIT, "at most 2^24 rays per call");
    ixg::FilmParams P;
    st = build_params(ctx, s, 0, 1, &P);
    if (st != IXG_OK) return st;
    st = upload_planes(ctx, s);
    if (st != IXG_OK) return st;
    float *d_rays = nullptr, *d_xyz = nullptr;
    ixg_path_events* d_ev = nullptr;
    hipError_t e = hipMalloc(&d_rays, (size_t)n * 8 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_xyz, (size_t)n * 3 * sizeof(float));
    if (e == hipSuccess && events) e = hipMalloc(&d_ev, (size_t)n * sizeof(ixg_path_events));
    if (e == hipSuccess) e = hipMemcpy(d_rays, rays, (size_t)n * 8 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = ixg_launch_rays(&P, ctx->d_planes, d_rays, n, d_xyz, d_ev, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out_xyz, d_xyz, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && events) e = hipMemcpy(events, d_ev, (size_t)n * sizeof(ixg_path_events), hipMemcpyDeviceToHost);
    if (d_rays) (void)hipFree(d_rays);
    if (d_xyz) (void)hipFree(d_xyz);
    if (d_ev) (void)hipFree(d_ev);
    if (e != hipSuccess) return hip_fail(ctx, e, "trace_rays");
    return IXG_OK;
}

ixg_status ixg_probe(ixg_ctx* ctx, const ixg_scene* s, int32_t fn, const float* in, uint32_t n, float* out) {
    ixg_status st = begin_call(ctx, s);
    if (st != IXG_OK) return st;
    uint32_t in_w = 0, out_w = 0;
    if (!ixg_probe_shape(fn, &in_w, &out_w)) return fail(ctx, IXG_INVALID, "unknown probe function");
    if (!in || !out) return fail(ctx, IXG_INVALID, "null in/out");
    if (n == 0) return IXG_OK;
    if (n > (1u << 24)) return fail(ctx, IXG_RESOURCE_LIMIT, "at most 2^24 records per call");
    ixg::FilmParams P;
    st = build_params(ctx, s, 0, 1, &P);
    if (st != IXG_OK) return st;
    st = upload_planes(ctx, s);
    if (st != IXG_OK) return st;
    float *d_in = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc(&d_in, (size_t)n * in_w * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_out, (size_t)n * out_w * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(d_in, in, (size_t)n * in_w * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = ixg_launch_probe(&P, ctx->d_planes, fn, d_in, n, d_out, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, (size_t)n * out_w * sizeof(float), hipMemcpyDeviceToHost);
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    if (e != hipSuccess) return hip_fail(ctx, e, "probe");
    return IXG_OK;
}

ixg_status ixg_count_events(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset, uint32_t spp,
                            uint64_t totals[IXG_EVENT_SLOTS]) {
    ixg_status st = begin_call(ctx, s);
    if (st != IXG_OK) return st;
    if (!totals) return fail(ctx, IXG_INVALID, "null totals");
    ixg::FilmParams P;
    st = build_params(ctx, s, sample_offset, spp, &P);
    if (st != IXG_OK) return st;
    st = upload_planes(ctx, s);
    if (st != IXG_OK) return st;
    unsigned long long* d_tot = nullptr;
    HIP_TRY(ctx, hipMalloc(&d_tot, ixg::EV_NUM * sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(d_tot, 0, ixg::EV_NUM * sizeof(unsigned long long), ctx->stream);
    const uint64_t paths = (uint64_t)s->width * s->height * spp;
    for (uint64_t first = 0; e == hipSuccess && first < paths; first += ctx->chunk_paths) {
        P.first_path = first;
        P.n_paths = std::min<uint64_t>(ctx->chunk_paths, paths - first);
        e = ixg_launch_film_events(&P, ctx->d_planes, nullptr, nullptr, d_tot, ctx->stream);
    }
    unsigned long long host[ixg::EV_NUM] = {};
    if (e == hipSuccess) e = hipMemcpyAsync(host, d_tot, sizeof host, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_tot);
    if (e != hipSuccess) return hip_fail(ctx, e, "count_events");
    for (int i = 0; i < ixg::EV_NUM; ++i) totals[i] = host[i];
    return IXG_OK;
}

// ixg_accumulate_device on caller-provided device buffers (the static
// scratch of the allocation-free entry point, film_noalloc.hip): no
// allocation on any path that returns IXG_OK.
ixg_status ixg_internal_accumulate_on(ixg_ctx* ctx, const ixg_scene* s, uint32_t sample_offset, uint32_t spp,
                                      double* d_accum_xyz, float4* d_planes, float* d_xyz, uint64_t xyz_paths,
                                      uint32_t* d_counter, float* d_hdr, size_t hdr_cap) {
    ixg_status st = begin_call(ctx, s);
    if (st != IXG_OK) return st;
    if (!d_accum_xyz) return fail(ctx, IXG_INVALID, "null accumulator");
    if (spp == 0) return IXG_OK;
    if (spp > xyz_paths) return fail(ctx, IXG_RESOURCE_LIMIT, "spp larger than the scratch path capacity");
    const FilmBuffers B{d_planes, d_xyz, xyz_paths, d_counter, d_hdr, hdr_cap};
    st = run_film_on(ctx, s, sample_offset, spp, d_accum_xyz, nullptr, true, B);
    if (st != IXG_OK) return st;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return IXG_OK;
}

// internals shared with the other host translation units of libixg
hipStream_t ixg_internal_stream(ixg_ctx* ctx) { return ctx->stream; }
int ixg_internal_device(ixg_ctx* ctx) { return ctx->device; }
ixg_status ixg_internal_fail(ixg_ctx* ctx, ixg_status st, const char* what) { return fail(ctx, st, what); }
ixg_status ixg_internal_hip_fail(ixg_ctx* ctx, hipError_t e, const char* where) { return hip_fail(ctx, e, where); }
int ixg_internal_lost(ixg_ctx* ctx) { return ctx->lost ? 1 : 0; }
void** ixg_internal_metric_plan(ixg_ctx* ctx, void (*free_fn)(void*)) {
    if (free_fn) ctx->metric_plan_free = free_fn;
    return &ctx->metric_plan;
}

void* ixg_stream(ixg_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

double ixg_last_kernel_ms(const ixg_ctx* ctx) {
    if (!ctx || ctx->last_kernel_launches == 0) return 0.0;
    return ctx->last_kernel_ms_total / ctx->last_kernel_launches;
}

uint32_t ixg_last_kernel_launches(const ixg_ctx* ctx) { return ctx ? ctx->last_kernel_launches : 0; }

}  // extern "C"
