// probe.hip -- per-function probes: one device function of the film path,
// chosen by id, evaluated on a batch of input records.
//
// The reference checks its GPU transport function by function against the
// CPU tracer, each with a stated ULP budget (renderer/gpu/ulp.rs, the
// *_check.rs modules under renderer/gpu/, their last run in
// crates/indicatrix/docs/gpu_harness_last_run.json:24-73).  This kernel is
// the device side of that tier: every case calls the same __forceinline__
// function the film kernels call (film_paths.h, uniaxial.h, device_common.h),
// so a probe and a path cannot drift apart.  Record layouts (floats; integer
// fields carry their bits) are listed with the ids in ixg.h.
#include "film_paths.h"

namespace ixg {
namespace {

struct ProbeShape {
    uint8_t in, out;
};

// in / out floats per record, indexed by ixg_probe_fn.
constexpr ProbeShape kProbeShapes[IXG_PROBE_COUNT] = {
    {2, 12},  // SAMPLE_DRAWS
    {2, 6},   // CAMERA_RAY
    {1, 3},   // CIE_CMF
    {1, 1},   // ILLUMINANT
    {7, 1},   // ENVIRONMENT
    {6, 4},   // FRAME_ROTATION
    {8, 5},   // FRESNEL_REFLECTION
    {8, 6},   // FRESNEL_TRANSMISSION
    {7, 4},   // TIR_RETARDATION
    {8, 24},  // CHANNEL_OPTICS
    {19, 8},  // PLEOCHROIC_ALPHA
    {6, 3},   // ORDINARY_EIGEN
    {8, 3},   // EXTRAORDINARY_POYNTING
    {9, 1},   // THETA_C
    {3, 2},   // UNIAXIAL_INDICES
    {11, 18}, // ENTRY_SOLVE
    {13, 13}, // INTERNAL_SOLVE
    {15, 5},  // EXIT_TRANSMISSION
    {6, 3},   // HG_DIRECTION
    {5, 3},   // COSINE_HEMISPHERE
    {8, 3},   // SHADING_NORMAL
    {26, 3},  // FINISH
    {6, 2},   // INTERSECT
    {1, 1},   // EXP_LANE
    {1, 1},   // EXP
    {4, 2},   // BIAX_WAVE_INDICES
    {5, 3},   // BIAX_EIGEN
    {5, 3},   // BIAX_MODE_POYNTING
    {8, 3},   // BIAX_RESOLVE_ENTRY
    {8, 1},   // BIAX_ASSIGNED_ALPHA
};

__device__ __forceinline__ V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }
__device__ __forceinline__ void st3(float* p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
__device__ __forceinline__ float4 ld4(const float* p) { return make_float4(p[0], p[1], p[2], p[3]); }
__device__ __forceinline__ void st4(float* p, float4 s) {
    p[0] = s.x;
    p[1] = s.y;
    p[2] = s.z;
    p[3] = s.w;
}
__device__ __forceinline__ void stc(float* p, Cx c) {
    p[0] = c.re;
    p[1] = c.im;
}

// The uniaxial interface frame of an incident wave normal k on a facet
// normal that faces the incident side, as the bounce forms it
// (compute_bounce_refraction_geometry, geometry.rs:306-425).
__device__ __forceinline__ UFrame probe_frame(V3 k, V3 normal, V3 c) {
    const float cos_i = clampf(dot(neg(k), normal), 0.0f, 1.0f);
    const float sin_i = sqrt_rn(fmaxf(fma_(cos_i, -cos_i, 1.0f), 0.0f));
    return uframe_build(k, normal, c, cos_i, sin_i);
}

__global__ void __launch_bounds__(256) probe_kernel(FilmParams P, const float4* __restrict__ planes, int32_t fn,
                                                    const float* __restrict__ in, uint32_t in_w, uint32_t n,
                                                    float* __restrict__ out, uint32_t out_w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* a = in + (size_t)i * in_w;
    float* o = out + (size_t)i * out_w;
    switch (fn) {
    case IXG_PROBE_SAMPLE_DRAWS: {
        uint32_t seed;
        float jx, jy, hero, lambda[NCH];
        sample_draws(__float_as_uint(a[0]), __float_as_uint(a[1]), seed, jx, jy, hero);
        hero_comb(hero, lambda);
        o[0] = jx;
        o[1] = jy;
        o[2] = hero;
        o[3] = __uint_as_float(seed);
        for (int c = 0; c < NCH; ++c) o[4 + c] = lambda[c];
        break;
    }
    case IXG_PROBE_CAMERA_RAY: {
        V3 org, dir;
        uint32_t seed;
        float hero;
        camera_sample(P, __float_as_uint(a[0]), __float_as_uint(a[1]), org, dir, seed, hero);
        st3(o, org);
        st3(o + 3, dir);
        break;
    }
    case IXG_PROBE_CIE_CMF:
        cie_cmf(a[0], o[0], o[1], o[2]);
        break;
    case IXG_PROBE_ILLUMINANT:
        o[0] = spec_power(P.env, a[0]);
        break;
    case IXG_PROBE_ENVIRONMENT:
        o[0] = env_radiance<true>(P.env, ld3(a), ld3(a + 3), a[6]);
        break;
    case IXG_PROBE_FRAME_ROTATION: {
        float c2, s2;
        rotation_double_angle(a[0], a[1], c2, s2);
        st4(o, frame_rotate(c2, s2, ld4(a + 2)));
        break;
    }
    case IXG_PROBE_FRESNEL_REFLECTION: {
        float ru;
        st4(o, fresnel_reflect(a[0], a[1], a[2], a[3], ld4(a + 4), ru));
        o[4] = ru;
        break;
    }
    case IXG_PROBE_FRESNEL_TRANSMISSION: {
        float rs, rp;
        st4(o, fresnel_transmit(a[0], a[1], a[2], a[3], ld4(a + 4), rs, rp));
        o[4] = rs;
        o[5] = rp;
        break;
    }
    case IXG_PROBE_TIR_RETARDATION: {
        float cd, sd;
        tir_retardation(a[0], a[1], a[2], cd, sd);
        st4(o, mueller_tir(cd, sd, ld4(a + 3)));
        break;
    }
    case IXG_PROBE_CHANNEL_OPTICS: {
        float n_o[NCH], ao[NCH], ae[NCH];
        path_optics<true>(P, a, n_o, ao, ae);
        for (int c = 0; c < NCH; ++c) {
            o[c] = n_o[c];
            o[NCH + c] = ao[c];
            o[2 * NCH + c] = ae[c];
        }
        break;
    }
    case IXG_PROBE_PLEOCHROIC_ALPHA: {
        PathState S;
        for (int c = 0; c < NCH; ++c) {
            S.alpha_o[c] = a[3 + c];
            S.alpha_e[c] = a[3 + NCH + c];
        }
        float alpha[NCH];
        channel_alphas(P.abs, S, ld3(a), alpha);
        for (int c = 0; c < NCH; ++c) o[c] = alpha[c];
        break;
    }
    case IXG_PROBE_ORDINARY_EIGEN:
        st3(o, ordinary_eigen(ld3(a), ld3(a + 3)));
        break;
    case IXG_PROBE_EXTRAORDINARY_POYNTING:
        st3(o, extraordinary_poynting_dir(ld3(a), ld3(a + 3), a[6], a[7]));
        break;
    case IXG_PROBE_THETA_C: {
        const V3 k = ld3(a + 2), normal = ld3(a + 5);
        const float cos_i = clampf(dot(neg(k), normal), 0.0f, 1.0f);
        const V3 cax = v3(P.an.c_axis[0], P.an.c_axis[1], P.an.c_axis[2]);
        o[0] = theta_c_for_bounce(P, a[0], a[1], k, normal, cax, cos_i, a[8] != 0.0f);
        break;
    }
    case IXG_PROBE_UNIAXIAL_INDICES: {
        const float n_e = e_index_at(P, a[0], a[1]);
        o[0] = n_e;
        o[1] = effective_extraordinary_index(a[1], n_e, a[2]);
        break;
    }
    case IXG_PROBE_ENTRY_SOLVE: {
        const V3 c = ld3(a + 2);
        const UFrame F = probe_frame(ld3(a + 5), ld3(a + 8), c);
        EntryInc I;
        entry_incidence(F, I);
        EntrySol sol;
        entry_solve(I, a[0], a[1], c, F, sol, false);
        for (int p = 0; p < 2; ++p)
            for (int j = 0; j < 4; ++j) stc(o + 8 * p + 2 * j, sol.x[p][j]);
        o[16] = sol.flux_o;
        o[17] = sol.flux_e;
        break;
    }
    case IXG_PROBE_INTERNAL_SOLVE:
    case IXG_PROBE_EXIT_TRANSMISSION: {
        const V3 c = ld3(a + 3);
        const UFrame F = probe_frame(ld3(a + 6), ld3(a + 9), c);
        const InternalSol sol = internal_solve(a[0], a[1], a[2], c, F, a[12] != 0.0f);
        if (fn == IXG_PROBE_EXIT_TRANSMISSION) {
            float iu;
            st4(o, exit_transmission(sol, a[13], div_rn(1.0f, 1.0f - a[14]), iu));
            o[4] = iu;
            break;
        }
        stc(o, sol.r_o);
        stc(o + 2, sol.r_e);
        stc(o + 4, sol.t_s);
        stc(o + 6, sol.t_p);
        o[8] = sol.flux_ro;
        o[9] = sol.flux_re;
        o[10] = sol.flux_ts;
        o[11] = sol.flux_tp;
        o[12] = sol.flux_inc;
        break;
    }
    case IXG_PROBE_HG_DIRECTION:
        st3(o, hg_direction(a[0], a[1], a[2], ld3(a + 3)));
        break;
    case IXG_PROBE_COSINE_HEMISPHERE:
        st3(o, cosine_hemisphere(a[0], a[1], ld3(a + 2)));
        break;
    case IXG_PROBE_SHADING_NORMAL:
        st3(o, shading_normal_near_edge(planes, P.n_planes, ld3(a), (int)__float_as_uint(a[3]), ld3(a + 4), a[7]));
        break;
    case IXG_PROBE_FINISH: {
        const uint64_t compat = (uint64_t)__float_as_uint(a[0]) | ((uint64_t)__float_as_uint(a[1]) << 32);
        float xyz[3];
        finish_xyz(P, compat, a + 2, a + 2 + NCH, a + 2 + 2 * NCH, xyz);
        o[0] = xyz[0];
        o[1] = xyz[1];
        o[2] = xyz[2];
        break;
    }
    case IXG_PROBE_INTERSECT: {
        float t = 0.0f;
        const int f = intersect(planes, P.n_planes, ld3(a), ld3(a + 3), t);
        o[0] = __uint_as_float((uint32_t)f);
        o[1] = f >= 0 ? t : 0.0f;
        break;
    }
    case IXG_PROBE_EXP_LANE:
        o[0] = exp_lane(a[0]);
        break;
    case IXG_PROBE_EXP:
        o[0] = exp_rn(a[0]);
        break;
    case IXG_PROBE_BIAX_WAVE_INDICES: {
        const Biax B = biax_at(P, dispersion(P.disp_kind, P.disp, a[0]));
        biax_wave_indices(B, biax_frame(P), ld3(a + 1), o[0], o[1]);
        break;
    }
    case IXG_PROBE_BIAX_EIGEN:
        st3(o, biax_eigen_d(biax_at(P, dispersion(P.disp_kind, P.disp, a[0])), biax_frame(P), ld3(a + 1),
                            a[4] != 0.0f));
        break;
    case IXG_PROBE_BIAX_MODE_POYNTING:
        st3(o, biax_mode_poynting(biax_at(P, dispersion(P.disp_kind, P.disp, a[0])), biax_frame(P), ld3(a + 1),
                                  a[4] != 0.0f));
        break;
    case IXG_PROBE_BIAX_RESOLVE_ENTRY: {
        const float nb = dispersion(P.disp_kind, P.disp, a[0]);
        const V3 k = ld3(a + 1), normal = ld3(a + 4);
        const float cos_i = clampf(dot(neg(k), normal), 0.0f, 1.0f);
        st3(o, biax_resolve_entry(biax_at(P, nb), biax_frame(P), k, normal, cos_i, nb, a[7] != 0.0f));
        break;
    }
    case IXG_PROBE_BIAX_ASSIGNED_ALPHA: {
        const V3 l = biax_assigned_locals(P.abs, biax_at(P, dispersion(P.disp_kind, P.disp, a[0])), biax_frame(P),
                                          ld3(a + 1), a[4] != 0.0f);
        o[0] = biax_tensor_qf(l, a[5], a[6], a[7]);
        break;
    }
    default:
        break;
    }
}

}  // namespace
}  // namespace ixg

extern "C" int ixg_probe_shape(int32_t fn, uint32_t* in_floats, uint32_t* out_floats) {
    if (fn < 0 || fn >= IXG_PROBE_COUNT) return 0;
    if (in_floats) *in_floats = ixg::kProbeShapes[fn].in;
    if (out_floats) *out_floats = ixg::kProbeShapes[fn].out;
    return 1;
}

extern "C" hipError_t ixg_launch_probe(const ixg::FilmParams* P, const float4* d_planes, int32_t fn,
                                       const float* d_in, uint32_t n, float* d_out, hipStream_t stream) {
    if (fn < 0 || fn >= IXG_PROBE_COUNT) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(ixg::probe_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, *P, d_planes, fn, d_in,
                       (uint32_t)ixg::kProbeShapes[fn].in, n, d_out, (uint32_t)ixg::kProbeShapes[fn].out);
    return hipGetLastError();
}
