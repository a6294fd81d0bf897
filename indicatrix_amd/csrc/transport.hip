// transport.hip -- the film hot path on CDNA4 (gfx950).
//
// One lane traces one 8-lambda hero-wavelength path (pixel, absolute sample)
// through a convex facet set staged in LDS, with full Stokes-Mueller
// transport, exit-event spectral splitting and per-channel MIS families;
// FP32 throughout, FP64 only in the per-pixel reduction.
//
// The arithmetic mirrors the reference's CPU tracer operation for operation
// (the reference never contracts a*b+c; this file is compiled with
// -ffp-contract=off and writes __builtin_fmaf exactly where the reference
// writes mul_add), so integer outputs (facet sequence, bounce count, stop
// reason) match the CPU bit for bit and the radiance differs only through
// the device's atan2/sin/cos/atan/exp.
//
// Reference map (paths relative to crates/indicatrix/src/):
//   sampling draws          optics/raytracer/sampling.rs:66-218
//   bounce loop             optics/raytracer/transport/inner/bounce_body.rs:216-764
//   slab scan               simd/slab.rs:147-160, optics/raytracer/intersect.rs:101-132
//   Fresnel / TIR / refract optics/raytracer/refraction/{dispatch,tir,reflect_refract,exit_split}.rs
//   environment             optics/raytracer/environment/rig.rs
//   XYZ families            optics/raytracer/color.rs:145-172
//   per-pixel reduction     renderer/shaders/reduce_xyz.wgsl:55-90, optics/raytracer/sampling.rs:237
#include "device_common.h"

namespace ixg {

// ---------------------------------------------------------------------------
// Stokes/Mueller (sparse forms of polarization.rs:138-190; equal to the full
// column-by-column products except for the sign of exact zeros)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 mueller_fresnel(float a, float b, float c, float4 s) {
    return make_float4(a * s.x + b * s.y, b * s.x + a * s.y, c * s.z, c * s.w);
}
__device__ __forceinline__ float4 mueller_tir(float cd, float sd, float4 s) {
    return make_float4(s.x, s.y, cd * s.z + sd * s.w, (-sd) * s.z + cd * s.w);
}
__device__ __forceinline__ float4 scale4(float4 s, float f) {
    return make_float4(s.x * f, s.y * f, s.z * f, s.w * f);
}
// f32::midpoint for non-subnormal finite operands: round((a+b)/2) == 0.5*(a+b)
__device__ __forceinline__ float midpoint(float a, float b) { return 0.5f * (a + b); }

__device__ __forceinline__ void fresnel_reflection_coeffs(float rs, float rp, float& a, float& b,
                                                          float& c) {
    float rs2 = rs * rs, rp2 = rp * rp;
    a = midpoint(rs2, rp2);
    b = 0.5f * (rs2 - rp2);
    c = rs * rp;
}

__device__ __forceinline__ float tir_phase_delta(float n1k, float cos_i, float sin_i) {
    float th = (cos_i * sqrtf(fmaxf(fma_(n1k * n1k * sin_i, sin_i, -1.0f), 0.0f))) /
               fmaxf(n1k * sin_i * sin_i, 1e-6f);
    return 2.0f * atanf(th);
}

__device__ __forceinline__ void fresnel_rs_rp(float n1, float n2, float cos_i, float cos_t,
                                              float& rs, float& rp) {
    rs = fma_(n2, -cos_t, n1 * cos_i) / fma_(n2, cos_t, n1 * cos_i);
    rp = fma_(n1, -cos_t, n2 * cos_i) / fma_(n1, cos_t, n2 * cos_i);
}

__device__ __forceinline__ uint32_t fnv(uint32_t h, uint32_t v) { return (h ^ v) * 16777619u; }

// ---------------------------------------------------------------------------
// one path
// ---------------------------------------------------------------------------
struct PathOut {
    float xyz[3];
    ixg_path_events ev;
};

// One (pixel, absolute sample): sample_draws (sampling.rs:176-218) and
// Camera::generate_ray (camera.rs:149-167).
__device__ __forceinline__ void camera_sample(const FilmParams& P, uint32_t pixel, uint32_t sample,
                                              V3& org, V3& dir, uint32_t& seed, float& hero) {
    seed = hash_u32((pixel * 0x9e3779b9u) ^ (sample * 0x85ebca6bu));
    float rx = ld_base2(hash_u32(pixel ^ 0xA511E9B3u));
    float ry = ld_base2(hash_u32(pixel ^ 0x63D81B23u));
    float rh = ld_base2(hash_u32(pixel ^ 0x1B873593u));
    float jx = cp_rotate(ld_base2(sample), rx) - 0.5f;
    float jy = cp_rotate(radical_inverse(sample, 3), ry) - 0.5f;
    hero = cp_rotate(radical_inverse(sample, 5), rh);
    uint32_t px = pixel % P.width, py = pixel / P.width;
    float aspect = P.width_f / P.height_f;
    float u = (((float)px + jx) / P.width_f - 0.5f) * 2.0f * aspect * P.fov_tan;
    float v = (0.5f - ((float)py + jy) / P.height_f) * 2.0f * P.fov_tan;
    dir = normalize((v3(P.forward[0], P.forward[1], P.forward[2]) +
                     v3(P.right[0], P.right[1], P.right[2]) * u) +
                    v3(P.up[0], P.up[1], P.up[2]) * v);
    org = v3(P.cam_origin[0], P.cam_origin[1], P.cam_origin[2]);
}

// trace_spectral_ray (transport/mod.rs:118) for one explicit ray.
template <bool EVENTS>
__device__ __forceinline__ void trace_path(const FilmParams& P, const float4* __restrict__ planes,
                                           V3 org, V3 dir, const uint32_t seed, const float hero,
                                           PathOut& out, uint32_t* counts) {
    const V3 observer = neg(dir);

    // PathState::new (bounce_body.rs:258-356)
    float lambda[NCH], n_o[NCH], path_pdf[NCH], radiance[NCH], split[NCH];
    float4 st[NCH];
    uint32_t compat[NCH];
    const float lambda_hero = fma_(hero, 400.0f, 380.0f);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        float off = fma_((float)c, 50.0f, lambda_hero - 380.0f);
        float r = fmodf(off, 400.0f);
        r = r < 0.0f ? r + 400.0f : r;
        lambda[c] = 380.0f + r;
        n_o[c] = dispersion(P.disp_kind, P.disp, lambda[c]);
        path_pdf[c] = 1.0f;
        radiance[c] = 0.0f;
        split[c] = 0.0f;
        st[c] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
        compat[c] = 0xFFu;
    }
    V3 k = dir;
    bool inside = false, has_prev = false, escaped = false;
    V3 prev_plane = v3(0.0f, 0.0f, 0.0f);

    uint32_t bounces = P.max_bounces, term = 3u;
    uint32_t fh = 2166136261u;
    uint32_t n_tir = 0, n_refl = 0, n_refr = 0, n_probe = 0, n_rot = 0;

    for (uint32_t bounce = 0; bounce < P.max_bounces; ++bounce) {
        float t_hit = 0.0f;
        int facet = intersect(planes, P.n_planes, org, dir, t_hit);
        if (facet < 0) {
            // handle_miss (bounce_body/miss.rs:33-54)
            if (P.env.model == ENV_ISO_HEMISPHERE) {
                float f = iso_factor(P.env, dir, observer);
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    float env = f * (spec_power(P.env, lambda[c]) * P.env.exposure);
                    radiance[c] = fma_(fmaxf(st[c].x, 0.0f) * 1.0f, env, radiance[c]);
                }
            } else {
                DirLight L;
                direction_lighting(P.env, dir, observer, L);
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    float env = light_radiance(P.env, L, spec_power(P.env, lambda[c]));
                    radiance[c] = fma_(fmaxf(st[c].x, 0.0f) * 1.0f, env, radiance[c]);
                }
            }
            escaped = true;
            bounces = bounce + 1;
            term = 0u;
            break;
        }

        // facet_event (bounce_body.rs:611-707)
        const V3 hit = org + dir * t_hit;
        const float4 pf = planes[facet];
        V3 normal = v3(pf.x, pf.y, pf.z);
        // rotate_stokes_to_plane_of_incidence (absorption.rs:69-87)
        V3 cur_plane = normalize_or_zero(cross(k, normal));
        if (has_prev && dot(cur_plane, cur_plane) > 1e-6f && dot(prev_plane, prev_plane) > 1e-6f) {
            float cos_psi = clampf(dot(prev_plane, cur_plane), -1.0f, 1.0f);
            float sin_psi = dot(cross(prev_plane, cur_plane), normalize_or_zero(k));
            float psi = atan2f(sin_psi, cos_psi);
            float c2 = cosf(2.0f * psi), s2 = sinf(2.0f * psi);
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                float4 s = st[c];
                st[c] = make_float4(s.x, c2 * s.y + s2 * s.z, (-s2) * s.y + c2 * s.z, s.w);
            }
            if (EVENTS) ++n_rot;
        }
        prev_plane = cur_plane;
        has_prev = true;
        // apply_interior_segment: absorption-free isotropic media transmit
        // exp(-0) == 1 exactly, so only the normal flips.
        if (inside) normal = neg(normal);

        // compute_bounce_refraction_geometry (geometry.rs:306-425), isotropic
        const float cos_i = clampf(dot(neg(k), normal), 0.0f, 1.0f);
        const float sin_i = sqrtf(fmaxf(fma_(cos_i, -cos_i, 1.0f), 0.0f));
        const float one_m_c2 = fma_(cos_i, -cos_i, 1.0f);
        const float n1 = inside ? n_o[0] : 1.0f;
        const float n2 = inside ? 1.0f : n_o[0];
        const float eta = n1 / n2;
        const float sin2_t = eta * eta * one_m_c2;

        V3 new_k;
        bool new_inside = inside;
        if (sin2_t > 1.0f) {
            // apply_tir_bounce (tir.rs:56-160), scalar branch
            if (EVENTS) {
                ++n_tir;
                fh = fnv(fnv(fh, (uint32_t)facet), 1u);
            }
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                float n1k = inside ? n_o[c] : 1.0f;
                float n2k = inside ? 1.0f : n_o[c];
                float ek = n1k / n2k;
                float s2k = ek * ek * one_m_c2;
                if (s2k > 1.0f) {
                    float delta = tir_phase_delta(n1k, cos_i, sin_i);
                    st[c] = mueller_tir(cosf(delta), sinf(delta), st[c]);
                } else {
                    float ct = sqrtf(fmaxf(1.0f - s2k, 0.0f));
                    float rs, rp, a, b, cc;
                    fresnel_rs_rp(n1k, n2k, cos_i, ct, rs, rp);
                    fresnel_reflection_coeffs(rs, rp, a, b, cc);
                    st[c] = mueller_fresnel(a, b, cc, st[c]);
                    path_pdf[c] *= clampf(0.5f * fma_(rp, rp, rs * rs), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                }
            }
            new_k = k - normal * (2.0f * dot(k, normal));
        } else {
            // apply_partial_fresnel_bounce (dispatch.rs:307-380)
            const float cos_t = sqrtf(fmaxf(1.0f - sin2_t, 0.0f));
            float rs, rp;
            fresnel_rs_rp(n1, n2, cos_i, cos_t, rs, rp);
            const float r_unpol = clampf(0.5f * fma_(rp, rp, rs * rs), R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
            const float rng = unit_draw(seed, bounce, FRESNEL_BRANCH_STREAM);
            if (rng < r_unpol) {
                // apply_partial_reflect_bounce (reflect_refract.rs:33-71)
                if (EVENTS) {
                    ++n_refl;
                    fh = fnv(fnv(fh, (uint32_t)facet), 2u);
                }
                const float inv_r = 1.0f / r_unpol;
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    float n1k = inside ? n_o[c] : 1.0f;
                    float n2k = inside ? 1.0f : n_o[c];
                    float ek = n1k / n2k;
                    float s2k = ek * ek * one_m_c2;
                    float4 s;
                    if (s2k > 1.0f) {
                        float delta = tir_phase_delta(n1k, cos_i, sin_i);
                        s = mueller_tir(cosf(delta), sinf(delta), st[c]);
                    } else {
                        float ct = sqrtf(fmaxf(1.0f - s2k, 0.0f));
                        float a_, b_, a, b, cc;
                        fresnel_rs_rp(n1k, n2k, cos_i, ct, a_, b_);
                        path_pdf[c] *= clampf(0.5f * fma_(b_, b_, a_ * a_), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                        fresnel_reflection_coeffs(a_, b_, a, b, cc);
                        s = mueller_fresnel(a, b, cc, st[c]);
                    }
                    st[c] = scale4(s, inv_r);
                }
                new_k = k - normal * (2.0f * dot(k, normal));
            } else {
                // apply_refract_bounce (reflect_refract.rs:121-234), isotropic
                if (EVENTS) {
                    ++n_refr;
                    fh = fnv(fnv(fh, (uint32_t)facet), 3u);
                }
                const float s2d = fminf(eta * eta * one_m_c2, 1.0f);
                const float ct_dir = sqrtf(fmaxf(1.0f - s2d, 0.0f));
                const V3 refr = normalize(k * eta + normal * fma_(eta, cos_i, -ct_dir));
                const float inv_t = 1.0f / (1.0f - r_unpol);
                V3 dirs[NCH];
                uint32_t live = 0u, match_mask = 0u;
#pragma unroll
                for (int c = 0; c < NCH; ++c) {
                    // apply_refract_channel (reflect_refract.rs:517-591)
                    const float4 orig = st[c];
                    float n1k = inside ? n_o[c] : 1.0f;
                    float n2k = inside ? 1.0f : n_o[c];
                    float r = n1k / n2k;
                    float s2k = r * r * one_m_c2;
                    if (s2k > 1.0f) {
                        st[c] = scale4(st[c], 0.0f);
                        path_pdf[c] = 0.0f;
                        dirs[c] = v3(0.0f, 0.0f, 0.0f);
                        continue;
                    }
                    float ctk = sqrtf(fmaxf(1.0f - s2k, 0.0f));
                    V3 dk = normalize(k * r + normal * fma_(r, cos_i, -ctk));
                    bool match = dot(dk, refr) >= DIRECTION_MATCH_COS_TOL;
                    // compute_channel_transmission (exit_split.rs:69-101)
                    float num = 2.0f * n1k * cos_i;
                    float t_s = num / fma_(n2k, ctk, n1k * cos_i);
                    float t_p = num / fma_(n1k, ctk, n2k * cos_i);
                    float factor = (n2k * ctk) / fmaxf(n1k * cos_i, 1e-6f);
                    float ts2 = t_s * t_s * factor, tp2 = t_p * t_p * factor;
                    float4 tr = mueller_fresnel(midpoint(ts2, tp2), 0.5f * (ts2 - tp2), t_s * t_p * factor, orig);
                    tr = scale4(tr, inv_t);
                    float a_, b_;
                    fresnel_rs_rp(n1k, n2k, cos_i, ctk, a_, b_);
                    float ruk = clampf(0.5f * fma_(b_, b_, a_ * a_), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                    if (match) {
                        st[c] = tr;
                        path_pdf[c] *= 1.0f - ruk;
                        match_mask |= 1u << c;
                    } else {
                        // apply_channel_chromatic_termination (reflect_refract.rs:444-515)
                        float prefix = path_pdf[c];
                        st[c] = scale4(st[c], 0.0f);
                        path_pdf[c] = prefix * (1.0f - ruk);
                        if (inside && fmaxf(orig.x, 0.0f) > 0.0f) {
                            // try_split_exit_channel (exit_split.rs:128-163)
                            if (EVENTS) ++n_probe;
                            float tt;
                            if (intersect(planes, P.n_planes, hit + dk * 1e-4f, dk, tt) < 0) {
                                float env;
                                if (P.env.model == ENV_ISO_HEMISPHERE) {
                                    env = iso_factor(P.env, dk, observer) *
                                          (spec_power(P.env, lambda[c]) * P.env.exposure);
                                } else {
                                    DirLight L;
                                    direction_lighting(P.env, dk, observer, L);
                                    env = light_radiance(P.env, L, spec_power(P.env, lambda[c]));
                                }
                                split[c] = fma_(fmaxf(fmaxf(tr.x, 0.0f), 0.0f) * 1.0f, env, split[c]);
                            }
                        }
                    }
                    dirs[c] = dk;
                    live |= 1u << c;
                }
                if (!inside) {
                    // narrow_compat (refraction/context.rs:219-242)
#pragma unroll
                    for (int a = 0; a < NCH; ++a)
#pragma unroll
                        for (int b = a + 1; b < NCH; ++b) {
                            bool m;
                            if (a == 0) m = (match_mask >> b) & 1u;
                            else if (((live >> a) & 1u) && ((live >> b) & 1u))
                                m = dot(dirs[a], dirs[b]) >= DIRECTION_MATCH_COS_TOL;
                            else m = true;
                            if (!m) {
                                compat[a] &= ~(1u << b);
                                compat[b] &= ~(1u << a);
                            }
                        }
                }
                new_k = refr;
                new_inside = !inside;
            }
        }
        // dispatch_bounce tail (transport/bounce.rs:200-238)
        dir = new_k;
        org = hit + new_k * 1e-4f;
        k = new_k;
        inside = new_inside;
        if (bounce == 0 && !inside && P.surface_glare < 1.0f) {
#pragma unroll
            for (int c = 0; c < NCH; ++c) st[c] = scale4(st[c], P.surface_glare);
        }
        // apply_russian_roulette (transport/bounce.rs:355-375)
        if (bounce > 4) {
            float mx = 0.0f;
#pragma unroll
            for (int c = 0; c < NCH; ++c) mx = fmaxf(mx, fmaxf(st[c].x, 0.0f));
            float q = clampf(mx, 0.05f, 1.0f);
            float rr = unit_draw(seed, bounce, RUSSIAN_ROULETTE_STREAM);
            if (rr > q) {
                bounces = bounce + 1;
                term = 2u;
                break;
            }
            float inv_q = 1.0f / q;
#pragma unroll
            for (int c = 0; c < NCH; ++c) st[c] = scale4(st[c], inv_q);
#pragma unroll
            for (int c = 0; c < NCH; ++c) split[c] /= q;
        }
    }

    // PathState::finish (bounce_body.rs:724-764) + color.rs:145-172
    if (escaped) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) radiance[c] += split[c];
    }
    const float norm_factor = (400.0f / 8.0f) / 106.856f;
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        float family = -0.0f;
#pragma unroll
        for (int j = 0; j < NCH; ++j)
            if ((compat[c] >> j) & 1u) family += path_pdf[j];
        float w = family <= 1e-12f ? 1.0f : (8.0f * path_pdf[0]) / family;
        float cx, cy, cz;
        cie_cmf(lambda[c], cx, cy, cz);
        float s = (radiance[c] * w) * norm_factor;
        X = X + cx * s;
        Y = Y + cy * s;
        Z = Z + cz * s;
    }
    X = X + 0.0f;
    Y = Y + 0.0f;
    Z = Z + 0.0f;
    // apply_von_kries_white_balance (color.rs:220-254)
    float L_ = fma_(0.8951f, X, fma_(0.2664f, Y, -0.1614f * Z)) * P.env.wb[0];
    float M_ = fma_(-0.7502f, X, fma_(1.7135f, Y, 0.0367f * Z)) * P.env.wb[1];
    float S_ = fma_(0.0389f, X, fma_(-0.0685f, Y, 1.0296f * Z)) * P.env.wb[2];
    out.xyz[0] = fmaxf(fma_(0.986993f, L_, fma_(-0.147054f, M_, 0.159963f * S_)), 0.0f);
    out.xyz[1] = fmaxf(fma_(0.432305f, L_, fma_(0.518360f, M_, 0.049291f * S_)), 0.0f);
    out.xyz[2] = fmaxf(fma_(-0.008529f, L_, fma_(0.040043f, M_, 0.968487f * S_)), 0.0f);
    if (EVENTS) {
        out.ev.bounces = bounces;
        out.ev.termination = term;
        out.ev.facet_hash = fh;
        out.ev.n_tir = (uint16_t)n_tir;
        out.ev.n_reflect = (uint16_t)n_refl;
        out.ev.n_refract = (uint16_t)n_refr;
        out.ev.n_probe = (uint16_t)n_probe;
        if (counts) {
            counts[EV_BOUNCES] += bounces;
            counts[EV_TIR] += n_tir;
            counts[EV_REFLECT] += n_refl;
            counts[EV_REFRACT] += n_refr;
            counts[EV_PROBE] += n_probe;
            counts[EV_ESCAPED] += term == 0u ? 1u : 0u;
            counts[EV_ROTATIONS] += n_rot;
        }
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;

__device__ __forceinline__ void stage_planes(const FilmParams& P, const float4* __restrict__ g,
                                             float4* s) {
    for (uint32_t i = threadIdx.x; i < P.n_planes; i += blockDim.x) s[i] = g[i];
    __syncthreads();
}

// Film: one lane per path, raw per-sample XYZ to `out_xyz` (path order).
__global__ __launch_bounds__(kBlock) void film_transport_kernel(const FilmParams P,
                                                                const float4* __restrict__ planes,
                                                                float* __restrict__ out_xyz) {
    __shared__ float4 sp[IXG_MAX_PLANES];
    stage_planes(P, planes, sp);
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.n_paths) return;
    uint64_t p = P.first_path + i;
    uint32_t pixel = (uint32_t)(p / P.spp);
    uint32_t sample = P.sample_offset + (uint32_t)(p % P.spp);
    V3 org, dir;
    uint32_t seed;
    float hero;
    camera_sample(P, pixel, sample, org, dir, seed, hero);
    PathOut o;
    trace_path<false>(P, sp, org, dir, seed, hero, o, nullptr);
    out_xyz[3 * i + 0] = o.xyz[0];
    out_xyz[3 * i + 1] = o.xyz[1];
    out_xyz[3 * i + 2] = o.xyz[2];
}

// Diagnostics: per-path XYZ + integer side channel.
__global__ __launch_bounds__(kBlock) void film_events_kernel(const FilmParams P,
                                                             const float4* __restrict__ planes,
                                                             float* __restrict__ out_xyz,
                                                             ixg_path_events* __restrict__ ev,
                                                             unsigned long long* __restrict__ totals) {
    __shared__ float4 sp[IXG_MAX_PLANES];
    __shared__ unsigned long long block_tot[EV_NUM];
    if (threadIdx.x < EV_NUM) block_tot[threadIdx.x] = 0ull;
    stage_planes(P, planes, sp);
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t counts[EV_NUM] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < P.n_paths) {
        uint64_t p = P.first_path + i;
        uint32_t pixel = (uint32_t)(p / P.spp);
        uint32_t sample = P.sample_offset + (uint32_t)(p % P.spp);
        V3 org, dir;
        uint32_t seed;
        float hero;
        camera_sample(P, pixel, sample, org, dir, seed, hero);
        PathOut o;
        trace_path<true>(P, sp, org, dir, seed, hero, o, counts);
        counts[EV_PATHS] = 1;
        if (out_xyz) {
            out_xyz[3 * i + 0] = o.xyz[0];
            out_xyz[3 * i + 1] = o.xyz[1];
            out_xyz[3 * i + 2] = o.xyz[2];
        }
        if (ev) ev[i] = o.ev;
    }
    if (totals) {
#pragma unroll
        for (int s = 0; s < EV_NUM; ++s) {
            if (counts[s]) atomicAdd(&block_tot[s], (unsigned long long)counts[s]);
        }
        __syncthreads();
        if (threadIdx.x < EV_NUM && block_tot[threadIdx.x])
            atomicAdd(&totals[threadIdx.x], block_tot[threadIdx.x]);
    }
}

// Explicit rays (the reference's trace_spectral_ray entry): ray i starts at
// rays[i] = {ox, oy, oz, dx, dy, dz, hero_rand, seed bits}.
__global__ __launch_bounds__(kBlock) void film_rays_kernel(const FilmParams P,
                                                           const float4* __restrict__ planes,
                                                           const float* __restrict__ rays,
                                                           uint32_t n_rays,
                                                           float* __restrict__ out_xyz,
                                                           ixg_path_events* __restrict__ ev) {
    __shared__ float4 sp[IXG_MAX_PLANES];
    stage_planes(P, planes, sp);
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_rays) return;
    const float* r = rays + 8 * (size_t)i;
    PathOut o;
    trace_path<true>(P, sp, v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), __float_as_uint(r[7]), r[6],
                     o, nullptr);
    out_xyz[3 * i + 0] = o.xyz[0];
    out_xyz[3 * i + 1] = o.xyz[1];
    out_xyz[3 * i + 2] = o.xyz[2];
    if (ev) ev[i] = o.ev;
}

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

// ---------------------------------------------------------------------------
// launchers (internal to libixg)
// ---------------------------------------------------------------------------
extern "C" hipError_t ixg_launch_film(const ixg::FilmParams* P, const float4* d_planes, float* d_xyz,
                                      hipStream_t stream) {
    uint64_t blocks = (P->n_paths + ixg::kBlock - 1) / ixg::kBlock;
    hipLaunchKernelGGL(ixg::film_transport_kernel, dim3((uint32_t)blocks), dim3(ixg::kBlock), 0,
                       stream, *P, d_planes, d_xyz);
    return hipGetLastError();
}

extern "C" hipError_t ixg_launch_film_events(const ixg::FilmParams* P, const float4* d_planes,
                                             float* d_xyz, ixg_path_events* d_ev,
                                             unsigned long long* d_totals, hipStream_t stream) {
    uint64_t blocks = (P->n_paths + ixg::kBlock - 1) / ixg::kBlock;
    hipLaunchKernelGGL(ixg::film_events_kernel, dim3((uint32_t)blocks), dim3(ixg::kBlock), 0,
                       stream, *P, d_planes, d_xyz, d_ev, d_totals);
    return hipGetLastError();
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
    uint32_t blocks = (n_rays + ixg::kBlock - 1) / ixg::kBlock;
    hipLaunchKernelGGL(ixg::film_rays_kernel, dim3(blocks), dim3(ixg::kBlock), 0, stream, *P, d_planes,
                       d_rays, n_rays, d_xyz, d_ev);
    return hipGetLastError();
}
