// film_paths.h -- the film hot path on CDNA4 (gfx950): the path state, its
// init / bounce / finish steps and the film kernel templates.  The kernel
// instantiations are spread over several translation units (the film_*.hip
// files) so they compile in parallel; transport.hip holds the reduction and
// the launch dispatch.
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
#pragma once
#include <algorithm>
#include "device_common.h"
#include "uniaxial.h"
#include "biaxial.h"

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

__device__ __forceinline__ void fresnel_reflection_coeffs(float rs, float rp, float& a, float& b,
                                                          float& c) {
    float rs2 = rs * rs, rp2 = rp * rp;
    a = midpoint(rs2, rp2);
    b = 0.5f * (rs2 - rp2);
    c = rs * rp;
}

// cos and sin of the TIR retardation delta = 2 atan(a / b) (tir.rs:24-40),
// as the rational functions they equal: (b^2 - a^2) / (b^2 + a^2) and
// 2 a b / (b^2 + a^2), one reciprocal for both.  a / b is bounded (n sin_i > 1
// under TIR), and the values agree with cos / sin of the rounded angle to a
// few 1e-8 (absolute).
__device__ __forceinline__ void tir_retardation(float n1k, float cos_i, float sin_i, float& cd, float& sd) {
    const float a = cos_i * sqrt_rn(fmaxf(fma_(n1k * n1k * sin_i, sin_i, -1.0f), 0.0f));
    const float b = fmaxf(n1k * sin_i * sin_i, 1e-6f);
    const float a2 = a * a, b2 = b * b;
    const float inv = div_rn(1.0f, b2 + a2);
    cd = (b2 - a2) * inv;
    sd = ((2.0f * a) * b) * inv;
}

// cos 2 psi and sin 2 psi for psi = atan2(sin_psi, cos_psi) (the Stokes frame
// rotation, absorption.rs:69-87), from the double-angle identities on the
// unnormalised pair: (c^2 - s^2) / r^2 and 2 c s / r^2.  A zero pair is the
// zero angle, as atan2(0, 0) is.
__device__ __forceinline__ void rotation_double_angle(float cos_psi, float sin_psi, float& c2, float& s2) {
    const float r2 = fma_(cos_psi, cos_psi, sin_psi * sin_psi);
    if (r2 > 1e-30f) {
        const float inv = div_rn(1.0f, r2);
        c2 = fma_(cos_psi, cos_psi, -(sin_psi * sin_psi)) * inv;
        s2 = ((2.0f * cos_psi) * sin_psi) * inv;
    } else {
        c2 = 1.0f;
        s2 = 0.0f;
    }
}

__device__ __forceinline__ void fresnel_rs_rp(float n1, float n2, float cos_i, float cos_t,
                                              float& rs, float& rp) {
    rs = div_rn(fma_(n2, -cos_t, n1 * cos_i), fma_(n2, cos_t, n1 * cos_i));
    rp = div_rn(fma_(n1, -cos_t, n2 * cos_i), fma_(n1, cos_t, n2 * cos_i));
}

// The Stokes frame rotation by the doubled angle (MuellerMatrix::
// frame_rotation, polarization.rs:138, applied).
__device__ __forceinline__ float4 frame_rotate(float c2, float s2, float4 s) {
    return make_float4(s.x, c2 * s.y + s2 * s.z, (-s2) * s.y + c2 * s.z, s.w);
}

// One channel's partial reflection (apply_partial_reflect_bounce's per-channel
// step, reflect_refract.rs:33-71): MuellerMatrix::fresnel_reflection
// (polarization.rs:148-158) applied, and the clamped unpolarised reflectance
// that the channel's path pdf takes.
__device__ __forceinline__ float4 fresnel_reflect(float n1, float n2, float cos_i, float cos_t, float4 s,
                                                  float& r_unpol) {
    float rs, rp, a, b, c;
    fresnel_rs_rp(n1, n2, cos_i, cos_t, rs, rp);
    r_unpol = clampf(0.5f * fma_(rp, rp, rs * rs), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
    fresnel_reflection_coeffs(rs, rp, a, b, c);
    return mueller_fresnel(a, b, c, s);
}

// compute_channel_transmission (exit_split.rs:69-101) with
// MuellerMatrix::fresnel_transmission (polarization.rs:162-180) applied.  t_s
// and r_s share the denominator n1 cos_i + n2 cos_t, t_p and r_p the other:
// one reciprocal each, the four ratios within an ULP of the quotients; r_s and
// r_p come back for the channel's reflectance.
__device__ __forceinline__ float4 fresnel_transmit(float n1, float n2, float cos_i, float cos_t, float4 s,
                                                   float& rs, float& rp) {
    const float num = 2.0f * n1 * cos_i;
    const float inv_s = div_rn(1.0f, fma_(n2, cos_t, n1 * cos_i));
    const float inv_p = div_rn(1.0f, fma_(n1, cos_t, n2 * cos_i));
    const float t_s = num * inv_s;
    const float t_p = num * inv_p;
    const float factor = div_rn(n2 * cos_t, fmaxf(n1 * cos_i, 1e-6f));
    const float ts2 = t_s * t_s * factor, tp2 = t_p * t_p * factor;
    rs = fma_(n2, -cos_t, n1 * cos_i) * inv_s;
    rp = fma_(n1, -cos_t, n2 * cos_i) * inv_p;
    return mueller_fresnel(midpoint(ts2, tp2), 0.5f * (ts2 - tp2), t_s * t_p * factor, s);
}

__device__ __forceinline__ uint32_t fnv(uint32_t h, uint32_t v) { return (h ^ v) * 16777619u; }

// The frosted-facet bitmask rides after the facet set (IXG_MAX_PLANES float4,
// then 8 words) in global and LDS memory alike.
__device__ __forceinline__ bool facet_frosted(const float4* __restrict__ planes, int facet) {
    const uint32_t* mask = reinterpret_cast<const uint32_t*>(planes + IXG_MAX_PLANES);
    return (mask[facet >> 5] >> (facet & 31)) & 1u;
}

// cosine_weighted_hemisphere about unit n with frosted_orthonormal_basis
// (scattering/frosted.rs, scattering/mod.rs:110-115); the azimuth's sine and
// cosine correctly rounded.
__device__ __forceinline__ V3 cosine_hemisphere(float u1, float u2, V3 n) {
    const float r = sqrt_rn(u1);
    const float theta = (2.0f * 3.14159265358979323846f) * u2;
    const float sin_t = (float)sin((double)theta), cos_t = (float)cos((double)theta);
    const V3 a = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    const V3 t = normalize_or_zero(a - n * dot(n, a));
    const V3 b = cross(n, t);
    return normalize_or_zero((t * (r * cos_t) + b * (r * sin_t)) + n * sqrt_rn(fmaxf(1.0f - u1, 0.0f)));
}

// shading_normal_near_edge (RT/intersect.rs:196-233): within radius of the
// nearest other plane the facet normal eases (smoothstep) toward the bisector
// of the two normals.  Called only for radius > 0 (a launch-uniform branch),
// so sharp stones run none of it.
__device__ __noinline__ V3 shading_normal_near_edge(const float4* __restrict__ planes, uint32_t n, V3 hit,
                                                    int facet, V3 hn, float radius) {
    float nearest = __builtin_inff();
    V3 nn = hn;
    for (uint32_t i = 0; i < n; ++i) {
        if ((int)i == facet) continue;
        const float4 p = planes[i];
        const V3 pn = v3(p.x, p.y, p.z);
        const float dist = -(p.w + dot(pn, hit));
        if (dist < nearest) {
            nearest = dist;
            nn = pn;
        }
    }
    if (nearest >= radius) return hn;
    const float t = clampf(1.0f - div_rn(nearest, radius), 0.0f, 1.0f);
    const float st = t * t * fma_(2.0f, -t, 3.0f);
    const V3 bis = normalize_or_zero(hn + nn);
    return normalize_or_zero(hn * (1.0f - st) + bis * st);
}

// ---------------------------------------------------------------------------
// one path, as init / bounce / finish steps
// ---------------------------------------------------------------------------
struct PathOut {
    float xyz[3];
    ixg_path_events ev;
};

// sample_draws (sampling.rs:176-218): the path seed, the pixel jitter and the
// hero draw of one (pixel, absolute sample).
__device__ __forceinline__ void sample_draws(uint32_t pixel, uint32_t sample, uint32_t& seed, float& jx, float& jy,
                                             float& hero) {
    seed = hash_u32((pixel * 0x9e3779b9u) ^ (sample * 0x85ebca6bu));
    float rx = ld_base2(hash_u32(pixel ^ 0xA511E9B3u));
    float ry = ld_base2(hash_u32(pixel ^ 0x63D81B23u));
    float rh = ld_base2(hash_u32(pixel ^ 0x1B873593u));
    jx = cp_rotate(ld_base2(sample), rx) - 0.5f;
    jy = cp_rotate(radical_inverse<3>(sample), ry) - 0.5f;
    hero = cp_rotate(radical_inverse<5>(sample), rh);
}

// One (pixel, absolute sample): sample_draws and Camera::generate_ray
// (camera.rs:149-167).
__device__ __forceinline__ void camera_sample(const FilmParams& P, uint32_t pixel, uint32_t sample,
                                              V3& org, V3& dir, uint32_t& seed, float& hero) {
    float jx, jy;
    sample_draws(pixel, sample, seed, jx, jy, hero);
    uint32_t px = pixel % P.width, py = pixel / P.width;
    float aspect = div_rn(P.width_f, P.height_f);
    float u = (div_rn((float)px + jx, P.width_f) - 0.5f) * 2.0f * aspect * P.fov_tan;
    float v = (0.5f - div_rn((float)py + jy, P.height_f)) * 2.0f * P.fov_tan;
    dir = normalize((v3(P.forward[0], P.forward[1], P.forward[2]) +
                     v3(P.right[0], P.right[1], P.right[2]) * u) +
                    v3(P.up[0], P.up[1], P.up[2]) * v);
    org = v3(P.cam_origin[0], P.cam_origin[1], P.cam_origin[2]);
}

// Every per-path local of the reference's bounce loop (bounce_body.rs:147-205)
// that survives from one bounce to the next.  For an isotropic medium the wave
// normal k equals the ray direction, so one vector carries both.
struct PathState {
    float4 st[NCH];           // Stokes per channel
    float path_pdf[NCH];
    float split[NCH];         // exit-split staging radiance; an escaped path's final radiance
    float hero;               // the hero draw; channel c's wavelength is lambda_of(S, c)
    float n_o[NCH];
    float alpha_o[NCH], alpha_e[NCH];  // spectral_absorption, absorbing materials only
    float alpha_b[NCH];               // of the beta bands (trichroic biaxial stones)
    float nee[NCH];                   // light-sample deposits (EXTRA kernels: frosted exterior events)
    bool has_mis;                     // the light-sample MIS carry of an exterior frosted bounce ...
    float mis_phase;                  // ... and its BSDF density, for the next bounce's miss
    uint64_t compat;          // byte c = MIS family of channel c
    V3 org, dir, observer, prev_plane;
    V3 k;                     // wave normal (uniaxial materials; dir is the Poynting direction)
    bool is_extra;            // eigenmode assigned at the latest entry (uniaxial)
    uint32_t seed;
    uint32_t bounce;
    int32_t last_facet;       // facet of the previous event (-1 before the first)
    bool inside, has_prev;
    // integer side channel
    uint32_t fh, n_tir, n_refl, n_refr, n_probe, n_rot;
    uint32_t n_scan, n_probe_scan;  // slab scans actually run (bounce, exit probes)
};

__device__ __forceinline__ uint32_t compat_of(uint64_t c, int ch) { return (uint32_t)(c >> (8 * ch)) & 0xFFu; }

// wrapped_hero_wavelengths (transport/mod.rs:58-66): the 8-lambda comb at
// 50 nm from the hero draw, wrapped into [380, 780).
__device__ __forceinline__ void hero_comb(float hero, float lambda[NCH]) {
    const float lambda_hero = fma_(hero, 400.0f, 380.0f);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        float off = fma_((float)c, 50.0f, lambda_hero - 380.0f);
        // rem_euclid(400): for 0 <= off < 800 this is off or off - 400, both
        // exact (Sterbenz), which is what fmodf returns; a hero draw outside
        // [0, 1) (explicit rays) takes the exact remainder loop
        float r = off >= 0.0f && off < 800.0f ? (off >= 400.0f ? off - 400.0f : off) : fmod_exact(off, 400.0f);
        r = r < 0.0f ? r + 400.0f : r;
        lambda[c] = 380.0f + r;
    }
}

// Channel c's wavelength of the comb (hero_comb's element c).
__device__ __forceinline__ float comb_lambda(float hero, int c) {
    const float lambda_hero = fma_(hero, 400.0f, 380.0f);
    float off = fma_((float)c, 50.0f, lambda_hero - 380.0f);
    float r = off >= 0.0f && off < 800.0f ? (off >= 400.0f ? off - 400.0f : off) : fmod_exact(off, 400.0f);
    r = r < 0.0f ? r + 400.0f : r;
    return 380.0f + r;
}

// The per-channel optics of a new path: n(lambda) and, for absorbing
// materials, spectral_absorption (absorption.rs:33), the bands' sum in order.
template <bool ABSORB>
__device__ __forceinline__ void path_optics(const FilmParams& P, const float lambda[NCH], float n_o[NCH],
                                            float alpha_o[NCH], float alpha_e[NCH]) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        n_o[c] = dispersion(P.disp_kind, P.disp, lambda[c]);
        if (ABSORB) {
            float so = -0.0f, se = -0.0f;
            for (int32_t b = 0; b < P.abs.n_o; ++b) so += band_eval(P.abs.o[b], lambda[c]);
            for (int32_t b = 0; b < P.abs.n_e; ++b) se += band_eval(P.abs.e[b], lambda[c]);
            alpha_o[c] = so;
            alpha_e[c] = se;
        }
    }
}

// PathState::new (bounce_body.rs:258-356) with the comb and the optics
// already formed.
__device__ __forceinline__ void path_start(V3 org, V3 dir, uint32_t seed, PathState& S) {
    S.org = org;
    S.dir = dir;
    S.k = dir;
    S.is_extra = false;
    S.observer = neg(dir);
    S.prev_plane = v3(0.0f, 0.0f, 0.0f);
    S.seed = seed;
    S.bounce = 0;
    S.last_facet = -1;
    S.inside = false;
    S.has_prev = false;
    S.compat = ~0ull;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        S.path_pdf[c] = 1.0f;
        S.split[c] = 0.0f;
        S.st[c] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
    }
    S.has_mis = false;
    S.mis_phase = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; ++c) S.nee[c] = 0.0f;
    S.fh = 2166136261u;
    S.n_tir = S.n_refl = S.n_refr = S.n_probe = S.n_rot = 0;
    S.n_scan = S.n_probe_scan = 0;
}

// PathState::new (bounce_body.rs:258-356).
template <bool ABSORB = false, int ANISO = 0>
__device__ __forceinline__ void path_init(const FilmParams& P, V3 org, V3 dir, uint32_t seed,
                                          float hero, PathState& S) {
    path_start(org, dir, seed, S);
    S.hero = hero;
    float lambda[NCH];
    hero_comb(hero, lambda);
    path_optics<ABSORB>(P, lambda, S.n_o, S.alpha_o, S.alpha_e);
    if (ABSORB && ANISO == 2) {  // the trichroic beta bands (wavelength_cache.rs:38-51)
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            float sb = -0.0f;
            for (int32_t b = 0; b < P.abs.n_b; ++b) sb += band_eval(P.abs.b[b], lambda[c]);
            S.alpha_b[c] = sb;
        }
    }
}

// The environment radiance seen along dir.  The HDR panorama is compiled only
// into the EXTRA instantiations (its angles are correctly rounded through
// double precision), which the launch chooses for a panorama.
template <bool EXTRA>
__device__ __forceinline__ float env_radiance(const EnvParams& E, V3 dir, V3 observer, float lambda) {
    if (E.model == ENV_ASET) return aset_radiance(E, normalize(dir), lambda, E.exposure);
    if (EXTRA && E.model == ENV_HDR) {
        float rgb[3];
        hdr_rgb(E, dir, rgb);
        return rgb_to_spectral(rgb, lambda);
    }
    if (E.model != ENV_STUDIO)
        return lit_factor(E, dir, observer) * (spec_power(E, lambda) * E.exposure);
    DirLight L;
    direction_lighting(E, dir, observer, L);
    return light_radiance(E, L, spec_power(E, lambda));
}

// The analytic sun (environment/sun.rs): cone sampling of its disc, its pdf,
// and next-event estimation at an exterior frosted event
// (nee_contribution_frosted_exterior, scattering/frosted.rs:98-182) with
// sample_environment_for_nee's DaylightSun light sample (environment/nee.rs:81-120).
#define FROSTED_NEE_ENV_DIR_U_STREAM 0x6A09E667u
#define FROSTED_NEE_ENV_DIR_V_STREAM 0xBB67AE85u
#define SUN_DISC_COS 0.9999889f
#define SUN_ONE_MINUS_COS 1.1086464e-5f
#define SUN_SOLID_ANGLE 6.965831e-5f

__device__ __forceinline__ V3 sun_cone_direction(V3 key, float u0, float u1) {
    const float omc = u0 * SUN_ONE_MINUS_COS;
    const float cos_t = 1.0f - omc;
    const float sin_t = sqrt_rn(fmaxf(omc * (2.0f - omc), 0.0f));
    const float phi = (2.0f * 3.14159265358979323846f) * u1;
    const float sp = (float)sin((double)phi), cp = (float)cos((double)phi);
    const V3 a = fabsf(key.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    const V3 t = normalize_or_zero(a - key * dot(key, a));
    const V3 b = cross(key, t);
    return normalize_or_zero((t * (sin_t * cp) + b * (sin_t * sp)) + key * cos_t);
}

// balance_heuristic (scattering/mod.rs:102-105)
__device__ __forceinline__ float balance_heuristic(float a, float b) {
    const float denom = a + b;
    return denom > 1e-12f ? div_rn(a, denom) : 0.0f;
}

// environment_nee_pdf (environment/nee.rs:144-157): the sun disc's 1 / solid
// angle under DaylightSun, 0 for every other studio model
__device__ __forceinline__ float environment_nee_pdf(const EnvParams& E, V3 dir) {
    if (E.model != ENV_DAYLIGHT_SUN) return 0.0f;
    return dot(dir, v3(E.key[0], E.key[1], E.key[2])) >= SUN_DISC_COS ? div_rn(1.0f, SUN_SOLID_ANGLE) : 0.0f;
}

// apply_frosted_bounce's exterior side (frosted.rs:331-341, 372-382): the sun
// light sample deposited into S.nee, and the MIS carry for the next miss.
__device__ __noinline__ void frosted_exterior_light(const FilmParams& P, PathState& S, V3 ext_normal, V3 new_dir,
                                                    uint32_t bounce) {
    S.has_mis = true;
    S.mis_phase = div_rn(dot(new_dir, ext_normal), 3.14159265358979323846f);
    if (P.env.model != ENV_DAYLIGHT_SUN) return;  // no light sample
    const float u0 = unit_draw(S.seed, bounce, FROSTED_NEE_ENV_DIR_U_STREAM);
    const float u1 = unit_draw(S.seed, bounce, FROSTED_NEE_ENV_DIR_V_STREAM);
    const V3 dir = sun_cone_direction(v3(P.env.key[0], P.env.key[1], P.env.key[2]), u0, u1);
    const float pdf = div_rn(1.0f, SUN_SOLID_ANGLE);
    const float cos_light = dot(dir, ext_normal);
    if (cos_light <= 0.0f) return;
    const float brdf_pdf = div_rn(cos_light, 3.14159265358979323846f);
    const float mis = balance_heuristic(pdf, brdf_pdf);
    if (mis <= 0.0f) return;
    const float common = div_rn(brdf_pdf * mis, pdf);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const float env = P.env.sun * (spec_power(P.env, comb_lambda(S.hero, c)) * P.env.exposure);
        S.nee[c] = fma_(fmaxf(S.st[c].x, 0.0f) * common * env, 1.0f, S.nee[c]);
    }
}

// A ray leaving facet f of a convex polyhedron outward cannot meet it again:
// if its origin is already strictly outside f's half-space and it moves away
// from f, the exit plane gives t_far < 0 and the slab scan returns no hit --
// exactly what scanning every plane would have returned.
__device__ __forceinline__ bool leaves_convex(const float4* __restrict__ planes, int f, V3 o, V3 d) {
    float4 p = planes[f];
    float denom = (p.x * d.x + p.y * d.y) + p.z * d.z;
    float side = p.w + ((p.x * o.x + p.y * o.y) + p.z * o.z);
    return side > 0.0f && denom > 1e-7f;
}

enum StepResult { STEP_CONTINUE = 0, STEP_ESCAPED = 1, STEP_ROULETTE = 2, STEP_SCATTER_ABSORBED = 3 };

// AbsorptionTensor3::quadratic_form (birefringence/absorption_tensor.rs):
// alpha . (axes^T e)^2, axes = (a1, a2, c), alpha = (ao, ao, ae).
__device__ __forceinline__ float tensor_qf(const AbsorbParams& A, float ao, float ae, V3 e) {
    const float l0 = dot(v3(A.a1[0], A.a1[1], A.a1[2]), e);
    const float l1 = dot(v3(A.a2[0], A.a2[1], A.a2[2]), e);
    const float l2 = dot(v3(A.cn[0], A.cn[1], A.cn[2]), e);
    return (ao * (l0 * l0) + ao * (l1 * l1)) + ae * (l2 * l2);
}

// apply_absorption (absorption.rs:99-122) through
// channel_absorption_alphas_assigned's isotropic-by-symmetry branch
// (absorption.rs:205-222): the mean of the two eigen-directions' quadratic
// forms (f32::midpoint, computed in f64 like the reference), Beer-Lambert
// through exp_f32x8's lane polynomial over the scaled segment length.
// channel_absorption_alphas_assigned, isotropic arm (absorption.rs:197-222):
// per channel the mean of the tensor's quadratic form over the two
// eigen-polarisations.
__device__ __forceinline__ void channel_alphas(const AbsorbParams& A, const PathState& S, V3 k, float alpha[NCH]) {
    const V3 c = v3(A.c_raw[0], A.c_raw[1], A.c_raw[2]);
    V3 ea = cross(k, c);
    if (dot(ea, ea) > 1e-8f) {
        ea = normalize(ea);
    } else {
        const V3 n = normalize_or_zero(k);
        const V3 a = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
        ea = normalize_or_zero(a - n * dot(n, a));
    }
    const V3 eb = normalize_or_zero(cross(k, ea));
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        const float qa = tensor_qf(A, S.alpha_o[ch], S.alpha_e[ch], ea);
        const float qb = tensor_qf(A, S.alpha_o[ch], S.alpha_e[ch], eb);
        alpha[ch] = midpoint(qa, qb);
    }
}

__device__ __forceinline__ void apply_absorption(const AbsorbParams& A, PathState& S, V3 k, float path_len) {
    const V3 c = v3(A.c_raw[0], A.c_raw[1], A.c_raw[2]);
    V3 ea = cross(k, c);
    if (dot(ea, ea) > 1e-8f) {
        ea = normalize(ea);
    } else {
        // stable_orthonormal_basis(normalize_or_zero(k)).0 (birefringence/helpers.rs)
        const V3 n = normalize_or_zero(k);
        const V3 a = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
        ea = normalize_or_zero(a - n * dot(n, a));
    }
    const V3 eb = normalize_or_zero(cross(k, ea));
    const float scaled = path_len * A.path_scale;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        const float qa = tensor_qf(A, S.alpha_o[ch], S.alpha_e[ch], ea);
        const float qb = tensor_qf(A, S.alpha_o[ch], S.alpha_e[ch], eb);
        const float alpha = midpoint(qa, qb);
        S.st[ch] = scale4(S.st[ch], exp_lane(-alpha * scaled));
    }
}

// ---------------------------------------------------------------------------
// uniaxial materials (refraction/{geometry,dispatch,tir,uniaxial_entry,
// uniaxial_internal,reflect_refract}.rs, transport/bounce.rs)
// ---------------------------------------------------------------------------
#define BIREFRINGENT_SPLIT_STREAM 0x2545F491u
#define MODE_COUPLING_STREAM 0xCC9E2D51u

// GemMaterial::extraordinary_index_at (materials/optics.rs:42-47)
__device__ __forceinline__ float e_index_at(const FilmParams& P, float lambda, float n_o) {
    return P.an.has_e ? dispersion(P.an.e_kind, P.an.e_disp, lambda) : n_o + P.an.biref;
}

__device__ __forceinline__ V3 refract_dir(V3 k, V3 normal, float eta, float cos_i, float cos_t) {
    return normalize(k * eta + normal * fma_(eta, cos_i, -cos_t));
}

// theta_c_for_bounce (refraction/geometry.rs:75-120): the wave's angle to the
// optic axis; at entry two fixed-point steps on the refracted direction
// (seeded with the hero's ordinary index), inside the wave normal's own.
__device__ __forceinline__ float theta_c_for_bounce(const FilmParams& P, float lambda_hero, float n_o_seed, V3 k,
                                                    V3 normal, V3 cax, float cos_i, bool inside) {
    if (inside) return acosf(fabsf(clampf(dot(k, cax), -1.0f, 1.0f)));
    const float one_m_c2 = fma_(cos_i, -cos_i, 1.0f);
    const float n_e_seed = e_index_at(P, lambda_hero, n_o_seed);
    float n_guess = n_o_seed;
    float theta_c = 0.0f;
    for (int it = 0; it < 2; ++it) {
        const float eg = 1.0f / n_guess;
        const float s2g = eg * eg * one_m_c2;
        if (s2g > 1.0f) break;
        const float ctg = sqrt_rn(fmaxf(1.0f - s2g, 0.0f));
        const V3 wd = refract_dir(k, normal, eg, cos_i, ctg);
        theta_c = acosf(fabsf(clampf(dot(wd, cax), -1.0f, 1.0f)));
        n_guess = effective_extraordinary_index(n_o_seed, n_e_seed, theta_c);
    }
    return theta_c;
}

// compute_uniaxial_exit_transmission (refraction/uniaxial_internal.rs): the
// transmitted Stokes vector of a unit-flux-normalised internal solve for an
// incident intensity inc_i, scaled by inv_t = 1 / (1 - r_branch); i_unit is
// the unit-intensity transmittance.
__device__ __forceinline__ float4 exit_transmission(const InternalSol& sol, float inc_i, float inv_t, float& i_unit) {
    const float flux_inc = fmaxf(sol.flux_inc, 1e-12f);
    const Cx ts = cscale(sol.t_s, sqrt_rn(sol.flux_ts / flux_inc));
    const Cx tp = cscale(sol.t_p, sqrt_rn(sol.flux_tp / flux_inc));
    i_unit = cnorm2(ts) + cnorm2(tp);
    const float qu = cnorm2(ts) - cnorm2(tp);
    const Cx cr = cmul(ts, cconj(tp));
    return scale4(make_float4(inc_i * i_unit, inc_i * qu, inc_i * (2.0f * cr.re), inc_i * (-2.0f * cr.im)), inv_t);
}

// entry_eigenmode_selection (refraction/uniaxial_entry.rs:50-75)
__device__ __forceinline__ bool eigenmode_selection(V3 c, V3 cpn, V3 k, float4 h, float& p_o, float& c2,
                                                    float& s2) {
    if (h.x <= 1e-7f || dot(cpn, cpn) <= 1e-6f) return false;
    if (fma_(h.y, h.y, h.z * h.z) <= 1e-12f) return false;
    const V3 o_hat = ordinary_eigen(k, c);
    const V3 p_hat = cross(k, cpn);
    const float sc = dot(o_hat, cpn), pc = dot(o_hat, p_hat);
    c2 = fma_(sc, sc, -(pc * pc));
    s2 = 2.0f * sc * pc;
    p_o = clampf(0.5f + 0.5f * fma_(h.y, c2, h.z * s2) / h.x, 0.0f, 1.0f);
    return true;
}

// try_split_exit_channel (exit_split.rs:128-163) for channel c leaving along dk.
template <bool EVENTS, bool ISO, bool EXTRA>
__device__ __forceinline__ void split_probe(const FilmParams& P, const float4* __restrict__ planes, PathState& S,
                                            float* split, int facet, V3 hit, int c, V3 dk, float transmitted_i) {
    if (EVENTS) ++S.n_probe;
    const V3 po = hit + dk * 1e-4f;
    bool escapes = leaves_convex(planes, facet, po, dk);
    if (!escapes) {
        float tt;
        escapes = intersect(planes, P.n_planes, po, dk, tt) < 0;
        if (EVENTS) ++S.n_probe_scan;
    }
    if (!escapes) return;
    const float env = ISO ? lit_factor(P.env, dk, S.observer) * (spec_power(P.env, comb_lambda(S.hero, c)) * P.env.exposure)
                          : env_radiance<EXTRA>(P.env, dk, S.observer, comb_lambda(S.hero, c));
    split[c] = fma_(fmaxf(transmitted_i, 0.0f) * 1.0f, env, split[c]);
}

__device__ __forceinline__ V3 reflect_dir(V3 k, V3 normal) { return k - normal * (2.0f * dot(k, normal)); }

// One facet event of a uniaxial material, the same steps as the isotropic
// arm with compute_bounce_refraction_geometry's anisotropic indices, the
// closed-form entry / internal arms where the wave is not along the optic
// axis, the eigenmode-drawing scalar arms where it is, assigned-mode
// absorption and the internal o<->e re-coupling draw.
// The refracted directions of one entry event's channels, kept in LDS (a
// column per lane of the 256-lane film blocks) from the channel loop to
// narrow_compat, so 24 registers are not held across the loop (whose exit
// probes run slab scans of their own).
constexpr int kFilmLanes = 256;
struct ChannelDirs {
    float v[3][NCH][kFilmLanes];
};
__device__ __forceinline__ ChannelDirs& channel_dirs() {
    __shared__ ChannelDirs d;
    return d;
}
__device__ __forceinline__ void channel_dir_put(int c, V3 dir) {
    ChannelDirs& d = channel_dirs();
    d.v[0][c][threadIdx.x] = dir.x;
    d.v[1][c][threadIdx.x] = dir.y;
    d.v[2][c][threadIdx.x] = dir.z;
}
// narrow_compat (refraction/context.rs:219-242): channels whose refracted
// directions disagree leave each other's MIS family; channel 0 against the
// others by the hero-direction match, the rest pairwise from the stored
// directions of the live channels.
__device__ __forceinline__ void narrow_compat(uint64_t& compat, uint32_t match_mask, uint32_t live) {
    const ChannelDirs& d = channel_dirs();
    V3 dirs[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) dirs[c] = v3(d.v[0][c][threadIdx.x], d.v[1][c][threadIdx.x], d.v[2][c][threadIdx.x]);
#pragma unroll
    for (int a = 0; a < NCH; ++a)
#pragma unroll
        for (int b = a + 1; b < NCH; ++b) {
            bool m;
            if (a == 0) m = (match_mask >> b) & 1u;
            else if (((live >> a) & 1u) && ((live >> b) & 1u))
                m = dot(dirs[a], dirs[b]) >= DIRECTION_MATCH_COS_TOL;
            else m = true;
            if (!m) compat &= ~((1ull << (8 * a + b)) | (1ull << (8 * b + a)));
        }
}

// apply_frosted_bounce (scattering/frosted.rs:265-390), next-event estimation
// off, for an anisotropic stone: the bounce geometry's hero indices decide TIR
// and the Fresnel draw; a cosine-weighted hemisphere on the reflect or far
// side with every channel depolarised; a transmission into the stone draws its
// eigenmode 50/50 (upd).
template <bool EVENTS>
__device__ __forceinline__ V3 frosted_arm_aniso(const FilmParams& P, PathState& S, uint32_t bounce, int facet,
                                                V3 normal, float cos_i, float n1, float n2, float sin2_t,
                                                bool& new_inside, int& upd) {
    const float u1 = unit_draw(S.seed, bounce, FROSTED_DIR_U_STREAM);
    const float u2 = unit_draw(S.seed, bounce, FROSTED_DIR_V_STREAM);
    uint32_t kind = 4u;
    V3 dir;
    float pdf = 1.0f;
    if (sin2_t > 1.0f) {
        dir = cosine_hemisphere(u1, u2, normal);
    } else {
        float rs, rp;
        fresnel_rs_rp(n1, n2, cos_i, sqrt_rn(fmaxf(1.0f - sin2_t, 0.0f)), rs, rp);
        const float r_unpol = clampf(0.5f * fma_(rp, rp, rs * rs), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
        const bool through = !(unit_draw(S.seed, bounce, FRESNEL_BRANCH_STREAM) < r_unpol);
        dir = cosine_hemisphere(u1, u2, through ? neg(normal) : normal);
        pdf = through ? 1.0f - r_unpol : r_unpol;
        kind = through ? 6u : 5u;
        if (through) {
            if (!S.inside) upd = unit_draw(S.seed, bounce, BIREFRINGENT_SPLIT_STREAM) < 0.5f ? 1 : 0;
            new_inside = !S.inside;
        }
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        S.st[c] = make_float4(fmaxf(S.st[c].x, 0.0f), 0.0f, 0.0f, 0.0f);
        if (kind != 4u) S.path_pdf[c] *= pdf;
    }
    // the exterior side: reflected while outside, or transmitted out
    if (kind == 5u && !S.inside) frosted_exterior_light(P, S, normal, dir, bounce);
    if (kind == 6u && S.inside) frosted_exterior_light(P, S, neg(normal), dir, bounce);
    if (EVENTS) S.fh = fnv(fnv(S.fh, (uint32_t)facet), kind);
    return dir;
}

// assigned_mode_e_field_uniaxial (birefringence/absorption_tensor.rs): the E
// field of the path's assigned eigenmode at wave normal k.
__device__ __forceinline__ V3 uniaxial_assigned_e(const FilmParams& P, const PathState& S, V3 k) {
    const V3 cax = v3(P.an.c_axis[0], P.an.c_axis[1], P.an.c_axis[2]);
    if (!S.is_extra) return ordinary_eigen(k, cax);
    const float n_o_h = S.n_o[0];
    const float n_e_h = e_index_at(P, comb_lambda(S.hero, 0), n_o_h);
    const V3 kxc = cross(k, cax);
    if (dot(kxc, kxc) < 1e-8f) return normalize_or_zero(cross(k, ordinary_eigen(k, cax)));
    return normalize(cross(kxc, extraordinary_poynting_dir(k, cax, n_o_h, n_e_h)));
}

template <bool EVENTS, bool ISO, bool ABSORB, bool EXTRA = true>
__device__ __forceinline__ int facet_aniso(const FilmParams& P, const float4* __restrict__ planes, PathState& S,
                                        float t_hit, int facet) {
    const uint32_t bounce = S.bounce;
    // The per-channel state in local arrays for the event: the channel loops
    // around the boundary solves stay rolled and index it at run time, which
    // would otherwise pin the whole path state to private memory.
    float4 st_l[NCH];
    float pdf_l[NCH], n_o_l[NCH], split_l[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        st_l[c] = S.st[c];
        pdf_l[c] = S.path_pdf[c];
        n_o_l[c] = S.n_o[c];
        split_l[c] = S.split[c];
    }
    const V3 cax = v3(P.an.c_axis[0], P.an.c_axis[1], P.an.c_axis[2]);
    const V3 hit = S.org + S.dir * t_hit;
    const float4 pf = planes[facet];
    V3 normal = v3(pf.x, pf.y, pf.z);
    if (EXTRA && P.edge_radius > 0.0f)
        normal = shading_normal_near_edge(planes, P.n_planes, hit, facet, normal, P.edge_radius);
    const V3 cur_plane = normalize_or_zero(cross(S.k, normal));
    if (S.has_prev && dot(cur_plane, cur_plane) > 1e-6f && dot(S.prev_plane, S.prev_plane) > 1e-6f) {
        const float cos_psi = clampf(dot(S.prev_plane, cur_plane), -1.0f, 1.0f);
        const float sin_psi = dot(cross(S.prev_plane, cur_plane), normalize_or_zero(S.k));
        float c2, s2;
        rotation_double_angle(cos_psi, sin_psi, c2, s2);
#pragma unroll
        for (int c = 0; c < NCH; ++c) st_l[c] = frame_rotate(c2, s2, st_l[c]);
        if (EVENTS) ++S.n_rot;
    }
    S.prev_plane = cur_plane;
    S.has_prev = true;
    const V3 k = S.k;
    const bool inside = S.inside, is_extra = S.is_extra;
    if (ABSORB && inside && !(EXTRA && P.sigma_s > 0.0f)) {
        // apply_absorption with the assigned-mode coefficient (absorption.rs:224-249);
        // a scattering stone's segment was attenuated by the scatter step
        const V3 e = uniaxial_assigned_e(P, S, k);
        const float scaled = t_hit * P.abs.path_scale;
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            st_l[c] = scale4(st_l[c], exp_lane(-tensor_qf(P.abs, S.alpha_o[c], S.alpha_e[c], e) * scaled));
    }
    if (inside) normal = neg(normal);

    // compute_bounce_refraction_geometry (geometry.rs:306-425), uniaxial
    const float cos_i = clampf(dot(neg(k), normal), 0.0f, 1.0f);
    const float one_m_c2 = fma_(cos_i, -cos_i, 1.0f);
    const float sin_i = sqrt_rn(fmaxf(one_m_c2, 0.0f));
    const float theta_c = theta_c_for_bounce(P, comb_lambda(S.hero, 0), n_o_l[0], k, normal, cax, cos_i, inside);
    float n_e_ch[NCH], n_med[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        n_e_ch[c] = e_index_at(P, comb_lambda(S.hero, c), n_o_l[c]);
        const float n_eff = effective_extraordinary_index(n_o_l[c], n_e_ch[c], theta_c);
        n_med[c] = (inside && !is_extra) ? n_o_l[c] : n_eff;
    }
    const float n_o_hero = n_o_l[0];
    const float n_e_hero = n_o_hero + P.an.biref;
    const float n1 = inside ? n_med[0] : 1.0f;
    const float n2 = inside ? 1.0f : n_med[0];
    const float eta = div_rn(n1, n2);
    const float sin2_t = eta * eta * one_m_c2;
    const UFrame F = uframe_build(k, normal, cax, cos_i, sin_i);
    const bool nondegenerate = dot(cross(k, cax), cross(k, cax)) > 1e-6f;

    V3 new_k, new_s;
    bool new_inside = inside;
    int upd = -1;
    float exact_p_o = -1.0f;
    uint32_t match_bits = 0u;  // channels whose refracted direction matched the hero's
    const bool frosted = EXTRA && P.frosted && facet_frosted(planes, facet);
    if (frosted) {
        // dispatch_bounce's frosted arm (transport/bounce.rs:137-160)
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            S.st[c] = st_l[c];
            S.path_pdf[c] = pdf_l[c];
        }
        new_k = frosted_arm_aniso<EVENTS>(P, S, bounce, facet, normal, cos_i, n1, n2, sin2_t, new_inside, upd);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            st_l[c] = S.st[c];
            pdf_l[c] = S.path_pdf[c];
        }
        new_s = new_k;
    } else if (sin2_t > 1.0f) {
        // apply_tir_bounce (tir.rs:56-160)
        if (EVENTS) {
            ++S.n_tir;
            S.fh = fnv(fnv(S.fh, (uint32_t)facet), 1u);
        }
        if (nondegenerate) {
            // the solves over small local arrays (a loop the compiler may keep
            // rolled), then the path state updated with constant channel indices,
            // so the state need not live in indexable private memory
            float r_tot[NCH];
            for (int c = 0; c < NCH; ++c) {
                const InternalSol sol = internal_solve(n_med[c], n_o_l[c], n_e_ch[c], cax, F, !is_extra);
                const float flux_inc = fmaxf(sol.flux_inc, 1e-12f);
                const float ro = cnorm2(sol.r_o) * sol.flux_ro, re = cnorm2(sol.r_e) * sol.flux_re;
                r_tot[c] = fminf((ro + re) / flux_inc, 1.0f);
                if (c == 0) exact_p_o = ro / fmaxf(ro + re, 1e-12f);
            }
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                st_l[c] = scale4(st_l[c], r_tot[c]);
                const float ek = n_med[c] / 1.0f;
                if (ek * ek * one_m_c2 <= 1.0f) pdf_l[c] *= clampf(r_tot[c], R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            }
        } else {
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const float n1k = n_med[c], n2k = 1.0f;
                const float ek = div_rn(n1k, n2k);
                const float s2k = ek * ek * one_m_c2;
                if (s2k > 1.0f) {
                    float cd, sd;
                    tir_retardation(n1k, cos_i, sin_i, cd, sd);
                    st_l[c] = mueller_tir(cd, sd, st_l[c]);
                } else {
                    float ru;
                    st_l[c] = fresnel_reflect(n1k, n2k, cos_i, sqrt_rn(fmaxf(1.0f - s2k, 0.0f)), st_l[c], ru);
                    pdf_l[c] *= ru;
                }
            }
        }
        new_k = reflect_dir(k, normal);
        new_s = (inside && is_extra) ? extraordinary_poynting_dir(new_k, cax, n_o_hero, n_e_hero) : new_k;
    } else if (!inside && nondegenerate) {
        // apply_uniaxial_entry_bounce (uniaxial_entry.rs:108-198)
        EntryInc I;
        entry_incidence(F, I);
        EntrySol hsol;
        entry_solve(I, n_o_l[0], n_e_ch[0], cax, F, hsol, false);
        M4 Mh;
        jones_to_mueller(hsol.x[0][0], hsol.x[1][0], hsol.x[0][1], hsol.x[1][1], Mh);
        const float r_branch = clampf(Mh.m[0][0], R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        const float inc_flux = 0.5f * F.cos_i;
        const float p_o_raw = mode_power(hsol.x[0][2], hsol.x[1][2], hsol.flux_o, inc_flux, st_l[0]);
        const float p_e_raw = mode_power(hsol.x[0][3], hsol.x[1][3], hsol.flux_e, inc_flux, st_l[0]);
        const float p_o_hero =
            clampf(p_o_raw / fmaxf(p_o_raw + p_e_raw, 1e-12f), R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        const bool use_extra = unit_draw(S.seed, bounce, BIREFRINGENT_SPLIT_STREAM) < (1.0f - p_o_hero);
        const float p_mode_frac = use_extra ? 1.0f - p_o_hero : p_o_hero;
        const float n2_dir = use_extra ? n2 : n_o_hero;
        const float eta_dir = div_rn(n1, n2_dir);
        const float s2d = fminf(eta_dir * eta_dir * one_m_c2, 1.0f);
        const V3 refr = refract_dir(k, normal, eta_dir, cos_i, sqrt_rn(fmaxf(1.0f - s2d, 0.0f)));
        const V3 final_refr = use_extra ? extraordinary_poynting_dir(refr, cax, n_o_hero, n_e_hero) : refr;
        const V3 p_axis_t = cross(refr, F.s_axis);
        if (unit_draw(S.seed, bounce, FRESNEL_BRANCH_STREAM) < r_branch) {
            if (EVENTS) {
                ++S.n_refl;
                S.fh = fnv(fnv(S.fh, (uint32_t)facet), 2u);
            }
            const float inv_r = div_rn(1.0f, r_branch);
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                EntrySol sol;
                entry_solve(I, n_o_l[c], n_e_ch[c], cax, F, sol, false);
                M4 Mk;
                jones_to_mueller(sol.x[0][0], sol.x[1][0], sol.x[0][1], sol.x[1][1], Mk);
                const float ru = clampf(Mk.m[0][0], R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                st_l[c] = scale4(mat_apply(Mk, st_l[c]), inv_r);
                pdf_l[c] *= ru;
            }
            new_k = reflect_dir(k, normal);
            new_s = new_k;
        } else {
            if (EVENTS) {
                ++S.n_refr;
                S.fh = fnv(fnv(S.fh, (uint32_t)facet), 3u);
            }
            const float inv_t = div_rn(1.0f, 1.0f - r_branch);
            uint32_t live = 0u, match_mask = 0u;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const float n_o_k = n_o_l[c];
                const float n2k = use_extra ? n_med[c] : n_o_k;
                const float r = div_rn(n1, n2k);
                const float s2k = r * r * one_m_c2;
                if (s2k > 1.0f) {
                    st_l[c] = scale4(st_l[c], 0.0f);
                    pdf_l[c] = 0.0f;
                    continue;
                }
                const float ek = div_rn(n1, n2k);
                const V3 rk = refract_dir(k, normal, ek, cos_i, sqrt_rn(fmaxf(1.0f - s2k, 0.0f)));
                const V3 fk = use_extra ? extraordinary_poynting_dir(rk, cax, n_o_k, n_o_k + P.an.biref) : rk;
                channel_dir_put(c, fk);
                live |= 1u << c;
                const bool match = dot(fk, final_refr) >= DIRECTION_MATCH_COS_TOL;
                EntrySol sol;
                entry_solve(I, n_o_k, n_e_ch[c], cax, F, sol, match);
                M4 Mk;
                jones_to_mueller(sol.x[0][0], sol.x[1][0], sol.x[0][1], sol.x[1][1], Mk);
                const float t_unpol = clampf(1.0f - Mk.m[0][0], R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                if (!match) {
                    st_l[c] = scale4(st_l[c], 0.0f);
                    pdf_l[c] *= t_unpol;
                    continue;
                }
                match_mask |= 1u << c;
                match_bits |= 1u << c;
                const Cx tsr = use_extra ? sol.x[0][3] : sol.x[0][2];
                const Cx tpr = use_extra ? sol.x[1][3] : sol.x[1][2];
                const float flux_k = use_extra ? sol.flux_e : sol.flux_o;
                const V3 mode_dir = use_extra ? sol.e_hat : sol.o_hat;
                const float dep = mode_power(tsr, tpr, flux_k, inc_flux, st_l[c]) / p_mode_frac;
                float c2, s2;
                azimuth2_in_frame(mode_dir, F.s_axis, p_axis_t, c2, s2);
                st_l[c] = scale4(make_float4(dep, dep * c2, dep * s2, 0.0f), inv_t);
                pdf_l[c] *= t_unpol;
            }
            narrow_compat(S.compat, match_mask, live);
            new_k = refr;
            new_s = final_refr;
            new_inside = true;
            upd = use_extra ? 1 : 0;
        }
    } else if (inside && nondegenerate) {
        // apply_uniaxial_internal_bounce (uniaxial_internal.rs)
        const InternalSol sh = internal_solve(n_med[0], n_o_l[0], n_e_ch[0], cax, F, !is_extra);
        const float flux_inc_h = fmaxf(sh.flux_inc, 1e-12f);
        const float ro = cnorm2(sh.r_o) * sh.flux_ro, re = cnorm2(sh.r_e) * sh.flux_re;
        const float r_branch = clampf((ro + re) / flux_inc_h, R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        const float p_o_exact = ro / fmaxf(ro + re, 1e-12f);
        if (unit_draw(S.seed, bounce, FRESNEL_BRANCH_STREAM) < r_branch) {
            if (EVENTS) {
                ++S.n_refl;
                S.fh = fnv(fnv(S.fh, (uint32_t)facet), 2u);
            }
            float r_tot[NCH];
            for (int c = 0; c < NCH; ++c) {  // the solves, then the state (as in the TIR arm)
                const InternalSol sol = c == 0 ? sh : internal_solve(n_med[c], n_o_l[c], n_e_ch[c], cax, F, !is_extra);
                const float flux_inc = fmaxf(sol.flux_inc, 1e-12f);
                r_tot[c] = fminf(fma_(cnorm2(sol.r_e), sol.flux_re, cnorm2(sol.r_o) * sol.flux_ro) / flux_inc, 1.0f);
            }
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                st_l[c] = scale4(st_l[c], r_tot[c] / r_branch);
                pdf_l[c] *= clampf(r_tot[c], R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            }
            new_k = reflect_dir(k, normal);
            new_s = new_k;
            exact_p_o = p_o_exact;
        } else {
            if (EVENTS) {
                ++S.n_refr;
                S.fh = fnv(fnv(S.fh, (uint32_t)facet), 3u);
            }
            const float s2d = fminf(eta * eta * one_m_c2, 1.0f);
            const V3 hero_refr = refract_dir(k, normal, eta, cos_i, sqrt_rn(fmaxf(1.0f - s2d, 0.0f)));
            const float inv_t = div_rn(1.0f, 1.0f - r_branch);
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const float orig_i = st_l[c].x;
                const float n_inc = n_med[c];
                const float r = n_inc / 1.0f;
                const float s2k = r * r * one_m_c2;
                if (s2k > 1.0f) {
                    st_l[c] = scale4(st_l[c], 0.0f);
                    pdf_l[c] = 0.0f;
                    continue;
                }
                const float ek = n_inc / 1.0f;
                const V3 rk = refract_dir(k, normal, ek, cos_i, sqrt_rn(fmaxf(1.0f - s2k, 0.0f)));
                const bool match = dot(rk, hero_refr) >= DIRECTION_MATCH_COS_TOL;
                if (match) match_bits |= 1u << c;
                const InternalSol sol = c == 0 ? sh : internal_solve(n_inc, n_o_l[c], n_e_ch[c], cax, F, !is_extra);
                float iu;
                const float4 tr = exit_transmission(sol, orig_i, inv_t, iu);
                const float t_unpol = clampf(iu, R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                if (!match) {
                    const float prefix = pdf_l[c];
                    st_l[c] = scale4(st_l[c], 0.0f);
                    pdf_l[c] = prefix * t_unpol;
                    if (orig_i > 0.0f) split_probe<EVENTS, ISO, EXTRA>(P, planes, S, split_l, facet, hit, c, rk, fmaxf(tr.x, 0.0f));
                    continue;
                }
                st_l[c] = tr;
                pdf_l[c] *= t_unpol;
            }
            new_k = hero_refr;
            new_s = hero_refr;
            new_inside = false;
        }
    } else {
        // wave along the optic axis: the scalar arms with the eigenmode draw
        // (dispatch.rs:143-380, reflect_refract.rs)
        const bool entering = !inside;
        float p_o = 0.5f, c2o = 0.0f, s2o = 0.0f;
        bool have_sel = false;
        if (entering) {
            float p;
            have_sel = eigenmode_selection(cax, cur_plane, k, st_l[0], p, c2o, s2o);
            if (have_sel) p_o = p;
        }
        const float msr = unit_draw(S.seed, bounce, BIREFRINGENT_SPLIT_STREAM);
        const bool use_extra = entering ? (msr < (1.0f - p_o)) : is_extra;
        const float c2x = use_extra ? -c2o : c2o, s2x = use_extra ? -s2o : s2o;
        // compute_entry_reflect_probability (dispatch.rs:200-230)
        const bool ord_sel = entering && !use_extra;
        const float n2h = ord_sel ? n_o_hero : n2;
        float s2h;
        if (ord_sel) {
            const float e2 = div_rn(n1, n2h);
            s2h = e2 * e2 * one_m_c2;
        } else {
            s2h = sin2_t;
        }
        float rs_h, rp_h;
        fresnel_rs_rp(n1, n2h, cos_i, sqrt_rn(fmaxf(1.0f - s2h, 0.0f)), rs_h, rp_h);
        const float r_unpol = clampf(0.5f * fma_(rp_h, rp_h, rs_h * rs_h), R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        if (unit_draw(S.seed, bounce, FRESNEL_BRANCH_STREAM) < r_unpol) {
            if (EVENTS) {
                ++S.n_refl;
                S.fh = fnv(fnv(S.fh, (uint32_t)facet), 2u);
            }
            const float sc = div_rn(1.0f, r_unpol);
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const float n1k = inside ? n_med[c] : 1.0f, n2k = inside ? 1.0f : n_med[c];
                const float ek = div_rn(n1k, n2k);
                const float s2k = ek * ek * one_m_c2;
                float4 st;
                if (s2k > 1.0f) {
                    float cd, sd;
                    tir_retardation(n1k, cos_i, sin_i, cd, sd);
                    st = mueller_tir(cd, sd, st_l[c]);
                } else {
                    float ru;
                    st = fresnel_reflect(n1k, n2k, cos_i, sqrt_rn(fmaxf(1.0f - s2k, 0.0f)), st_l[c], ru);
                    pdf_l[c] *= ru;
                }
                st_l[c] = scale4(st, sc);
            }
            new_k = reflect_dir(k, normal);
            new_s = (inside && is_extra) ? extraordinary_poynting_dir(new_k, cax, n_o_hero, n_e_hero) : new_k;
        } else {
            if (EVENTS) {
                ++S.n_refr;
                S.fh = fnv(fnv(S.fh, (uint32_t)facet), 3u);
            }
            const float n2d = (entering && !use_extra) ? n_o_hero : n2;
            const float eta_dir = div_rn(n1, n2d);
            const float s2d = fminf(eta_dir * eta_dir * one_m_c2, 1.0f);
            const V3 refr = refract_dir(k, normal, eta_dir, cos_i, sqrt_rn(fmaxf(1.0f - s2d, 0.0f)));
            const V3 final_refr =
                (entering && use_extra) ? extraordinary_poynting_dir(refr, cax, n_o_hero, n_e_hero) : refr;
            const float inv_t = div_rn(1.0f, 1.0f - r_unpol);
            uint32_t live = 0u, match_mask = 0u;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const float4 orig = st_l[c];
                const float n1k = inside ? n_med[c] : 1.0f;
                const float n2k = (entering && !use_extra) ? n_o_l[c] : (inside ? 1.0f : n_med[c]);
                const float r = div_rn(n1k, n2k);
                const float s2k = r * r * one_m_c2;
                if (s2k > 1.0f) {
                    st_l[c] = scale4(st_l[c], 0.0f);
                    pdf_l[c] = 0.0f;
                    continue;
                }
                const float ctk = sqrt_rn(fmaxf(1.0f - s2k, 0.0f));
                const V3 rk = refract_dir(k, normal, div_rn(n1k, n2k), cos_i, ctk);
                const V3 fk = (entering && use_extra)
                                  ? extraordinary_poynting_dir(rk, cax, n_o_l[c], n_o_l[c] + P.an.biref)
                                  : rk;
                const bool match = dot(fk, final_refr) >= DIRECTION_MATCH_COS_TOL;
                // compute_channel_transmission (exit_split.rs:69-101)
                const float num = 2.0f * n1k * cos_i;
                const float t_s = div_rn(num, fma_(n2k, ctk, n1k * cos_i));
                const float t_p = div_rn(num, fma_(n1k, ctk, n2k * cos_i));
                const float factor = div_rn(n2k * ctk, fmaxf(n1k * cos_i, 1e-6f));
                const float ts2 = t_s * t_s * factor, tp2 = t_p * t_p * factor;
                const float4 incident =
                    (entering && have_sel) ? make_float4(orig.x, orig.x * c2x, orig.x * s2x, 0.0f) : orig;
                const float4 tr =
                    scale4(mueller_fresnel(midpoint(ts2, tp2), 0.5f * (ts2 - tp2), t_s * t_p * factor, incident), inv_t);
                float a_, b_;
                fresnel_rs_rp(n1k, n2k, cos_i, ctk, a_, b_);
                const float ruk = clampf(0.5f * fma_(b_, b_, a_ * a_), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                if (match) {
                    st_l[c] = tr;
                    pdf_l[c] *= 1.0f - ruk;
                    match_mask |= 1u << c;
                    match_bits |= 1u << c;
                } else {
                    const float prefix = pdf_l[c];
                    st_l[c] = scale4(st_l[c], 0.0f);
                    pdf_l[c] = prefix * (1.0f - ruk);
                    if (inside && fmaxf(orig.x, 0.0f) > 0.0f)
                        split_probe<EVENTS, ISO, EXTRA>(P, planes, S, split_l, facet, hit, c, rk, fmaxf(tr.x, 0.0f));
                }
                if (!inside) channel_dir_put(c, fk);
                live |= 1u << c;
            }
            if (!inside) narrow_compat(S.compat, match_mask, live);
            new_k = refr;
            new_s = final_refr;
            new_inside = !inside;
            if (entering) upd = use_extra ? 1 : 0;
        }
    }
    // dispatch_bounce tail (transport/bounce.rs:200-238)
    const bool was_internal_reflection = inside && upd < 0 && new_inside == inside;
    S.dir = new_s;
    S.org = hit + new_s * 1e-4f;
    S.k = new_k;
    S.inside = new_inside;
    S.last_facet = facet;
    if (upd >= 0) S.is_extra = upd == 1;
    if (was_internal_reflection) {
        // apply_internal_mode_coupling (transport/bounce.rs:439-470)
        float p_o = 0.5f;
        if (exact_p_o >= 0.0f) {
            p_o = exact_p_o;
        } else {
            float p, c2, s2;
            if (eigenmode_selection(cax, cur_plane, new_k, st_l[0], p, c2, s2)) p_o = p;
        }
        S.is_extra = unit_draw(S.seed, bounce, MODE_COUPLING_STREAM) >= p_o;
    }
    // the side channel also records the eigenmode after every event
    if (EVENTS) S.fh = fnv(fnv(S.fh, 16u + (S.is_extra ? 1u : 0u)), 256u + match_bits);
    if (bounce == 0 && !new_inside && P.surface_glare < 1.0f && !frosted) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) st_l[c] = scale4(st_l[c], P.surface_glare);
    }
    if (bounce > 4) {
        float mx = 0.0f;
#pragma unroll
        for (int c = 0; c < NCH; ++c) mx = fmaxf(mx, fmaxf(st_l[c].x, 0.0f));
        const float q = clampf(mx, 0.05f, 1.0f);
        if (unit_draw(S.seed, bounce, RUSSIAN_ROULETTE_STREAM) > q) {
    #pragma unroll
        for (int c = 0; c < NCH; ++c) {
            S.st[c] = st_l[c];
            S.path_pdf[c] = pdf_l[c];
            S.split[c] = split_l[c];
        }
            return STEP_ROULETTE;
        }
        const float inv_q = div_rn(1.0f, q);
#pragma unroll
        for (int c = 0; c < NCH; ++c) st_l[c] = scale4(st_l[c], inv_q);
#pragma unroll
        for (int c = 0; c < NCH; ++c) split_l[c] = div_rn(split_l[c], q);
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        S.st[c] = st_l[c];
        S.path_pdf[c] = pdf_l[c];
        S.split[c] = split_l[c];
    }
    return STEP_CONTINUE;
}

// The assigned biaxial mode's E field (assigned_mode_e_field, biaxial.rs:432-436:
// mode B = the slow D-eigenvector) in the absorption tensor's frame
// (axes^T e, absorption_tensor.rs:80-83).
__device__ __forceinline__ V3 biax_assigned_locals(const AbsorbParams& A, const Biax& hero, const BiaxFrame& BF, V3 k,
                                                   bool is_extra) {
    const V3 e = biax_d_to_e(hero, BF, biax_eigen_d(hero, BF, k, is_extra));
    return v3(dot(v3(A.a1[0], A.a1[1], A.a1[2]), e), dot(v3(A.a2[0], A.a2[1], A.a2[2]), e),
              dot(v3(A.cn[0], A.cn[1], A.cn[2]), e));
}
// AbsorptionTensor3::quadratic_form with alpha = (o, beta, e) (assigned_mode_alpha,
// absorption_tensor.rs:115-117): alpha . (l * l)
__device__ __forceinline__ float biax_tensor_qf(V3 l, float ao, float ab, float ae) {
    return (ao * (l.x * l.x) + ab * (l.y * l.y)) + ae * (l.z * l.z);
}
// The film's per-channel indicatrix: n_beta from the dispersion curve
// (GemMaterial::biaxial_indicatrix, materials/optics.rs:16-27).
__device__ __forceinline__ Biax biax_at(const FilmParams& P, float n_beta) {
    const float na = n_beta - P.an.dba;
    return Biax{na, n_beta, na + P.an.biref};
}
__device__ __forceinline__ BiaxFrame biax_frame(const FilmParams& P) {
    return BiaxFrame{v3(P.an.bx[0], P.an.bx[1], P.an.bx[2]), v3(P.an.by[0], P.an.by[1], P.an.by[2]),
                     v3(P.an.bz[0], P.an.bz[1], P.an.bz[2])};
}

// One facet event of a biaxial material (biaxial_delta_beta_alpha is Some):
// compute_bounce_refraction_geometry's biaxial block (geometry.rs:199-270,
// 306-425: mode A = the fast root, mode B = the slow root, each channel's
// indicatrix at the hero's wave normals, resolved at entry by two fixed-point
// steps per mode), the scalar TIR and partial-Fresnel arms with a 50/50 mode
// draw at entry (tir.rs, dispatch.rs:143-262), the refract arm's per-channel
// biaxial walk-off (reflect_refract.rs:122-330), the assigned mode's
// pleochroic absorption (absorption.rs:224-249) and the 50/50 internal
// re-coupling draw (transport/bounce.rs:439-470).
template <bool EVENTS, bool ISO, bool ABSORB, bool EXTRA = true>
__device__ __forceinline__ int facet_biax(const FilmParams& P, const float4* __restrict__ planes, PathState& S,
                                          float t_hit, int facet) {
    const uint32_t bounce = S.bounce;
    const BiaxFrame BF = biax_frame(P);
    const V3 hit = S.org + S.dir * t_hit;
    const float4 pf = planes[facet];
    V3 normal = v3(pf.x, pf.y, pf.z);
    if (EXTRA && P.edge_radius > 0.0f)
        normal = shading_normal_near_edge(planes, P.n_planes, hit, facet, normal, P.edge_radius);
    const V3 cur_plane = normalize_or_zero(cross(S.k, normal));
    if (S.has_prev && dot(cur_plane, cur_plane) > 1e-6f && dot(S.prev_plane, S.prev_plane) > 1e-6f) {
        const float cos_psi = clampf(dot(S.prev_plane, cur_plane), -1.0f, 1.0f);
        const float sin_psi = dot(cross(S.prev_plane, cur_plane), normalize_or_zero(S.k));
        float c2, s2;
        rotation_double_angle(cos_psi, sin_psi, c2, s2);
#pragma unroll
        for (int c = 0; c < NCH; ++c) S.st[c] = frame_rotate(c2, s2, S.st[c]);
        if (EVENTS) ++S.n_rot;
    }
    S.prev_plane = cur_plane;
    S.has_prev = true;
    const V3 k = S.k;
    const bool inside = S.inside, is_extra = S.is_extra;
    const Biax hero = biax_at(P, S.n_o[0]);
    if (ABSORB && inside && !(EXTRA && P.sigma_s > 0.0f)) {
        // channel_absorption_alphas_assigned (absorption.rs:224-249); a
        // scattering stone's segment was attenuated by the scatter step
        const V3 l = biax_assigned_locals(P.abs, hero, BF, k, is_extra);
        const float scaled = t_hit * P.abs.path_scale;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const float ab = P.an.trichroic ? S.alpha_b[c] : S.alpha_o[c];
            S.st[c] = scale4(S.st[c], exp_lane(-biax_tensor_qf(l, S.alpha_o[c], ab, S.alpha_e[c]) * scaled));
        }
    }
    if (inside) normal = neg(normal);

    const float cos_i = clampf(dot(neg(k), normal), 0.0f, 1.0f);
    const float one_m_c2 = fma_(cos_i, -cos_i, 1.0f);
    const float sin_i = sqrt_rn(fmaxf(one_m_c2, 0.0f));
    // hero_biaxial_wave_dirs / per_channel_biaxial_indices (geometry.rs:216-270)
    V3 dir_a = k, dir_b = k;
    if (!inside) {
        dir_a = biax_resolve_entry(hero, BF, k, normal, cos_i, S.n_o[0], false);
        dir_b = biax_resolve_entry(hero, BF, k, normal, cos_i, S.n_o[0], true);
    }
    float n_a[NCH], n_b[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const Biax B = biax_at(P, S.n_o[c]);
        float ns, nf;
        biax_wave_indices(B, BF, dir_a, ns, nf);
        n_a[c] = nf;
        n_b[c] = ns;
        if (!inside) {
            biax_wave_indices(B, BF, dir_b, ns, nf);
            n_b[c] = ns;
        }
    }
    const bool mode_a = inside && !is_extra;
    const float n_med0 = mode_a ? n_a[0] : n_b[0];
    const float n1 = inside ? n_med0 : 1.0f;
    const float n2 = inside ? 1.0f : n_med0;
    const float eta = div_rn(n1, n2);
    const float sin2_t = eta * eta * one_m_c2;

    V3 new_k, new_s;
    bool new_inside = inside;
    int upd = -1;
    uint32_t match_bits = 0u;
    const bool frosted = EXTRA && P.frosted && facet_frosted(planes, facet);
    if (frosted) {
        // dispatch_bounce's frosted arm (transport/bounce.rs:137-160)
        new_k = frosted_arm_aniso<EVENTS>(P, S, bounce, facet, normal, cos_i, n1, n2, sin2_t, new_inside, upd);
        new_s = new_k;
    } else if (sin2_t > 1.0f) {
        // apply_tir_bounce, scalar arm (tir.rs:56-160)
        if (EVENTS) {
            ++S.n_tir;
            S.fh = fnv(fnv(S.fh, (uint32_t)facet), 1u);
        }
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const float nm = mode_a ? n_a[c] : n_b[c];
            const float n1k = inside ? nm : 1.0f, n2k = inside ? 1.0f : nm;
            const float ek = div_rn(n1k, n2k);
            const float s2k = ek * ek * one_m_c2;
            if (s2k > 1.0f) {
                float cd, sd;
                tir_retardation(n1k, cos_i, sin_i, cd, sd);
                S.st[c] = mueller_tir(cd, sd, S.st[c]);
            } else {
                float ru;
                S.st[c] = fresnel_reflect(n1k, n2k, cos_i, sqrt_rn(fmaxf(1.0f - s2k, 0.0f)), S.st[c], ru);
                S.path_pdf[c] *= ru;
            }
        }
        new_k = reflect_dir(k, normal);
        new_s = inside ? biax_mode_poynting(hero, BF, new_k, is_extra) : new_k;
    } else {
        // apply_partial_fresnel_bounce (dispatch.rs:143-262): no polarisation-
        // weighted selection for a biaxial stone, so p_o = 0.5
        const bool entering = !inside;
        const float msr = unit_draw(S.seed, bounce, BIREFRINGENT_SPLIT_STREAM);
        const bool use_extra = entering ? (msr < (1.0f - 0.5f)) : is_extra;
        float rs_h, rp_h;
        fresnel_rs_rp(n1, n2, cos_i, sqrt_rn(fmaxf(1.0f - sin2_t, 0.0f)), rs_h, rp_h);
        const float r_unpol = clampf(0.5f * fma_(rp_h, rp_h, rs_h * rs_h), R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        if (unit_draw(S.seed, bounce, FRESNEL_BRANCH_STREAM) < r_unpol) {
            if (EVENTS) {
                ++S.n_refl;
                S.fh = fnv(fnv(S.fh, (uint32_t)facet), 2u);
            }
            const float sc = div_rn(1.0f, r_unpol);
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const float nm = mode_a ? n_a[c] : n_b[c];
                const float n1k = inside ? nm : 1.0f, n2k = inside ? 1.0f : nm;
                const float ek = div_rn(n1k, n2k);
                const float s2k = ek * ek * one_m_c2;
                float4 st;
                if (s2k > 1.0f) {
                    float cd, sd;
                    tir_retardation(n1k, cos_i, sin_i, cd, sd);
                    st = mueller_tir(cd, sd, S.st[c]);
                } else {
                    float ru;
                    st = fresnel_reflect(n1k, n2k, cos_i, sqrt_rn(fmaxf(1.0f - s2k, 0.0f)), S.st[c], ru);
                    S.path_pdf[c] *= ru;
                }
                S.st[c] = scale4(st, sc);
            }
            new_k = reflect_dir(k, normal);
            new_s = inside ? biax_mode_poynting(hero, BF, new_k, is_extra) : new_k;
        } else {
            if (EVENTS) {
                ++S.n_refr;
                S.fh = fnv(fnv(S.fh, (uint32_t)facet), 3u);
            }
            const float n2d = entering ? (use_extra ? n_b[0] : n_a[0]) : n2;
            const float eta_dir = div_rn(n1, n2d);
            const float s2d = fminf(eta_dir * eta_dir * one_m_c2, 1.0f);
            const V3 refr = refract_dir(k, normal, eta_dir, cos_i, sqrt_rn(fmaxf(1.0f - s2d, 0.0f)));
            const V3 final_refr = entering ? biax_mode_poynting(hero, BF, refr, use_extra) : refr;
            const float inv_t = div_rn(1.0f, 1.0f - r_unpol);
            uint32_t live = 0u, match_mask = 0u;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const float4 orig = S.st[c];
                const float nm = mode_a ? n_a[c] : n_b[c];
                const float n1k = inside ? nm : 1.0f;
                const float n2k = (entering && !use_extra) ? n_a[c] : (inside ? 1.0f : nm);
                const float r = div_rn(n1k, n2k);
                const float s2k = r * r * one_m_c2;
                if (s2k > 1.0f) {
                    S.st[c] = scale4(S.st[c], 0.0f);
                    S.path_pdf[c] = 0.0f;
                    continue;
                }
                const float ctk = sqrt_rn(fmaxf(1.0f - s2k, 0.0f));
                const V3 rk = refract_dir(k, normal, div_rn(n1k, n2k), cos_i, ctk);
                V3 fk = rk;
                if (entering) {
                    const Biax B = biax_at(P, S.n_o[c]);
                    fk = biax_mode_poynting(B, BF, rk, use_extra);
                }
                const bool match = dot(fk, final_refr) >= DIRECTION_MATCH_COS_TOL;
                // compute_channel_transmission (exit_split.rs:69-101)
                const float num = 2.0f * n1k * cos_i;
                const float t_s = div_rn(num, fma_(n2k, ctk, n1k * cos_i));
                const float t_p = div_rn(num, fma_(n1k, ctk, n2k * cos_i));
                const float factor = div_rn(n2k * ctk, fmaxf(n1k * cos_i, 1e-6f));
                const float ts2 = t_s * t_s * factor, tp2 = t_p * t_p * factor;
                const float4 tr =
                    scale4(mueller_fresnel(midpoint(ts2, tp2), 0.5f * (ts2 - tp2), t_s * t_p * factor, orig), inv_t);
                float a_, b_;
                fresnel_rs_rp(n1k, n2k, cos_i, ctk, a_, b_);
                const float ruk = clampf(0.5f * fma_(b_, b_, a_ * a_), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                if (match) {
                    S.st[c] = tr;
                    S.path_pdf[c] *= 1.0f - ruk;
                    match_mask |= 1u << c;
                    match_bits |= 1u << c;
                } else {
                    const float prefix = S.path_pdf[c];
                    S.st[c] = scale4(S.st[c], 0.0f);
                    S.path_pdf[c] = prefix * (1.0f - ruk);
                    if (inside && fmaxf(orig.x, 0.0f) > 0.0f)
                        split_probe<EVENTS, ISO, EXTRA>(P, planes, S, S.split, facet, hit, c, rk, fmaxf(tr.x, 0.0f));
                }
                if (!inside) channel_dir_put(c, fk);
                live |= 1u << c;
            }
            if (!inside) narrow_compat(S.compat, match_mask, live);
            new_k = refr;
            new_s = final_refr;
            new_inside = !inside;
            if (entering) upd = use_extra ? 1 : 0;
        }
    }
    // dispatch_bounce tail (transport/bounce.rs:200-238)
    const bool was_internal_reflection = inside && upd < 0 && new_inside == inside;
    S.dir = new_s;
    S.org = hit + new_s * 1e-4f;
    S.k = new_k;
    S.inside = new_inside;
    S.last_facet = facet;
    if (upd >= 0) S.is_extra = upd == 1;
    if (was_internal_reflection) S.is_extra = unit_draw(S.seed, bounce, MODE_COUPLING_STREAM) >= 0.5f;
    if (EVENTS) S.fh = fnv(fnv(S.fh, 16u + (S.is_extra ? 1u : 0u)), 256u + match_bits);
    if (bounce == 0 && !new_inside && P.surface_glare < 1.0f && !frosted) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) S.st[c] = scale4(S.st[c], P.surface_glare);
    }
    if (bounce > 4) {
        float mx = 0.0f;
#pragma unroll
        for (int c = 0; c < NCH; ++c) mx = fmaxf(mx, fmaxf(S.st[c].x, 0.0f));
        const float q = clampf(mx, 0.05f, 1.0f);
        if (unit_draw(S.seed, bounce, RUSSIAN_ROULETTE_STREAM) > q) return STEP_ROULETTE;
        const float inv_q = div_rn(1.0f, q);
#pragma unroll
        for (int c = 0; c < NCH; ++c) S.st[c] = scale4(S.st[c], inv_q);
#pragma unroll
        for (int c = 0; c < NCH; ++c) S.split[c] = div_rn(S.split[c], q);
    }
    return STEP_CONTINUE;
}

// One iteration of the bounce loop (bounce_body.rs:361-707): intersection, and
// either the escape lookup (added to the split staging) or the facet event.
// apply_russian_roulette (transport/bounce.rs:355-375): false when the path dies.
__device__ __forceinline__ bool russian_roulette(PathState& S, uint32_t bounce) {
    float mx = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; ++c) mx = fmaxf(mx, fmaxf(S.st[c].x, 0.0f));
    const float q = clampf(mx, 0.05f, 1.0f);
    if (unit_draw(S.seed, bounce, RUSSIAN_ROULETTE_STREAM) > q) return false;
    const float inv_q = div_rn(1.0f, q);
#pragma unroll
    for (int c = 0; c < NCH; ++c) S.st[c] = scale4(S.st[c], inv_q);
#pragma unroll
    for (int c = 0; c < NCH; ++c) S.split[c] = div_rn(S.split[c], q);
    return true;
}

// sample_henyey_greenstein_direction (scattering/hg.rs:636-651) about fwd,
// the azimuth's sine and cosine correctly rounded.
__device__ __forceinline__ V3 hg_direction(float u1, float u2, float g, V3 fwd) {
    float cos_theta;
    if (fabsf(g) < 1e-3f) {
        cos_theta = fma_(2.0f, -u1, 1.0f);
    } else {
        const float one_minus_g2 = fma_(g, -g, 1.0f);
        const float denom = fma_(2.0f * g, u1, 1.0f - g);
        const float sq = div_rn(one_minus_g2, denom);
        cos_theta = div_rn(fma_(sq, -sq, fma_(g, g, 1.0f)), 2.0f * g);
    }
    cos_theta = clampf(cos_theta, -1.0f, 1.0f);
    const float sin_theta = sqrt_rn(fmaxf(fma_(cos_theta, -cos_theta, 1.0f), 0.0f));
    const float phi = (2.0f * 3.14159265358979323846f) * u2;
    const float sin_p = (float)sin((double)phi), cos_p = (float)cos((double)phi);
    const V3 a = fabsf(fwd.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    const V3 t = normalize_or_zero(a - fwd * dot(fwd, a));
    const V3 b = cross(fwd, t);
    return normalize_or_zero((t * (sin_theta * cos_p) + b * (sin_theta * sin_p)) + fwd * cos_theta);
}

// try_scatter_step / maybe_scatter_or_extinguish (scattering/hg.rs:91-160,
// 435-600) for an isotropic, unzoned stone with next-event estimation off: a
// free-flight distance against the hero channel's extinction; short of the
// facet the path scatters (Henyey-Greenstein, every channel depolarised and
// reweighted by its own transmittance) and the iteration ends there, else it
// reaches the facet with each channel's survival relative to the hero's.
// Returns a StepResult, or -1 for "reached the facet".
template <bool EVENTS, bool ABSORB, int ANISO = 0>
__device__ __forceinline__ int scatter_step(const FilmParams& P, PathState& S, float hit_t) {
    const uint32_t bounce = S.bounce;
    float alpha[NCH];
    if (ABSORB && ANISO == 2) {
        // channel_absorption_alphas_assigned (absorption.rs:224-249), biaxial
        const V3 l = biax_assigned_locals(P.abs, biax_at(P, S.n_o[0]), biax_frame(P), S.k, S.is_extra);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            alpha[c] = biax_tensor_qf(l, S.alpha_o[c], P.an.trichroic ? S.alpha_b[c] : S.alpha_o[c], S.alpha_e[c]);
    } else if (ABSORB && ANISO) {
        // the same, uniaxial: the assigned o / e mode's E field
        const V3 e = uniaxial_assigned_e(P, S, S.k);
#pragma unroll
        for (int c = 0; c < NCH; ++c) alpha[c] = tensor_qf(P.abs, S.alpha_o[c], S.alpha_e[c], e);
    } else if (ABSORB) {
        channel_alphas(P.abs, S, S.dir, alpha);
    } else {
#pragma unroll
        for (int c = 0; c < NCH; ++c) alpha[c] = 0.0f;
    }
    const float path_scale = P.abs.path_scale;
    const float sigma_s = div_rn(P.sigma_s, path_scale);
    const float sigma_t_hero = alpha[0] + sigma_s;
    const float one_minus_u = fmaxf(1.0f - unit_draw(S.seed, bounce, DISTANCE_SAMPLE_STREAM), 1e-7f);
    const float t_free = div_rn(-(float)log((double)one_minus_u), sigma_t_hero);
    const float hit_t_scaled = hit_t * path_scale;
    if (t_free < hit_t_scaled) {
        const float pdf_hero = sigma_t_hero * one_minus_u;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const float sigma_t = alpha[c] + sigma_s;
            const float tr = exp_lane(-(alpha[c] + sigma_s) * t_free);
            const float weight = div_rn(tr * sigma_s, pdf_hero);
            S.st[c] = make_float4(fmaxf(S.st[c].x, 0.0f) * weight, 0.0f, 0.0f, 0.0f);
            S.path_pdf[c] *= sigma_t * tr;
        }
        const float u1 = unit_draw(S.seed, bounce, PHASE_DIR_U_STREAM);
        const float u2 = unit_draw(S.seed, bounce, PHASE_DIR_V_STREAM);
        const V3 old_dir = S.dir;
        const V3 new_dir = hg_direction(u1, u2, P.scat_g, old_dir);
        S.org = S.org + old_dir * div_rn(t_free, path_scale);
        S.dir = new_dir;
        S.k = new_dir;       // the wave normal of an anisotropic stone follows the new direction
        S.has_prev = false;  // depolarised: no previous plane of incidence
        if (EVENTS) S.fh = fnv(fnv(S.fh, 0xFFFFu), 7u);
        if (bounce > 4 && !russian_roulette(S, bounce)) return STEP_SCATTER_ABSORBED;
        return STEP_CONTINUE;
    }
    float survive[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) survive[c] = exp_lane(-(alpha[c] + sigma_s) * hit_t_scaled);
    const float hero = fmaxf(survive[0], 1e-30f);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        S.st[c] = scale4(S.st[c], div_rn(survive[c], hero));
        S.path_pdf[c] *= survive[c];
    }
    return -1;
}

// handle_miss (bounce_body/miss.rs:33-54): the environment seen along the
// path's direction, weighted by each channel's intensity, added to its
// exit-split staging.
template <bool ISO, bool EXTRA>
__device__ __forceinline__ void path_miss(const FilmParams& P, PathState& S) {
    // the light-sample MIS carry of an exterior frosted bounce (bounce_body/miss.rs:22-26):
    // the balance heuristic against the environment's light-sample density
    float w = 1.0f;
    if (EXTRA && S.has_mis) w = balance_heuristic(S.mis_phase, environment_nee_pdf(P.env, S.dir));
    if (EXTRA) S.has_mis = false;
    if (EXTRA && !ISO && P.env.model == ENV_HDR) {  // the panorama: one texel, per-channel spectrum
        float rgb[3];
        hdr_rgb(P.env, S.dir, rgb);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            float env = rgb_to_spectral(rgb, comb_lambda(S.hero, c));
            S.split[c] = fma_(fmaxf(S.st[c].x, 0.0f) * (EXTRA ? w : 1.0f), env, EXTRA ? S.nee[c] : 0.0f) + S.split[c];
        }
    } else if (!ISO && P.env.model == ENV_ASET) {
        const V3 d = normalize(S.dir);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            float env = aset_radiance(P.env, d, comb_lambda(S.hero, c), P.env.exposure);
            S.split[c] = fma_(fmaxf(S.st[c].x, 0.0f) * (EXTRA ? w : 1.0f), env, EXTRA ? S.nee[c] : 0.0f) + S.split[c];
        }
    } else if (ISO || P.env.model != ENV_STUDIO) {
        const float f = lit_factor(P.env, S.dir, S.observer);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            float env = f * (spec_power(P.env, comb_lambda(S.hero, c)) * P.env.exposure);
            S.split[c] = fma_(fmaxf(S.st[c].x, 0.0f) * (EXTRA ? w : 1.0f), env, EXTRA ? S.nee[c] : 0.0f) + S.split[c];
        }
    } else {
        DirLight L;
        direction_lighting(P.env, S.dir, S.observer, L);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            float env = light_radiance(P.env, L, spec_power(P.env, comb_lambda(S.hero, c)));
            S.split[c] = fma_(fmaxf(S.st[c].x, 0.0f) * (EXTRA ? w : 1.0f), env, EXTRA ? S.nee[c] : 0.0f) + S.split[c];
        }
    }
}

// EXTRA: the rarely used surface and medium features (edge-rounded normals,
// frosted facets, volume scattering) are compiled in; the persistent film
// kernel is also built without them, so a stone that uses none runs the
// lean bounce.
template <bool EVENTS, bool ISO = false, bool ABSORB = false, int ANISO = 0, bool EXTRA = true>
__device__ __forceinline__ int path_bounce(const FilmParams& P, const float4* __restrict__ planes,
                                           PathState& S) {
    const uint32_t bounce = S.bounce;
    float t_hit = 0.0f;
    int facet;
    if (!S.inside && S.last_facet >= 0 && leaves_convex(planes, S.last_facet, S.org, S.dir))
        facet = -1;
    else {
        facet = intersect(planes, P.n_planes, S.org, S.dir, t_hit);
        if (EVENTS) ++S.n_scan;
    }
    if (facet < 0) {
        path_miss<ISO, EXTRA>(P, S);
        return STEP_ESCAPED;
    }
    if (EXTRA) S.has_mis = false;  // the carry is this bounce's only (bounce_body.rs:372)

    if (EXTRA && P.sigma_s > 0.0f && S.inside) {  // run_bounce (bounce_body.rs:410-416)
        const int r = scatter_step<EVENTS, ABSORB, ANISO>(P, S, t_hit);
        if (r >= 0) return r;
    }
    if (ANISO == 2) return facet_biax<EVENTS, ISO, ABSORB, EXTRA>(P, planes, S, t_hit, facet);
    if (ANISO) return facet_aniso<EVENTS, ISO, ABSORB, EXTRA>(P, planes, S, t_hit, facet);
    // facet_event (bounce_body.rs:611-707)
    const V3 k = S.dir;
    const V3 hit = S.org + k * t_hit;
    const float4 pf = planes[facet];
    V3 normal = v3(pf.x, pf.y, pf.z);
    // bounce_body.rs:620-629: the edge-rounded shading normal replaces the facet's
    if (EXTRA && P.edge_radius > 0.0f)
        normal = shading_normal_near_edge(planes, P.n_planes, hit, facet, normal, P.edge_radius);
    // rotate_stokes_to_plane_of_incidence (absorption.rs:69-87)
    const V3 cur_plane = normalize_or_zero(cross(k, normal));
    if (S.has_prev && dot(cur_plane, cur_plane) > 1e-6f && dot(S.prev_plane, S.prev_plane) > 1e-6f) {
        float cos_psi = clampf(dot(S.prev_plane, cur_plane), -1.0f, 1.0f);
        float sin_psi = dot(cross(S.prev_plane, cur_plane), normalize_or_zero(k));
        float c2, s2;
        rotation_double_angle(cos_psi, sin_psi, c2, s2);
#pragma unroll
        for (int c = 0; c < NCH; ++c) S.st[c] = frame_rotate(c2, s2, S.st[c]);
        if (EVENTS) ++S.n_rot;
    }
    S.prev_plane = cur_plane;
    S.has_prev = true;
    // apply_interior_segment (transport/inner.rs:161-178); without bands every
    // channel transmits exp(-0) == 1 exactly, so only the normal flips.
    const bool inside = S.inside;
    if (ABSORB && inside && !(EXTRA && P.sigma_s > 0.0f)) apply_absorption(P.abs, S, k, t_hit);  // scattering: done above
    if (inside) normal = neg(normal);

    // compute_bounce_refraction_geometry (geometry.rs:306-425), isotropic
    const float cos_i = clampf(dot(neg(k), normal), 0.0f, 1.0f);
    const float one_m_c2 = fma_(cos_i, -cos_i, 1.0f);
    const float sin_i = sqrt_rn(fmaxf(one_m_c2, 0.0f));
    const float n1 = inside ? S.n_o[0] : 1.0f;
    const float n2 = inside ? 1.0f : S.n_o[0];
    const float eta = div_rn(n1, n2);
    const float sin2_t = eta * eta * one_m_c2;

    // hero decision: forced TIR (tir.rs:56) or the reflect-vs-transmit draw
    // against the clamped unpolarised reflectance (dispatch.rs:230-257, 345-378)
    const bool hero_tir = sin2_t > 1.0f;
    const float cos_t = sqrt_rn(fmaxf(1.0f - sin2_t, 0.0f));
    float rs_h, rp_h;
    fresnel_rs_rp(n1, n2, cos_i, cos_t, rs_h, rp_h);
    const float r_unpol = clampf(0.5f * fma_(rp_h, rp_h, rs_h * rs_h), R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
    const float rng = unit_draw(S.seed, bounce, FRESNEL_BRANCH_STREAM);
    const bool reflect = hero_tir || rng < r_unpol;

    V3 new_k;
    bool new_inside = inside;
    bool frosted_hit = false;
    if (EXTRA && P.frosted && facet_frosted(planes, facet)) {
        // apply_frosted_bounce (scattering/frosted.rs), next-event estimation
        // off: a cosine-weighted hemisphere on the reflect side (always under
        // TIR) or, transmitted, the far side; every channel depolarised
        frosted_hit = true;
        const float u1 = unit_draw(S.seed, bounce, FROSTED_DIR_U_STREAM);
        const float u2 = unit_draw(S.seed, bounce, FROSTED_DIR_V_STREAM);
        bool through = false;
        float pdf = 1.0f;
        if (!hero_tir) {
            const float r_pdf = clampf(0.5f * fma_(rp_h, rp_h, rs_h * rs_h), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            through = !(rng < r_pdf);
            pdf = through ? 1.0f - r_pdf : r_pdf;
        }
        new_k = cosine_hemisphere(u1, u2, through ? neg(normal) : normal);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            S.st[c] = make_float4(fmaxf(S.st[c].x, 0.0f), 0.0f, 0.0f, 0.0f);
            if (!hero_tir) S.path_pdf[c] *= pdf;
        }
        // the exterior side: reflected while outside, or transmitted out
        if (!hero_tir && through == inside) frosted_exterior_light(P, S, through ? neg(normal) : normal, new_k, bounce);
        if (through) new_inside = !inside;
        if (EVENTS) S.fh = fnv(fnv(S.fh, (uint32_t)facet), hero_tir ? 4u : (through ? 6u : 5u));
    } else if (reflect) {
        // apply_tir_bounce's scalar branch (tir.rs:139-160) and
        // apply_partial_reflect_bounce (reflect_refract.rs:33-71) share every
        // per-channel step; only the partial reflection divides by r_unpol
        // (a multiplication by exactly 1 for the forced TIR).
        if (EVENTS) {
            if (hero_tir) ++S.n_tir;
            else ++S.n_refl;
            S.fh = fnv(fnv(S.fh, (uint32_t)facet), hero_tir ? 1u : 2u);
        }
        const float scale = hero_tir ? 1.0f : div_rn(1.0f, r_unpol);
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            float n1k = inside ? S.n_o[c] : 1.0f;
            float n2k = inside ? 1.0f : S.n_o[c];
            float ek = S.n_o[c];  // n1k / n2k: n / 1 is n, only an entering ray divides
            if (!inside) ek = div_rn(1.0f, ek);
            float s2k = ek * ek * one_m_c2;
            float4 s;
            if (s2k > 1.0f) {
                float cd, sd;
                tir_retardation(n1k, cos_i, sin_i, cd, sd);
                s = mueller_tir(cd, sd, S.st[c]);
            } else {
                float ru;
                s = fresnel_reflect(n1k, n2k, cos_i, sqrt_rn(fmaxf(1.0f - s2k, 0.0f)), S.st[c], ru);
                S.path_pdf[c] *= ru;
            }
            S.st[c] = scale4(s, scale);
        }
        new_k = k - normal * (2.0f * dot(k, normal));
    } else {
        // apply_refract_bounce (reflect_refract.rs:121-234), isotropic
        if (EVENTS) {
            ++S.n_refr;
            S.fh = fnv(fnv(S.fh, (uint32_t)facet), 3u);
        }
        const float s2d = fminf(eta * eta * one_m_c2, 1.0f);
        const float ct_dir = sqrt_rn(fmaxf(1.0f - s2d, 0.0f));
        const V3 refr = normalize(k * eta + normal * fma_(eta, cos_i, -ct_dir));
        const float inv_t = div_rn(1.0f, 1.0f - r_unpol);
        uint32_t live = 0u, match_mask = 0u;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            // apply_refract_channel (reflect_refract.rs:517-591)
            const float4 orig = S.st[c];
            float n1k = inside ? S.n_o[c] : 1.0f;
            float n2k = inside ? 1.0f : S.n_o[c];
            float r = S.n_o[c];  // n1k / n2k, as above
            if (!inside) r = div_rn(1.0f, r);
            float s2k = r * r * one_m_c2;
            if (s2k > 1.0f) {
                S.st[c] = scale4(S.st[c], 0.0f);
                S.path_pdf[c] = 0.0f;
                continue;
            }
            float ctk = sqrt_rn(fmaxf(1.0f - s2k, 0.0f));
            V3 dk = normalize(k * r + normal * fma_(r, cos_i, -ctk));
            bool match = dot(dk, refr) >= DIRECTION_MATCH_COS_TOL;
            float a_, b_;
            const float4 tr = scale4(fresnel_transmit(n1k, n2k, cos_i, ctk, orig, a_, b_), inv_t);
            float ruk = clampf(0.5f * fma_(b_, b_, a_ * a_), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            if (match) {
                S.st[c] = tr;
                S.path_pdf[c] *= 1.0f - ruk;
                match_mask |= 1u << c;
            } else {
                // apply_channel_chromatic_termination (reflect_refract.rs:444-515)
                float prefix = S.path_pdf[c];
                S.st[c] = scale4(S.st[c], 0.0f);
                S.path_pdf[c] = prefix * (1.0f - ruk);
                if (inside && fmaxf(orig.x, 0.0f) > 0.0f) {
                    // try_split_exit_channel (exit_split.rs:128-163): the
                    // re-entry probe of a convex stone resolves from the exit
                    // facet alone in all but near-grazing cases.
                    if (EVENTS) ++S.n_probe;
                    const V3 po = hit + dk * 1e-4f;
                    bool escapes = leaves_convex(planes, facet, po, dk);
                    if (!escapes) {
                        float tt;
                        escapes = intersect(planes, P.n_planes, po, dk, tt) < 0;
                        if (EVENTS) ++S.n_probe_scan;
                    }
                    if (escapes) {
                        float env = ISO ? lit_factor(P.env, dk, S.observer) *
                                              (spec_power(P.env, comb_lambda(S.hero, c)) * P.env.exposure)
                                        : env_radiance<EXTRA>(P.env, dk, S.observer, comb_lambda(S.hero, c));
                        S.split[c] = fma_(fmaxf(fmaxf(tr.x, 0.0f), 0.0f) * 1.0f, env, S.split[c]);
                    }
                }
            }
            if (!inside) channel_dir_put(c, dk);
            live |= 1u << c;
        }
        if (!inside) narrow_compat(S.compat, match_mask, live);
        new_k = refr;
        new_inside = !inside;
    }
    // dispatch_bounce tail (transport/bounce.rs:200-238)
    S.dir = new_k;
    S.org = hit + new_k * 1e-4f;
    S.inside = new_inside;
    S.last_facet = facet;
    if (bounce == 0 && !new_inside && P.surface_glare < 1.0f && !frosted_hit) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) S.st[c] = scale4(S.st[c], P.surface_glare);
    }
    // apply_russian_roulette (transport/bounce.rs:355-375)
    if (bounce > 4) {
        float mx = 0.0f;
#pragma unroll
        for (int c = 0; c < NCH; ++c) mx = fmaxf(mx, fmaxf(S.st[c].x, 0.0f));
        float q = clampf(mx, 0.05f, 1.0f);
        float rr = unit_draw(S.seed, bounce, RUSSIAN_ROULETTE_STREAM);
        if (rr > q) return STEP_ROULETTE;
        float inv_q = div_rn(1.0f, q);
#pragma unroll
        for (int c = 0; c < NCH; ++c) S.st[c] = scale4(S.st[c], inv_q);
#pragma unroll
        for (int c = 0; c < NCH; ++c) S.split[c] = div_rn(S.split[c], q);
    }
    return STEP_CONTINUE;
}

// PathState::finish (bounce_body.rs:724-764) + color.rs:145-172 + von Kries,
// from the path's final radiance (the exit-split staging already added for
// an escaped path), MIS families, pdfs and wavelengths.
__device__ __forceinline__ void finish_xyz(const FilmParams& P, uint64_t compat, const float path_pdf[NCH],
                                           const float lambda[NCH], const float radiance[NCH], float xyz[3]) {
    const float norm_factor = (400.0f / 8.0f) / 106.856f;
    float X = 0.0f, Y = 0.0f, Z = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const uint32_t fam = compat_of(compat, c);
        float family = -0.0f;
#pragma unroll
        for (int j = 0; j < NCH; ++j)
            if ((fam >> j) & 1u) family += path_pdf[j];
        float w = family <= 1e-12f ? 1.0f : div_rn(8.0f * path_pdf[0], family);
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
    float L_ = fma_(0.8951f, X, fma_(0.2664f, Y, -0.1614f * Z)) * P.env.wb[0];
    float M_ = fma_(-0.7502f, X, fma_(1.7135f, Y, 0.0367f * Z)) * P.env.wb[1];
    float S_ = fma_(0.0389f, X, fma_(-0.0685f, Y, 1.0296f * Z)) * P.env.wb[2];
    xyz[0] = fmaxf(fma_(0.986993f, L_, fma_(-0.147054f, M_, 0.159963f * S_)), 0.0f);
    xyz[1] = fmaxf(fma_(0.432305f, L_, fma_(0.518360f, M_, 0.049291f * S_)), 0.0f);
    xyz[2] = fmaxf(fma_(-0.008529f, L_, fma_(0.040043f, M_, 0.968487f * S_)), 0.0f);
}

// An escaped path's radiance is in S.split (the miss radiance added to the
// exit-split staging when the path escaped); any other ending has none.
template <bool EXTRA = true>
__device__ __forceinline__ void path_finish(const FilmParams& P, const PathState& S, bool escaped, float xyz[3]) {
    float radiance[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) radiance[c] = escaped ? S.split[c] : (EXTRA ? S.nee[c] : 0.0f);
    float lambda[NCH];
    hero_comb(S.hero, lambda);
    finish_xyz(P, S.compat, S.path_pdf, lambda, radiance, xyz);
}

// trace_spectral_ray (transport/mod.rs:118) for one explicit ray, start to end.
template <bool EVENTS, bool ABSORB, int ANISO>
__device__ __forceinline__ void trace_path(const FilmParams& P, const float4* __restrict__ planes,
                                           V3 org, V3 dir, const uint32_t seed, const float hero,
                                           PathOut& out, uint32_t* counts) {
    PathState S;
    path_init<ABSORB, ANISO>(P, org, dir, seed, hero, S);
    uint32_t bounces = P.max_bounces, term = 3u;
    bool escaped = false;
    for (uint32_t b = 0; b < P.max_bounces; ++b) {
        S.bounce = b;
        int r = path_bounce<EVENTS, false, ABSORB, ANISO>(P, planes, S);
        if (r != STEP_CONTINUE) {
            bounces = b + 1;
            term = r == STEP_ESCAPED ? 0u : (r == STEP_SCATTER_ABSORBED ? 4u : 2u);
            escaped = r == STEP_ESCAPED;
            break;
        }
    }
    path_finish(P, S, escaped, out.xyz);
    if (EVENTS) {
        out.ev.bounces = bounces;
        out.ev.termination = term;
        out.ev.facet_hash = S.fh;
        out.ev.n_tir = (uint16_t)S.n_tir;
        out.ev.n_reflect = (uint16_t)S.n_refl;
        out.ev.n_refract = (uint16_t)S.n_refr;
        out.ev.n_probe = (uint16_t)S.n_probe;
        if (counts) {
            counts[EV_BOUNCES] += bounces;
            counts[EV_TIR] += S.n_tir;
            counts[EV_REFLECT] += S.n_refl;
            counts[EV_REFRACT] += S.n_refr;
            counts[EV_PROBE] += S.n_probe;
            counts[EV_ESCAPED] += term == 0u ? 1u : 0u;
            counts[EV_ROTATIONS] += S.n_rot;
            counts[EV_SCANS] += S.n_scan;
            counts[EV_PROBE_SCANS] += S.n_probe_scan;
        }
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;

// Occupancy target of the persistent film kernel (waves per SIMD); the
// register budget follows from it (512 VGPRs / waves).  The uniaxial
// instantiations run at 2 (256 registers): measured on the C2 oval, 2 waves
// beat 3 by 2 % and 1, 4 and 5 lose 20-35 % (fewer spills do not pay for the
// lost latency hiding at 1).
#ifndef IXG_FILM_WAVES_PER_EU
#define IXG_FILM_WAVES_PER_EU 3
#endif
#ifndef IXG_FILM_WAVES_PER_EU_ANISO
#define IXG_FILM_WAVES_PER_EU_ANISO 2
#endif
#define IXG_WAVES_ATTR \
    __attribute__((amdgpu_waves_per_eu(ANISO ? IXG_FILM_WAVES_PER_EU_ANISO : IXG_FILM_WAVES_PER_EU)))
// Idle lanes a wave collects before it finishes and regenerates paths.
#ifndef IXG_FILM_REGEN_MIN
#define IXG_FILM_REGEN_MIN 32
#endif

__device__ __forceinline__ void stage_planes(const FilmParams& P, const float4* __restrict__ g,
                                             float4* s) {
    for (uint32_t i = threadIdx.x; i < P.n_planes; i += blockDim.x) s[i] = g[i];
    if (P.frosted && threadIdx.x < 2) s[IXG_MAX_PLANES + threadIdx.x] = g[IXG_MAX_PLANES + threadIdx.x];
    __syncthreads();
}

// Film, persistent: each lane loops one bounce at a time and, the moment its
// path ends, takes the next path index from a wave-wide grab on a global
// counter.  Lanes therefore never idle behind a longer path of their wave
// (paths run 1..max_bounces iterations, mostly 2-6), and the per-path results
// are the same as film_transport_kernel's because a path is a pure function
// of its (pixel, sample) index.
template <bool ISO, bool ABSORB, int ANISO, bool EXTRA>
__global__ __launch_bounds__(kBlock) IXG_WAVES_ATTR void film_persistent_kernel(const FilmParams P,
                                                                 const float4* __restrict__ planes,
                                                                 float* __restrict__ out_xyz,
                                                                 uint32_t* __restrict__ counter) {
    __shared__ float4 sp[IXG_MAX_PLANES + 2];  // + the frosted bitmask
    stage_planes(P, planes, sp);
    const uint32_t n = (uint32_t)P.n_paths;
    const uint32_t lane = __lane_id();
    PathState S;
    uint32_t idx = 0;
    bool active = false, done = false;
    bool pending = false, pending_escaped = false;  // finished, XYZ not yet written
    while (true) {
        // Regenerate (finish the lanes that ended, then take new paths) once
        // IXG_FILM_REGEN_MIN lanes are idle, or when no lane is active: the
        // finish / camera / dispersion code then runs once for a batch of
        // lanes instead of on nearly every bounce.
        const bool need = !active && !done;
        const uint64_t need_mask = __ballot(need);
        const uint64_t active_mask = __ballot(active);
        if (need_mask && (__popcll(need_mask) >= IXG_FILM_REGEN_MIN || active_mask == 0ull)) {
            if (pending) {
                float xyz[3];
                path_finish<EXTRA>(P, S, pending_escaped, xyz);
                out_xyz[3 * (size_t)idx + 0] = xyz[0];
                out_xyz[3 * (size_t)idx + 1] = xyz[1];
                out_xyz[3 * (size_t)idx + 2] = xyz[2];
                pending = false;
            }
            const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)need_mask) - 1);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(need_mask));
            base = __shfl(base, (int)leader);
            if (need) {
                idx = base + (uint32_t)__popcll(need_mask & ((1ull << lane) - 1ull));
                if (idx < n) {
                    const uint64_t p = P.first_path + idx;
                    V3 org, dir;
                    uint32_t seed;
                    float hero;
                    camera_sample(P, (uint32_t)(p / P.spp), P.sample_offset + (uint32_t)(p % P.spp),
                                  org, dir, seed, hero);
                    path_init<ABSORB, ANISO>(P, org, dir, seed, hero, S);
                    active = true;
                } else {
                    done = true;
                }
            }
        }
        if (__ballot(active) == 0ull) break;
        if (active) {
            int r = STEP_CONTINUE;
            if (S.bounce < P.max_bounces) r = path_bounce<false, ISO, ABSORB, ANISO, EXTRA>(P, sp, S);
            // A path that now leaves the convex stone through the facet it
            // just crossed would find no facet on its next bounce: that
            // bounce's miss is taken here, so the lane does not sit out the
            // other lanes' slab scans for an iteration.
            if (r == STEP_CONTINUE && ++S.bounce < P.max_bounces && !S.inside && S.last_facet >= 0 &&
                leaves_convex(sp, S.last_facet, S.org, S.dir)) {
                path_miss<ISO, EXTRA>(P, S);
                r = STEP_ESCAPED;
            }
            if (r != STEP_CONTINUE || S.bounce >= P.max_bounces) {
                active = false;
                pending = true;
                pending_escaped = r == STEP_ESCAPED;
            }
        }
    }
}

// Diagnostics: per-path XYZ + integer side channel.
template <bool ABSORB, int ANISO>
__global__ __launch_bounds__(kBlock) void film_events_kernel(const FilmParams P,
                                                             const float4* __restrict__ planes,
                                                             float* __restrict__ out_xyz,
                                                             ixg_path_events* __restrict__ ev,
                                                             unsigned long long* __restrict__ totals) {
    __shared__ float4 sp[IXG_MAX_PLANES + 2];  // + the frosted bitmask
    __shared__ unsigned long long block_tot[EV_NUM];
    if (threadIdx.x < EV_NUM) block_tot[threadIdx.x] = 0ull;
    stage_planes(P, planes, sp);
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    uint32_t counts[EV_NUM] = {};
    if (i < P.n_paths) {
        uint64_t p = P.first_path + i;
        uint32_t pixel = (uint32_t)(p / P.spp);
        uint32_t sample = P.sample_offset + (uint32_t)(p % P.spp);
        V3 org, dir;
        uint32_t seed;
        float hero;
        camera_sample(P, pixel, sample, org, dir, seed, hero);
        PathOut o;
        trace_path<true, ABSORB, ANISO>(P, sp, org, dir, seed, hero, o, counts);
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
template <bool ABSORB, int ANISO>
__global__ __launch_bounds__(kBlock) void film_rays_kernel(const FilmParams P,
                                                           const float4* __restrict__ planes,
                                                           const float* __restrict__ rays,
                                                           uint32_t n_rays,
                                                           float* __restrict__ out_xyz,
                                                           ixg_path_events* __restrict__ ev) {
    __shared__ float4 sp[IXG_MAX_PLANES + 2];  // + the frosted bitmask
    stage_planes(P, planes, sp);
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_rays) return;
    const float* r = rays + 8 * (size_t)i;
    PathOut o;
    trace_path<true, ABSORB, ANISO>(P, sp, v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), __float_as_uint(r[7]),
                             r[6], o, nullptr);
    out_xyz[3 * i + 0] = o.xyz[0];
    out_xyz[3 * i + 1] = o.xyz[1];
    out_xyz[3 * i + 2] = o.xyz[2];
    if (ev) ev[i] = o.ev;
}

// Resident blocks per CU of the lean isotropic persistent kernel (film_lean_iso.hip).
extern "C" int ixg_lean_blocks_per_cu(void);

// Launch one instantiation of the persistent kernel with as many blocks as
// stay resident for it.  grid_blocks is the residency of the lean isotropic
// instantiation (3 blocks per CU) times the CUs; the uniaxial ones fit 2 per
// CU, and blocks beyond residency would only start once the path counter is
// spent.  Residency per CU is a property of the compiled kernel, read once.
template <bool ISO, bool ABSORB, int ANISO, bool EXTRA>
hipError_t launch_persistent(const FilmParams* P, const float4* d_planes, float* d_xyz, uint32_t* d_counter,
                             uint32_t grid_blocks, hipStream_t stream) {
    static const int per_cu = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, film_persistent_kernel<ISO, ABSORB, ANISO, EXTRA>,
                                                         kBlock, 0) != hipSuccess)
            n = 0;
        const int lean = ixg_lean_blocks_per_cu();
        return (n > 0 && lean > 0) ? n * 1024 / lean : 1024;  // this kernel's residency relative to the lean one, x1024
    }();
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, (uint64_t)grid_blocks * (uint32_t)per_cu / 1024u);
    hipLaunchKernelGGL((film_persistent_kernel<ISO, ABSORB, ANISO, EXTRA>), dim3(grid), dim3(kBlock), 0, stream, *P,
                       d_planes, d_xyz, d_counter);
    return hipGetLastError();
}

// The persistent kernel's instantiations for one (uniaxial, extra-features)
// pair, chosen at launch by the environment class and the absorbing flag.
template <int ANISO, bool EXTRA>
hipError_t launch_persistent_subset(const FilmParams* P, const float4* d_planes, float* d_xyz, uint32_t* d_counter,
                                    uint32_t grid_blocks, hipStream_t stream, bool iso) {
    if (iso) {
        if (P->absorb) return launch_persistent<true, true, ANISO, EXTRA>(P, d_planes, d_xyz, d_counter, grid_blocks, stream);
        return launch_persistent<true, false, ANISO, EXTRA>(P, d_planes, d_xyz, d_counter, grid_blocks, stream);
    }
    if (P->absorb) return launch_persistent<false, true, ANISO, EXTRA>(P, d_planes, d_xyz, d_counter, grid_blocks, stream);
    return launch_persistent<false, false, ANISO, EXTRA>(P, d_planes, d_xyz, d_counter, grid_blocks, stream);
}

// The diagnostic kernels (per-path events, explicit rays) for one material class.
template <int ANISO>
hipError_t launch_events_subset(const FilmParams* P, const float4* d_planes, float* d_xyz, ixg_path_events* d_ev,
                                unsigned long long* d_totals, hipStream_t stream) {
    const uint64_t blocks = (P->n_paths + kBlock - 1) / kBlock;
    if (P->absorb)
        hipLaunchKernelGGL((film_events_kernel<true, ANISO>), dim3((uint32_t)blocks), dim3(kBlock), 0, stream, *P,
                           d_planes, d_xyz, d_ev, d_totals);
    else
        hipLaunchKernelGGL((film_events_kernel<false, ANISO>), dim3((uint32_t)blocks), dim3(kBlock), 0, stream, *P,
                           d_planes, d_xyz, d_ev, d_totals);
    return hipGetLastError();
}

template <int ANISO>
hipError_t launch_rays_subset(const FilmParams* P, const float4* d_planes, const float* d_rays, uint32_t n_rays,
                              float* d_xyz, ixg_path_events* d_ev, hipStream_t stream) {
    const uint32_t blocks = (n_rays + kBlock - 1) / kBlock;
    if (P->absorb)
        hipLaunchKernelGGL((film_rays_kernel<true, ANISO>), dim3(blocks), dim3(kBlock), 0, stream, *P, d_planes,
                           d_rays, n_rays, d_xyz, d_ev);
    else
        hipLaunchKernelGGL((film_rays_kernel<false, ANISO>), dim3(blocks), dim3(kBlock), 0, stream, *P, d_planes,
                           d_rays, n_rays, d_xyz, d_ev);
    return hipGetLastError();
}

}  // namespace ixg
