// biaxial.h -- the biaxial optical indicatrix on the device
// (crates/indicatrix/src/optics/birefringence/biaxial.rs): Fresnel's equation
// of the wave normals as a quadratic in x = 1/n^2 with its uniaxial and
// isotropic closed forms below the degeneracy tolerance, the D-eigenvector as
// the sign-aligned sum of the row-pair cross products of Gamma - x I (root
// refined through the exact discriminant reformulation), D -> E through
// eps^-1, the Poynting direction, and the two-step fixed point of an entry
// mode's refracted wave normal.
//
// Operation order as the reference (glam scalar Vec3, mul_add as fma, the
// Mat3 products column-accumulated); every division and square root is the
// correctly rounded one, so the general branch -- no transcendental -- gives
// the CPU's bits; the degenerate branches carry the device acos / sin / cos.
#pragma once
#include "device_common.h"
#include "uniaxial.h"

namespace ixg {

// BiaxialIndicatrix: principal indices and the principal frame (columns
// alpha, beta, gamma).  The film's materials share one frame -- gamma along
// the normalised optic axis, alpha / beta its stable orthonormal completion
// (from_gamma_axis, biaxial.rs:61-65) -- and differ per channel only in the
// indices (GemMaterial::biaxial_indicatrix, materials/optics.rs:16-27).
struct Biax {
    float na, nb, ng;
};
struct BiaxFrame {
    V3 ax, ay, az;
};

__device__ __forceinline__ V3 biax_local(const BiaxFrame& F, V3 v) { return v3(dot(F.ax, v), dot(F.ay, v), dot(F.az, v)); }
__device__ __forceinline__ V3 biax_world(const BiaxFrame& F, V3 v) { return (F.ax * v.x + F.ay * v.y) + F.az * v.z; }

// indices_are_degenerate (biaxial.rs:181-184); sqrt(f32::EPSILON) rounded to f32
__device__ __forceinline__ bool biax_degenerate(float a, float b) {
    const float scale = fmaxf(fmaxf(fabsf(a), fabsf(b)), 1.0f);
    return fabsf(a - b) <= 3.45266977e-4f * scale;
}

// uniaxial_wave_indices (biaxial.rs:163-167)
__device__ __forceinline__ void biax_uniaxial_indices(V3 k, V3 c, float n_o, float n_e, float& n_slow,
                                                      float& n_fast) {
    const float theta = acosf(fabsf(clampf(dot(k, c), -1.0f, 1.0f)));
    const float n_eff = effective_extraordinary_index(n_o, n_e, theta);
    n_slow = fmaxf(n_o, n_eff);
    n_fast = fminf(n_o, n_eff);
}

// wave_indices (biaxial.rs:108-156): (n_slow, n_fast)
__device__ __forceinline__ void biax_wave_indices(const Biax& B, const BiaxFrame& F, V3 wave_normal, float& n_slow,
                                                  float& n_fast) {
    const V3 k = normalize_or_zero(wave_normal);
    const bool ab = biax_degenerate(B.na, B.nb), bg = biax_degenerate(B.nb, B.ng);
    if (ab && bg) {
        n_slow = n_fast = div_rn((B.na + B.nb) + B.ng, 3.0f);
        return;
    }
    if (ab) {
        biax_uniaxial_indices(k, F.az, midpoint(B.na, B.nb), B.ng, n_slow, n_fast);
        return;
    }
    if (bg) {
        biax_uniaxial_indices(k, F.ax, midpoint(B.nb, B.ng), B.na, n_slow, n_fast);
        return;
    }
    const V3 l = biax_local(F, k);
    const float a2 = l.x * l.x, b2 = l.y * l.y, g2 = l.z * l.z;
    const float b0 = div_rn(1.0f, B.na * B.na), b1 = div_rn(1.0f, B.nb * B.nb), b3 = div_rn(1.0f, B.ng * B.ng);
    const float big_b = fma_(a2, b1 + b3, fma_(b2, b0 + b3, g2 * (b0 + b1)));
    const float big_c = fma_(a2, b1 * b3, fma_(b2, b0 * b3, (g2 * b0) * b1));
    const float disc = sqrt_rn(fmaxf(fma_(big_b, big_b, -4.0f * big_c), 0.0f));
    const float x_lo = 0.5f * (big_b - disc);
    const float x_hi = midpoint(big_b, disc);
    n_slow = div_rn(1.0f, sqrt_rn(fmaxf(x_lo, 1e-12f)));
    n_fast = div_rn(1.0f, sqrt_rn(fmaxf(x_hi, 1e-12f)));
}

__device__ __forceinline__ V3 cross_fma(V3 a, V3 b) {
    return v3(fma_(a.y, b.z, -(a.z * b.y)), fma_(a.z, b.x, -(a.x * b.z)), fma_(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ float dot_fma(V3 a, V3 b) { return fma_(a.x, b.x, fma_(a.y, b.y, a.z * b.z)); }

// eigen_polarizations (biaxial.rs:197-208) for one root, through
// eigenvector_world (:244-288) and precise_root_near (:296-325):
// want_slow = the n_slow D-eigenvector.
__device__ __noinline__ V3 biax_eigen_d(const Biax& B, const BiaxFrame& F, V3 wave_normal, bool want_slow) {
    const V3 k = normalize_or_zero(wave_normal);
    const V3 l = biax_local(F, k);
    const float b0 = div_rn(1.0f, B.na * B.na), b1 = div_rn(1.0f, B.nb * B.nb), b2 = div_rn(1.0f, B.ng * B.ng);
    float ns, nf;
    biax_wave_indices(B, F, wave_normal, ns, nf);
    const float n = want_slow ? ns : nf;
    const float x_in = div_rn(1.0f, n * n);
    // precise_root_near
    const float a = l.x * l.x, bb = l.y * l.y, cc = l.z * l.z;
    const float big_b = fma_(a, b1 + b2, fma_(bb, b0 + b2, cc * (b0 + b1)));
    const float xd = b0 - b1, yd = b2 - b0;
    const float apc = a + cc, apb = a + bb;
    const float tab = 2.0f * fma_(bb, -cc, a);
    const float disc_sq = fma_(apc * apc, xd * xd, fma_(apb * apb, yd * yd, (tab * xd) * yd));
    const float disc = sqrt_rn(fmaxf(disc_sq, 0.0f));
    const float x_lo = 0.5f * (big_b - disc), x_hi = midpoint(big_b, disc);
    const float x = fabsf(x_in - x_lo) <= fabsf(x_in - x_hi) ? x_lo : x_hi;
    // eigenvector_world
    const float s = fma_(b0, l.x * l.x, fma_(b1, l.y * l.y, (b2 * l.z) * l.z));
    const float m00 = fma_(l.x * l.x, fma_(2.0f, -b0, s), b0 - x);
    const float m11 = fma_(l.y * l.y, fma_(2.0f, -b1, s), b1 - x);
    const float m22 = fma_(l.z * l.z, fma_(2.0f, -b2, s), b2 - x);
    const float m01 = (l.x * l.y) * ((s - b0) - b1);
    const float m02 = (l.x * l.z) * ((s - b0) - b2);
    const float m12 = (l.y * l.z) * ((s - b1) - b2);
    const V3 r0 = v3(m00, m01, m02), r1 = v3(m01, m11, m12), r2 = v3(m02, m12, m22);
    const V3 c01 = cross_fma(r0, r1), c02 = cross_fma(r0, r2), c12 = cross_fma(r1, r2);
    const float s02 = dot_fma(c01, c02) < 0.0f ? -1.0f : 1.0f;
    const float s12 = dot_fma(c01, c12) < 0.0f ? -1.0f : 1.0f;
    const V3 vw = biax_world(F, (c01 + c02 * s02) + c12 * s12);
    if (dot(vw, vw) > 1e-12f) {
        // canonicalize_eigenvector_sign (helpers.rs): largest component >= 0
        const float ax = fabsf(vw.x), ay = fabsf(vw.y), az = fabsf(vw.z);
        const float largest = (ax >= ay && ax >= az) ? vw.x : (ay >= az ? vw.y : vw.z);
        return normalize(largest < 0.0f ? neg(vw) : vw);
    }
    // stable_orthonormal_basis(k).0
    const V3 av = fabsf(k.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    return normalize_or_zero(av - k * dot(k, av));
}

// d_to_e_direction (biaxial.rs:333-341)
__device__ __forceinline__ V3 biax_d_to_e(const Biax& B, const BiaxFrame& F, V3 d) {
    const V3 dl = biax_local(F, d);
    const V3 el = v3(div_rn(dl.x, B.na * B.na), div_rn(dl.y, B.nb * B.nb), div_rn(dl.z, B.ng * B.ng));
    return normalize_or_zero(biax_world(F, el));
}

// poynting_direction (biaxial.rs:463-473)
__device__ __forceinline__ V3 poynting_direction(V3 wave_normal, V3 e_field) {
    const V3 k = normalize_or_zero(wave_normal);
    const V3 e = normalize_or_zero(e_field);
    const V3 perp = k - e * dot(e, k);
    const float len2 = dot(perp, perp);
    if (len2 > 1e-10f) {
        const float len = sqrt_rn(len2);
        return v3(div_rn(perp.x, len), div_rn(perp.y, len), div_rn(perp.z, len));
    }
    return k;
}

// mode_poynting_dir (biaxial.rs:416-421)
__device__ __forceinline__ V3 biax_mode_poynting(const Biax& B, const BiaxFrame& F, V3 wave_normal, bool want_slow) {
    return poynting_direction(wave_normal, biax_d_to_e(B, F, biax_eigen_d(B, F, wave_normal, want_slow)));
}

// resolve_entry_mode (biaxial.rs:371-394): the wave normal after two
// fixed-point steps (the index is not needed by the transport).
__device__ __forceinline__ V3 biax_resolve_entry(const Biax& B, const BiaxFrame& F, V3 incident, V3 normal,
                                                 float cos_i, float n_seed, bool want_slow) {
    float n_guess = n_seed;
    V3 wave_dir = incident;
    const float one_m_c2 = fma_(cos_i, -cos_i, 1.0f);
    for (int it = 0; it < 2; ++it) {
        const float eg = div_rn(1.0f, n_guess);
        const float s2 = eg * eg * one_m_c2;
        if (s2 > 1.0f) break;
        const float ct = sqrt_rn(fmaxf(1.0f - s2, 0.0f));
        wave_dir = normalize(incident * eg + normal * fma_(eg, cos_i, -ct));
        float ns, nf;
        biax_wave_indices(B, F, wave_dir, ns, nf);
        n_guess = want_slow ? ns : nf;
    }
    return wave_dir;
}

}  // namespace ixg
