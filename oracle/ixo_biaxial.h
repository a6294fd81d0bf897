/*
 * ixo_biaxial.h -- the biaxial optical indicatrix of the CPU oracle (test
 * infrastructure only; included by ixo.c after ixo_uniaxial.h).
 *
 * A restatement of crates/indicatrix/src/optics/birefringence/biaxial.rs:
 * Fresnel's equation of the wave normals as a quadratic in x = 1/n^2, its
 * uniaxial / isotropic closed forms below the degeneracy tolerance, the
 * D-eigenvector as the sign-aligned sum of the three row-pair cross products
 * of Gamma - x I with the root refined by the exact discriminant
 * reformulation, D -> E through eps^-1, the Poynting direction S ~ E x (k x E),
 * the two-step fixed point of an entry mode's refracted wave normal, and the
 * assigned mode's E field for pleochroic absorption.  glam 0.34 scalar order:
 * dot ((x x' + y y') + z z'), Mat3 * v column-accumulated, mul_add as fmaf.
 *
 * Parity: pinned by the reference's biaxial_reduction.rs properties (uniaxial
 * and isotropic reductions, orthonormal transverse eigenvectors, the
 * eigen-equation residual over a dense sweep for the built-in biaxial
 * materials, walk-off facts) and by closed forms in tests/test_biaxial.py
 * (principal-axis index pairs, coincident roots along the optic axes at the
 * textbook angle V, the fixed point's Snell residual).
 */
#ifndef IXO_BIAXIAL_H
#define IXO_BIAXIAL_H

/* BiaxialIndicatrix (biaxial.rs:19-28): columns ax / ay / az are the alpha,
 * beta, gamma principal directions. */
typedef struct {
    float na, nb, ng;
    v3 ax, ay, az;
} biax_t;

/* BiaxialIndicatrix::from_gamma_axis (biaxial.rs:61-65): gamma along
 * normalize_or_zero(c), alpha / beta its stable orthonormal completion. */
static biax_t biax_from_gamma_axis(float na, float nb, float ng, v3 gamma_axis) {
    biax_t B;
    B.na = na;
    B.nb = nb;
    B.ng = ng;
    B.az = vnormalize_or_zero(gamma_axis);
    stable_basis(B.az, &B.ax, &B.ay);
    return B;
}

/* axes^T v: each component the glam dot of one column with v */
static v3 biax_local(const biax_t* B, v3 v) { return mk(vdot(B->ax, v), vdot(B->ay, v), vdot(B->az, v)); }

/* axes * v (glam Mat3::mul_vec3: x_axis * v.x, + y_axis * v.y, + z_axis * v.z) */
static v3 biax_world(const biax_t* B, v3 v) {
    return vadd(vadd(vmul(B->ax, v.x), vmul(B->ay, v.y)), vmul(B->az, v.z));
}

/* b_coeffs (biaxial.rs:69-75) */
static void biax_b(const biax_t* B, float b[3]) {
    b[0] = 1.0f / (B->na * B->na);
    b[1] = 1.0f / (B->nb * B->nb);
    b[2] = 1.0f / (B->ng * B->ng);
}

/* indices_are_degenerate (biaxial.rs:181-184): closer than sqrt(f32::EPSILON)
 * of their own scale (at least 1). */
static int biax_degenerate(float a, float b) {
    float scale = maxf(maxf(fabsf(a), fabsf(b)), 1.0f);
    return fabsf(a - b) <= sqrtf((float)FLT_EPSILON) * scale;
}

/* uniaxial_wave_indices (biaxial.rs:163-167) */
static void biax_uniaxial_indices(v3 k, v3 c, float n_o, float n_e, float* n_slow, float* n_fast) {
    float theta = acosf(fabsf(clampf(vdot(k, c), -1.0f, 1.0f)));
    float n_eff = effective_extraordinary_index(n_o, n_e, theta);
    *n_slow = maxf(n_o, n_eff);
    *n_fast = minf(n_o, n_eff);
}

/* wave_indices (biaxial.rs:108-156): (n_slow, n_fast), n_slow >= n_fast. */
static void biax_wave_indices(const biax_t* B, v3 wave_normal, float* n_slow, float* n_fast) {
    v3 k = vnormalize_or_zero(wave_normal);
    int ab = biax_degenerate(B->na, B->nb), bg = biax_degenerate(B->nb, B->ng);
    if (ab && bg) {
        float n = ((B->na + B->nb) + B->ng) / 3.0f;
        *n_slow = n;
        *n_fast = n;
        return;
    }
    if (ab) {
        biax_uniaxial_indices(k, B->az, midpointf(B->na, B->nb), B->ng, n_slow, n_fast);
        return;
    }
    if (bg) {
        biax_uniaxial_indices(k, B->ax, midpointf(B->nb, B->ng), B->na, n_slow, n_fast);
        return;
    }
    v3 l = biax_local(B, k);
    float a2 = l.x * l.x, b2 = l.y * l.y, g2 = l.z * l.z;
    float bc[3];
    biax_b(B, bc);
    float big_b = fmaf(a2, bc[1] + bc[2], fmaf(b2, bc[0] + bc[2], g2 * (bc[0] + bc[1])));
    float big_c = fmaf(a2, bc[1] * bc[2], fmaf(b2, bc[0] * bc[2], (g2 * bc[0]) * bc[1]));
    float disc = sqrtf(maxf(fmaf(big_b, big_b, -4.0f * big_c), 0.0f));
    float x_lo = 0.5f * (big_b - disc);
    float x_hi = midpointf(big_b, disc);
    *n_slow = 1.0f / sqrtf(maxf(x_lo, 1e-12f));
    *n_fast = 1.0f / sqrtf(maxf(x_hi, 1e-12f));
}

/* precise_root_near (biaxial.rs:296-325): the root of the exactly
 * reformulated discriminant nearest x. */
static float biax_precise_root_near(v3 l, const float b[3], float x) {
    float a = l.x * l.x, bb = l.y * l.y, cc = l.z * l.z;
    float big_b = fmaf(a, b[1] + b[2], fmaf(bb, b[0] + b[2], cc * (b[0] + b[1])));
    float xdiff = b[0] - b[1];
    float ydiff = b[2] - b[0];
    float a_plus_c = a + cc;
    float a_plus_bb = a + bb;
    float two_a_minus_bc = 2.0f * fmaf(bb, -cc, a);
    float disc_sq = fmaf(a_plus_c * a_plus_c, xdiff * xdiff,
                         fmaf(a_plus_bb * a_plus_bb, ydiff * ydiff, (two_a_minus_bc * xdiff) * ydiff));
    float disc = sqrtf(maxf(disc_sq, 0.0f));
    float x_lo = 0.5f * (big_b - disc);
    float x_hi = midpointf(big_b, disc);
    return fabsf(x - x_lo) <= fabsf(x - x_hi) ? x_lo : x_hi;
}

/* cross_fma / dot_fma (birefringence/helpers.rs) */
static v3 cross_fma(v3 a, v3 b) {
    return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
static float dot_fma(v3 a, v3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

/* canonicalize_eigenvector_sign (helpers.rs): largest-magnitude component >= 0 */
static v3 canonical_sign(v3 v) {
    float ax = fabsf(v.x), ay = fabsf(v.y), az = fabsf(v.z);
    float largest = (ax >= ay && ax >= az) ? v.x : (ay >= az ? v.y : v.z);
    return largest < 0.0f ? vneg(v) : v;
}

/* eigenvector_world (biaxial.rs:244-288) for the root x at local = axes^T k */
static v3 biax_eigenvector(const biax_t* B, v3 l, const float b[3], float x_in, v3 k_hat) {
    float x = biax_precise_root_near(l, b, x_in);
    float s = fmaf(b[0], l.x * l.x, fmaf(b[1], l.y * l.y, (b[2] * l.z) * l.z));
    float m00 = fmaf(l.x * l.x, fmaf(2.0f, -b[0], s), b[0] - x);
    float m11 = fmaf(l.y * l.y, fmaf(2.0f, -b[1], s), b[1] - x);
    float m22 = fmaf(l.z * l.z, fmaf(2.0f, -b[2], s), b[2] - x);
    float m01 = (l.x * l.y) * ((s - b[0]) - b[1]);
    float m02 = (l.x * l.z) * ((s - b[0]) - b[2]);
    float m12 = (l.y * l.z) * ((s - b[1]) - b[2]);
    v3 r0 = mk(m00, m01, m02), r1 = mk(m01, m11, m12), r2 = mk(m02, m12, m22);
    v3 c01 = cross_fma(r0, r1), c02 = cross_fma(r0, r2), c12 = cross_fma(r1, r2);
    float s02 = dot_fma(c01, c02) < 0.0f ? -1.0f : 1.0f;
    float s12 = dot_fma(c01, c12) < 0.0f ? -1.0f : 1.0f;
    v3 vl = vadd(vadd(c01, vmul(c02, s02)), vmul(c12, s12));
    v3 vw = biax_world(B, vl);
    if (vlen_sq(vw) > 1e-12f) return vnormalize(canonical_sign(vw));
    v3 t, bb;
    stable_basis(k_hat, &t, &bb);
    return t;
}

/* eigen_polarizations (biaxial.rs:197-208), one root: want_slow picks the
 * n_slow D-eigenvector, else the n_fast one. */
static v3 biax_eigen_d(const biax_t* B, v3 wave_normal, int want_slow) {
    v3 k = vnormalize_or_zero(wave_normal);
    v3 l = biax_local(B, k);
    float b[3];
    biax_b(B, b);
    float ns, nf;
    biax_wave_indices(B, wave_normal, &ns, &nf);
    float n = want_slow ? ns : nf;
    return biax_eigenvector(B, l, b, 1.0f / (n * n), k);
}

/* d_to_e_direction (biaxial.rs:333-341): E = eps^-1 D, normalised */
static v3 biax_d_to_e(const biax_t* B, v3 d) {
    v3 dl = biax_local(B, d);
    v3 el = mk(dl.x / (B->na * B->na), dl.y / (B->nb * B->nb), dl.z / (B->ng * B->ng));
    return vnormalize_or_zero(biax_world(B, el));
}

/* poynting_direction (biaxial.rs:463-473): k without its component along E */
static v3 poynting_direction(v3 wave_normal, v3 e_field) {
    v3 k = vnormalize_or_zero(wave_normal);
    v3 e = vnormalize_or_zero(e_field);
    v3 perp = vsub(k, vmul(e, vdot(e, k)));
    float len2 = vlen_sq(perp);
    if (len2 > 1e-10f) {
        float len = sqrtf(len2);
        return mk(perp.x / len, perp.y / len, perp.z / len);
    }
    return k;
}

/* mode_poynting_dir (biaxial.rs:416-421) */
static v3 biax_mode_poynting(const biax_t* B, v3 wave_normal, int want_slow) {
    return poynting_direction(wave_normal, biax_d_to_e(B, biax_eigen_d(B, wave_normal, want_slow)));
}

/* assigned_mode_e_field (biaxial.rs:432-436): mode B (is_extra) = slow */
static v3 biax_assigned_e(const biax_t* B, v3 wave_normal, int is_extra) {
    return biax_d_to_e(B, biax_eigen_d(B, wave_normal, is_extra));
}

/* resolve_entry_mode (biaxial.rs:371-394): two fixed-point steps of the
 * refracted wave normal against the requested root's index, seeded with
 * n_seed; returns the index and writes the direction. */
static float biax_resolve_entry(const biax_t* B, v3 incident, v3 normal, float cos_i, float n_seed, int want_slow,
                                v3* wave_dir_out) {
    float n_guess = n_seed;
    v3 wave_dir = incident;
    for (int it = 0; it < 2; ++it) {
        float eg = 1.0f / n_guess;
        float s2 = eg * eg * fmaf(cos_i, -cos_i, 1.0f);
        if (s2 > 1.0f) break;
        float ct = sqrtf(maxf(1.0f - s2, 0.0f));
        wave_dir = vnormalize(vadd(vmul(incident, eg), vmul(normal, fmaf(eg, cos_i, -ct))));
        float ns, nf;
        biax_wave_indices(B, wave_dir, &ns, &nf);
        n_guess = want_slow ? ns : nf;
    }
    *wave_dir_out = wave_dir;
    return n_guess;
}

#endif
