// uniaxial.h -- the closed-form uniaxial interface solution on the device.
//
// The same operations, in the same order, as the reference's CPU solver
// (paths relative to crates/indicatrix/src/optics/):
//   Cplx / CVec3                      raytracer/uniaxial_fresnel/complex.rs
//   UniaxialFrame, q roots, fields,
//   Poynting flux                     raytracer/uniaxial_fresnel/frame.rs
//   solve4 / solve4_two_rhs           raytracer/uniaxial_fresnel/linalg.rs
//   entry_incidence_frame / solves    raytracer/uniaxial_fresnel/entry.rs
//   internal_solve                    raytracer/uniaxial_fresnel/internal.rs
//   jones_to_mueller, mode_power,
//   azimuth2_in_frame                 raytracer/uniaxial_fresnel/mueller.rs
//   BirefringenceParams               birefringence/uniaxial.rs
// mul_add is __builtin_fmaf, everything else plain f32 with contraction off.
#pragma once
#include "device_common.h"

namespace ixg {

struct Cx {
    float re, im;
};
__device__ __forceinline__ Cx cx(float re, float im) { return Cx{re, im}; }
__device__ __forceinline__ Cx cadd(Cx a, Cx b) { return cx(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ Cx csub(Cx a, Cx b) { return cx(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ Cx cneg0(Cx a) { return csub(cx(0.0f, 0.0f), a); }
__device__ __forceinline__ Cx cmul(Cx a, Cx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__device__ __forceinline__ Cx cscale(Cx a, float s) { return cx(a.re * s, a.im * s); }
__device__ __forceinline__ Cx cconj(Cx a) { return cx(a.re, -a.im); }
__device__ __forceinline__ float cnorm2(Cx a) { return fma_(a.im, a.im, a.re * a.re); }
__device__ __forceinline__ Cx cdiv(Cx a, Cx o) {
    const float d = fmaxf(fma_(o.im, o.im, o.re * o.re), 1e-20f);
    return cx(fma_(a.im, o.im, a.re * o.re) / d, fma_(a.re, -o.im, a.im * o.re) / d);
}
// sqrt_forward_branch: Re >= 0, a pure negative real straight to i*sqrt(-re)
__device__ __forceinline__ Cx csqrt_fwd(Cx z) {
    if (z.im == 0.0f && z.re < 0.0f) return cx(0.0f, sqrt_rn(-z.re));
    const float r = sqrt_rn(sqrt_rn(cnorm2(z)));
    const float th = atan2f(z.im, z.re) * 0.5f;
    Cx o = cx(r * cosf(th), r * sinf(th));
    if (o.re < 0.0f) o = cx(-o.re, -o.im);
    if (fabsf(o.re) < 1e-9f && o.im < 0.0f) o = cx(-o.re, -o.im);
    return o;
}

struct CV3 {
    V3 re, im;
};
__device__ __forceinline__ CV3 cvr(V3 v) { return CV3{v, v3(0.0f, 0.0f, 0.0f)}; }
__device__ __forceinline__ CV3 cv_wave(V3 that, V3 zhat, float tangential, Cx q) {
    return CV3{that * tangential + zhat * q.re, zhat * q.im};
}
__device__ __forceinline__ CV3 cv_add(CV3 a, CV3 b) { return CV3{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ CV3 cv_scale_real(CV3 a, float s) { return CV3{a.re * s, a.im * s}; }
__device__ __forceinline__ CV3 cv_scale_cplx(CV3 a, Cx s) {
    return CV3{a.re * s.re - a.im * s.im, a.re * s.im + a.im * s.re};
}
__device__ __forceinline__ CV3 cv_cross_real(CV3 a, V3 r) { return CV3{cross(a.re, r), cross(a.im, r)}; }
__device__ __forceinline__ CV3 cv_cross(CV3 a, CV3 o) {
    return CV3{cross(a.re, o.re) - cross(a.im, o.im), cross(a.re, o.im) + cross(a.im, o.re)};
}
__device__ __forceinline__ Cx cv_dot_real(CV3 a, V3 r) { return cx(dot(a.re, r), dot(a.im, r)); }
__device__ __forceinline__ CV3 cv_conj(CV3 a) { return CV3{a.re, neg(a.im)}; }

struct UFrame {
    V3 that, zhat, s_axis;
    float cos_i, sin_i, alpha, beta, gamma;
};

__device__ __forceinline__ V3 vdiv(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }

__device__ __forceinline__ UFrame uframe_build(V3 k, V3 normal, V3 c, float cos_i, float sin_i) {
    UFrame F;
    F.zhat = neg(normal);
    if (sin_i > 1e-5f) {
        F.that = normalize_or_zero(vdiv(k + normal * cos_i, sin_i));
    } else {
        const V3 fb = fabsf(F.zhat.x) < 0.9f ? v3(1.0f, 0.0f, 0.0f) : v3(0.0f, 1.0f, 0.0f);
        F.that = normalize_or_zero(fb - F.zhat * dot(F.zhat, fb));
    }
    const V3 s = normalize_or_zero(cross(k, normal));
    F.s_axis = dot(s, s) > 1e-8f ? s : normalize_or_zero(cross(F.zhat, F.that));
    F.cos_i = cos_i;
    F.sin_i = sin_i;
    F.alpha = dot(c, F.that);
    F.beta = dot(c, F.s_axis);
    F.gamma = dot(c, F.zhat);
    return F;
}

struct QRoots {
    Cx qo_plus, qo_minus, qe_plus, qe_minus;
};

__device__ __forceinline__ QRoots q_roots(float n_o, float n_e, const UFrame& F, float tangential) {
    QRoots R;
    const float n_o2 = n_o * n_o;
    const float k2 = tangential * tangential;
    R.qo_plus = csqrt_fwd(cx(n_o2 - k2, 0.0f));
    R.qo_minus = cneg0(R.qo_plus);
    if (n_o == n_e) {
        R.qe_plus = R.qo_plus;
        R.qe_minus = R.qo_minus;
        return R;
    }
    const float n_e2 = n_e * n_e;
    const float deleps = n_e2 - n_o2;
    const float denom = fma_(F.gamma * F.gamma, deleps, n_o2);
    const float d_val = n_o2 * fma_(fma_(F.beta * F.beta, -deleps, n_e2), -k2, n_e2 * fma_(F.gamma * F.gamma, deleps, n_o2));
    const Cx sqrt_d = csqrt_fwd(cx(d_val, 0.0f));
    const Cx shift = cx(F.alpha * F.gamma * tangential * deleps, 0.0f);
    const Cx denom_c = cx(fmaxf(denom, 1e-12f), 0.0f);
    R.qe_plus = cdiv(csub(sqrt_d, shift), denom_c);
    R.qe_minus = cdiv(csub(cneg0(sqrt_d), shift), denom_c);
    return R;
}

struct MFields {
    CV3 e, h;
};

__device__ __forceinline__ MFields ordinary_fields(float n_o, V3 c, const UFrame& F, float tangential, Cx q) {
    const CV3 k = cv_wave(F.that, F.zhat, tangential, q);
    MFields f;
    f.e = cv_scale_real(cv_cross_real(k, c), 1.0f / (n_o * n_o));
    f.h = cv_cross(k, f.e);
    return f;
}

__device__ __forceinline__ MFields extraordinary_fields(float n_o, float n_e, V3 c, const UFrame& F, float tangential,
                                                        Cx q) {
    const float n_o2 = n_o * n_o, n_e2 = n_e * n_e;
    const float deleps = n_e2 - n_o2;
    const CV3 k = cv_wave(F.that, F.zhat, tangential, q);
    const CV3 d_o = cv_cross_real(k, c);
    const CV3 d_e = cv_cross(k, d_o);
    const Cx c_dot_de = cv_dot_real(d_e, c);
    MFields f;
    f.e = cv_add(cv_scale_real(d_e, 1.0f / n_o2), cv_scale_cplx(cvr(c), cscale(c_dot_de, -(deleps / (n_o2 * n_e2)))));
    f.h = cv_cross(k, f.e);
    return f;
}

__device__ __forceinline__ float poynting_z(const MFields& f, V3 zhat) {
    const CV3 e = cv_scale_cplx(f.e, cx(1.0f, 0.0f));
    const CV3 h = cv_scale_cplx(f.h, cx(1.0f, 0.0f));
    return 0.5f * cv_dot_real(cv_cross(e, cv_conj(h)), zhat).re;
}

__device__ __forceinline__ void tangential(const MFields& f, const UFrame& F, Cx out[4]) {
    out[0] = cv_dot_real(f.e, F.that);
    out[1] = cv_dot_real(f.e, F.s_axis);
    out[2] = cv_dot_real(f.h, F.that);
    out[3] = cv_dot_real(f.h, F.s_axis);
}

// Gauss-Jordan with partial pivoting on |a|^2 (linalg.rs), NRHS right-hand sides.
template <int NRHS>
__device__ __forceinline__ void solve4(Cx a[4][4], Cx b[NRHS][4]) {
#pragma unroll
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        float piv_mag = cnorm2(a[col][col]);
#pragma unroll
        for (int row = col + 1; row < 4; ++row) {
            const float mag = cnorm2(a[row][col]);
            if (mag > piv_mag) {
                piv = row;
                piv_mag = mag;
            }
        }
        // row swap by predicated selects (no dynamic register indexing)
#pragma unroll
        for (int row = col + 1; row < 4; ++row) {
            if (row == piv) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const Cx t = a[col][j];
                    a[col][j] = a[row][j];
                    a[row][j] = t;
                }
#pragma unroll
                for (int r = 0; r < NRHS; ++r) {
                    const Cx t = b[r][col];
                    b[r][col] = b[r][row];
                    b[r][row] = t;
                }
            }
        }
        const Cx pivot = a[col][col];
#pragma unroll
        for (int j = col; j < 4; ++j) a[col][j] = cdiv(a[col][j], pivot);
#pragma unroll
        for (int r = 0; r < NRHS; ++r) b[r][col] = cdiv(b[r][col], pivot);
#pragma unroll
        for (int row = 0; row < 4; ++row) {
            if (row == col) continue;
            const Cx factor = a[row][col];
            if (cnorm2(factor) == 0.0f) continue;
#pragma unroll
            for (int j = col; j < 4; ++j) a[row][j] = csub(a[row][j], cmul(factor, a[col][j]));
#pragma unroll
            for (int r = 0; r < NRHS; ++r) b[r][row] = csub(b[r][row], cmul(factor, b[r][col]));
        }
    }
}

// The air side of an entry (entry.rs entry_incidence_frame), n1 = 1.
struct EntryInc {
    Cx inc_s[4], inc_p[4], refl_s[4], refl_p[4];
};

__device__ __forceinline__ void entry_incidence(const UFrame& F, EntryInc& I) {
    const float n1 = 1.0f;
    const float k_tan = n1 * F.sin_i;
    const CV3 k_fwd = cv_wave(F.that, F.zhat, k_tan, cx(n1 * F.cos_i, 0.0f));
    const CV3 k_bwd = cv_wave(F.that, F.zhat, k_tan, cx(-n1 * F.cos_i, 0.0f));
    const CV3 s_c = cvr(F.s_axis);
    const CV3 ep_bwd = cv_scale_real(cv_cross(k_bwd, s_c), -1.0f / (n1 * n1));
    const CV3 ep_fwd = cv_scale_real(cv_cross(k_fwd, s_c), -1.0f / (n1 * n1));
    tangential(MFields{s_c, cv_cross(k_fwd, s_c)}, F, I.inc_s);
    tangential(MFields{ep_fwd, cv_cross(k_fwd, ep_fwd)}, F, I.inc_p);
    tangential(MFields{s_c, cv_cross(k_bwd, s_c)}, F, I.refl_s);
    tangential(MFields{ep_bwd, cv_cross(k_bwd, ep_bwd)}, F, I.refl_p);
}

// entry_solve_pair_with_incidence: x[0] for s incidence, x[1] for p, each
// {r_s, r_p, t_o, t_e}.
struct EntrySol {
    Cx x[2][4];
    float flux_o, flux_e;
    V3 o_hat, e_hat;
};

__device__ __forceinline__ void entry_solve(const EntryInc& I, float n_o, float n_e, V3 c, const UFrame& F,
                                            EntrySol& out, bool want_axes) {
    const float k_tan = 1.0f * F.sin_i;
    const QRoots R = q_roots(n_o, n_e, F, k_tan);
    const MFields fo = ordinary_fields(n_o, c, F, k_tan, R.qo_plus);
    const MFields fe = extraordinary_fields(n_o, n_e, c, F, k_tan, R.qe_plus);
    Cx oc[4], ec[4];
    tangential(fo, F, oc);
    tangential(fe, F, ec);
    Cx a[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        a[i][0] = I.refl_s[i];
        a[i][1] = I.refl_p[i];
        a[i][2] = cneg0(oc[i]);
        a[i][3] = cneg0(ec[i]);
        out.x[0][i] = cneg0(I.inc_s[i]);
        out.x[1][i] = cneg0(I.inc_p[i]);
    }
    solve4<2>(a, out.x);
    out.flux_o = fabsf(poynting_z(fo, F.zhat));
    out.flux_e = fabsf(poynting_z(fe, F.zhat));
    if (want_axes) {
        out.o_hat = normalize_or_zero(fo.e.re);
        out.e_hat = normalize_or_zero(fe.e.re);
    }
}

struct InternalSol {
    Cx r_o, r_e, t_s, t_p;
    float flux_ro, flux_re, flux_ts, flux_tp, flux_inc;
};

__device__ __forceinline__ InternalSol internal_solve(float n_mode_inc, float n_o, float n_e, V3 c, const UFrame& F,
                                                      bool incident_is_ordinary) {
    const float k_tan = n_mode_inc * F.sin_i;
    const QRoots R = q_roots(n_o, n_e, F, k_tan);
    const MFields f_inc = incident_is_ordinary ? ordinary_fields(n_o, c, F, k_tan, R.qo_plus)
                                               : extraordinary_fields(n_o, n_e, c, F, k_tan, R.qe_plus);
    const MFields f_o = ordinary_fields(n_o, c, F, k_tan, R.qo_minus);
    const MFields f_e = extraordinary_fields(n_o, n_e, c, F, k_tan, R.qe_minus);
    const float n2 = 1.0f;
    const float zeta = k_tan / n2;
    const Cx cos_t = fabsf(zeta) <= 1.0f ? cx(sqrt_rn(fmaxf(fma_(zeta, -zeta, 1.0f), 0.0f)), 0.0f)
                                         : cx(0.0f, sqrt_rn(fma_(zeta, zeta, -1.0f)));
    const CV3 k_iso = cv_wave(F.that, F.zhat, k_tan, cscale(cos_t, n2));
    const CV3 s_c = cvr(F.s_axis);
    const CV3 ep_t = cv_scale_real(cv_cross(k_iso, s_c), -1.0f / (n2 * n2));
    const MFields f_s{s_c, cv_cross(k_iso, s_c)};
    const MFields f_p{ep_t, cv_cross(k_iso, ep_t)};
    Cx inc_c[4], eo[4], ee[4], es[4], ep[4];
    tangential(f_inc, F, inc_c);
    tangential(f_o, F, eo);
    tangential(f_e, F, ee);
    tangential(f_s, F, es);
    tangential(f_p, F, ep);
    Cx a[4][4], b[1][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        a[i][0] = eo[i];
        a[i][1] = ee[i];
        a[i][2] = cneg0(es[i]);
        a[i][3] = cneg0(ep[i]);
        b[0][i] = cneg0(inc_c[i]);
    }
    solve4<1>(a, b);
    InternalSol S;
    S.r_o = b[0][0];
    S.r_e = b[0][1];
    S.t_s = b[0][2];
    S.t_p = b[0][3];
    S.flux_ro = fabsf(poynting_z(f_o, F.zhat));
    S.flux_re = fabsf(poynting_z(f_e, F.zhat));
    S.flux_ts = fabsf(poynting_z(f_s, F.zhat));
    S.flux_tp = fabsf(poynting_z(f_p, F.zhat));
    S.flux_inc = fabsf(poynting_z(f_inc, F.zhat));
    return S;
}

// jones_to_mueller (mueller.rs), column-major m[col][row], and glam's
// Mat4 * Vec4 column accumulation.
struct M4 {
    float m[4][4];
};

__device__ __forceinline__ void jones_to_mueller(Cx j_ss, Cx j_sp, Cx j_ps, Cx j_pp, M4& M) {
    const float m_ss = cnorm2(j_ss), m_sp = cnorm2(j_sp), m_ps = cnorm2(j_ps), m_pp = cnorm2(j_pp);
    const Cx a_ssp = cmul(j_ss, cconj(j_sp));
    const Cx a_pep = cmul(j_ps, cconj(j_pp));
    const Cx a_plus = cadd(a_ssp, a_pep), a_minus = csub(a_ssp, a_pep);
    const Cx b_sp = cmul(j_ss, cconj(j_ps));
    const Cx b_pp = cmul(j_sp, cconj(j_pp));
    const Cx c_sspp = cmul(j_ss, cconj(j_pp));
    const Cx c_spps = cmul(j_sp, cconj(j_ps));
    M.m[0][0] = 0.5f * (((m_ss + m_ps) + m_sp) + m_pp);
    M.m[0][1] = 0.5f * (((m_ss - m_ps) + m_sp) - m_pp);
    M.m[0][2] = b_sp.re + b_pp.re;
    M.m[0][3] = -(b_sp.im + b_pp.im);
    M.m[1][0] = 0.5f * (((m_ss + m_ps) - m_sp) - m_pp);
    M.m[1][1] = midpoint((m_ss - m_ps) - m_sp, m_pp);
    M.m[1][2] = b_sp.re - b_pp.re;
    M.m[1][3] = -(b_sp.im) + b_pp.im;
    M.m[2][0] = a_plus.re;
    M.m[2][1] = a_minus.re;
    M.m[2][2] = c_sspp.re + c_spps.re;
    M.m[2][3] = -(c_sspp.im + c_spps.im);
    M.m[3][0] = a_plus.im;
    M.m[3][1] = a_minus.im;
    M.m[3][2] = c_sspp.im - c_spps.im;
    M.m[3][3] = c_sspp.re - c_spps.re;
}

__device__ __forceinline__ float4 mat_apply(const M4& M, float4 s) {
    float r0 = M.m[0][0] * s.x, r1 = M.m[0][1] * s.x, r2 = M.m[0][2] * s.x, r3 = M.m[0][3] * s.x;
    r0 = r0 + M.m[1][0] * s.y;
    r1 = r1 + M.m[1][1] * s.y;
    r2 = r2 + M.m[1][2] * s.y;
    r3 = r3 + M.m[1][3] * s.y;
    r0 = r0 + M.m[2][0] * s.z;
    r1 = r1 + M.m[2][1] * s.z;
    r2 = r2 + M.m[2][2] * s.z;
    r3 = r3 + M.m[2][3] * s.z;
    r0 = r0 + M.m[3][0] * s.w;
    r1 = r1 + M.m[3][1] * s.w;
    r2 = r2 + M.m[3][2] * s.w;
    r3 = r3 + M.m[3][3] * s.w;
    return make_float4(r0, r1, r2, r3);
}

__device__ __forceinline__ float mode_power(Cx t_s, Cx t_p, float mode_flux, float inc_flux, float4 st) {
    const float m_s = cnorm2(t_s), m_p = cnorm2(t_p);
    const Cx cr = cmul(t_s, cconj(t_p));
    const float raw = fma_(cr.im, st.w, fma_(cr.re, st.z, 0.5f * fma_(m_s, st.x + st.y, m_p * (st.x - st.y))));
    return (mode_flux / fmaxf(inc_flux, 1e-12f)) * raw;
}

__device__ __forceinline__ void azimuth2_in_frame(V3 dir, V3 s_axis, V3 p_axis, float& c2, float& s2) {
    const float s = dot(dir, s_axis), p = dot(dir, p_axis);
    const float norm = fmaxf(fma_(p, p, s * s), 1e-12f);
    c2 = fma_(p, -p, s * s) / norm;
    s2 = 2.0f * s * p / norm;
}

// ---- BirefringenceParams (birefringence/uniaxial.rs) ----
__device__ __forceinline__ float effective_extraordinary_index(float n_o, float n_e, float theta) {
    if (fabsf(n_o - n_e) < 1e-5f) return n_o;
    const float st = sinf(theta), ct = cosf(theta);
    const float nos = n_o * st;
    const float denom_sq = fma_(n_e * ct, n_e * ct, nos * nos);
    if (denom_sq <= 1e-8f) return n_o;
    return (n_o * n_e) / sqrt_rn(denom_sq);
}

__device__ __forceinline__ V3 extraordinary_poynting_dir(V3 w, V3 c, float n_o, float n_e) {
    const float cos_theta = clampf(dot(w, c), -1.0f, 1.0f);
    const float ac = fabsf(cos_theta);
    const float st = sqrt_rn(fmaxf((1.0f - ac) * (1.0f + ac), 0.0f));
    if (fabsf(n_o - n_e) < 1e-5f) return w;
    const float n_o2 = n_o * n_o, n_e2 = n_e * n_e;
    const float tan_rho = ((n_o2 - n_e2) * st * ac) / fmaxf(fma_(n_e2 * ac, ac, n_o2 * st * st), 1e-6f);
    const float delta = atanf(tan_rho);
    if (fabsf(delta) < 1e-5f) return w;
    const V3 c_proj = normalize_or_zero(c - w * cos_theta);
    if (dot(c_proj, c_proj) < 1e-6f) return w;
    const float sign = cos_theta >= 0.0f ? 1.0f : -1.0f;
    return normalize(w * cosf(delta) - (c_proj * sign) * sinf(delta));
}

// BirefringenceParams::ordinary_eigen_polarization
__device__ __forceinline__ V3 ordinary_eigen(V3 k, V3 c) {
    const V3 x = cross(k, c);
    if (dot(x, x) > 1e-8f) return normalize(x);
    const V3 n = normalize_or_zero(k);
    const V3 a = fabsf(n.x) > 0.9f ? v3(0.0f, 1.0f, 0.0f) : v3(1.0f, 0.0f, 0.0f);
    return normalize_or_zero(a - n * dot(n, a));
}

}  // namespace ixg
