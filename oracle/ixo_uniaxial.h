/* ixo_uniaxial.h -- CPU oracle (test infrastructure only): the closed-form
 * uniaxial interface solution and the birefringence helpers the uniaxial
 * transport arms use.  Included by ixo.c after its vector / Stokes / Mueller
 * helpers; every function restates the reference operation for operation
 * (mul_add as fmaf, no contraction, glam's left-associated products).
 *
 * Reference map (paths relative to crates/indicatrix/src/optics/):
 *   Cplx / CVec3                raytracer/uniaxial_fresnel/complex.rs
 *   UniaxialFrame, q roots,
 *   mode fields, Poynting flux  raytracer/uniaxial_fresnel/frame.rs
 *   solve4 / solve4_two_rhs     raytracer/uniaxial_fresnel/linalg.rs
 *   entry_* solves              raytracer/uniaxial_fresnel/entry.rs
 *   internal_solve              raytracer/uniaxial_fresnel/internal.rs
 *   jones_to_mueller, mode_power,
 *   azimuth2_in_frame           raytracer/uniaxial_fresnel/mueller.rs
 *   BirefringenceParams         birefringence/uniaxial.rs, birefringence/helpers.rs
 *   assigned_mode_e_field_uniaxial birefringence/absorption_tensor.rs
 */
#ifndef IXO_UNIAXIAL_H
#define IXO_UNIAXIAL_H

/* ---- complex scalars and vectors (complex.rs) ---- */
typedef struct { float re, im; } cplx;
static inline cplx cx(float re, float im) { cplx c = {re, im}; return c; }
static inline cplx cadd(cplx a, cplx b) { return cx(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return cx(a.re - b.re, a.im - b.im); }
static inline cplx cneg0(cplx a) { return csub(cx(0.0f, 0.0f), a); } /* Cplx::ZERO.sub(a) */
static inline cplx cmul(cplx a, cplx b) {
    return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx cscale(cplx a, float s) { return cx(a.re * s, a.im * s); }
static inline cplx cconj(cplx a) { return cx(a.re, -a.im); }
static inline float cnorm2(cplx a) { return fmaf(a.im, a.im, a.re * a.re); }
static inline cplx cdiv(cplx a, cplx o) {
    float d = maxf(fmaf(o.im, o.im, o.re * o.re), 1e-20f);
    return cx(fmaf(a.im, o.im, a.re * o.re) / d, fmaf(a.re, -o.im, a.im * o.re) / d);
}
/* sqrt_forward_branch: the root with Re >= 0 (Im >= 0 on the cut); a pure
 * negative real goes straight to i*sqrt(-re). */
static cplx csqrt_fwd(cplx z) {
    if (z.im == 0.0f && z.re < 0.0f) return cx(0.0f, sqrtf(-z.re));
    float r = sqrtf(sqrtf(cnorm2(z)));
    float th = TF(atan2f(z.im, z.re), z.re) * 0.5f;
    cplx o = cx(r * TF(cosf(th), th + 1.0f), r * TF(sinf(th), th));
    if (o.re < 0.0f) o = cx(-o.re, -o.im);
    if (fabsf(o.re) < 1e-9f && o.im < 0.0f) o = cx(-o.re, -o.im);
    return o;
}

typedef struct { v3 re, im; } cv3;
static inline cv3 cv(v3 re, v3 im) { cv3 c = {re, im}; return c; }
static inline cv3 cv_real(v3 v) { return cv(v, mk(0.0f, 0.0f, 0.0f)); }
static inline cv3 cv_wave(v3 that, v3 zhat, float tangential, cplx q) {
    return cv(vadd(vmul(that, tangential), vmul(zhat, q.re)), vmul(zhat, q.im));
}
static inline cv3 cv_add(cv3 a, cv3 b) { return cv(vadd(a.re, b.re), vadd(a.im, b.im)); }
static inline cv3 cv_scale_real(cv3 a, float s) { return cv(vmul(a.re, s), vmul(a.im, s)); }
static inline cv3 cv_scale_cplx(cv3 a, cplx s) {
    return cv(vsub(vmul(a.re, s.re), vmul(a.im, s.im)), vadd(vmul(a.re, s.im), vmul(a.im, s.re)));
}
static inline cv3 cv_cross_real(cv3 a, v3 r) { return cv(vcross(a.re, r), vcross(a.im, r)); }
static inline cv3 cv_cross(cv3 a, cv3 o) {
    return cv(vsub(vcross(a.re, o.re), vcross(a.im, o.im)), vadd(vcross(a.re, o.im), vcross(a.im, o.re)));
}
static inline cplx cv_dot_real(cv3 a, v3 r) { return cx(vdot(a.re, r), vdot(a.im, r)); }
static inline cv3 cv_conj(cv3 a) { return cv(a.re, vneg(a.im)); }

/* ---- incidence frame (frame.rs UniaxialFrame::build) ---- */
typedef struct {
    v3 that, zhat, s_axis, p_axis;
    float cos_i, sin_i, alpha, beta, gamma;
} uframe;

static inline v3 vdiv(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }

static uframe uframe_build(v3 k, v3 normal, v3 c, float cos_i, float sin_i) {
    uframe F;
    F.zhat = vneg(normal);
    if (sin_i > 1e-5f) {
        F.that = vnormalize_or_zero(vdiv(vadd(k, vmul(normal, cos_i)), sin_i));
    } else {
        v3 fb = fabsf(F.zhat.x) < 0.9f ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f);
        F.that = vnormalize_or_zero(vsub(fb, vmul(F.zhat, vdot(F.zhat, fb))));
    }
    v3 s = vnormalize_or_zero(vcross(k, normal));
    F.s_axis = vlen_sq(s) > 1e-8f ? s : vnormalize_or_zero(vcross(F.zhat, F.that));
    F.p_axis = vcross(k, F.s_axis);
    F.cos_i = cos_i;
    F.sin_i = sin_i;
    F.alpha = vdot(c, F.that);
    F.beta = vdot(c, F.s_axis);
    F.gamma = vdot(c, F.zhat);
    return F;
}

/* uniaxial_q_roots: normal wavevector components of the o and e waves. */
typedef struct { cplx qo_plus, qo_minus, qe_plus, qe_minus; } qroots;

static qroots uniaxial_q_roots(float n_o, float n_e, const uframe* F, float tangential) {
    qroots R;
    float n_o2 = n_o * n_o;
    float k2 = tangential * tangential;
    R.qo_plus = csqrt_fwd(cx(n_o2 - k2, 0.0f));
    R.qo_minus = cneg0(R.qo_plus);
    if (n_o == n_e) {
        R.qe_plus = R.qo_plus;
        R.qe_minus = R.qo_minus;
        return R;
    }
    float n_e2 = n_e * n_e;
    float deleps = n_e2 - n_o2;
    float denom = fmaf(F->gamma * F->gamma, deleps, n_o2);
    float d_val = n_o2 * fmaf(fmaf(F->beta * F->beta, -deleps, n_e2), -k2,
                              n_e2 * fmaf(F->gamma * F->gamma, deleps, n_o2));
    cplx sqrt_d = csqrt_fwd(cx(d_val, 0.0f));
    cplx shift = cx(F->alpha * F->gamma * tangential * deleps, 0.0f);
    cplx denom_c = cx(maxf(denom, 1e-12f), 0.0f);
    R.qe_plus = cdiv(csub(sqrt_d, shift), denom_c);
    R.qe_minus = cdiv(csub(cneg0(sqrt_d), shift), denom_c);
    return R;
}

typedef struct { cv3 k, e, h; } mfields;

static mfields ordinary_mode_fields(float n_o, v3 c, v3 that, v3 zhat, float tangential, cplx q) {
    mfields f;
    f.k = cv_wave(that, zhat, tangential, q);
    cv3 d_o = cv_cross_real(f.k, c);
    f.e = cv_scale_real(d_o, 1.0f / (n_o * n_o));
    f.h = cv_cross(f.k, f.e);
    return f;
}

static mfields extraordinary_mode_fields(float n_o, float n_e, v3 c, v3 that, v3 zhat, float tangential,
                                         cplx q) {
    mfields f;
    float n_o2 = n_o * n_o, n_e2 = n_e * n_e;
    float deleps = n_e2 - n_o2;
    f.k = cv_wave(that, zhat, tangential, q);
    cv3 d_o = cv_cross_real(f.k, c);
    cv3 d_e = cv_cross(f.k, d_o);
    cplx c_dot_de = cv_dot_real(d_e, c);
    f.e = cv_add(cv_scale_real(d_e, 1.0f / n_o2),
                 cv_scale_cplx(cv_real(c), cscale(c_dot_de, -(deleps / (n_o2 * n_e2)))));
    f.h = cv_cross(f.k, f.e);
    return f;
}

/* 0.5 Re((E x H*) . zhat) of a unit-amplitude mode. */
static float poynting_z(const mfields* f, v3 zhat) {
    cv3 e = cv_scale_cplx(f->e, cx(1.0f, 0.0f));
    cv3 h = cv_scale_cplx(f->h, cx(1.0f, 0.0f));
    cv3 s = cv_cross(e, cv_conj(h));
    return 0.5f * cv_dot_real(s, zhat).re;
}

static void tangential_components(const mfields* f, v3 that, v3 s_axis, cplx out[4]) {
    out[0] = cv_dot_real(f->e, that);
    out[1] = cv_dot_real(f->e, s_axis);
    out[2] = cv_dot_real(f->h, that);
    out[3] = cv_dot_real(f->h, s_axis);
}

/* linalg.rs: Gauss-Jordan with partial pivoting on |a|^2, one or two RHS. */
static void solve4_n(cplx a[4][4], cplx* b1, cplx* b2) {
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        float piv_mag = cnorm2(a[col][col]);
        for (int row = col + 1; row < 4; ++row) {
            float mag = cnorm2(a[row][col]);
            if (mag > piv_mag) {
                piv = row;
                piv_mag = mag;
            }
        }
        if (piv != col) {
            for (int j = 0; j < 4; ++j) {
                cplx t = a[col][j];
                a[col][j] = a[piv][j];
                a[piv][j] = t;
            }
            cplx t = b1[col];
            b1[col] = b1[piv];
            b1[piv] = t;
            if (b2) {
                t = b2[col];
                b2[col] = b2[piv];
                b2[piv] = t;
            }
        }
        cplx pivot = a[col][col];
        for (int j = col; j < 4; ++j) a[col][j] = cdiv(a[col][j], pivot);
        b1[col] = cdiv(b1[col], pivot);
        if (b2) b2[col] = cdiv(b2[col], pivot);
        for (int row = 0; row < 4; ++row) {
            if (row == col) continue;
            cplx factor = a[row][col];
            if (cnorm2(factor) == 0.0f) continue;
            for (int j = col; j < 4; ++j) a[row][j] = csub(a[row][j], cmul(factor, a[col][j]));
            b1[row] = csub(b1[row], cmul(factor, b1[col]));
            if (b2) b2[row] = csub(b2[row], cmul(factor, b2[col]));
        }
    }
}

/* entry.rs: the isotropic (incidence) side, shared by all channels. */
typedef struct { cplx inc_s[4], inc_p[4], refl_s[4], refl_p[4]; } entry_inc;

static entry_inc entry_incidence_frame(float n1, const uframe* F) {
    entry_inc I;
    float k_tan = n1 * F->sin_i;
    cv3 k_fwd = cv_wave(F->that, F->zhat, k_tan, cx(n1 * F->cos_i, 0.0f));
    cv3 k_bwd = cv_wave(F->that, F->zhat, k_tan, cx(-n1 * F->cos_i, 0.0f));
    cv3 s_c = cv_real(F->s_axis);
    cv3 ep_bwd = cv_scale_real(cv_cross(k_bwd, s_c), -1.0f / (n1 * n1));
    cv3 ep_fwd = cv_scale_real(cv_cross(k_fwd, s_c), -1.0f / (n1 * n1));
    mfields inc_s = {k_fwd, s_c, cv_cross(k_fwd, s_c)};
    mfields inc_p = {k_fwd, ep_fwd, cv_cross(k_fwd, ep_fwd)};
    mfields refl_s = {k_bwd, s_c, cv_cross(k_bwd, s_c)};
    mfields refl_p = {k_bwd, ep_bwd, cv_cross(k_bwd, ep_bwd)};
    tangential_components(&inc_s, F->that, F->s_axis, I.inc_s);
    tangential_components(&inc_p, F->that, F->s_axis, I.inc_p);
    tangential_components(&refl_s, F->that, F->s_axis, I.refl_s);
    tangential_components(&refl_p, F->that, F->s_axis, I.refl_p);
    return I;
}

typedef struct {
    cplx r_s, r_p, t_o, t_e;
    float flux_o, flux_e;
    v3 o_hat, e_hat;
} entry_sol;

static void entry_solve_pair(const entry_inc* I, float n1, float n_o, float n_e, v3 c, const uframe* F,
                             entry_sol* s_sol, entry_sol* p_sol) {
    float k_tan = n1 * F->sin_i;
    qroots R = uniaxial_q_roots(n_o, n_e, F, k_tan);
    mfields fo = ordinary_mode_fields(n_o, c, F->that, F->zhat, k_tan, R.qo_plus);
    mfields fe = extraordinary_mode_fields(n_o, n_e, c, F->that, F->zhat, k_tan, R.qe_plus);
    cplx oc[4], ec[4];
    tangential_components(&fo, F->that, F->s_axis, oc);
    tangential_components(&fe, F->that, F->s_axis, ec);
    cplx a[4][4], bs[4], bp[4];
    for (int i = 0; i < 4; ++i) {
        a[i][0] = I->refl_s[i];
        a[i][1] = I->refl_p[i];
        a[i][2] = cneg0(oc[i]);
        a[i][3] = cneg0(ec[i]);
        bs[i] = cneg0(I->inc_s[i]);
        bp[i] = cneg0(I->inc_p[i]);
    }
    solve4_n(a, bs, bp);
    float flux_o = fabsf(poynting_z(&fo, F->zhat));
    float flux_e = fabsf(poynting_z(&fe, F->zhat));
    v3 o_hat = vnormalize_or_zero(fo.e.re), e_hat = vnormalize_or_zero(fe.e.re);
    s_sol->r_s = bs[0]; s_sol->r_p = bs[1]; s_sol->t_o = bs[2]; s_sol->t_e = bs[3];
    p_sol->r_s = bp[0]; p_sol->r_p = bp[1]; p_sol->t_o = bp[2]; p_sol->t_e = bp[3];
    s_sol->flux_o = p_sol->flux_o = flux_o;
    s_sol->flux_e = p_sol->flux_e = flux_e;
    s_sol->o_hat = p_sol->o_hat = o_hat;
    s_sol->e_hat = p_sol->e_hat = e_hat;
}

/* internal.rs: a mode inside the crystal meeting the crystal/air interface. */
typedef struct {
    cplx r_o, r_e, t_s, t_p;
    float flux_ro, flux_re, flux_ts, flux_tp, flux_inc;
} internal_sol;

static internal_sol internal_solve(float n_mode_inc, float n_o, float n_e, v3 c, const uframe* F,
                                   int incident_is_ordinary) {
    internal_sol S;
    float k_tan = n_mode_inc * F->sin_i;
    qroots R = uniaxial_q_roots(n_o, n_e, F, k_tan);
    mfields f_inc = incident_is_ordinary
                        ? ordinary_mode_fields(n_o, c, F->that, F->zhat, k_tan, R.qo_plus)
                        : extraordinary_mode_fields(n_o, n_e, c, F->that, F->zhat, k_tan, R.qe_plus);
    mfields f_o = ordinary_mode_fields(n_o, c, F->that, F->zhat, k_tan, R.qo_minus);
    mfields f_e = extraordinary_mode_fields(n_o, n_e, c, F->that, F->zhat, k_tan, R.qe_minus);
    const float n2 = 1.0f;
    float zeta = k_tan / n2;
    cplx cos_t = fabsf(zeta) <= 1.0f ? cx(sqrtf(maxf(fmaf(zeta, -zeta, 1.0f), 0.0f)), 0.0f)
                                     : cx(0.0f, sqrtf(fmaf(zeta, zeta, -1.0f)));
    cv3 k_iso = cv_wave(F->that, F->zhat, k_tan, cscale(cos_t, n2));
    cv3 s_c = cv_real(F->s_axis);
    cv3 ep_t = cv_scale_real(cv_cross(k_iso, s_c), -1.0f / (n2 * n2));
    mfields f_s = {k_iso, s_c, cv_cross(k_iso, s_c)};
    mfields f_p = {k_iso, ep_t, cv_cross(k_iso, ep_t)};
    cplx inc_c[4], eo[4], ee[4], es[4], ep[4];
    tangential_components(&f_inc, F->that, F->s_axis, inc_c);
    tangential_components(&f_o, F->that, F->s_axis, eo);
    tangential_components(&f_e, F->that, F->s_axis, ee);
    tangential_components(&f_s, F->that, F->s_axis, es);
    tangential_components(&f_p, F->that, F->s_axis, ep);
    cplx a[4][4], b[4];
    for (int i = 0; i < 4; ++i) {
        a[i][0] = eo[i];
        a[i][1] = ee[i];
        a[i][2] = cneg0(es[i]);
        a[i][3] = cneg0(ep[i]);
        b[i] = cneg0(inc_c[i]);
    }
    solve4_n(a, b, NULL);
    S.r_o = b[0]; S.r_e = b[1]; S.t_s = b[2]; S.t_p = b[3];
    S.flux_ro = fabsf(poynting_z(&f_o, F->zhat));
    S.flux_re = fabsf(poynting_z(&f_e, F->zhat));
    S.flux_ts = fabsf(poynting_z(&f_s, F->zhat));
    S.flux_tp = fabsf(poynting_z(&f_p, F->zhat));
    S.flux_inc = fabsf(poynting_z(&f_inc, F->zhat));
    return S;
}

/* mueller.rs */
static mat4 jones_to_mueller(cplx j_ss, cplx j_sp, cplx j_ps, cplx j_pp) {
    float m_ss = cnorm2(j_ss), m_sp = cnorm2(j_sp), m_ps = cnorm2(j_ps), m_pp = cnorm2(j_pp);
    cplx a_ssp = cmul(j_ss, cconj(j_sp));
    cplx a_pep = cmul(j_ps, cconj(j_pp));
    cplx a_plus = cadd(a_ssp, a_pep), a_minus = csub(a_ssp, a_pep);
    cplx b_sp = cmul(j_ss, cconj(j_ps));
    cplx b_pp = cmul(j_sp, cconj(j_pp));
    cplx c_sspp = cmul(j_ss, cconj(j_pp));
    cplx c_spps = cmul(j_sp, cconj(j_ps));
    float m[16] = {
        0.5f * (((m_ss + m_ps) + m_sp) + m_pp),                      /* I <- I */
        0.5f * (((m_ss - m_ps) + m_sp) - m_pp),                      /* Q <- I */
        b_sp.re + b_pp.re,                                           /* U <- I */
        -(b_sp.im + b_pp.im),                                        /* V <- I */
        0.5f * (((m_ss + m_ps) - m_sp) - m_pp),                      /* I <- Q */
        midpointf((m_ss - m_ps) - m_sp, m_pp),                       /* Q <- Q */
        b_sp.re - b_pp.re,                                           /* U <- Q */
        -(b_sp.im) + b_pp.im,                                        /* V <- Q */
        a_plus.re, a_minus.re, c_sspp.re + c_spps.re, -(c_sspp.im + c_spps.im),
        a_plus.im, a_minus.im, c_sspp.im - c_spps.im, c_sspp.re - c_spps.re,
    };
    return mat_cols(m);
}

static float mode_power(cplx t_s, cplx t_p, float mode_flux, float inc_flux, stokes st) {
    float m_s = cnorm2(t_s), m_p = cnorm2(t_p);
    cplx cr = cmul(t_s, cconj(t_p));
    float raw = fmaf(cr.im, st.v, fmaf(cr.re, st.u, 0.5f * fmaf(m_s, st.i + st.q, m_p * (st.i - st.q))));
    return (mode_flux / maxf(inc_flux, 1e-12f)) * raw;
}

static void azimuth2_in_frame(v3 dir, v3 s_axis, v3 p_axis, float* c2, float* s2) {
    float s = vdot(dir, s_axis), p = vdot(dir, p_axis);
    float norm = maxf(fmaf(p, p, s * s), 1e-12f);
    *c2 = fmaf(p, -p, s * s) / norm;
    *s2 = 2.0f * s * p / norm;
}

/* ---- BirefringenceParams (birefringence/uniaxial.rs) ---- */
static float effective_extraordinary_index(float n_o, float n_e, float theta) {
    if (fabsf(n_o - n_e) < 1e-5f) return n_o;
    float st = TF(sinf(theta), theta), ct = TF(cosf(theta), theta + 1.0f);
    float nos = n_o * st;
    float denom_sq = fmaf(n_e * ct, n_e * ct, nos * nos);
    if (denom_sq <= 1e-8f) return n_o;
    return (n_o * n_e) / sqrtf(denom_sq);
}

static float walk_off_from_sin_cos(float n_o, float n_e, float st, float ct) {
    if (fabsf(n_o - n_e) < 1e-5f) return 0.0f;
    float n_o2 = n_o * n_o, n_e2 = n_e * n_e;
    float tan_rho = ((n_o2 - n_e2) * st * ct) / maxf(fmaf(n_e2 * ct, ct, n_o2 * st * st), 1e-6f);
    return TF(atanf(tan_rho), tan_rho);
}

static v3 extraordinary_poynting_dir(v3 w, v3 c, float n_o, float n_e) {
    float cos_theta = clampf(vdot(w, c), -1.0f, 1.0f);
    float ac = fabsf(cos_theta);
    float st = sqrtf(maxf((1.0f - ac) * (1.0f + ac), 0.0f));
    float delta = walk_off_from_sin_cos(n_o, n_e, st, ac);
    if (fabsf(delta) < 1e-5f) return w;
    v3 c_proj = vnormalize_or_zero(vsub(c, vmul(w, cos_theta)));
    if (vlen_sq(c_proj) < 1e-6f) return w;
    float sign = cos_theta >= 0.0f ? 1.0f : -1.0f;
    return vnormalize(vsub(vmul(w, TF(cosf(delta), delta + 1.0f)), vmul(vmul(c_proj, sign), TF(sinf(delta), delta))));
}

/* assigned_mode_e_field_uniaxial (birefringence/absorption_tensor.rs) */
static v3 assigned_mode_e_field_uniaxial(v3 k, v3 c, int is_extra, float n_o_hero, float n_e_hero) {
    if (!is_extra) return ordinary_eigen(k, c);
    v3 kxc = vcross(k, c);
    if (vlen_sq(kxc) < 1e-8f) return vnormalize_or_zero(vcross(k, ordinary_eigen(k, c)));
    v3 s = extraordinary_poynting_dir(k, c, n_o_hero, n_e_hero);
    return vnormalize(vcross(kxc, s));
}

#endif
