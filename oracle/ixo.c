/*
 * ixo.c -- CPU oracle (test infrastructure only; see ixo.h).
 *
 * An FP32 restatement of the reference's isotropic spectral tracer, written
 * against the reference's evaluation order: Rust never contracts a*b+c, so
 * this file is built with -ffp-contract=off and uses fmaf() exactly where the
 * reference writes mul_add.  Vector helpers follow glam 0.34's scalar Vec3
 * order (left-associated dot, multiply-then-add) and Mat4::mul_vec4's
 * column-by-column accumulation.  Transcendentals are glibc libm, the Linux
 * binding of Rust's f32 methods.
 *
 * Parity status: pinned by the reference's golden tables (trace bits for
 * Diamond, studio-environment bits) and closed forms -- see tests/.
 */
#define _GNU_SOURCE
#include "ixo.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define IX_TABLE_QUAL static const
#include "spectral_tables.h"

/* The reference's AVX2 kernels (simd/slab.rs:381, simd/exp_poly.rs:122): an
 * 8-wide slab scan over an SoA copy of the facet set and the 8-lane
 * exponential, each the scalar function's IEEE operations lane by lane (no
 * contraction; FMA exactly where the scalar writes fmaf), so every result is
 * bit-identical to the scalar path.  The FP64 twin and the FLOP-counting
 * build use the scalar path. */
#if defined(__AVX2__) && defined(__FMA__) && !defined(IXO_SCALAR_ONLY)
#include <immintrin.h>
#define IXO_SIMD 1
#endif

/* The FLOP-counting build (ixo_flops.cpp) counts the f32 operations this
 * file performs; IXO_FLOPS(n) books the ones it writes through double libm
 * calls (correctly rounded sin / cos / atan2, the f64 midpoints). */
#ifndef IXO_FLOPS
#define IXO_FLOPS(n) ((void)0)
#endif
/* A build for the tests only (-DIXO_ULP_JITTER) moves the uniaxial
 * transport's transcendental results (acos, sin, cos, atan, atan2) by one
 * ULP, a hash of the argument choosing down / none / up, as another
 * library's functions may differ from glibc's, to measure how far such
 * differences carry through the transport.  Off, TF(y, x) is y. */
#ifdef IXO_ULP_JITTER
static float ulp_jitter(float y, float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u = (u ^ (u >> 13)) * 0x9E3779B1u;
    const uint32_t d = u >> 30;
    if (d == 1u) return nextafterf(y, INFINITY);
    if (d == 2u) return nextafterf(y, -INFINITY);
    return y;
}
#define TF(y, x) ulp_jitter((y), (x))
#else
#define TF(y, x) (y)
#endif

/* a value stored into an ABI struct's f32 field */
#define ABI_F32(x) ((ixo_abi_f32)(x))

#define NCH 8
#define F32_PI 3.14159265358979323846f

/* ------------------------------------------------------------------ */
/* glam-order vector helpers                                            */
/* ------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 vneg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float vdot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline float vlen_sq(v3 a) { return vdot(a, a); }
static inline v3 vnormalize(v3 a) {
    float r = 1.0f / sqrtf(vdot(a, a));
    return vmul(a, r);
}
static inline v3 vnormalize_or_zero(v3 a) {
    float r = 1.0f / sqrtf(vdot(a, a));
    if (isfinite(r) && r > 0.0f) return vmul(a, r);
    return mk(0.0f, 0.0f, 0.0f);
}
static inline v3 vfrom(const ixo_abi_f32* p) { return mk(p[0], p[1], p[2]); }  /* an ABI float[3] */
static inline v3 vfromr(const float* p) { return mk(p[0], p[1], p[2]); }      /* an arithmetic float[3] */
/* Rust f32::clamp: NaN propagates. */
static inline float clampf(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
/* Rust f32::max/min: the non-NaN operand wins (fmaxf/fminf). */
static inline float maxf(float a, float b) { return fmaxf(a, b); }
static inline float minf(float a, float b) { return fminf(a, b); }
/* powi by binary exponentiation, LLVM's expansion order. */
static float powi_f(float x, int n) {
    float res = 0.0f;
    int have = 0;
    float sq = x;
    while (n) {
        if (n & 1) {
            if (have) res = res * sq;
            else { res = sq; have = 1; }
        }
        n >>= 1;
        if (n) sq = sq * sq;
    }
    return have ? res : 1.0f;
}

/* Stokes vectors and Mueller matrices (IX/optics/polarization.rs). */
typedef struct { float i, q, u, v; } stokes;
/* Column-major 4x4: m[c][r]. */
typedef struct { float m[4][4]; } mat4;

static inline stokes st_apply(const mat4* M, stokes s) {
    /* glam Mat4::mul_vec4: res = c0*x; res += c1*y; res += c2*z; res += c3*w. */
    float in[4] = {s.i, s.q, s.u, s.v};
    float r[4];
    for (int k = 0; k < 4; ++k) r[k] = M->m[0][k] * in[0];
    for (int c = 1; c < 4; ++c)
        for (int k = 0; k < 4; ++k) r[k] = r[k] + M->m[c][k] * in[c];
    stokes o = {r[0], r[1], r[2], r[3]};
    return o;
}
static inline stokes st_scale(stokes s, float f) {
    stokes o = {s.i * f, s.q * f, s.u * f, s.v * f};
    return o;
}
static inline float st_intensity(stokes s) { return maxf(s.i, 0.0f); }
static inline stokes st_unpolarized(stokes s) { /* StokesVector::unpolarized(s.intensity()) */
    stokes o = {st_intensity(s), 0.0f, 0.0f, 0.0f};
    return o;
}

static mat4 mat_cols(const float a[16]) {
    mat4 M;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) M.m[c][r] = a[c * 4 + r];
    return M;
}
/* f32::midpoint: correctly rounded (a+b)/2; equals 0.5*(a+b) for the
 * magnitudes reached here, computed in double exactly like the std impl. */
static inline float midpointf(float a, float b) { IXO_FLOPS(2); return (float)(((double)a + (double)b) / 2.0); }

/* MuellerMatrix::frame_rotation (polarization.rs:138). */
static mat4 mueller_frame_rotation(float psi) {
    float c2 = cosf(2.0f * psi), s2 = sinf(2.0f * psi);
    float a[16] = {1, 0, 0, 0, 0, c2, -s2, 0, 0, s2, c2, 0, 0, 0, 0, 1};
    return mat_cols(a);
}
/* MuellerMatrix::fresnel_reflection (polarization.rs:148-158). */
static mat4 mueller_fresnel_reflection(float r_s, float r_p) {
    float rs2 = r_s * r_s, rp2 = r_p * r_p;
    float a = midpointf(rs2, rp2), b = 0.5f * (rs2 - rp2), c = r_s * r_p;
    float m[16] = {a, b, 0, 0, b, a, 0, 0, 0, 0, c, 0, 0, 0, 0, c};
    return mat_cols(m);
}
/* MuellerMatrix::fresnel_transmission (polarization.rs:162-180). */
static mat4 mueller_fresnel_transmission(float n1, float n2, float cos_i, float cos_t, float t_s,
                                         float t_p) {
    float factor = (n2 * cos_t) / maxf(n1 * cos_i, 1e-6f);
    float ts2 = t_s * t_s * factor, tp2 = t_p * t_p * factor;
    float a = midpointf(ts2, tp2), b = 0.5f * (ts2 - tp2), c = t_s * t_p * factor;
    float m[16] = {a, b, 0, 0, b, a, 0, 0, 0, 0, c, 0, 0, 0, 0, c};
    return mat_cols(m);
}
/* MuellerMatrix::tir_retardation (polarization.rs:184-190). */
static mat4 mueller_tir_retardation(float delta) {
    float c = cosf(delta), s = sinf(delta);
    float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, c, -s, 0, 0, s, c};
    return mat_cols(m);
}

/* ------------------------------------------------------------------ */
/* RNG (RT/sampling.rs)                                                 */
/* ------------------------------------------------------------------ */
#define FRESNEL_BRANCH_STREAM 0x9E3779B1u
#define FROSTED_DIR_U_STREAM 0x27D4EB2Fu /* sampling.rs:24-32 */
#define DISTANCE_SAMPLE_STREAM 0xA24BAED4u
#define PHASE_DIR_U_STREAM 0x9FB21C65u
#define PHASE_DIR_V_STREAM 0x1CE4E5B9u
#define FROSTED_DIR_V_STREAM 0x165667B1u
#define RUSSIAN_ROULETTE_STREAM 0x517CC1B7u
#define FROSTED_NEE_ENV_DIR_U_STREAM 0x6A09E667u /* sampling.rs:52-53 */
#define FROSTED_NEE_ENV_DIR_V_STREAM 0xBB67AE85u
#define PIXEL_JITTER_X_ROTATION_STREAM 0xA511E9B3u
#define PIXEL_JITTER_Y_ROTATION_STREAM 0x63D81B23u
#define HERO_WAVELENGTH_ROTATION_STREAM 0x1B873593u

uint32_t ixo_hash_u32(uint32_t x) { /* sampling.rs:66 */
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}
static uint32_t bit_reverse(uint32_t v) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) { r = (r << 1) | (v & 1u); v >>= 1; }
    return r;
}
float ixo_low_discrepancy_base2(uint32_t n) { /* sampling.rs:110 */
    return (float)bit_reverse(n) / 4294967296.0f;
}
float ixo_radical_inverse_base(uint32_t n, uint32_t base) { /* sampling.rs:127 */
    uint32_t remaining = n;
    float val = 0.0f;
    float inv_base = 1.0f / (float)base;
    while (remaining > 0) {
        uint32_t digit = remaining % base;
        val = fmaf((float)digit, inv_base, val);
        inv_base /= (float)base;
        remaining /= base;
    }
    return val;
}
static float cp_rotate(float x, float offset) { /* sampling.rs:148 */
    float sum = x + offset;
    return sum - floorf(sum);
}
void ixo_sample_draws(uint32_t pixel, uint32_t sample, uint32_t* seed, float* jx, float* jy,
                      float* hero) { /* sampling.rs:176-218 */
    float rx = ixo_low_discrepancy_base2(ixo_hash_u32(pixel ^ PIXEL_JITTER_X_ROTATION_STREAM));
    float ry = ixo_low_discrepancy_base2(ixo_hash_u32(pixel ^ PIXEL_JITTER_Y_ROTATION_STREAM));
    float rh = ixo_low_discrepancy_base2(ixo_hash_u32(pixel ^ HERO_WAVELENGTH_ROTATION_STREAM));
    *seed = ixo_hash_u32((pixel * 0x9e3779b9u) ^ (sample * 0x85ebca6bu));
    *jx = cp_rotate(ixo_low_discrepancy_base2(sample), rx) - 0.5f;
    *jy = cp_rotate(ixo_radical_inverse_base(sample, 3), ry) - 0.5f;
    *hero = cp_rotate(ixo_radical_inverse_base(sample, 5), rh);
}
static inline float unit_draw(uint32_t seed, uint32_t bounce, uint32_t stream) {
    /* bounce_body.rs:71-73 */
    return (float)ixo_hash_u32(seed ^ ixo_hash_u32(bounce ^ stream)) / 4294967295.0f;
}
void ixo_wrapped_hero_wavelengths(float hero_rand, float out[8]) { /* transport/mod.rs:58-66 */
    const float span = 780.0f - 380.0f;
    float channel_width = span / 8.0f;
    float lambda_hero = fmaf(hero_rand, span, 380.0f);
    for (int k = 0; k < 8; ++k) {
        float offset = fmaf((float)k, channel_width, lambda_hero - 380.0f);
        float r = fmodf(offset, span);
        if (r < 0.0f) r = r + fabsf(span);
        out[k] = 380.0f + r;
    }
}

/* ------------------------------------------------------------------ */
/* Optical primitives                                                   */
/* ------------------------------------------------------------------ */
static float dispersion_unclamped(const ixg_dispersion* d, float lambda_nm) {
    /* dispersion.rs:181-210 */
    float um = lambda_nm * 1e-3f;
    float l2 = um * um;
    const ixo_abi_f32* c = d->coeff;
    switch (d->kind) {
    case IXG_DISP_SELLMEIER1: {
        float n2 = 1.0f + (c[0] * l2) / (l2 - c[1]);
        return sqrtf(maxf(n2, 1.0f));
    }
    case IXG_DISP_SELLMEIER3: {
        float n2 = 1.0f;
        n2 += (c[0] * l2) / (l2 - c[3]);
        n2 += (c[1] * l2) / (l2 - c[4]);
        n2 += (c[2] * l2) / (l2 - c[5]);
        return sqrtf(maxf(n2, 1.0f));
    }
    default: {
        float l4 = l2 * l2;
        return maxf((c[0] + (c[1] / l2)) + (c[2] / l4), 1.0f);
    }
    }
}
float ixo_dispersion_evaluate(const ixg_dispersion* d, float lambda_nm) {
    /* dispersion.rs:170-178; the sub-380 nm clamp only serves fluorescence
     * (not on this path), so wavelengths below 380 evaluate at 380. */
    if (lambda_nm < 380.0f) lambda_nm = 380.0f;
    return dispersion_unclamped(d, lambda_nm);
}

void ixo_cie_cmf(float lambda_nm, float out[3]) { /* color/cie1931.rs:184-209 */
    const float end = 380.0f + (float)(IX_CMF_ROWS - 1) * 5.0f;
    if (!(lambda_nm >= 380.0f && lambda_nm <= end)) {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    float position = (lambda_nm - 380.0f) / 5.0f;
    size_t i0 = (size_t)floorf(position);
    if (i0 > IX_CMF_ROWS - 2) i0 = IX_CMF_ROWS - 2;
    float t = position - (float)i0;
    for (int c = 0; c < 3; ++c) {
        float lo = ix_cmf_table[i0][c], hi = ix_cmf_table[i0 + 1][c];
        out[c] = lo + (hi - lo) * t;
    }
}
float ixo_d65(float lambda_nm) { /* environment/spectral.rs:49-76 (>= 380 nm part) */
    float clamped = clampf(lambda_nm, 380.0f, fmaf(10.0f, (float)(IX_D65_ROWS - 1), 380.0f));
    float position = (clamped - 380.0f) / 10.0f;
    size_t i0 = (size_t)floorf(position);
    if (i0 > IX_D65_ROWS - 2) i0 = IX_D65_ROWS - 2;
    float frac = position - (float)i0;
    float v0 = ix_d65_table[i0], v1 = ix_d65_table[i0 + 1];
    return fmaf(frac, v1 - v0, v0) / 100.0f;
}
typedef struct { float t_k, denom_560; } bb_norm;
static bb_norm bb_norm_new(float temp_k) { /* spectral.rs:101-107 */
    bb_norm n;
    n.t_k = maxf(temp_k, 1000.0f);
    float e = expf(minf(14388000.0f / (560.0f * n.t_k), 80.0f));
    n.denom_560 = maxf(e - 1.0f, 1e-6f);
    return n;
}
static float bb_with(bb_norm n, float lambda_nm) { /* spectral.rs:110-116 */
    float e = expf(minf(14388000.0f / (lambda_nm * n.t_k), 80.0f));
    float denom = maxf(e - 1.0f, 1e-6f);
    float ratio = n.denom_560 / denom;
    return clampf(powi_f(560.0f / lambda_nm, 5) * ratio, 0.01f, 20.0f);
}
float ixo_blackbody(float lambda_nm, float temp_k) { return bb_with(bb_norm_new(temp_k), lambda_nm); }

static const float BRADFORD_XYZ_TO_LMS[3][3] = {
    {0.8951f, 0.2664f, -0.1614f}, {-0.7502f, 1.7135f, 0.0367f}, {0.0389f, -0.0685f, 1.0296f}};
static const float BRADFORD_LMS_TO_XYZ[3][3] = {{0.986993f, -0.147054f, 0.159963f},
                                                {0.432305f, 0.518360f, 0.049291f},
                                                {-0.008529f, 0.040043f, 0.968487f}};
static v3 mat3_fma(const float m[3][3], v3 a) { /* color.rs:220-240 */
    return mk(fmaf(m[0][0], a.x, fmaf(m[0][1], a.y, m[0][2] * a.z)),
              fmaf(m[1][0], a.x, fmaf(m[1][1], a.y, m[1][2] * a.z)),
              fmaf(m[2][0], a.x, fmaf(m[2][1], a.y, m[2][2] * a.z)));
}
void ixo_white_balance(float temp_k, float out[3]) { /* color.rs:268-305 */
    v3 w = mk(0, 0, 0);
    for (int step = 0; step <= 400; ++step) {
        float lambda = 380.0f + (float)step;
        float cmf[3];
        ixo_cie_cmf(lambda, cmf);
        float bb = ixo_blackbody(lambda, temp_k);
        w = vadd(w, mk(cmf[0] * bb, cmf[1] * bb, cmf[2] * bb));
    }
    float ty = maxf(w.y, 1e-6f);
    const float X = 0.3127f, Y = 0.3290f;
    v3 target = mk((X / Y) * ty, ty, ((1.0f - X - Y) / Y) * ty);
    v3 ls = mat3_fma(BRADFORD_XYZ_TO_LMS, w), lt = mat3_fma(BRADFORD_XYZ_TO_LMS, target);
    out[0] = ls.x > 1e-6f ? lt.x / ls.x : 1.0f;
    out[1] = ls.y > 1e-6f ? lt.y / ls.y : 1.0f;
    out[2] = ls.z > 1e-6f ? lt.z / ls.z : 1.0f;
}
static v3 von_kries(v3 xyz, v3 scale) { /* color.rs:252-254 */
    v3 lms = mat3_fma(BRADFORD_XYZ_TO_LMS, xyz);
    lms = mk(lms.x * scale.x, lms.y * scale.y, lms.z * scale.z);
    return mat3_fma(BRADFORD_LMS_TO_XYZ, lms);
}

float ixo_exp_lane(float x) { /* simd/exp_poly.rs:37-60 */
#ifdef IXO_FP64_TWIN
    return exp(x);  /* the twin: the function the lane polynomial approximates */
#endif
    const float LOG2E = 1.44269504088896340736f;
    union { uint32_t u; float f; } c1 = {0x3F318000u};
    const float C2 = -2.1219444e-4f;
    const float P0 = 1.9875691e-4f, P1 = 1.3981999e-3f, P2 = 8.333452e-3f, P3 = 4.1665796e-2f,
                P4 = 1.6666665e-1f, P5 = 5.0000003e-1f;
    const float HI = 88.02875f, LO = -87.33654f;
    float upper = x < HI ? x : HI;
    x = upper > LO ? upper : LO;
    float n = floorf(fmaf(x, LOG2E, 0.5f));
    x = fmaf(n, -c1.f, x);
    x = fmaf(n, -C2, x);
    float z = x * x;
    float p = P0;
    p = fmaf(p, x, P1);
    p = fmaf(p, x, P2);
    p = fmaf(p, x, P3);
    p = fmaf(p, x, P4);
    p = fmaf(p, x, P5);
    p = fmaf(p, z, x) + 1.0f;
    int32_t e = (int32_t)n;
    union { uint32_t u; float f; } pw = {(uint32_t)((e + 127) << 23)};
    return p * pw.f;
}

/* exp_f32x8_avx2 (simd/exp_poly.rs:71, dispatched from :122): ixo_exp_lane on eight lanes. */
static void exp_lane8(const float x[8], float out[8]) {
#ifdef IXO_SIMD
    const __m256 HI = _mm256_set1_ps(88.02875f), LO = _mm256_set1_ps(-87.33654f);
    __m256 v = _mm256_loadu_ps(x);
    v = _mm256_max_ps(_mm256_min_ps(v, HI), LO);  /* NaN -> HI, as the scalar selects */
    const __m256 n = _mm256_floor_ps(_mm256_fmadd_ps(v, _mm256_set1_ps(1.44269504088896340736f), _mm256_set1_ps(0.5f)));
    const __m256 c1 = _mm256_castsi256_ps(_mm256_set1_epi32(0x3F318000));
    v = _mm256_fmadd_ps(n, _mm256_xor_ps(c1, _mm256_set1_ps(-0.0f)), v);
    v = _mm256_fmadd_ps(n, _mm256_set1_ps(2.1219444e-4f), v);
    const __m256 z = _mm256_mul_ps(v, v);
    __m256 p = _mm256_set1_ps(1.9875691e-4f);
    p = _mm256_fmadd_ps(p, v, _mm256_set1_ps(1.3981999e-3f));
    p = _mm256_fmadd_ps(p, v, _mm256_set1_ps(8.333452e-3f));
    p = _mm256_fmadd_ps(p, v, _mm256_set1_ps(4.1665796e-2f));
    p = _mm256_fmadd_ps(p, v, _mm256_set1_ps(1.6666665e-1f));
    p = _mm256_fmadd_ps(p, v, _mm256_set1_ps(5.0000003e-1f));
    p = _mm256_add_ps(_mm256_fmadd_ps(p, z, v), _mm256_set1_ps(1.0f));
    const __m256i e = _mm256_slli_epi32(_mm256_add_epi32(_mm256_cvttps_epi32(n), _mm256_set1_epi32(127)), 23);
    _mm256_storeu_ps(out, _mm256_mul_ps(p, _mm256_castsi256_ps(e)));
#else
    for (int c = 0; c < 8; ++c) out[c] = ixo_exp_lane(x[c]);
#endif
}

void ixo_exp_lane8(const float x[8], float out[8]) { exp_lane8(x, out); }

float ixo_band_evaluate(const ixg_band* b, float lambda_nm) { /* optics/absorption.rs:108-138 */
    if (b->shape == 0) {
        float t = (lambda_nm - b->center_nm) / b->width;
        return b->peak * expf(-0.5f * t * t);
    }
    float delta = b->center_nm - lambda_nm;
    float nu = 1.0e7f * delta / (lambda_nm * b->center_nm);
    float t = nu / b->width;
    return b->peak * expf(-0.5f * t * t);
}

/* ------------------------------------------------------------------ */
/* Environment (RT/environment/)                                        */
/* ------------------------------------------------------------------ */
enum { MODEL_STUDIO = 0, MODEL_ISO = 1, MODEL_TENT = 2, MODEL_DOME = 3, MODEL_ASET = 4, MODEL_SUN = 5 };  /* gpu_id */
typedef struct { float walls, cards, spark, ground, flat; } tent_t;  /* TentParams (mod.rs:185-215) */
typedef struct {
    int model;
    int d65;            /* 1: tabulated D65 illuminant, 0: blackbody */
    bb_norm bb;
    float exposure, spot_mult;
    float shadow[2];
    v3 key_dir, fill_dir, ring[16];
    v3 wb;              /* von Kries scale (identity for D65 presets) */
    float tray;         /* lit backdrop radiance at exposure 1 (0 = as lit) */
    tent_t tent;
    float sun;          /* sun_radiance_factor (environment/sun.rs) for DaylightSun */
} env_t;

static int preset_params(int32_t preset, int* model, int* d65, float* temp_k, float* spot_mult,
                         int* white_balance, tent_t* tent) { /* environment/mod.rs:321-410,519-590 */
    static const tent_t tent_default = {1.0f, 1.0f, 1.0f, 0.02f, 0.0f};
    *tent = tent_default;
    switch (preset) {
    case IXG_PRESET_INCANDESCENT: *model = MODEL_STUDIO; *d65 = 0; *temp_k = 3200.0f; *spot_mult = 1.2f; *white_balance = 1; return 1;
    case IXG_PRESET_RING_LIGHTS: *model = MODEL_STUDIO; *d65 = 0; *temp_k = 5000.0f; *spot_mult = 1.6f; *white_balance = 1; return 1;
    case IXG_PRESET_DARK_SPOTLIGHT: *model = MODEL_STUDIO; *d65 = 0; *temp_k = 6000.0f; *spot_mult = 2.4f; *white_balance = 1; return 1;
    case IXG_PRESET_ILLUMINANT_A: *model = MODEL_STUDIO; *d65 = 0; *temp_k = 2856.0f; *spot_mult = 1.2f; *white_balance = 1; return 1;
    case IXG_PRESET_DAYLIGHT: *model = MODEL_STUDIO; *d65 = 1; *temp_k = 6500.0f; *spot_mult = 1.0f; *white_balance = 0; return 1;
    case IXG_PRESET_ISO_HEMISPHERE: *model = MODEL_ISO; *d65 = 1; *temp_k = 6500.0f; *spot_mult = 1.0f; *white_balance = 0; return 1;
    case IXG_PRESET_LIGHT_TENT: *model = MODEL_TENT; *d65 = 0; *temp_k = 5000.0f; *spot_mult = 1.0f; *white_balance = 1; return 1;
    case IXG_PRESET_WHITE_TRAY:
        *model = MODEL_TENT; *d65 = 1; *temp_k = 5000.0f; *spot_mult = 0.0f; *white_balance = 0;
        *tent = (tent_t){5.0f, 0.0f, 0.0f, 0.8f, 1.0f};
        return 1;
    case IXG_PRESET_SHOP_LIGHTS:
        *model = MODEL_TENT; *d65 = 1; *temp_k = 5000.0f; *spot_mult = 1.0f; *white_balance = 0;
        *tent = (tent_t){3.0f, 0.0f, 1.0f, 0.05f, 0.0f};
        return 1;
    case IXG_PRESET_WINDOW_DAYLIGHT:
        *model = MODEL_TENT; *d65 = 1; *temp_k = 5000.0f; *spot_mult = 1.0f; *white_balance = 0;
        *tent = (tent_t){0.4f, 0.0f, 0.0f, 0.03f, 0.0f};
        return 1;
    case IXG_PRESET_DAYLIGHT_DOME: *model = MODEL_DOME; *d65 = 1; *temp_k = 6500.0f; *spot_mult = 1.0f; *white_balance = 0; return 1;
    case IXG_PRESET_DAYLIGHT_SUN: *model = MODEL_SUN; *d65 = 1; *temp_k = 6500.0f; *spot_mult = 1.0f; *white_balance = 0; return 1;
    case IXG_PRESET_ASET: *model = MODEL_ASET; *d65 = 1; *temp_k = 6500.0f; *spot_mult = 1.0f; *white_balance = 0; return 1;
    default: return 0;
    }
}

static void head_shadow_cosines(float deg, float out[2]) { /* environment/mod.rs:124-137 */
    if (isnan(deg) || deg == 16.0f) { out[0] = 0.9510565f; out[1] = 0.9702957f; return; }
    if (deg <= 0.0f) { out[0] = 2.0f; out[1] = 3.0f; return; }
    deg = clampf(deg, 3.0f, 88.0f);
    const float to_rad = F32_PI / 180.0f;
    out[0] = cosf((deg + 2.0f) * to_rad);
    out[1] = cosf((deg - 2.0f) * to_rad);
}

static void studio_rig_new(float yaw, float pitch, env_t* e) { /* optics/studio_rig.rs:47-80 */
    float cp = cosf(pitch), sp = sinf(pitch), cy = cosf(yaw), sy = sinf(yaw);
    e->key_dir = vnormalize(mk(cp * sy, sp, cp * cy));
    float fill_yaw = fmaf(F32_PI, 0.78f, yaw);
    float fill_pitch = clampf(pitch * 0.65f, 0.15f, 1.2f);
    e->fill_dir = vnormalize(mk(cosf(fill_pitch) * sinf(fill_yaw), sinf(fill_pitch),
                                cosf(fill_pitch) * cosf(fill_yaw)));
    for (int i = 0; i < 16; ++i) {
        float a = fmaf((float)i, F32_PI * 2.0f / 16.0f, yaw);
        e->ring[i] = vnormalize(mk(sinf(a) * 0.75f, sp * 0.8f, cosf(a) * 0.75f));
    }
}

static int env_build(const ixg_environment* src, env_t* e) {
    float temp_k, spot;
    int wb;
    if (!preset_params(src->preset, &e->model, &e->d65, &temp_k, &spot, &wb, &e->tent)) return 0;
    e->tray = src->backdrop;
    e->bb = bb_norm_new(temp_k);
    e->exposure = src->exposure;
    e->spot_mult = spot;
    head_shadow_cosines(src->head_shadow_deg, e->shadow);
    studio_rig_new(src->light_yaw, src->light_pitch, e);
    e->sun = 0.0f;
    if (e->model == MODEL_SUN) { /* SUN_RADIANCE * horizon_blend(key_dir), sun.rs */
        float t = clampf((e->key_dir.y - -0.05f) / (0.05f - -0.05f), 0.0f, 1.0f);
        e->sun = 40000.0f * (t * t * fmaf(-2.0f, t, 3.0f));
    }
    if (wb) {
        float s[3];
        ixo_white_balance(temp_k, s);
        e->wb = mk(s[0], s[1], s[2]);
    } else {
        e->wb = mk(1.0f, 1.0f, 1.0f);
    }
    return 1;
}

static float smoothstep_f(float e0, float e1, float x) { /* rig.rs:86-89 */
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * fmaf(-2.0f, t, 3.0f);
}

/* Wavelength-independent lighting of one direction (rig.rs:379-405). */
typedef struct {
    int model;
    float bg, softbox, fill, ring[16];
    int has_softbox, has_fill, ring_len;
    int has_ground;
    float horizon, ground;
    float scaled;
    float red, green, blue, tray_ground; /* ASET zone amplitudes (rig.rs:137-142) */
} dir_light;

/* lit_ground (rig.rs:230-236): the tray replaces the model's own ground. */
static float lit_ground(float horizon, float vis, float tray, float own_ground) {
    return tray > 0.0f ? ((1.0f - horizon) * vis) * tray : own_ground * (1.0f - horizon);
}

static dir_light direction_lighting(const env_t* e, v3 dir, v3 observer) {
    dir_light L;
    memset(&L, 0, sizeof L);
    v3 d = vnormalize(dir);
    L.model = e->model;
    if (e->model == MODEL_ASET) { /* aset_radiance, rig.rs:439-459 */
        const float horizon = smoothstep_f(-0.05f, 0.05f, d.y);
        const float s45 = smoothstep_f(0.6871068f, 0.7271068f, d.y);
        const float s75 = smoothstep_f(0.9459258f, 0.9859258f, d.y);
        L.red = (s45 * (1.0f - s75)) * horizon;
        L.green = (1.0f - s45) * horizon;
        L.blue = s75 * horizon;
        L.tray_ground = e->tray > 0.0f ? (1.0f - horizon) * e->tray : 0.0f;
        return L;
    }
    if (e->model != MODEL_STUDIO) {
        const float horizon = smoothstep_f(-0.05f, 0.05f, d.y);             /* rig.rs:98-100 */
        const float vis = 1.0f - smoothstep_f(e->shadow[0], e->shadow[1], vdot(d, observer));
        if (e->model == MODEL_ISO) { /* rig.rs:238-252 */
            L.scaled = e->tray > 0.0f ? fmaf((1.0f - horizon) * vis, e->tray, horizon * vis) : horizon * vis;
        } else if (e->model == MODEL_TENT) { /* light_tent_lighting, rig.rs:254-294 */
            const tent_t* T = &e->tent;
            float walls = fmaf(0.08f, maxf(d.y, 0.0f), 0.14f) * T->walls;
            if (T->flat > 0.0f) walls = fmaf(0.18f * T->walls, T->flat, walls * (1.0f - T->flat));
            float card = 0.0f;
            static const int slots[3] = {4, 8, 12};
            for (int k = 0; k < 3; ++k)
                card = maxf(card, smoothstep_f(0.898794f, 0.9612617f, vdot(d, e->ring[slots[k]])));
            walls *= fmaf(card * T->cards, -0.9f, 1.0f);
            const float key = smoothstep_f(0.7660444f, 0.9396926f, vdot(d, e->key_dir)) * (1.4f * e->spot_mult);
            const float spark =
                smoothstep_f(0.9961947f, 0.9993908f, vdot(d, e->fill_dir)) * (5.0f * e->spot_mult) * T->spark;
            const float above = ((walls + key) + spark) * (horizon * vis);
            L.scaled = above + lit_ground(horizon, vis, e->tray, T->ground);
        } else { /* daylight_dome_lighting, rig.rs:296-320 */
            const float sun_dot = vdot(d, e->key_dir);
            const float sky = fmaf(0.08f, 1.0f - maxf(d.y, 0.0f), 0.10f);
            const float glow = maxf(sun_dot, 0.0f);
            const float glow2 = glow * glow;
            const float glow4 = glow2 * glow2;
            const float aureole = (glow4 * glow4) * 0.30f;
            const float above = (sky + aureole) * (horizon * vis);
            L.scaled = above + lit_ground(horizon, vis, e->tray, 0.04f);
            if (e->model == MODEL_SUN) /* daylight_sun_radiance, rig.rs:466-489: SUN_DISC_COS */
                L.scaled = L.scaled + (sun_dot >= 0.9999889f ? e->sun : 0.0f);
        }
        return L;
    }
    /* studio_rig_lighting, rig.rs:322-370 (ambient fill on) */
    L.bg = maxf(fmaf(0.012f, fmaf(d.y, 0.5f, 0.5f), 0.015f), 0.005f) * e->exposure;
    float key_dot = maxf(vdot(d, e->key_dir), 0.0f);
    if (key_dot > 0.0f) {
        L.has_softbox = 1;
        L.softbox = powi_f(key_dot, 28) * 12.0f * e->spot_mult * e->exposure;
    }
    float fill_dot = maxf(vdot(d, e->fill_dir), 0.0f);
    if (fill_dot > 0.0f) {
        L.has_fill = 1;
        L.fill = powi_f(fill_dot, 18) * 4.5f * e->exposure;
    }
    for (int i = 0; i < 16; ++i) {
        float rd = maxf(vdot(d, e->ring[i]), 0.0f);
        if (rd > 0.96f) {
            float spark = (rd - 0.96f) / 0.04f;
            L.ring[L.ring_len++] = powi_f(spark, 6) * 22.0f * e->spot_mult * e->exposure;
        }
    }
    if (e->tray > 0.0f) { /* the tray under the rig, rig.rs:358-364 */
        L.has_ground = 1;
        L.horizon = smoothstep_f(-0.05f, 0.05f, d.y);
        L.ground = ((1.0f - L.horizon) * e->tray) * e->exposure;
    }
    return L;
}
/* studio_spec_power (rig.rs:579-585): the illuminant's power, or for ASET the
 * wavelength itself, which its banded radiance reads. */
static float spec_power(const env_t* e, float lambda) {
    if (e->model == MODEL_ASET) return lambda;
    return e->d65 ? ixo_d65(lambda) : bb_with(e->bb, lambda);
}
static float aset_band(float lambda, float centre) { /* rig.rs:173-176 */
    float z = (lambda - centre) / 8.493218f;
    return expf(-0.5f * z * z);
}
static float light_radiance(const dir_light* L, float sp, float exposure) { /* rig.rs:178-222 */
    if (L->model == MODEL_ASET) {
        const float lambda = sp;
        const float r = (L->red * aset_band(lambda, 610.0f)) * 8.0f;
        const float g = (L->green * aset_band(lambda, 540.0f)) * 5.0f;
        const float b = (L->blue * aset_band(lambda, 460.0f)) * 16.0f;
        const float zones = ((r + g) + b) * exposure;
        return L->tray_ground > 0.0f ? fmaf(L->tray_ground * ixo_d65(lambda), exposure, zones) : zones;
    }
    if (L->model != MODEL_STUDIO) return L->scaled * (sp * exposure);
    float r = L->bg * sp;
    if (L->has_softbox) r = fmaf(L->softbox, sp, r);
    if (L->has_fill) r = fmaf(L->fill, sp, r);
    for (int i = 0; i < L->ring_len; ++i) r = fmaf(L->ring[i], sp, r);
    if (L->has_ground) r = fmaf(L->horizon, r, L->ground * sp);
    return r;
}

float ixo_sample_studio_environment(const float dir[3], float lambda_nm, int32_t preset,
                                    float exposure, float light_yaw, float light_pitch) {
    ixg_environment src = {preset, ABI_F32(exposure), ABI_F32(light_yaw), ABI_F32(light_pitch), 0.0f, 1.0f, 16.0f};
    env_t e;
    if (!env_build(&src, &e)) return NAN;
    dir_light L = direction_lighting(&e, vfromr(dir), mk(0, 0, 0));
    return light_radiance(&L, spec_power(&e, lambda_nm), exposure);
}

/* ------------------------------------------------------------------ */
/* Camera and geometry                                                  */
/* ------------------------------------------------------------------ */
void ixo_camera_new(float yaw, float pitch, float distance, float fov_deg, uint32_t width,
                    uint32_t height, uint32_t spp, ixg_camera* out) { /* camera.rs:108-147 */
    float cp = cosf(pitch), sp = sinf(pitch), cy = cosf(yaw), sy = sinf(yaw);
    v3 origin = mk(distance * cp * sy, distance * sp, distance * cp * cy);
    v3 forward = vnormalize(vneg(origin));
    v3 world_up;
    if (fabsf(cp) < 1e-4f) world_up = mk(0, 0, -1);
    else if (cp < 0.0f) world_up = mk(0, -1, 0);
    else world_up = mk(0, 1, 0);
    v3 right = vnormalize(vcross(forward, world_up));
    v3 up = vnormalize(vcross(right, forward));
    float fov_tan = tanf(fov_deg * (F32_PI / 180.0f) * 0.5f);
    out->origin[0] = ABI_F32(origin.x); out->origin[1] = ABI_F32(origin.y); out->origin[2] = ABI_F32(origin.z);
    out->fov_tan = ABI_F32(fov_tan);
    out->forward[0] = ABI_F32(forward.x); out->forward[1] = ABI_F32(forward.y); out->forward[2] = ABI_F32(forward.z);
    out->width = ABI_F32((float)width);
    out->right[0] = ABI_F32(right.x); out->right[1] = ABI_F32(right.y); out->right[2] = ABI_F32(right.z);
    out->height = ABI_F32((float)height);
    out->up[0] = ABI_F32(up.x); out->up[1] = ABI_F32(up.y); out->up[2] = ABI_F32(up.z);
    out->num_samples = spp;
}
void ixo_generate_ray(const ixg_camera* c, float sx, float sy, float jx, float jy, float o[3],
                      float d[3]) { /* camera.rs:149-167 */
    float width = c->width, height = c->height;
    float aspect = width / height;
    float u = ((sx + jx) / width - 0.5f) * 2.0f * aspect * c->fov_tan;
    float v = (0.5f - (sy + jy) / height) * 2.0f * c->fov_tan;
    v3 dir = vnormalize(vadd(vadd(vfrom(c->forward), vmul(vfrom(c->right), u)),
                             vmul(vfrom(c->up), v)));
    o[0] = c->origin[0]; o[1] = c->origin[1]; o[2] = c->origin[2];
    d[0] = dir.x; d[1] = dir.y; d[2] = dir.z;
}

static void plane_new(v3 n, float d, ixg_facet_plane* p) { /* geometry/plane.rs:47-54 */
    v3 u = vnormalize(n);
    p->normal[0] = ABI_F32(u.x); p->normal[1] = ABI_F32(u.y); p->normal[2] = ABI_F32(u.z);
    p->d = ABI_F32(d);
}
/* The cut's sines and cosines are the correctly rounded f32 values (double
 * evaluation, one rounding).  The reference's pinned metrics
 * (golden_regression.rs:207-233) were produced where sin(42.5 deg) rounds
 * correctly; glibc 2.35's sinf returns the f32 one ULP above it, which moves
 * the 16 lower-girdle normals and, through them, the Fire index. */
static float sin_cr(float x) { IXO_FLOPS(1); return (float)sin((double)x); }
/* Correctly rounded f32 atan2 (double evaluation, one rounding) for the
 * Stokes frame rotation angle.  glibc 2.35's atan2f misrounds some of these
 * arguments by one ULP; with the correctly rounded angle the reference's
 * whole-frame pin (renderer/gpu/pin_tests.rs:128-141, offset 0) matches bit
 * for bit, with glibc's it does not. */
static float atan2_cr(float y, float x) { IXO_FLOPS(1); return (float)atan2((double)y, (double)x); }
static float cos_cr(float x) { IXO_FLOPS(1); return (float)cos((double)x); }
static float index_to_azimuth(float index, float teeth) { /* cuts/round_brilliant.rs:65 */
    return 2.0f * F32_PI * (index + 0.0f) / teeth;
}
static uint32_t tier(ixg_facet_plane* out, uint32_t at, float angle_rad, float sign_y,
                     const float* idx, int count, float d) {
    for (int i = 0; i < count; ++i) {
        float phi = index_to_azimuth(idx[i], 96.0f);
        v3 n = mk(sin_cr(angle_rad) * cos_cr(phi), sign_y * cos_cr(angle_rad), sin_cr(angle_rad) * sin_cr(phi));
        plane_new(n, d, &out[at++]);
    }
    return at;
}
static const float STAR_IDX[8] = {6, 18, 30, 42, 54, 66, 78, 90};
static const float MAIN_IDX[8] = {0, 12, 24, 36, 48, 60, 72, 84};
static const float BREAK_IDX[16] = {95, 1, 11, 13, 23, 25, 35, 37, 47, 49, 59, 61, 71, 73, 83, 85};
static const float TO_RAD = F32_PI / 180.0f;

static uint32_t round_brilliant(float table_d, float crown_main_deg, float crown_main_d,
                                float pav_main_deg, float pav_main_d, ixg_facet_plane* out) {
    uint32_t n = 0;
    plane_new(mk(0, 1, 0), table_d, &out[n++]);
    n = tier(out, n, 15.0f * TO_RAD, 1.0f, STAR_IDX, 8, -0.45f);
    n = tier(out, n, crown_main_deg * TO_RAD, 1.0f, MAIN_IDX, 8, crown_main_d);
    n = tier(out, n, 41.0f * TO_RAD, 1.0f, BREAK_IDX, 16, -0.67f);
    for (int i = 0; i < 16; ++i) {
        float phi = 2.0f * F32_PI * (float)i / 16.0f;
        plane_new(mk(cos_cr(phi), 0.0f, sin_cr(phi)), -1.0f, &out[n++]);
    }
    n = tier(out, n, pav_main_deg * TO_RAD, -1.0f, MAIN_IDX, 8, pav_main_d);
    n = tier(out, n, 42.5f * TO_RAD, -1.0f, BREAK_IDX, 16, -0.68f);
    plane_new(mk(0, -1, 0), -0.88f, &out[n++]);
    return n;
}
uint32_t ixo_cut_standard_round_brilliant(ixg_facet_plane* out) { /* round_brilliant.rs:71-155 */
    return round_brilliant(-0.32f, 34.5f, -0.59f, 41.0f, -0.67f, out);
}
uint32_t ixo_cut_round_brilliant(float crown_deg, float pavilion_deg, float table_ratio,
                                 ixg_facet_plane* out) {
    /* The standard cut's tiers with the crown/pavilion main angles moved and
     * the table plane placed where it cuts the crown mains at radius
     * table_ratio (girdle radius 1).  Mains keep the standard cut's girdle
     * heights (crown +0.02866, pavilion -0.01847) so the girdle band stays. */
    double ca = (double)crown_deg * 3.14159265358979323846 / 180.0;
    double pa = (double)pavilion_deg * 3.14159265358979323846 / 180.0;
    double crown_d = sin(ca) + cos(ca) * 0.028660;
    double pav_d = sin(pa) + cos(pa) * 0.018470;
    double table_y = (crown_d - sin(ca) * (double)table_ratio) / cos(ca);
    return round_brilliant((float)-table_y, crown_deg, (float)-crown_d, pavilion_deg,
                           (float)-pav_d, out);
}

/* slab_scan + intersect_polyhedron_soa (simd/slab.rs:147-160, RT/intersect.rs:101-132). */
int32_t ixo_intersect(const ixg_facet_plane* planes, uint32_t n, const float o[3], const float d[3],
                      float* t_out) {
    v3 origin = vfromr(o), dir = vfromr(d);
    float t_near = -1e30f, t_far = 1e30f;
    int32_t near_idx = -1, far_idx = -1;
    for (uint32_t i = 0; i < n; ++i) {
        v3 nrm = vfrom(planes[i].normal);
        float denom = vdot(nrm, dir);
        float side = planes[i].d + vdot(nrm, origin);
        float t = (-side) / denom;
        if (fabsf(denom) > 1e-7f) {
            if (denom < 0.0f) {
                if (t > t_near) { t_near = t; near_idx = (int32_t)i; }
            } else if (t < t_far) {
                t_far = t;
                far_idx = (int32_t)i;
            }
        } else if (side > 0.0f) {
            return -1;
        }
    }
    if (t_near > t_far) return -1;
    if (t_near > 1e-4f) { *t_out = t_near; return near_idx < 0 ? 0 : near_idx; }
    if (t_far > 1e-4f) { *t_out = t_far; return far_idx < 0 ? 0 : far_idx; }
    return -1;
}

/* ------------------------------------------------------------------ */
/* Transport (RT/transport/, RT/refraction/)                            */
/* ------------------------------------------------------------------ */
#define DIRECTION_MATCH_COS_TOL (1.0f - 1e-6f)
#define R_UNPOL_PDF_MIN 1e-4f
#define R_UNPOL_PDF_MAX (1.0f - 1e-4f)
#define R_UNPOL_SELECT_MIN 0.02f
#define R_UNPOL_SELECT_MAX 0.98f

/* A facet set laid out for the 8-wide slab scan: planes rounded up to 8,
 * structure of arrays; padding planes (n = 0, d = -1) are parallel to every
 * ray and seen from inside, so they never hit and never reject.  n_pad = 0
 * (the scalar builds, or more planes than the ABI's maximum) scans the
 * array-of-structs planes with ixo_intersect. */
typedef struct {
    uint32_t n_pad;
#ifdef IXO_SIMD
    float soa[4][IXG_MAX_PLANES + 8];
#endif
} slab_t;

static void slab_fill(slab_t* L, const ixg_facet_plane* planes, uint32_t n) {
    L->n_pad = 0;
#ifdef IXO_SIMD
    if (n > IXG_MAX_PLANES) return;
    L->n_pad = (n + 7u) & ~7u;
    for (uint32_t i = 0; i < L->n_pad; ++i) {
        const int real = i < n;
        L->soa[0][i] = real ? planes[i].normal[0] : 0.0f;
        L->soa[1][i] = real ? planes[i].normal[1] : 0.0f;
        L->soa[2][i] = real ? planes[i].normal[2] : 0.0f;
        L->soa[3][i] = real ? planes[i].d : -1.0f;
    }
#else
    (void)planes;
    (void)n;
#endif
}

typedef struct {
    const ixg_facet_plane* planes;
    uint32_t n_planes;
    const ixg_material* mat;
    uint32_t max_bounces;
    env_t env;
    float surface_glare;
    int absorbing;      /* any absorption band (o or e) */
    int aniso;          /* RayMaterialContext::is_anisotropic */
    int biax;           /* genuinely biaxial (biaxial_delta_beta_alpha is Some) */
    int trichroic;      /* biaxial with a beta_ray band set: tensor (o, beta, e) */
    v3 c_raw;           /* material c_axis as given */
    v3 ta1, ta2, tcn;   /* AbsorptionTensor3 axes: basis of normalize_or_zero(c_axis) */
    const uint8_t* finish; /* FacetFinish per plane, NULL = all polished */
    const ixg_hdr_map* hdr; /* EnvironmentSource::HdrMap, or NULL */
    slab_t slab;            /* the facet set for the 8-wide scan */
} scene_t;

#ifdef IXO_SIMD
/* slab_scan_avx2 (simd/slab.rs:193, dispatched from :381) over 8 planes at a time: per lane the scalar
 * scan's operations and strict updates, so each lane keeps its lowest-index
 * best; the lanes then combine by the same rule (largest entering t, smallest
 * exiting t, ties to the lowest index).  A parallel plane seen from outside
 * misses wherever it sits, as the scalar early return does. */
static int32_t intersect_simd(const slab_t* S, const float o[3], const float d[3], float* t_out) {
    const __m256 ox = _mm256_set1_ps(o[0]), oy = _mm256_set1_ps(o[1]), oz = _mm256_set1_ps(o[2]);
    const __m256 dx = _mm256_set1_ps(d[0]), dy = _mm256_set1_ps(d[1]), dz = _mm256_set1_ps(d[2]);
    const __m256 sign = _mm256_set1_ps(-0.0f), zero = _mm256_setzero_ps(), eps = _mm256_set1_ps(1e-7f);
    __m256 tn = _mm256_set1_ps(-1e30f), tf = _mm256_set1_ps(1e30f);
    __m256i in = _mm256_set1_epi32(-1), iff = _mm256_set1_epi32(-1);
    __m256i idx = _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7);
    const __m256i eight = _mm256_set1_epi32(8);
    int outside = 0;
    for (uint32_t i = 0; i < S->n_pad; i += 8) {
        const __m256 nx = _mm256_loadu_ps(&S->soa[0][i]), ny = _mm256_loadu_ps(&S->soa[1][i]);
        const __m256 nz = _mm256_loadu_ps(&S->soa[2][i]), w = _mm256_loadu_ps(&S->soa[3][i]);
        const __m256 denom = _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(nx, dx), _mm256_mul_ps(ny, dy)), _mm256_mul_ps(nz, dz));
        const __m256 side = _mm256_add_ps(w, _mm256_add_ps(_mm256_add_ps(_mm256_mul_ps(nx, ox), _mm256_mul_ps(ny, oy)),
                                                           _mm256_mul_ps(nz, oz)));
        const __m256 t = _mm256_div_ps(_mm256_xor_ps(side, sign), denom);
        const __m256 valid = _mm256_cmp_ps(_mm256_andnot_ps(sign, denom), eps, _CMP_GT_OQ);
        outside |= _mm256_movemask_ps(_mm256_andnot_ps(valid, _mm256_cmp_ps(side, zero, _CMP_GT_OQ)));
        const __m256 entering = _mm256_and_ps(valid, _mm256_cmp_ps(denom, zero, _CMP_LT_OQ));
        const __m256 up_n = _mm256_and_ps(entering, _mm256_cmp_ps(t, tn, _CMP_GT_OQ));
        const __m256 up_f = _mm256_andnot_ps(entering, _mm256_and_ps(valid, _mm256_cmp_ps(t, tf, _CMP_LT_OQ)));
        tn = _mm256_blendv_ps(tn, t, up_n);
        tf = _mm256_blendv_ps(tf, t, up_f);
        in = _mm256_castps_si256(_mm256_blendv_ps(_mm256_castsi256_ps(in), _mm256_castsi256_ps(idx), up_n));
        iff = _mm256_castps_si256(_mm256_blendv_ps(_mm256_castsi256_ps(iff), _mm256_castsi256_ps(idx), up_f));
        idx = _mm256_add_epi32(idx, eight);
    }
    if (outside) return -1;
    float btn[8], btf[8];
    int32_t bin[8], bif[8];
    _mm256_storeu_ps(btn, tn);
    _mm256_storeu_ps(btf, tf);
    _mm256_storeu_si256((__m256i*)bin, in);
    _mm256_storeu_si256((__m256i*)bif, iff);
    float t_near = -1e30f, t_far = 1e30f;
    int32_t near_idx = -1, far_idx = -1;
    for (int k = 0; k < 8; ++k) {
        if (bin[k] >= 0 && (btn[k] > t_near || (btn[k] == t_near && near_idx >= 0 && bin[k] < near_idx))) {
            t_near = btn[k];
            near_idx = bin[k];
        }
        if (bif[k] >= 0 && (btf[k] < t_far || (btf[k] == t_far && far_idx >= 0 && bif[k] < far_idx))) {
            t_far = btf[k];
            far_idx = bif[k];
        }
    }
    if (t_near > t_far) return -1;
    if (t_near > 1e-4f) { *t_out = t_near; return near_idx < 0 ? 0 : near_idx; }
    if (t_far > 1e-4f) { *t_out = t_far; return far_idx < 0 ? 0 : far_idx; }
    return -1;
}
#endif

/* The slab scan: the 8-wide twin where it is built, else the scalar one. */
static int32_t slab_scan(const slab_t* L, const ixg_facet_plane* planes, uint32_t n, const float o[3],
                         const float d[3], float* t_out) {
#ifdef IXO_SIMD
    if (L->n_pad) return intersect_simd(L, o, d, t_out);
#else
    (void)L;
#endif
    return ixo_intersect(planes, n, o, d, t_out);
}

int32_t ixo_slab_scan(const ixg_facet_plane* planes, uint32_t n, const float o[3], const float d[3], float* t_out) {
    static slab_t L; /* single-threaded test entry; the scene and metric paths own theirs */
    slab_fill(&L, planes, n);
    return slab_scan(&L, planes, n, o, d, t_out);
}

static int32_t scene_intersect(const scene_t* S, const float o[3], const float d[3], float* t_out) {
    return slab_scan(&S->slab, S->planes, S->n_planes, o, d, t_out);
}

typedef struct {
    float lambdas[NCH];
    stokes st[NCH];
    float radiance[NCH];
    float split[NCH];
    float path_pdf[NCH];
    float n_o[NCH];
    float alpha_o[NCH], alpha_e[NCH]; /* spectral_absorption of o / e bands */
    float alpha_b[NCH];               /* of the beta bands (trichroic biaxial) */
    int has_mis;        /* the light-sample MIS carry of an exterior frosted bounce ... */
    float mis_phase;    /* ... and its BSDF density, for the next bounce's miss */
    uint8_t compat[NCH];
    v3 org, dir, k;     /* ray origin, ray (Poynting) direction S, wave normal k */
    int inside;
    int is_extra;       /* eigenmode assigned at the latest entry (uniaxial) */
    int has_prev;
    v3 prev_plane;
    int escaped;
    v3 observer;
} path_t;

/* frosted_orthonormal_basis (scattering/mod.rs:110-115) and
 * cosine_weighted_hemisphere (scattering/frosted.rs): the azimuth's sine and
 * cosine correctly rounded. */
static v3 cosine_weighted_hemisphere(float u1, float u2, v3 n) {
    const float r = sqrtf(u1);
    const float theta = (2.0f * F32_PI) * u2;
    IXO_FLOPS(2);
    const float sin_t = (float)sin((double)theta), cos_t = (float)cos((double)theta);
    const v3 a = fabsf(n.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    const v3 t = vnormalize_or_zero(vsub(a, vmul(n, vdot(n, a))));
    const v3 b = vcross(n, t);
    return vnormalize_or_zero(vadd(vadd(vmul(t, r * cos_t), vmul(b, r * sin_t)), vmul(n, sqrtf(maxf(1.0f - u1, 0.0f)))));
}

/* The analytic sun (environment/sun.rs): its disc, cone sampling and pdf,
 * and next-event estimation at an exterior frosted event
 * (nee_contribution_frosted_exterior, scattering/frosted.rs:98-182) with the
 * light sample of sample_environment_for_nee (environment/nee.rs:81-120) --
 * DaylightSun's disc; the panorama's importance sampling is not carried. */
#define SUN_DISC_COS 0.9999889f
#define SUN_ONE_MINUS_COS 1.1086464e-5f
#define SUN_SOLID_ANGLE 6.965831e-5f

static v3 sun_cone_direction(v3 key, float u0, float u1) {
    const float omc = u0 * SUN_ONE_MINUS_COS;
    const float cos_t = 1.0f - omc;
    const float sin_t = sqrtf(maxf(omc * (2.0f - omc), 0.0f));
    const float phi = 2.0f * F32_PI * u1;
    const float sp = sin_cr(phi), cp = cos_cr(phi);
    const v3 a = fabsf(key.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    const v3 t = vnormalize_or_zero(vsub(a, vmul(key, vdot(key, a))));
    const v3 b = vcross(key, t);
    return vnormalize_or_zero(vadd(vadd(vmul(t, sin_t * cp), vmul(b, sin_t * sp)), vmul(key, cos_t)));
}

void ixo_sun_cone_direction(const float key[3], float u0, float u1, float out[3]) {
    const v3 d = sun_cone_direction(vfromr(key), u0, u1);
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}

/* balance_heuristic (scattering/mod.rs:102-105) */
static float balance_heuristic(float a, float b) {
    const float denom = a + b;
    return denom > 1e-12f ? a / denom : 0.0f;
}

/* environment_nee_pdf (environment/nee.rs:144-157): the sun disc's 1 / solid
 * angle under DaylightSun, 0 for every other studio model */
static float environment_nee_pdf(const scene_t* S, v3 dir) {
    if (S->hdr || S->env.model != MODEL_SUN) return 0.0f;
    return vdot(dir, S->env.key_dir) >= SUN_DISC_COS ? 1.0f / SUN_SOLID_ANGLE : 0.0f;
}

static void nee_frosted_exterior(const scene_t* S, path_t* P, v3 ext_normal, uint32_t seed, uint32_t bounce) {
    if (S->hdr || S->env.model != MODEL_SUN) return; /* no light sample */
    const float u0 = unit_draw(seed, bounce, FROSTED_NEE_ENV_DIR_U_STREAM);
    const float u1 = unit_draw(seed, bounce, FROSTED_NEE_ENV_DIR_V_STREAM);
    const v3 dir = sun_cone_direction(S->env.key_dir, u0, u1);
    const float pdf = 1.0f / SUN_SOLID_ANGLE;
    const float cos_light = vdot(dir, ext_normal);
    if (cos_light <= 0.0f) return;
    const float brdf_pdf = cos_light / F32_PI;
    const float mis = balance_heuristic(pdf, brdf_pdf);
    if (mis <= 0.0f) return;
    const float common = brdf_pdf * mis / pdf;
    for (int c = 0; c < NCH; ++c) {
        const float env = S->env.sun * (spec_power(&S->env, P->lambdas[c]) * S->env.exposure);
        P->radiance[c] = fmaf(st_intensity(P->st[c]) * common * env, 1.0f, P->radiance[c]);
    }
}

/* apply_frosted_bounce's exterior side (frosted.rs:331-341, 372-382): the
 * light sample, and the MIS carry the next bounce's miss weighs against */
/* The reference's enable_nee switch (transport/inner.rs:240-290), on in every
 * render; the tests turn it off to check the estimator without light samples. */
int ixo_enable_nee = 1;

static void frosted_exterior_light(const scene_t* S, path_t* P, v3 ext_normal, v3 new_dir, uint32_t seed,
                                   uint32_t bounce) {
    if (!ixo_enable_nee) return;
    nee_frosted_exterior(S, P, ext_normal, seed, bounce);
    P->has_mis = 1;
    P->mis_phase = vdot(new_dir, ext_normal) / F32_PI;
}

void ixo_cosine_hemisphere(float u1, float u2, const float n[3], float out[3]) {
    const v3 r = cosine_weighted_hemisphere(u1, u2, vfromr(n));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

static void ev_facet(ixg_path_events* ev, int32_t facet, uint32_t kind) {
    if (!ev) return;
    uint32_t h = ev->facet_hash;
    h = (h ^ (uint32_t)facet) * 16777619u;
    h = (h ^ kind) * 16777619u;
    ev->facet_hash = h;
}

/* stable_orthonormal_basis (optics/birefringence/helpers.rs) */
static void stable_basis(v3 n, v3* t, v3* b) {
    v3 a = fabsf(n.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    *t = vnormalize_or_zero(vsub(a, vmul(n, vdot(n, a))));
    *b = vcross(n, *t);
}

/* BirefringenceParams::ordinary_eigen_polarization (optics/birefringence/uniaxial.rs) */
static v3 ordinary_eigen(v3 k, v3 c) {
    v3 x = vcross(k, c);
    if (vlen_sq(x) > 1e-8f) return vnormalize(x);
    v3 t, b;
    stable_basis(vnormalize_or_zero(k), &t, &b);
    return t;
}

/* spectral_absorption (RT/absorption.rs:33): the bands' sum in order */
static float band_sum(const ixg_band* bands, int32_t n, float lambda) {
    float s = -0.0f;
    for (int32_t i = 0; i < n; ++i) s += ixo_band_evaluate(&bands[i], lambda);
    return s;
}

/* AbsorptionTensor3::quadratic_form (optics/birefringence/absorption_tensor.rs):
 * alpha . (axes^T e)^2 with axes = (a1, a2, c) and alpha = (ao, ao, ae) */
static float tensor_qf(const scene_t* S, float ao, float ae, v3 e) {
    float l0 = vdot(S->ta1, e), l1 = vdot(S->ta2, e), l2 = vdot(S->tcn, e);
    return (ao * (l0 * l0) + ao * (l1 * l1)) + ae * (l2 * l2);
}

/* apply_absorption (RT/absorption.rs:99-122) with
 * channel_absorption_alphas_assigned's isotropic-by-symmetry branch
 * (:205-222): the unpolarised mean of the two eigen-directions' quadratic
 * forms, Beer-Lambert through exp_f32x8 over the scaled segment length. */
static void apply_absorption(const scene_t* S, path_t* P, float path_len) {
    v3 ea = ordinary_eigen(P->k, S->c_raw);
    v3 eb = vnormalize_or_zero(vcross(P->k, ea));
    float scaled = path_len * S->mat->absorption_path_scale;
    float arg[NCH], tr[NCH];
    for (int c = 0; c < NCH; ++c) {
        float qa = tensor_qf(S, P->alpha_o[c], P->alpha_e[c], ea);
        float qb = tensor_qf(S, P->alpha_o[c], P->alpha_e[c], eb);
        arg[c] = -midpointf(qa, qb) * scaled;
    }
    exp_lane8(arg, tr);
    for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], tr[c]);
}

#include "ixo_uniaxial.h"
#include "ixo_biaxial.h"

/* sample_henyey_greenstein_direction (scattering/hg.rs:636-651); the
 * azimuth's sine and cosine correctly rounded. */
static v3 hg_direction(float u1, float u2, float g, v3 fwd) {
    float cos_theta;
    if (fabsf(g) < 1e-3f) {
        cos_theta = fmaf(2.0f, -u1, 1.0f);
    } else {
        const float one_minus_g2 = fmaf(g, -g, 1.0f);
        const float denom = fmaf(2.0f * g, u1, 1.0f - g);
        const float sq = one_minus_g2 / denom;
        cos_theta = fmaf(sq, -sq, fmaf(g, g, 1.0f)) / (2.0f * g);
    }
    cos_theta = clampf(cos_theta, -1.0f, 1.0f);
    const float sin_theta = sqrtf(maxf(fmaf(cos_theta, -cos_theta, 1.0f), 0.0f));
    const float phi = (2.0f * F32_PI) * u2;
    IXO_FLOPS(2);
    const float sin_p = (float)sin((double)phi), cos_p = (float)cos((double)phi);
    const v3 a = fabsf(fwd.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    const v3 t = vnormalize_or_zero(vsub(a, vmul(fwd, vdot(fwd, a))));
    const v3 b = vcross(fwd, t);
    return vnormalize_or_zero(vadd(vadd(vmul(t, sin_theta * cos_p), vmul(b, sin_theta * sin_p)), vmul(fwd, cos_theta)));
}

void ixo_hg_direction(float u1, float u2, float g, const float fwd[3], float out[3]) {
    const v3 r = hg_direction(u1, u2, g, vfromr(fwd));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* try_scatter_step / maybe_scatter_or_extinguish (scattering/hg.rs:91-160,
 * 435-600), isotropic and unzoned, next-event estimation off.  Returns 0
 * scattered (the iteration ends at the scatter point), 3 killed by the
 * roulette at the scatter, -1 reached the facet. */
static void scatter_alphas_assigned(const scene_t* S, path_t* P, float alpha[NCH]);

static int scatter_step(const scene_t* S, path_t* P, uint32_t bounce, uint32_t seed, float hit_t,
                        ixg_path_events* ev) {
    float alpha[NCH];
    if (S->aniso) { /* channel_absorption_alphas_assigned, the assigned mode (absorption.rs:224-249) */
        scatter_alphas_assigned(S, P, alpha);
    } else { /* its isotropic-by-symmetry arm (absorption.rs:206-222) */
        v3 ea = ordinary_eigen(P->k, S->c_raw);
        v3 eb = vnormalize_or_zero(vcross(P->k, ea));
        for (int c = 0; c < NCH; ++c) {
            if (!S->absorbing) { alpha[c] = 0.0f; continue; }
            const float qa = tensor_qf(S, P->alpha_o[c], P->alpha_e[c], ea);
            const float qb = tensor_qf(S, P->alpha_o[c], P->alpha_e[c], eb);
            IXO_FLOPS(2);
            alpha[c] = (float)(((double)qa + (double)qb) / 2.0);
        }
    }
    const float path_scale = S->mat->absorption_path_scale;
    const float sigma_s = S->mat->scattering_sigma_s / path_scale;
    const float sigma_t_hero = alpha[0] + sigma_s;
    const float one_minus_u = maxf(1.0f - unit_draw(seed, bounce, DISTANCE_SAMPLE_STREAM), 1e-7f);
    const float t_free = -logf(one_minus_u) / sigma_t_hero;
    const float hit_t_scaled = hit_t * path_scale;
    if (t_free < hit_t_scaled) {
        const float pdf_hero = sigma_t_hero * one_minus_u;
        for (int c = 0; c < NCH; ++c) {
            const float sigma_t = alpha[c] + sigma_s;
            const float tr = ixo_exp_lane(-(alpha[c] + sigma_s) * t_free);
            const float weight = tr * sigma_s / pdf_hero;
            P->st[c] = (stokes){st_intensity(P->st[c]) * weight, 0.0f, 0.0f, 0.0f};
            P->path_pdf[c] *= sigma_t * tr;
        }
        const float u1 = unit_draw(seed, bounce, PHASE_DIR_U_STREAM);
        const float u2 = unit_draw(seed, bounce, PHASE_DIR_V_STREAM);
        const v3 old_dir = P->dir;
        const v3 new_dir = hg_direction(u1, u2, S->mat->scattering_g, old_dir);
        P->org = vadd(P->org, vmul(old_dir, t_free / path_scale));
        P->dir = new_dir;
        P->k = new_dir;
        P->has_prev = 0;
        ev_facet(ev, 0xFFFF, 7);
        if (bounce > 4) { /* apply_russian_roulette (transport/bounce.rs:355-375) */
            float mx = 0.0f;
            for (int c = 0; c < NCH; ++c) mx = maxf(mx, st_intensity(P->st[c]));
            const float q = clampf(mx, 0.05f, 1.0f);
            if (unit_draw(seed, bounce, RUSSIAN_ROULETTE_STREAM) > q) return 3;
            for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], 1.0f / q);
            for (int c = 0; c < NCH; ++c) P->split[c] /= q;
        }
        return 0;
    }
    float survive[NCH];
    for (int c = 0; c < NCH; ++c) survive[c] = ixo_exp_lane(-(alpha[c] + sigma_s) * hit_t_scaled);
    const float hero = maxf(survive[0], 1e-30f);
    for (int c = 0; c < NCH; ++c) {
        P->st[c] = st_scale(P->st[c], survive[c] / hero);
        P->path_pdf[c] *= survive[c];
    }
    return -1;
}

static float tir_phase_delta(float n1k, float cos_i, float sin_i) { /* refraction/tir.rs:33-37 */
    float tan_half = (cos_i * sqrtf(maxf(fmaf(n1k * n1k * sin_i, sin_i, -1.0f), 0.0f))) /
                     maxf(n1k * sin_i * sin_i, 1e-6f);
    return 2.0f * atanf(tan_half);
}

/* EnvironmentMap::radiance_rgb (renderer/env_map.rs:303-417), the angles
 * correctly rounded. */
static float rem_euclid_f(float x, float m) {
    float r = fmodf(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
}
static void hdr_rgb(const ixg_hdr_map* H, v3 dir, float out[3]) {
    const v3 d = vnormalize_or_zero(dir);
    IXO_FLOPS(2);
    const float theta = (float)acos((double)clampf(d.y, -1.0f, 1.0f));
    const float phi = (float)atan2((double)d.x, (double)d.z);
    const float v = theta / F32_PI;
    const float u = rem_euclid_f(phi / (2.0f * F32_PI), 1.0f);
    if (H->width == 1 && H->height == 1) {
        out[0] = H->rgb[0]; out[1] = H->rgb[1]; out[2] = H->rgb[2];
        return;
    }
    const float fx = fmaf(rem_euclid_f(u, 1.0f), (float)H->width, -0.5f);
    const float fy = fmaf(clampf(v, 0.0f, 1.0f), (float)H->height, -0.5f);
    const float x0 = floorf(fx), y0 = floorf(fy);
    const float tx = fx - x0, ty = fy - y0;
    const int64_t w = H->width, h = H->height, xc = (int64_t)x0, yc = (int64_t)y0;
    const int64_t x0i = ((xc % w) + w) % w, x1i = (((xc + 1) % w) + w) % w;
    const int64_t y0i = yc < 0 ? 0 : (yc > h - 1 ? h - 1 : yc);
    const int64_t y1i = yc + 1 < 0 ? 0 : (yc + 1 > h - 1 ? h - 1 : yc + 1);
    const ixo_abi_f32* p00 = H->rgb + (y0i * w + x0i) * 3;
    const ixo_abi_f32* p10 = H->rgb + (y0i * w + x1i) * 3;
    const ixo_abi_f32* p01 = H->rgb + (y1i * w + x0i) * 3;
    const ixo_abi_f32* p11 = H->rgb + (y1i * w + x1i) * 3;
    for (int c = 0; c < 3; ++c) {
        const float top = fmaf(p10[c], tx, p00[c] * (1.0f - tx));
        const float bottom = fmaf(p11[c], tx, p01[c] * (1.0f - tx));
        out[c] = fmaf(bottom, ty, top * (1.0f - ty));
    }
}
/* rgb_to_spectral_radiance (renderer/env_map_spectrum.rs:84-125). */
static float asym_gauss(float x, float mu, float lo, float hi) {
    const float t = (x - mu) / (x < mu ? lo : hi);
    return expf(-0.5f * t * t);
}
static const float NEUTRAL_TO_BUMP[3][3] = {{0.82398987f, -0.30506778f, 0.26391557f},
                                            {-0.2971578f, 1.2528949f, -0.4611476f},
                                            {0.078837544f, -0.122806296f, 1.2317826f}};
static const float CHROMA_TO_BUMP[3][3] = {{1.0995291f, 0.08711245f, 0.01771266f},
                                           {0.007455747f, 1.3345385f, 0.014020192f},
                                           {0.02487475f, 0.058899768f, 1.2609048f}};
static float row_dot(const float m[3], const float v[3]) { return fmaf(m[0], v[0], fmaf(m[1], v[1], m[2] * v[2])); }
float ixo_rgb_to_spectral(const float rgb_in[3], float lambda) {
    const float r = maxf(rgb_in[0], 0.0f), g = maxf(rgb_in[1], 0.0f), b = maxf(rgb_in[2], 0.0f);
    const float w = minf(minf(r, g), b);
    const float chroma[3] = {r - w, g - w, b - w}, grey[3] = {w, w, w};
    const float neutral = fmaf(row_dot(NEUTRAL_TO_BUMP[0], grey), asym_gauss(lambda, 615.0f, 45.0f, 65.0f),
                               fmaf(row_dot(NEUTRAL_TO_BUMP[1], grey), asym_gauss(lambda, 545.0f, 45.0f, 45.0f),
                                    row_dot(NEUTRAL_TO_BUMP[2], grey) * asym_gauss(lambda, 465.0f, 40.0f, 45.0f)));
    const float saturated = fmaf(row_dot(CHROMA_TO_BUMP[0], chroma), asym_gauss(lambda, 635.0f, 28.0f, 39.2f),
                                 fmaf(row_dot(CHROMA_TO_BUMP[1], chroma), asym_gauss(lambda, 540.0f, 28.0f, 28.0f),
                                      row_dot(CHROMA_TO_BUMP[2], chroma) * asym_gauss(lambda, 450.0f, 25.2f, 28.0f)));
    return maxf(neutral + saturated, 0.0f);
}

static float env_channel(const scene_t* S, v3 dir, float lambda, v3 observer) {
    if (S->hdr) {
        float rgb[3];
        hdr_rgb(S->hdr, dir, rgb);
        return ixo_rgb_to_spectral(rgb, lambda);
    }
    dir_light L = direction_lighting(&S->env, dir, observer);
    return light_radiance(&L, spec_power(&S->env, lambda), S->env.exposure);
}

/* The hit point and geometry of one facet event. */
typedef struct {
    float cos_i, sin_i, n1, n2, sin2_t;
    float n1_ch[NCH], n2_ch[NCH], sin2_t_ch[NCH];
} geo_t;

static void fresnel_rs_rp(float n1, float n2, float cos_i, float cos_t, float* rs, float* rp) {
    *rs = fmaf(n2, -cos_t, n1 * cos_i) / fmaf(n2, cos_t, n1 * cos_i);
    *rp = fmaf(n1, -cos_t, n2 * cos_i) / fmaf(n1, cos_t, n2 * cos_i);
}

static void facet_event_aniso(const scene_t* S, path_t* P, uint32_t bounce, uint32_t seed, float t_hit,
                              int32_t facet, ixg_path_events* ev, int* stop);
static void facet_event_biax(const scene_t* S, path_t* P, uint32_t bounce, uint32_t seed, float t_hit,
                             int32_t facet, ixg_path_events* ev, int* stop);

/* shading_normal_near_edge (RT/intersect.rs:196-233): within rounding_radius
 * of the nearest other plane the facet normal eases (smoothstep) toward the
 * bisector of the two normals; radius <= 0 returns it unchanged. */
static v3 shading_normal_near_edge(const ixg_facet_plane* planes, uint32_t n, v3 hit, int32_t facet, v3 hn,
                                   float radius) {
    if (radius <= 0.0f) return hn;
    float nearest = INFINITY;
    v3 nn = hn;
    for (uint32_t i = 0; i < n; ++i) {
        if ((int32_t)i == facet) continue;
        const v3 pn = vfrom(planes[i].normal);
        const float dist = -(planes[i].d + vdot(pn, hit));
        if (dist < nearest) { nearest = dist; nn = pn; }
    }
    if (nearest >= radius) return hn;
    const float t = clampf(1.0f - nearest / radius, 0.0f, 1.0f);
    const float st = t * t * fmaf(2.0f, -t, 3.0f);
    const v3 bis = vnormalize_or_zero(vadd(hn, nn));
    return vnormalize_or_zero(vadd(vmul(hn, 1.0f - st), vmul(bis, st)));
}

void ixo_shading_normal_near_edge(const ixg_facet_plane* planes, uint32_t n, const float hit[3], int32_t facet,
                                  const float hn[3], float radius, float out[3]) {
    const v3 r = shading_normal_near_edge(planes, n, vfromr(hit), facet, vfromr(hn), radius);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

static void facet_event(const scene_t* S, path_t* P, uint32_t bounce, uint32_t seed, float t_hit,
                        int32_t facet, ixg_path_events* ev, int* stop) {
    if (S->biax) {
        facet_event_biax(S, P, bounce, seed, t_hit, facet, ev, stop);
        return;
    }
    if (S->aniso) {
        facet_event_aniso(S, P, bounce, seed, t_hit, facet, ev, stop);
        return;
    }
    v3 hit_point = vadd(P->org, vmul(P->dir, t_hit));
    v3 normal = shading_normal_near_edge(S->planes, S->n_planes, hit_point, facet, vfrom(S->planes[facet].normal),
                                         S->mat->edge_rounding_radius); /* bounce_body.rs:620-629 */
    /* rotate_stokes_to_plane_of_incidence (RT/absorption.rs:69-87) */
    v3 cur_plane = vnormalize_or_zero(vcross(P->k, normal));
    if (P->has_prev && vlen_sq(cur_plane) > 1e-6f && vlen_sq(P->prev_plane) > 1e-6f) {
        float cos_psi = clampf(vdot(P->prev_plane, cur_plane), -1.0f, 1.0f);
        float sin_psi = vdot(vcross(P->prev_plane, cur_plane), vnormalize_or_zero(P->k));
        float psi = atan2_cr(sin_psi, cos_psi);
        mat4 R = mueller_frame_rotation(psi);
        for (int c = 0; c < NCH; ++c) P->st[c] = st_apply(&R, P->st[c]);
    }
    P->prev_plane = cur_plane;
    P->has_prev = 1;
    /* apply_interior_segment (transport/inner.rs:161-178); without bands
     * every channel transmits exp(-0) == 1 exactly, a no-op. */
    if (P->inside && S->absorbing && !(S->mat->scattering_sigma_s > 0.0f)) apply_absorption(S, P, t_hit);
    if (P->inside) normal = vneg(normal);

    /* compute_bounce_refraction_geometry (refraction/geometry.rs:306-425), isotropic */
    geo_t G;
    G.cos_i = clampf(vdot(vneg(P->k), normal), 0.0f, 1.0f);
    G.sin_i = sqrtf(maxf(fmaf(G.cos_i, -G.cos_i, 1.0f), 0.0f));
    float nh = P->n_o[0];
    G.n1 = P->inside ? nh : 1.0f;
    G.n2 = P->inside ? 1.0f : nh;
    float eta = G.n1 / G.n2;
    G.sin2_t = eta * eta * fmaf(G.cos_i, -G.cos_i, 1.0f);
    for (int c = 0; c < NCH; ++c) {
        float n1k = P->inside ? P->n_o[c] : 1.0f;
        float n2k = P->inside ? 1.0f : P->n_o[c];
        float etak = n1k / n2k;
        G.n1_ch[c] = n1k;
        G.n2_ch[c] = n2k;
        G.sin2_t_ch[c] = etak * etak * fmaf(G.cos_i, -G.cos_i, 1.0f);
    }

    v3 new_k, new_s;
    int new_inside = P->inside;
    v3 kk = P->k;
    const int frosted = S->finish != NULL && S->finish[facet] != 0;
    if (frosted) {
        /* apply_frosted_bounce (scattering/frosted.rs), next-event estimation off:
         * a cosine-weighted hemisphere on the reflect (or, transmitted, the far)
         * side, the Stokes vectors depolarised, the hero Fresnel draw */
        const float u1 = unit_draw(seed, bounce, FROSTED_DIR_U_STREAM);
        const float u2 = unit_draw(seed, bounce, FROSTED_DIR_V_STREAM);
        if (G.sin2_t > 1.0f) {
            new_k = cosine_weighted_hemisphere(u1, u2, normal);
            for (int c = 0; c < NCH; ++c) P->st[c] = st_unpolarized(P->st[c]);
            ev_facet(ev, facet, 4);
        } else {
            const float cos_t = sqrtf(maxf(1.0f - G.sin2_t, 0.0f));
            float rs, rp;
            fresnel_rs_rp(G.n1, G.n2, G.cos_i, cos_t, &rs, &rp);
            const float r_unpol = clampf(0.5f * fmaf(rp, rp, rs * rs), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            if (unit_draw(seed, bounce, FRESNEL_BRANCH_STREAM) < r_unpol) {
                new_k = cosine_weighted_hemisphere(u1, u2, normal);
                for (int c = 0; c < NCH; ++c) {
                    P->st[c] = st_unpolarized(P->st[c]);
                    P->path_pdf[c] *= r_unpol;
                }
                if (!P->inside) frosted_exterior_light(S, P, normal, new_k, seed, bounce);
                ev_facet(ev, facet, 5);
            } else {
                new_k = cosine_weighted_hemisphere(u1, u2, vneg(normal));
                const float t_unpol = 1.0f - r_unpol;
                for (int c = 0; c < NCH; ++c) {
                    P->st[c] = st_unpolarized(P->st[c]);
                    P->path_pdf[c] *= t_unpol;
                }
                if (P->inside) frosted_exterior_light(S, P, vneg(normal), new_k, seed, bounce);
                new_inside = !P->inside;
                ev_facet(ev, facet, 6);
            }
        }
        new_s = new_k;
    } else if (G.sin2_t > 1.0f) {
        /* apply_tir_bounce (refraction/tir.rs:56-160), scalar branch */
        if (ev) ev->n_tir++;
        ev_facet(ev, facet, 1);
        for (int c = 0; c < NCH; ++c) {
            float n1k = G.n1_ch[c], n2k = G.n2_ch[c];
            if (G.sin2_t_ch[c] > 1.0f) {
                mat4 M = mueller_tir_retardation(tir_phase_delta(n1k, G.cos_i, G.sin_i));
                P->st[c] = st_apply(&M, P->st[c]);
            } else {
                float cos_t = sqrtf(maxf(1.0f - G.sin2_t_ch[c], 0.0f));
                float rs, rp;
                fresnel_rs_rp(n1k, n2k, G.cos_i, cos_t, &rs, &rp);
                mat4 M = mueller_fresnel_reflection(rs, rp);
                P->st[c] = st_apply(&M, P->st[c]);
                float ru = clampf(0.5f * fmaf(rp, rp, rs * rs), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                P->path_pdf[c] *= ru;
            }
        }
        new_k = vsub(kk, vmul(normal, 2.0f * vdot(kk, normal)));
        new_s = new_k;
    } else {
        /* apply_partial_fresnel_bounce (refraction/dispatch.rs:307-380) */
        float cos_t = sqrtf(maxf(1.0f - G.sin2_t, 0.0f));
        float rs, rp;
        fresnel_rs_rp(G.n1, G.n2, G.cos_i, cos_t, &rs, &rp);
        float r_unpol = clampf(0.5f * fmaf(rp, rp, rs * rs), R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        float rng = unit_draw(seed, bounce, FRESNEL_BRANCH_STREAM);
        if (rng < r_unpol) {
            /* apply_partial_reflect_bounce (reflect_refract.rs:33-71) */
            if (ev) ev->n_reflect++;
            ev_facet(ev, facet, 2);
            for (int c = 0; c < NCH; ++c) {
                float n1k = G.n1_ch[c], n2k = G.n2_ch[c];
                mat4 M;
                if (G.sin2_t_ch[c] > 1.0f) {
                    M = mueller_tir_retardation(tir_phase_delta(n1k, G.cos_i, G.sin_i));
                } else {
                    float ct = sqrtf(maxf(1.0f - G.sin2_t_ch[c], 0.0f));
                    float a, b;
                    fresnel_rs_rp(n1k, n2k, G.cos_i, ct, &a, &b);
                    float ru = clampf(0.5f * fmaf(b, b, a * a), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                    P->path_pdf[c] *= ru;
                    M = mueller_fresnel_reflection(a, b);
                }
                P->st[c] = st_scale(st_apply(&M, P->st[c]), 1.0f / r_unpol);
            }
            new_k = vsub(kk, vmul(normal, 2.0f * vdot(kk, normal)));
            new_s = new_k;
        } else {
            /* apply_refract_bounce (reflect_refract.rs:121-234), isotropic */
            if (ev) ev->n_refract++;
            ev_facet(ev, facet, 3);
            float eta_dir = G.n1 / G.n2;
            float s2 = minf(eta_dir * eta_dir * fmaf(G.cos_i, -G.cos_i, 1.0f), 1.0f);
            float ct_dir = sqrtf(maxf(1.0f - s2, 0.0f));
            v3 refr = vnormalize(vadd(vmul(kk, eta_dir), vmul(normal, fmaf(eta_dir, G.cos_i, -ct_dir))));
            v3 dirs[NCH];
            int has_dir[NCH], hero_match[NCH];
            for (int c = 0; c < NCH; ++c) {
                /* apply_refract_channel (reflect_refract.rs:517-591) */
                stokes orig = P->st[c];
                float n1k = G.n1_ch[c], n2k = G.n2_ch[c];
                float r = n1k / n2k;
                float s2k = r * r * fmaf(G.cos_i, -G.cos_i, 1.0f);
                if (s2k > 1.0f) {
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    P->path_pdf[c] = 0.0f;
                    has_dir[c] = 0;
                    hero_match[c] = 0;
                    continue;
                }
                float ctk = sqrtf(maxf(1.0f - s2k, 0.0f));
                float ek = n1k / n2k;
                v3 dk = vnormalize(vadd(vmul(kk, ek), vmul(normal, fmaf(ek, G.cos_i, -ctk))));
                int match = vdot(dk, refr) >= DIRECTION_MATCH_COS_TOL;
                /* transmission of channel c (exit_split.rs:69-101) */
                float num = 2.0f * n1k * G.cos_i;
                float t_s = num / fmaf(n2k, ctk, n1k * G.cos_i);
                float t_p = num / fmaf(n1k, ctk, n2k * G.cos_i);
                mat4 T = mueller_fresnel_transmission(n1k, n2k, G.cos_i, ctk, t_s, t_p);
                stokes transmitted = st_scale(st_apply(&T, orig), 1.0f / (1.0f - r_unpol));
                float a, b;
                fresnel_rs_rp(n1k, n2k, G.cos_i, ctk, &a, &b);
                float ruk = clampf(0.5f * fmaf(b, b, a * a), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                if (match) {
                    P->st[c] = transmitted;
                    P->path_pdf[c] *= 1.0f - ruk;
                } else {
                    /* apply_channel_chromatic_termination (reflect_refract.rs:444-515) */
                    float prefix = P->path_pdf[c];
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    P->path_pdf[c] = prefix * (1.0f - ruk);
                    if (P->inside && st_intensity(orig) > 0.0f) {
                        /* try_split_exit_channel (exit_split.rs:128-163) */
                        float po[3], pd[3], tt;
                        v3 probe_o = vadd(hit_point, vmul(dk, 1e-4f));
                        po[0] = probe_o.x; po[1] = probe_o.y; po[2] = probe_o.z;
                        pd[0] = dk.x; pd[1] = dk.y; pd[2] = dk.z;
                        if (ev) ev->n_probe++;
                        if (scene_intersect(S, po, pd, &tt) < 0) {
                            float env = env_channel(S, dk, P->lambdas[c], P->observer);
                            P->split[c] = fmaf(maxf(st_intensity(transmitted), 0.0f) * 1.0f, env,
                                               P->split[c]);
                        }
                    }
                }
                dirs[c] = dk;
                has_dir[c] = 1;
                hero_match[c] = match;
            }
            if (!P->inside) { /* narrow_compat (refraction/context.rs:219-242) */
                for (int a = 0; a < NCH; ++a)
                    for (int b = a + 1; b < NCH; ++b) {
                        int m;
                        if (a == 0) m = hero_match[b];
                        else if (has_dir[a] && has_dir[b]) m = vdot(dirs[a], dirs[b]) >= DIRECTION_MATCH_COS_TOL;
                        else m = 1;
                        if (!m) {
                            P->compat[a] &= (uint8_t)~(1u << b);
                            P->compat[b] &= (uint8_t)~(1u << a);
                        }
                    }
            }
            new_k = refr;
            new_s = refr;
            new_inside = !P->inside;
        }
    }
    /* dispatch_bounce tail (transport/bounce.rs:200-238) */
    P->dir = new_s;
    P->org = vadd(hit_point, vmul(new_s, 1e-4f));
    P->k = new_k;
    P->inside = new_inside;
    /* surface glare (bounce_body.rs:686-692), not after a frosted facet */
    if (bounce == 0 && !P->inside && S->surface_glare < 1.0f && !frosted)
        for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], S->surface_glare);
    /* apply_russian_roulette (transport/bounce.rs:355-375) */
    if (bounce > 4) {
        float mx = 0.0f;
        for (int c = 0; c < NCH; ++c) mx = maxf(mx, st_intensity(P->st[c]));
        float q = clampf(mx, 0.05f, 1.0f);
        float rr = unit_draw(seed, bounce, RUSSIAN_ROULETTE_STREAM);
        if (rr > q) { *stop = 1; return; }
        for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], 1.0f / q);
        for (int c = 0; c < NCH; ++c) P->split[c] /= q;
    }
}

/* ------------------------------------------------------------------ */
/* Uniaxial facet event (RT/refraction/{geometry,dispatch,tir,         */
/* uniaxial_entry,uniaxial_internal,reflect_refract,exit_split}.rs,     */
/* RT/transport/bounce.rs)                                              */
/* ------------------------------------------------------------------ */
#define BIREFRINGENT_SPLIT_STREAM 0x2545F491u
#define MODE_COUPLING_STREAM 0xCC9E2D51u

/* GemMaterial::extraordinary_index_at (IX/optics/materials/optics.rs:42-47) */
static float e_index_at(const scene_t* S, float lambda, float n_o) {
    return S->mat->has_e_dispersion ? ixo_dispersion_evaluate(&S->mat->e_dispersion, lambda)
                                    : n_o + S->mat->birefringence_delta;
}

/* theta_c_for_bounce (refraction/geometry.rs:75-120), uniaxial: a two-step
 * fixed point for the refracted wave's angle to the optic axis at entry. */
static float theta_c_for_bounce(const scene_t* S, const path_t* P, v3 normal, v3 k, float cos_i, int inside,
                                float n_o_seed) {
    v3 c = S->c_raw;
    if (!inside) {
        float n_e_seed = e_index_at(S, P->lambdas[0], n_o_seed);
        float n_guess = n_o_seed, theta = 0.0f;
        for (int it = 0; it < 2; ++it) {
            float eg = 1.0f / n_guess;
            float s2g = eg * eg * fmaf(cos_i, -cos_i, 1.0f);
            if (s2g > 1.0f) break;
            float ctg = sqrtf(maxf(1.0f - s2g, 0.0f));
            v3 wd = vnormalize(vadd(vmul(k, eg), vmul(normal, fmaf(eg, cos_i, -ctg))));
            theta = TF(acosf(fabsf(clampf(vdot(wd, c), -1.0f, 1.0f))), vdot(wd, c));
            n_guess = effective_extraordinary_index(n_o_seed, n_e_seed, theta);
        }
        return theta;
    }
    return TF(acosf(fabsf(clampf(vdot(k, c), -1.0f, 1.0f))), vdot(k, c));
}

/* entry_eigenmode_selection (refraction/uniaxial_entry.rs:50-75): the
 * Malus-law probability of the ordinary eigenmode and its doubled azimuth. */
static int entry_eigenmode_selection(v3 c, v3 cpn, v3 k, stokes h, float* p_o, float* c2, float* s2) {
    if (h.i <= 1e-7f || vlen_sq(cpn) <= 1e-6f) return 0;
    if (fmaf(h.q, h.q, h.u * h.u) <= 1e-12f) return 0;
    v3 o_hat = ordinary_eigen(k, c);
    v3 p_hat = vcross(k, cpn);
    float sc = vdot(o_hat, cpn), pc = vdot(o_hat, p_hat);
    *c2 = fmaf(sc, sc, -(pc * pc));
    *s2 = 2.0f * sc * pc;
    *p_o = clampf(0.5f + 0.5f * fmaf(h.q, *c2, h.u * *s2) / h.i, 0.0f, 1.0f);
    return 1;
}

/* apply_absorption with the assigned-mode coefficient
 * (RT/absorption.rs:224-249): alpha = e_mode . A . e_mode for the eigenmode
 * the path entered in. */
static void apply_absorption_assigned(const scene_t* S, path_t* P, float path_len) {
    float n_o_h = P->n_o[0];
    float n_e_h = e_index_at(S, P->lambdas[0], n_o_h);
    v3 e = assigned_mode_e_field_uniaxial(P->k, S->c_raw, P->is_extra, n_o_h, n_e_h);
    float scaled = path_len * S->mat->absorption_path_scale;
    float arg[NCH], tr[NCH];
    for (int c = 0; c < NCH; ++c) arg[c] = -tensor_qf(S, P->alpha_o[c], P->alpha_e[c], e) * scaled;
    exp_lane8(arg, tr);
    for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], tr[c]);
}

/* try_split_exit_channel (refraction/exit_split.rs:128-163) */
static void split_probe(const scene_t* S, path_t* P, v3 hit, int c, v3 dk, float transmitted_i,
                        ixg_path_events* ev) {
    v3 o = vadd(hit, vmul(dk, 1e-4f));
    float po[3] = {o.x, o.y, o.z}, pd[3] = {dk.x, dk.y, dk.z}, tt;
    if (ev) ev->n_probe++;
    if (scene_intersect(S, po, pd, &tt) >= 0) return;
    float env = env_channel(S, dk, P->lambdas[c], P->observer);
    P->split[c] = fmaf(maxf(transmitted_i, 0.0f) * 1.0f, env, P->split[c]);
}

/* narrow_compat (refraction/context.rs:219-242) */
static void narrow_compat_dirs(path_t* P, const v3* dirs, const int* has_dir, const int* hero_match) {
    for (int a = 0; a < NCH; ++a)
        for (int b = a + 1; b < NCH; ++b) {
            int m;
            if (a == 0) m = hero_match[b];
            else if (has_dir[a] && has_dir[b]) m = vdot(dirs[a], dirs[b]) >= DIRECTION_MATCH_COS_TOL;
            else m = 1;
            if (!m) {
                P->compat[a] &= (uint8_t)~(1u << b);
                P->compat[b] &= (uint8_t)~(1u << a);
            }
        }
}

typedef struct {
    float cos_i, sin_i, n1, n2, sin2_t, n_o_hero, n_e_hero;
    float n1_ch[NCH], n2_ch[NCH], sin2_t_ch[NCH];
    uframe F;
} ageo_t;

static v3 refract_dir(v3 k, v3 normal, float eta, float cos_i, float cos_t) {
    return vnormalize(vadd(vmul(k, eta), vmul(normal, fmaf(eta, cos_i, -cos_t))));
}

/* poynting_dir_for_mode (refraction/geometry.rs) */
static v3 poynting_for_mode(const scene_t* S, const ageo_t* G, v3 k, int inside, int is_extra) {
    if (!inside || !is_extra) return k;
    return extraordinary_poynting_dir(k, S->c_raw, G->n_o_hero, G->n_e_hero);
}

/* apply_frosted_bounce (scattering/frosted.rs:265-390), next-event estimation
 * off, for an anisotropic stone: the bounce geometry's hero indices decide TIR
 * and the Fresnel draw, the new direction is a cosine-weighted hemisphere on
 * the reflect or far side with every channel depolarised, and a transmission
 * into the stone draws its eigenmode 50/50.  *upd receives the entry's
 * eigenmode (or stays -1). */
static v3 frosted_arm_aniso(const scene_t* S, path_t* P, uint32_t bounce, uint32_t seed, int32_t facet, v3 normal,
                            float cos_i, float n1, float n2, float sin2_t, int* new_inside, int* upd,
                            ixg_path_events* ev) {
    const float u1 = unit_draw(seed, bounce, FROSTED_DIR_U_STREAM);
    const float u2 = unit_draw(seed, bounce, FROSTED_DIR_V_STREAM);
    if (sin2_t > 1.0f) {
        for (int c = 0; c < NCH; ++c) P->st[c] = st_unpolarized(P->st[c]);
        ev_facet(ev, facet, 4);
        return cosine_weighted_hemisphere(u1, u2, normal);
    }
    const float cos_t = sqrtf(maxf(1.0f - sin2_t, 0.0f));
    float rs, rp;
    fresnel_rs_rp(n1, n2, cos_i, cos_t, &rs, &rp);
    const float r_unpol = clampf(0.5f * fmaf(rp, rp, rs * rs), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
    if (unit_draw(seed, bounce, FRESNEL_BRANCH_STREAM) < r_unpol) {
        const v3 dir = cosine_weighted_hemisphere(u1, u2, normal);
        for (int c = 0; c < NCH; ++c) {
            P->st[c] = st_unpolarized(P->st[c]);
            P->path_pdf[c] *= r_unpol;
        }
        if (!P->inside) frosted_exterior_light(S, P, normal, dir, seed, bounce);
        ev_facet(ev, facet, 5);
        return dir;
    }
    const v3 dir = cosine_weighted_hemisphere(u1, u2, vneg(normal));
    const float t_unpol = 1.0f - r_unpol;
    for (int c = 0; c < NCH; ++c) {
        P->st[c] = st_unpolarized(P->st[c]);
        P->path_pdf[c] *= t_unpol;
    }
    if (P->inside) frosted_exterior_light(S, P, vneg(normal), dir, seed, bounce);
    if (!P->inside) *upd = unit_draw(seed, bounce, BIREFRINGENT_SPLIT_STREAM) < 0.5f;
    *new_inside = !P->inside;
    ev_facet(ev, facet, 6);
    return dir;
}

/* compute_uniaxial_exit_transmission (refraction/uniaxial_internal.rs) */
static stokes uniaxial_exit_transmission(const internal_sol* sol, float r_branch, float inc_i, float* i_unit) {
    float flux_inc = maxf(sol->flux_inc, 1e-12f);
    cplx ts = cscale(sol->t_s, sqrtf(sol->flux_ts / flux_inc));
    cplx tp = cscale(sol->t_p, sqrtf(sol->flux_tp / flux_inc));
    float iu = cnorm2(ts) + cnorm2(tp);
    float qu = cnorm2(ts) - cnorm2(tp);
    cplx cr = cmul(ts, cconj(tp));
    float uu = 2.0f * cr.re, vu = -2.0f * cr.im;
    stokes t = {inc_i * iu, inc_i * qu, inc_i * uu, inc_i * vu};
    *i_unit = iu;
    return st_scale(t, 1.0f / (1.0f - r_branch));
}

/* One facet event of a uniaxial material: the bounce body's facet_event
 * (transport/inner/bounce_body.rs:611-707) with compute_bounce_refraction_geometry's
 * anisotropic indices, dispatch_bounce's TIR / partial-Fresnel arms
 * (transport/bounce.rs:108-238, refraction/dispatch.rs) and the internal
 * o<->e re-coupling draw (transport/bounce.rs:439-500). */
static void facet_event_aniso(const scene_t* S, path_t* P, uint32_t bounce, uint32_t seed, float t_hit,
                              int32_t facet, ixg_path_events* ev, int* stop) {
    const ixg_material* M = S->mat;
    const v3 cax = S->c_raw;
    v3 hit_point = vadd(P->org, vmul(P->dir, t_hit));
    v3 normal = shading_normal_near_edge(S->planes, S->n_planes, hit_point, facet, vfrom(S->planes[facet].normal),
                                         M->edge_rounding_radius);
    v3 cur_plane = vnormalize_or_zero(vcross(P->k, normal));
    if (P->has_prev && vlen_sq(cur_plane) > 1e-6f && vlen_sq(P->prev_plane) > 1e-6f) {
        float cos_psi = clampf(vdot(P->prev_plane, cur_plane), -1.0f, 1.0f);
        float sin_psi = vdot(vcross(P->prev_plane, cur_plane), vnormalize_or_zero(P->k));
        mat4 R = mueller_frame_rotation(atan2_cr(sin_psi, cos_psi));
        for (int c = 0; c < NCH; ++c) P->st[c] = st_apply(&R, P->st[c]);
    }
    P->prev_plane = cur_plane;
    P->has_prev = 1;
    if (P->inside && S->absorbing && !(S->mat->scattering_sigma_s > 0.0f)) apply_absorption_assigned(S, P, t_hit);
    if (P->inside) normal = vneg(normal);

    const v3 k = P->k;
    const int inside = P->inside, is_extra = P->is_extra;
    ageo_t G;
    G.cos_i = clampf(vdot(vneg(k), normal), 0.0f, 1.0f);
    G.sin_i = sqrtf(maxf(fmaf(G.cos_i, -G.cos_i, 1.0f), 0.0f));
    float theta_c = theta_c_for_bounce(S, P, normal, k, G.cos_i, inside, P->n_o[0]);
    float n_eff[NCH], n_e_ch[NCH];
    for (int c = 0; c < NCH; ++c) {
        n_e_ch[c] = e_index_at(S, P->lambdas[c], P->n_o[c]);
        n_eff[c] = effective_extraordinary_index(P->n_o[c], n_e_ch[c], theta_c);
    }
    G.n_o_hero = P->n_o[0];
    G.n_e_hero = G.n_o_hero + M->birefringence_delta;
    const float* n_med = (inside && !is_extra) ? P->n_o : n_eff;
    G.n1 = inside ? n_med[0] : 1.0f;
    G.n2 = inside ? 1.0f : n_med[0];
    float eta = G.n1 / G.n2;
    G.sin2_t = eta * eta * fmaf(G.cos_i, -G.cos_i, 1.0f);
    for (int c = 0; c < NCH; ++c) {
        float n1k = inside ? n_med[c] : 1.0f, n2k = inside ? 1.0f : n_med[c];
        float ek = n1k / n2k;
        G.n1_ch[c] = n1k;
        G.n2_ch[c] = n2k;
        G.sin2_t_ch[c] = ek * ek * fmaf(G.cos_i, -G.cos_i, 1.0f);
    }
    G.F = uframe_build(k, normal, cax, G.cos_i, G.sin_i);
    const int nondegenerate = vlen_sq(vcross(k, cax)) > 1e-6f;
    const float one_m_c2 = fmaf(G.cos_i, -G.cos_i, 1.0f);

    v3 new_k, new_s;
    int new_inside = inside, upd = -1;
    float exact_p_o = -1.0f;
    uint32_t match_bits = 0u; /* channels whose refracted direction matched the hero's */
    const int frosted = S->finish != NULL && S->finish[facet] != 0;
    if (frosted) { /* dispatch_bounce's frosted arm (transport/bounce.rs:137-160) */
        new_k = frosted_arm_aniso(S, P, bounce, seed, facet, normal, G.cos_i, G.n1, G.n2, G.sin2_t, &new_inside, &upd, ev);
        new_s = new_k;
    } else if (G.sin2_t > 1.0f) {
        /* apply_tir_bounce (refraction/tir.rs:56-160) */
        if (ev) ev->n_tir++;
        ev_facet(ev, facet, 1);
        if (nondegenerate) {
            for (int c = 0; c < NCH; ++c) {
                internal_sol sol = internal_solve(G.n1_ch[c], P->n_o[c], n_e_ch[c], cax, &G.F, !is_extra);
                float flux_inc = maxf(sol.flux_inc, 1e-12f);
                float ro = cnorm2(sol.r_o) * sol.flux_ro, re = cnorm2(sol.r_e) * sol.flux_re;
                float r_total = minf((ro + re) / flux_inc, 1.0f);
                P->st[c] = st_scale(P->st[c], r_total);
                if (G.sin2_t_ch[c] <= 1.0f) P->path_pdf[c] *= clampf(r_total, R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                if (c == 0) exact_p_o = ro / maxf(ro + re, 1e-12f);
            }
        } else {
            for (int c = 0; c < NCH; ++c) {
                float n1k = G.n1_ch[c], n2k = G.n2_ch[c];
                if (G.sin2_t_ch[c] > 1.0f) {
                    mat4 Mt = mueller_tir_retardation(tir_phase_delta(n1k, G.cos_i, G.sin_i));
                    P->st[c] = st_apply(&Mt, P->st[c]);
                } else {
                    float ct = sqrtf(maxf(1.0f - G.sin2_t_ch[c], 0.0f));
                    float rs, rp;
                    fresnel_rs_rp(n1k, n2k, G.cos_i, ct, &rs, &rp);
                    mat4 Mr = mueller_fresnel_reflection(rs, rp);
                    P->st[c] = st_apply(&Mr, P->st[c]);
                    P->path_pdf[c] *= clampf(0.5f * fmaf(rp, rp, rs * rs), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                }
            }
        }
        new_k = vsub(k, vmul(normal, 2.0f * vdot(k, normal)));
        new_s = poynting_for_mode(S, &G, new_k, inside, is_extra);
    } else if (!inside && nondegenerate) {
        /* apply_uniaxial_entry_bounce (refraction/uniaxial_entry.rs:108-198) */
        entry_inc I = entry_incidence_frame(1.0f, &G.F);
        entry_sol hs, hp;
        entry_solve_pair(&I, 1.0f, P->n_o[0], n_e_ch[0], cax, &G.F, &hs, &hp);
        mat4 Mh = jones_to_mueller(hs.r_s, hp.r_s, hs.r_p, hp.r_p);
        float r_branch = clampf(Mh.m[0][0], R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        float inc_flux = 0.5f * G.F.cos_i;
        float p_o_raw = mode_power(hs.t_o, hp.t_o, hs.flux_o, inc_flux, P->st[0]);
        float p_e_raw = mode_power(hs.t_e, hp.t_e, hs.flux_e, inc_flux, P->st[0]);
        float p_o_hero = clampf(p_o_raw / maxf(p_o_raw + p_e_raw, 1e-12f), R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        float msr = unit_draw(seed, bounce, BIREFRINGENT_SPLIT_STREAM);
        int use_extra = msr < (1.0f - p_o_hero);
        float p_mode_frac = use_extra ? 1.0f - p_o_hero : p_o_hero;
        float n2_dir = use_extra ? G.n2 : G.n_o_hero;
        float eta_dir = G.n1 / n2_dir;
        float s2d = minf(eta_dir * eta_dir * one_m_c2, 1.0f);
        float ctd = sqrtf(maxf(1.0f - s2d, 0.0f));
        v3 refr = refract_dir(k, normal, eta_dir, G.cos_i, ctd);
        v3 final_refr = use_extra ? extraordinary_poynting_dir(refr, cax, G.n_o_hero, G.n_e_hero) : refr;
        v3 p_axis_t = vcross(refr, G.F.s_axis);
        float rng = unit_draw(seed, bounce, FRESNEL_BRANCH_STREAM);
        if (rng < r_branch) {
            if (ev) ev->n_reflect++;
            ev_facet(ev, facet, 2);
            for (int c = 0; c < NCH; ++c) {
                entry_sol ss, sp;
                entry_solve_pair(&I, 1.0f, P->n_o[c], n_e_ch[c], cax, &G.F, &ss, &sp);
                mat4 Mk = jones_to_mueller(ss.r_s, sp.r_s, ss.r_p, sp.r_p);
                float ru = clampf(Mk.m[0][0], R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                P->st[c] = st_scale(st_apply(&Mk, P->st[c]), 1.0f / r_branch);
                P->path_pdf[c] *= ru;
            }
            new_k = vsub(k, vmul(normal, 2.0f * vdot(k, normal)));
            new_s = new_k;
        } else {
            if (ev) ev->n_refract++;
            ev_facet(ev, facet, 3);
            v3 dirs[NCH];
            int has_dir[NCH], hero_match[NCH];
            for (int c = 0; c < NCH; ++c) {
                float n_o_k = P->n_o[c];
                float n2k = use_extra ? G.n2_ch[c] : n_o_k;
                float r = G.n1 / n2k;
                float s2k = r * r * one_m_c2;
                has_dir[c] = 0;
                hero_match[c] = 0;
                if (s2k > 1.0f) {
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    P->path_pdf[c] = 0.0f;
                    continue;
                }
                float ctk = sqrtf(maxf(1.0f - s2k, 0.0f));
                float ek = G.n1 / n2k;
                v3 rk = refract_dir(k, normal, ek, G.cos_i, ctk);
                v3 fk = use_extra ? extraordinary_poynting_dir(rk, cax, n_o_k, n_o_k + M->birefringence_delta) : rk;
                dirs[c] = fk;
                has_dir[c] = 1;
                int match = vdot(fk, final_refr) >= DIRECTION_MATCH_COS_TOL;
                hero_match[c] = match;
                if (match) match_bits |= 1u << c;
                entry_sol ss, sp;
                entry_solve_pair(&I, 1.0f, n_o_k, n_e_ch[c], cax, &G.F, &ss, &sp);
                if (!match) {
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    mat4 Mk = jones_to_mueller(ss.r_s, sp.r_s, ss.r_p, sp.r_p);
                    P->path_pdf[c] *= clampf(1.0f - Mk.m[0][0], R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                    continue;
                }
                cplx tsr = use_extra ? ss.t_e : ss.t_o, tpr = use_extra ? sp.t_e : sp.t_o;
                float flux_k = use_extra ? ss.flux_e : ss.flux_o;
                v3 mode_dir = use_extra ? ss.e_hat : ss.o_hat;
                float p_mode_k = mode_power(tsr, tpr, flux_k, inc_flux, P->st[c]);
                float dep = p_mode_k / p_mode_frac;
                float c2, s2;
                azimuth2_in_frame(mode_dir, G.F.s_axis, p_axis_t, &c2, &s2);
                stokes t = {dep, dep * c2, dep * s2, 0.0f};
                P->st[c] = st_scale(t, 1.0f / (1.0f - r_branch));
                mat4 Mk = jones_to_mueller(ss.r_s, sp.r_s, ss.r_p, sp.r_p);
                P->path_pdf[c] *= clampf(1.0f - Mk.m[0][0], R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            }
            narrow_compat_dirs(P, dirs, has_dir, hero_match);
            new_k = refr;
            new_s = final_refr;
            new_inside = 1;
            upd = use_extra;
        }
    } else if (inside && nondegenerate) {
        /* apply_uniaxial_internal_bounce (refraction/uniaxial_internal.rs) */
        internal_sol sh = internal_solve(G.n1_ch[0], P->n_o[0], n_e_ch[0], cax, &G.F, !is_extra);
        float flux_inc_h = maxf(sh.flux_inc, 1e-12f);
        float ro = cnorm2(sh.r_o) * sh.flux_ro, re = cnorm2(sh.r_e) * sh.flux_re;
        float r_branch = clampf((ro + re) / flux_inc_h, R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        float p_o_exact = ro / maxf(ro + re, 1e-12f);
        float rng = unit_draw(seed, bounce, FRESNEL_BRANCH_STREAM);
        if (rng < r_branch) {
            if (ev) ev->n_reflect++;
            ev_facet(ev, facet, 2);
            for (int c = 0; c < NCH; ++c) {
                internal_sol sol = c == 0 ? sh : internal_solve(G.n1_ch[c], P->n_o[c], n_e_ch[c], cax, &G.F, !is_extra);
                float flux_inc = maxf(sol.flux_inc, 1e-12f);
                float r_total = minf(fmaf(cnorm2(sol.r_e), sol.flux_re, cnorm2(sol.r_o) * sol.flux_ro) / flux_inc, 1.0f);
                P->st[c] = st_scale(P->st[c], r_total / r_branch);
                P->path_pdf[c] *= clampf(r_total, R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            }
            new_k = vsub(k, vmul(normal, 2.0f * vdot(k, normal)));
            new_s = new_k;
            exact_p_o = p_o_exact;
        } else {
            if (ev) ev->n_refract++;
            ev_facet(ev, facet, 3);
            float eta_dir = G.n1 / G.n2;
            float s2d = minf(eta_dir * eta_dir * one_m_c2, 1.0f);
            float ctd = sqrtf(maxf(1.0f - s2d, 0.0f));
            v3 hero_refr = refract_dir(k, normal, eta_dir, G.cos_i, ctd);
            for (int c = 0; c < NCH; ++c) {
                float orig_i = P->st[c].i;
                float n_inc = G.n1_ch[c], n2k = G.n2_ch[c];
                float r = n_inc / n2k;
                float s2k = r * r * one_m_c2;
                if (s2k > 1.0f) {
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    P->path_pdf[c] = 0.0f;
                    continue;
                }
                float ctk = sqrtf(maxf(1.0f - s2k, 0.0f));
                float ek = n_inc / n2k;
                v3 rk = refract_dir(k, normal, ek, G.cos_i, ctk);
                int match = vdot(rk, hero_refr) >= DIRECTION_MATCH_COS_TOL;
                if (match) match_bits |= 1u << c;
                internal_sol sol = c == 0 ? sh : internal_solve(n_inc, P->n_o[c], n_e_ch[c], cax, &G.F, !is_extra);
                float iu;
                if (!match) {
                    float prefix = P->path_pdf[c];
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    stokes tr = uniaxial_exit_transmission(&sol, r_branch, orig_i, &iu);
                    P->path_pdf[c] = prefix * clampf(iu, R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                    if (orig_i > 0.0f) split_probe(S, P, hit_point, c, rk, st_intensity(tr), ev);
                    continue;
                }
                P->st[c] = uniaxial_exit_transmission(&sol, r_branch, orig_i, &iu);
                P->path_pdf[c] *= clampf(iu, R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            }
            new_k = hero_refr;
            new_s = hero_refr;
            new_inside = 0;
        }
    } else {
        /* k along the optic axis: the scalar-Fresnel arms with the eigenmode
         * draw (refraction/dispatch.rs:143-380, reflect_refract.rs) */
        const int entering = !inside;
        float p_o = 0.5f, c2o = 0.0f, s2o = 0.0f;
        int have_sel = 0;
        if (entering) {
            float p;
            have_sel = entry_eigenmode_selection(cax, cur_plane, k, P->st[0], &p, &c2o, &s2o);
            if (have_sel) p_o = p;
        }
        float msr = unit_draw(seed, bounce, BIREFRINGENT_SPLIT_STREAM);
        int use_extra = entering ? (msr < (1.0f - p_o)) : is_extra;
        float c2x = use_extra ? -c2o : c2o, s2x = use_extra ? -s2o : s2o;
        /* compute_entry_reflect_probability */
        int ord_sel = entering && !use_extra;
        float n2h = ord_sel ? G.n_o_hero : G.n2;
        float s2h;
        if (ord_sel) {
            float e2 = G.n1 / n2h;
            s2h = e2 * e2 * one_m_c2;
        } else {
            s2h = G.sin2_t;
        }
        float cth = sqrtf(maxf(1.0f - s2h, 0.0f));
        float rs, rp;
        fresnel_rs_rp(G.n1, n2h, G.cos_i, cth, &rs, &rp);
        float r_unpol = clampf(0.5f * fmaf(rp, rp, rs * rs), R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        float rng = unit_draw(seed, bounce, FRESNEL_BRANCH_STREAM);
        if (rng < r_unpol) {
            if (ev) ev->n_reflect++;
            ev_facet(ev, facet, 2);
            for (int c = 0; c < NCH; ++c) {
                float n1k = G.n1_ch[c], n2k = G.n2_ch[c];
                mat4 Mr;
                if (G.sin2_t_ch[c] > 1.0f) {
                    Mr = mueller_tir_retardation(tir_phase_delta(n1k, G.cos_i, G.sin_i));
                } else {
                    float ct = sqrtf(maxf(1.0f - G.sin2_t_ch[c], 0.0f));
                    float a, b;
                    fresnel_rs_rp(n1k, n2k, G.cos_i, ct, &a, &b);
                    P->path_pdf[c] *= clampf(0.5f * fmaf(b, b, a * a), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                    Mr = mueller_fresnel_reflection(a, b);
                }
                P->st[c] = st_scale(st_apply(&Mr, P->st[c]), 1.0f / r_unpol);
            }
            new_k = vsub(k, vmul(normal, 2.0f * vdot(k, normal)));
            new_s = poynting_for_mode(S, &G, new_k, inside, is_extra);
        } else {
            if (ev) ev->n_refract++;
            ev_facet(ev, facet, 3);
            float n2d = (entering && !use_extra) ? G.n_o_hero : G.n2;
            float eta_dir = G.n1 / n2d;
            float s2d = minf(eta_dir * eta_dir * one_m_c2, 1.0f);
            float ctd = sqrtf(maxf(1.0f - s2d, 0.0f));
            v3 refr = refract_dir(k, normal, eta_dir, G.cos_i, ctd);
            v3 final_refr = (entering && use_extra) ? extraordinary_poynting_dir(refr, cax, G.n_o_hero, G.n_e_hero) : refr;
            v3 dirs[NCH];
            int has_dir[NCH], hero_match[NCH];
            for (int c = 0; c < NCH; ++c) {
                stokes orig = P->st[c];
                float n1k = G.n1_ch[c];
                float n2k = (entering && !use_extra) ? P->n_o[c] : G.n2_ch[c];
                float r = n1k / n2k;
                float s2k = r * r * one_m_c2;
                has_dir[c] = 0;
                hero_match[c] = 0;
                if (s2k > 1.0f) {
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    P->path_pdf[c] = 0.0f;
                    continue;
                }
                float ctk = sqrtf(maxf(1.0f - s2k, 0.0f));
                float ek = n1k / n2k;
                v3 rk = refract_dir(k, normal, ek, G.cos_i, ctk);
                v3 fk = (entering && use_extra)
                            ? extraordinary_poynting_dir(rk, cax, P->n_o[c], P->n_o[c] + M->birefringence_delta)
                            : rk;
                int match = vdot(fk, final_refr) >= DIRECTION_MATCH_COS_TOL;
                if (match) match_bits |= 1u << c;
                /* compute_channel_transmission (exit_split.rs:69-101) */
                float num = 2.0f * n1k * G.cos_i;
                float t_s = num / fmaf(n2k, ctk, n1k * G.cos_i);
                float t_p = num / fmaf(n1k, ctk, n2k * G.cos_i);
                mat4 T = mueller_fresnel_transmission(n1k, n2k, G.cos_i, ctk, t_s, t_p);
                stokes incident = orig;
                if (entering && have_sel) {
                    stokes o2 = {orig.i, orig.i * c2x, orig.i * s2x, 0.0f};
                    incident = o2;
                }
                stokes transmitted = st_scale(st_apply(&T, incident), 1.0f / (1.0f - r_unpol));
                float a, b;
                fresnel_rs_rp(n1k, n2k, G.cos_i, ctk, &a, &b);
                float ruk = clampf(0.5f * fmaf(b, b, a * a), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                if (match) {
                    P->st[c] = transmitted;
                    P->path_pdf[c] *= 1.0f - ruk;
                } else {
                    float prefix = P->path_pdf[c];
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    P->path_pdf[c] = prefix * (1.0f - ruk);
                    if (inside && !entering && st_intensity(orig) > 0.0f)
                        split_probe(S, P, hit_point, c, rk, st_intensity(transmitted), ev);
                }
                dirs[c] = fk;
                has_dir[c] = 1;
                hero_match[c] = match;
            }
            if (!inside) narrow_compat_dirs(P, dirs, has_dir, hero_match);
            new_k = refr;
            new_s = final_refr;
            new_inside = !inside;
            if (entering) upd = use_extra;
        }
    }
    /* dispatch_bounce tail (transport/bounce.rs:200-238) */
    int was_internal_reflection = inside && upd < 0 && new_inside == inside;
    P->dir = new_s;
    P->org = vadd(hit_point, vmul(new_s, 1e-4f));
    P->k = new_k;
    P->inside = new_inside;
    if (upd >= 0) P->is_extra = upd;
    if (was_internal_reflection) {
        /* apply_internal_mode_coupling (transport/bounce.rs:439-470) */
        float p_o = 0.5f;
        if (exact_p_o >= 0.0f) {
            p_o = exact_p_o;
        } else {
            float p, c2, s2;
            if (entry_eigenmode_selection(cax, cur_plane, new_k, P->st[0], &p, &c2, &s2)) p_o = p;
        }
        P->is_extra = unit_draw(seed, bounce, MODE_COUPLING_STREAM) >= p_o;
    }
    /* the side channel also records the eigenmode after every event */
    if (ev) {
        ev->facet_hash = (ev->facet_hash ^ (16u + (uint32_t)P->is_extra)) * 16777619u;
        ev->facet_hash = (ev->facet_hash ^ (256u + match_bits)) * 16777619u;
    }
    if (bounce == 0 && !P->inside && S->surface_glare < 1.0f && !frosted)
        for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], S->surface_glare);
    if (bounce > 4) {
        float mx = 0.0f;
        for (int c = 0; c < NCH; ++c) mx = maxf(mx, st_intensity(P->st[c]));
        float q = clampf(mx, 0.05f, 1.0f);
        float rr = unit_draw(seed, bounce, RUSSIAN_ROULETTE_STREAM);
        if (rr > q) { *stop = 1; return; }
        for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], 1.0f / q);
        for (int c = 0; c < NCH; ++c) P->split[c] /= q;
    }
}

/* One facet event of a biaxial material (biaxial_delta_beta_alpha is Some):
 * the bounce body's facet_event with compute_bounce_refraction_geometry's
 * biaxial block (refraction/geometry.rs:199-270, 306-425: mode A = the fast
 * root, mode B = the slow root, per channel at the hero's wave normals,
 * resolved at entry by two fixed-point steps per mode), the scalar TIR and
 * partial-Fresnel arms (tir.rs, dispatch.rs:143-262 with no
 * polarisation-weighted selection: p_o = 0.5), the refract arm's biaxial
 * walk-off per channel (reflect_refract.rs:122-330), the assigned mode's
 * pleochroic absorption (absorption.rs:224-249 through
 * assigned_mode_e_field) and the 50/50 internal re-coupling draw
 * (transport/bounce.rs:439-470). */
/* The assigned mode's E field (assigned_mode_e_field, biaxial.rs:432-436)
 * in the absorption tensor's frame (axes^T e, absorption_tensor.rs:80-83),
 * and AbsorptionTensor3::quadratic_form with alpha = (o, beta, e)
 * (assigned_mode_alpha, absorption_tensor.rs:115-117). */
static v3 biax_assigned_locals(const scene_t* S, const biax_t* hero, v3 k, int is_extra) {
    v3 e = biax_assigned_e(hero, k, is_extra);
    return mk(vdot(S->ta1, e), vdot(S->ta2, e), vdot(S->tcn, e));
}
static float biax_tensor_qf(v3 l, float ao, float ab, float ae) {
    return (ao * (l.x * l.x) + ab * (l.y * l.y)) + ae * (l.z * l.z);
}

/* channel_absorption_alphas_assigned (absorption.rs:224-249) for the scatter
 * step of an anisotropic stone: the assigned mode's E field at the current
 * wave normal -- the uniaxial o / e field, or the biaxial slow / fast one of
 * the hero's indicatrix -- through each channel's tensor. */
static void scatter_alphas_assigned(const scene_t* S, path_t* P, float alpha[NCH]) {
    if (!S->absorbing) {
        for (int c = 0; c < NCH; ++c) alpha[c] = 0.0f;
        return;
    }
    if (S->biax) {
        float nb = P->n_o[0], na = nb - S->mat->biaxial_delta_beta_alpha;
        biax_t hero = biax_from_gamma_axis(na, nb, na + S->mat->birefringence_delta, S->c_raw);
        v3 l = biax_assigned_locals(S, &hero, P->k, P->is_extra);
        for (int c = 0; c < NCH; ++c)
            alpha[c] = biax_tensor_qf(l, P->alpha_o[c], S->trichroic ? P->alpha_b[c] : P->alpha_o[c], P->alpha_e[c]);
        return;
    }
    float n_o_h = P->n_o[0];
    float n_e_h = e_index_at(S, P->lambdas[0], n_o_h);
    v3 e = assigned_mode_e_field_uniaxial(P->k, S->c_raw, P->is_extra, n_o_h, n_e_h);
    for (int c = 0; c < NCH; ++c) alpha[c] = tensor_qf(S, P->alpha_o[c], P->alpha_e[c], e);
}

static void facet_event_biax(const scene_t* S, path_t* P, uint32_t bounce, uint32_t seed, float t_hit,
                             int32_t facet, ixg_path_events* ev, int* stop) {
    const ixg_material* M = S->mat;
    v3 hit_point = vadd(P->org, vmul(P->dir, t_hit));
    v3 normal = shading_normal_near_edge(S->planes, S->n_planes, hit_point, facet, vfrom(S->planes[facet].normal),
                                         M->edge_rounding_radius);
    v3 cur_plane = vnormalize_or_zero(vcross(P->k, normal));
    if (P->has_prev && vlen_sq(cur_plane) > 1e-6f && vlen_sq(P->prev_plane) > 1e-6f) {
        float cos_psi = clampf(vdot(P->prev_plane, cur_plane), -1.0f, 1.0f);
        float sin_psi = vdot(vcross(P->prev_plane, cur_plane), vnormalize_or_zero(P->k));
        mat4 R = mueller_frame_rotation(atan2_cr(sin_psi, cos_psi));
        for (int c = 0; c < NCH; ++c) P->st[c] = st_apply(&R, P->st[c]);
    }
    P->prev_plane = cur_plane;
    P->has_prev = 1;
    const float dba = M->biaxial_delta_beta_alpha, bd = M->birefringence_delta;
    biax_t ind[NCH]; /* GemMaterial::biaxial_indicatrix per channel (materials/optics.rs:16-27) */
    for (int c = 0; c < NCH; ++c) {
        float nb = P->n_o[c], na = nb - dba;
        ind[c] = biax_from_gamma_axis(na, nb, na + bd, S->c_raw);
    }
    if (P->inside && S->absorbing && !(M->scattering_sigma_s > 0.0f)) {
        /* channel_absorption_alphas_assigned (absorption.rs:224-249); a
         * scattering stone's segment was attenuated by the scatter step */
        v3 l = biax_assigned_locals(S, &ind[0], P->k, P->is_extra);
        float scaled = t_hit * M->absorption_path_scale;
        float arg[NCH], tr[NCH];
        for (int c = 0; c < NCH; ++c) {
            float ab = S->trichroic ? P->alpha_b[c] : P->alpha_o[c];
            arg[c] = -biax_tensor_qf(l, P->alpha_o[c], ab, P->alpha_e[c]) * scaled;
        }
        exp_lane8(arg, tr);
        for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], tr[c]);
    }
    if (P->inside) normal = vneg(normal);

    const v3 k = P->k;
    const int inside = P->inside, is_extra = P->is_extra;
    const float cos_i = clampf(vdot(vneg(k), normal), 0.0f, 1.0f);
    const float one_m_c2 = fmaf(cos_i, -cos_i, 1.0f);
    const float sin_i = sqrtf(maxf(one_m_c2, 0.0f));
    /* hero_biaxial_wave_dirs / per_channel_biaxial_indices (geometry.rs:216-270) */
    v3 dir_a = k, dir_b = k;
    if (!inside) {
        (void)biax_resolve_entry(&ind[0], k, normal, cos_i, P->n_o[0], 0, &dir_a);
        (void)biax_resolve_entry(&ind[0], k, normal, cos_i, P->n_o[0], 1, &dir_b);
    }
    float n_a[NCH], n_b[NCH], n_med[NCH];
    for (int c = 0; c < NCH; ++c) {
        float ns, nf;
        biax_wave_indices(&ind[c], dir_a, &ns, &nf);
        n_a[c] = nf;
        biax_wave_indices(&ind[c], dir_b, &ns, &nf);
        n_b[c] = ns;
        n_med[c] = (inside && !is_extra) ? n_a[c] : n_b[c];
    }
    const float n1 = inside ? n_med[0] : 1.0f, n2 = inside ? 1.0f : n_med[0];
    const float eta = n1 / n2;
    const float sin2_t = eta * eta * one_m_c2;
    float n1_ch[NCH], n2_ch[NCH], sin2_t_ch[NCH];
    for (int c = 0; c < NCH; ++c) {
        n1_ch[c] = inside ? n_med[c] : 1.0f;
        n2_ch[c] = inside ? 1.0f : n_med[c];
        float ek = n1_ch[c] / n2_ch[c];
        sin2_t_ch[c] = ek * ek * one_m_c2;
    }

    v3 new_k, new_s;
    int new_inside = inside, upd = -1;
    uint32_t match_bits = 0u;
    const int frosted = S->finish != NULL && S->finish[facet] != 0;
    if (frosted) { /* dispatch_bounce's frosted arm (transport/bounce.rs:137-160) */
        new_k = frosted_arm_aniso(S, P, bounce, seed, facet, normal, cos_i, n1, n2, sin2_t, &new_inside, &upd, ev);
        new_s = new_k;
    } else if (sin2_t > 1.0f) {
        /* apply_tir_bounce, scalar arm (refraction/tir.rs:56-160) */
        if (ev) ev->n_tir++;
        ev_facet(ev, facet, 1);
        for (int c = 0; c < NCH; ++c) {
            if (sin2_t_ch[c] > 1.0f) {
                mat4 Mt = mueller_tir_retardation(tir_phase_delta(n1_ch[c], cos_i, sin_i));
                P->st[c] = st_apply(&Mt, P->st[c]);
            } else {
                float ct = sqrtf(maxf(1.0f - sin2_t_ch[c], 0.0f));
                float rs, rp;
                fresnel_rs_rp(n1_ch[c], n2_ch[c], cos_i, ct, &rs, &rp);
                mat4 Mr = mueller_fresnel_reflection(rs, rp);
                P->st[c] = st_apply(&Mr, P->st[c]);
                P->path_pdf[c] *= clampf(0.5f * fmaf(rp, rp, rs * rs), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            }
        }
        new_k = vsub(k, vmul(normal, 2.0f * vdot(k, normal)));
        new_s = inside ? biax_mode_poynting(&ind[0], new_k, is_extra) : new_k;
    } else {
        /* apply_partial_fresnel_bounce (refraction/dispatch.rs:143-262) */
        const int entering = !inside;
        float msr = unit_draw(seed, bounce, BIREFRINGENT_SPLIT_STREAM);
        int use_extra = entering ? (msr < (1.0f - 0.5f)) : is_extra;
        float cth = sqrtf(maxf(1.0f - sin2_t, 0.0f));
        float rs, rp;
        fresnel_rs_rp(n1, n2, cos_i, cth, &rs, &rp);
        float r_unpol = clampf(0.5f * fmaf(rp, rp, rs * rs), R_UNPOL_SELECT_MIN, R_UNPOL_SELECT_MAX);
        float rng = unit_draw(seed, bounce, FRESNEL_BRANCH_STREAM);
        if (rng < r_unpol) {
            if (ev) ev->n_reflect++;
            ev_facet(ev, facet, 2);
            for (int c = 0; c < NCH; ++c) {
                mat4 Mr;
                if (sin2_t_ch[c] > 1.0f) {
                    Mr = mueller_tir_retardation(tir_phase_delta(n1_ch[c], cos_i, sin_i));
                } else {
                    float ct = sqrtf(maxf(1.0f - sin2_t_ch[c], 0.0f));
                    float a, b;
                    fresnel_rs_rp(n1_ch[c], n2_ch[c], cos_i, ct, &a, &b);
                    P->path_pdf[c] *= clampf(0.5f * fmaf(b, b, a * a), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                    Mr = mueller_fresnel_reflection(a, b);
                }
                P->st[c] = st_scale(st_apply(&Mr, P->st[c]), 1.0f / r_unpol);
            }
            new_k = vsub(k, vmul(normal, 2.0f * vdot(k, normal)));
            new_s = inside ? biax_mode_poynting(&ind[0], new_k, is_extra) : new_k;
        } else {
            if (ev) ev->n_refract++;
            ev_facet(ev, facet, 3);
            float n2d = entering ? (use_extra ? n_b[0] : n_a[0]) : n2;
            float eta_dir = n1 / n2d;
            float s2d = minf(eta_dir * eta_dir * one_m_c2, 1.0f);
            float ctd = sqrtf(maxf(1.0f - s2d, 0.0f));
            v3 refr = refract_dir(k, normal, eta_dir, cos_i, ctd);
            v3 final_refr = entering ? biax_mode_poynting(&ind[0], refr, use_extra) : refr;
            v3 dirs[NCH];
            int has_dir[NCH], hero_match[NCH];
            for (int c = 0; c < NCH; ++c) {
                stokes orig = P->st[c];
                float n1k = n1_ch[c];
                float n2k = (entering && !use_extra) ? n_a[c] : n2_ch[c];
                float r = n1k / n2k;
                float s2k = r * r * one_m_c2;
                has_dir[c] = 0;
                hero_match[c] = 0;
                if (s2k > 1.0f) {
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    P->path_pdf[c] = 0.0f;
                    continue;
                }
                float ctk = sqrtf(maxf(1.0f - s2k, 0.0f));
                float ek = n1k / n2k;
                v3 rk = refract_dir(k, normal, ek, cos_i, ctk);
                v3 fk = entering ? biax_mode_poynting(&ind[c], rk, use_extra) : rk;
                int match = vdot(fk, final_refr) >= DIRECTION_MATCH_COS_TOL;
                if (match) match_bits |= 1u << c;
                float num = 2.0f * n1k * cos_i;
                float t_s = num / fmaf(n2k, ctk, n1k * cos_i);
                float t_p = num / fmaf(n1k, ctk, n2k * cos_i);
                mat4 T = mueller_fresnel_transmission(n1k, n2k, cos_i, ctk, t_s, t_p);
                stokes transmitted = st_scale(st_apply(&T, orig), 1.0f / (1.0f - r_unpol));
                float a, b;
                fresnel_rs_rp(n1k, n2k, cos_i, ctk, &a, &b);
                float ruk = clampf(0.5f * fmaf(b, b, a * a), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
                if (match) {
                    P->st[c] = transmitted;
                    P->path_pdf[c] *= 1.0f - ruk;
                } else {
                    float prefix = P->path_pdf[c];
                    P->st[c] = st_scale(P->st[c], 0.0f);
                    P->path_pdf[c] = prefix * (1.0f - ruk);
                    if (inside && st_intensity(orig) > 0.0f)
                        split_probe(S, P, hit_point, c, rk, st_intensity(transmitted), ev);
                }
                dirs[c] = fk;
                has_dir[c] = 1;
                hero_match[c] = match;
            }
            if (!inside) narrow_compat_dirs(P, dirs, has_dir, hero_match);
            new_k = refr;
            new_s = final_refr;
            new_inside = !inside;
            if (entering) upd = use_extra;
        }
    }
    /* dispatch_bounce tail (transport/bounce.rs:200-238) */
    int was_internal_reflection = inside && upd < 0 && new_inside == inside;
    P->dir = new_s;
    P->org = vadd(hit_point, vmul(new_s, 1e-4f));
    P->k = new_k;
    P->inside = new_inside;
    if (upd >= 0) P->is_extra = upd;
    if (was_internal_reflection) /* apply_internal_mode_coupling, biaxial: p_o = 0.5 */
        P->is_extra = unit_draw(seed, bounce, MODE_COUPLING_STREAM) >= 0.5f;
    if (ev) {
        ev->facet_hash = (ev->facet_hash ^ (16u + (uint32_t)P->is_extra)) * 16777619u;
        ev->facet_hash = (ev->facet_hash ^ (256u + match_bits)) * 16777619u;
    }
    if (bounce == 0 && !P->inside && S->surface_glare < 1.0f && !frosted)
        for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], S->surface_glare);
    if (bounce > 4) {
        float mx = 0.0f;
        for (int c = 0; c < NCH; ++c) mx = maxf(mx, st_intensity(P->st[c]));
        float q = clampf(mx, 0.05f, 1.0f);
        float rr = unit_draw(seed, bounce, RUSSIAN_ROULETTE_STREAM);
        if (rr > q) { *stop = 1; return; }
        for (int c = 0; c < NCH; ++c) P->st[c] = st_scale(P->st[c], 1.0f / q);
        for (int c = 0; c < NCH; ++c) P->split[c] /= q;
    }
}

static biax_t biax_from_abi(float na, float nb, float ng, const float* axes) {
    biax_t B;
    B.na = na;
    B.nb = nb;
    B.ng = ng;
    B.ax = vfromr(axes);
    B.ay = vfromr(axes + 3);
    B.az = vfromr(axes + 6);
    return B;
}
static void v3_out(v3 v, float* o) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}
void ixo_biax_from_gamma_axis(const float gamma_axis[3], float axes_out[9]) {
    biax_t B = biax_from_gamma_axis(1.0f, 1.0f, 1.0f, vfromr(gamma_axis));
    v3_out(B.ax, axes_out);
    v3_out(B.ay, axes_out + 3);
    v3_out(B.az, axes_out + 6);
}
void ixo_biax_wave_indices(float na, float nb, float ng, const float axes[9], const float k[3], float out[2]) {
    biax_t B = biax_from_abi(na, nb, ng, axes);
    biax_wave_indices(&B, vfromr(k), &out[0], &out[1]);
}
void ixo_biax_eigen_d(float na, float nb, float ng, const float axes[9], const float k[3], int want_slow,
                      float out[3]) {
    biax_t B = biax_from_abi(na, nb, ng, axes);
    v3_out(biax_eigen_d(&B, vfromr(k), want_slow), out);
}
void ixo_biax_d_to_e(float na, float nb, float ng, const float axes[9], const float d[3], float out[3]) {
    biax_t B = biax_from_abi(na, nb, ng, axes);
    v3_out(biax_d_to_e(&B, vfromr(d)), out);
}
void ixo_biax_mode_poynting(float na, float nb, float ng, const float axes[9], const float k[3], int want_slow,
                            float out[3]) {
    biax_t B = biax_from_abi(na, nb, ng, axes);
    v3_out(biax_mode_poynting(&B, vfromr(k), want_slow), out);
}
float ixo_biax_resolve_entry(float na, float nb, float ng, const float axes[9], const float incident[3],
                             const float normal[3], float cos_i, float n_seed, int want_slow, float dir_out[3]) {
    biax_t B = biax_from_abi(na, nb, ng, axes);
    v3 d;
    float n = biax_resolve_entry(&B, vfromr(incident), vfromr(normal), cos_i, n_seed, want_slow, &d);
    v3_out(d, dir_out);
    return n;
}
void ixo_poynting_direction(const float k[3], const float e[3], float out[3]) {
    v3_out(poynting_direction(vfromr(k), vfromr(e)), out);
}

/* Energy bookkeeping of one uniaxial interface solve, for the tests:
 * out = {R, T} as fractions of the incident Poynting flux along the normal.
 * kind 0: air -> crystal entry, incident polarisation pol 0 = s, 1 = p;
 * kind 1: crystal -> air, incident mode pol 0 = ordinary, 1 = extraordinary
 * (n_inc is that mode's index along k).  k is the incident wave normal and
 * normal faces the incident side ((-k).normal >= 0). */
void ixo_uniaxial_energy(int kind, float n_inc, float n_o, float n_e, const float c[3], const float kk[3],
                         const float nn[3], int pol, float out[2]) {
    v3 k = vnormalize(vfromr(kk)), normal = vnormalize(vfromr(nn)), cax = vnormalize(vfromr(c));
    float cos_i = clampf(vdot(vneg(k), normal), 0.0f, 1.0f);
    float sin_i = sqrtf(maxf(fmaf(cos_i, -cos_i, 1.0f), 0.0f));
    uframe F = uframe_build(k, normal, cax, cos_i, sin_i);
    if (kind == 0) {
        entry_inc I = entry_incidence_frame(1.0f, &F);
        entry_sol ss, sp;
        entry_solve_pair(&I, 1.0f, n_o, n_e, cax, &F, &ss, &sp);
        const entry_sol* so = pol == 0 ? &ss : &sp;
        float k_tan = F.sin_i;
        cv3 k_bwd = cv_wave(F.that, F.zhat, k_tan, cx(-F.cos_i, 0.0f));
        cv3 s_c = cv_real(F.s_axis);
        cv3 ep_bwd = cv_scale_real(cv_cross(k_bwd, s_c), -1.0f);
        mfields rs = {k_bwd, s_c, cv_cross(k_bwd, s_c)}, rp = {k_bwd, ep_bwd, cv_cross(k_bwd, ep_bwd)};
        float inc = 0.5f * F.cos_i;
        float R = (cnorm2(so->r_s) * fabsf(poynting_z(&rs, F.zhat)) + cnorm2(so->r_p) * fabsf(poynting_z(&rp, F.zhat))) / inc;
        float T = (cnorm2(so->t_o) * so->flux_o + cnorm2(so->t_e) * so->flux_e) / inc;
        out[0] = R;
        out[1] = T;
    } else {
        internal_sol sol = internal_solve(n_inc, n_o, n_e, cax, &F, pol == 0);
        float inc = maxf(sol.flux_inc, 1e-12f);
        out[0] = (cnorm2(sol.r_o) * sol.flux_ro + cnorm2(sol.r_e) * sol.flux_re) / inc;
        out[1] = (cnorm2(sol.t_s) * sol.flux_ts + cnorm2(sol.t_p) * sol.flux_tp) / inc;
    }
}

/* The closed-form solution's reflection amplitudes at entry for s or p
 * incidence: out = {re, im} of (r_ss, r_sp) for pol 0, (r_ps, r_pp) for 1. */
void ixo_uniaxial_entry_r(float n_o, float n_e, const float c[3], const float kk[3], const float nn[3], int pol,
                          float out[4]) {
    v3 k = vnormalize(vfromr(kk)), normal = vnormalize(vfromr(nn)), cax = vnormalize(vfromr(c));
    float cos_i = clampf(vdot(vneg(k), normal), 0.0f, 1.0f);
    float sin_i = sqrtf(maxf(fmaf(cos_i, -cos_i, 1.0f), 0.0f));
    uframe F = uframe_build(k, normal, cax, cos_i, sin_i);
    entry_inc I = entry_incidence_frame(1.0f, &F);
    entry_sol ss, sp;
    entry_solve_pair(&I, 1.0f, n_o, n_e, cax, &F, &ss, &sp);
    const entry_sol* so = pol == 0 ? &ss : &sp;
    out[0] = so->r_s.re; out[1] = so->r_s.im; out[2] = so->r_p.re; out[3] = so->r_p.im;
}

static int scene_build(const ixg_scene* sc, scene_t* S) {
    slab_fill(&S->slab, sc->planes, sc->n_planes);
    S->planes = sc->planes;
    S->finish = sc->facet_finishes;
    S->n_planes = sc->n_planes;
    S->mat = sc->material;
    S->max_bounces = sc->max_bounces;
    float g = sc->environment.surface_glare;
    S->surface_glare = (g >= 1.0f || isnan(g)) ? 1.0f : (g > 0.0f ? g : 0.0f);
    S->aniso = !sc->material->cubic && fabsf(sc->material->birefringence_delta) > 1e-4f;
    S->biax = sc->material->biaxial != 0;
    S->trichroic = S->biax && sc->material->has_beta_bands;
    if (S->biax && !S->aniso) return 0; /* a biaxial stone is anisotropic; this form is declined */
    S->absorbing = sc->material->n_bands_o > 0 || sc->material->n_bands_e > 0 ||
                   (S->trichroic && sc->material->n_bands_beta > 0);
    S->c_raw = vfrom(sc->material->c_axis);
    S->tcn = vnormalize_or_zero(S->c_raw);
    stable_basis(S->tcn, &S->ta1, &S->ta2);
    S->hdr = sc->hdr_map;
    if (S->hdr) { /* the panorama: no exposure, observer, tray or white balance */
        memset(&S->env, 0, sizeof S->env);
        S->env.exposure = 1.0f;
        S->env.wb = mk(1.0f, 1.0f, 1.0f);
        S->env.model = -1;
        return 1;
    }
    return env_build(&sc->environment, &S->env);
}

/* PathState::finish's estimate (bounce_body.rs:724-764) from the path's final
 * radiance (the exit-split staging already added for an escaped path): the
 * balance-heuristic weight over each channel's MIS family (color.rs:145-172),
 * the CIE integration and von Kries white balance. */
static void path_finish_xyz(const scene_t* S, const path_t* P, float xyz[3]) {
    const float norm_factor = (400.0f / 8.0f) / 106.856f;
    v3 acc = mk(0, 0, 0);
    for (int c = 0; c < NCH; ++c) { /* color.rs:145-172 */
        float family = -0.0f;
        for (int j = 0; j < NCH; ++j)
            if (P->compat[c] & (1u << j)) family += P->path_pdf[j];
        float w = family <= 1e-12f ? 1.0f : (8.0f * P->path_pdf[0]) / family;
        float cmf[3];
        ixo_cie_cmf(P->lambdas[c], cmf);
        float s = (P->radiance[c] * w) * norm_factor;
        acc = vadd(acc, mk(cmf[0] * s, cmf[1] * s, cmf[2] * s));
    }
    acc = vadd(acc, mk(0, 0, 0)); /* + nee_xyz (zero on this path) */
    acc = von_kries(acc, S->env.wb);
    xyz[0] = maxf(acc.x, 0.0f);
    xyz[1] = maxf(acc.y, 0.0f);
    xyz[2] = maxf(acc.z, 0.0f);
}

static void trace_core(const scene_t* S, v3 origin, v3 dir, uint32_t seed, float hero_rand,
                       float xyz[3], ixg_path_events* ev) {
    path_t P;
    memset(&P, 0, sizeof P);
    ixo_wrapped_hero_wavelengths(hero_rand, P.lambdas);
    for (int c = 0; c < NCH; ++c) {
        P.st[c].i = 1.0f;
        P.path_pdf[c] = 1.0f;
        P.compat[c] = 0xFF;
        P.n_o[c] = ixo_dispersion_evaluate(&S->mat->dispersion, P.lambdas[c]);
        if (S->absorbing) {
            P.alpha_o[c] = band_sum(S->mat->bands_o, S->mat->n_bands_o, P.lambdas[c]);
            P.alpha_e[c] = band_sum(S->mat->bands_e, S->mat->n_bands_e, P.lambdas[c]);
            if (S->trichroic) P.alpha_b[c] = band_sum(S->mat->bands_beta, S->mat->n_bands_beta, P.lambdas[c]);
        }
    }
    P.org = origin;
    P.dir = dir;
    P.k = dir;
    P.observer = vneg(dir);
    if (ev) { memset(ev, 0, sizeof *ev); ev->facet_hash = 2166136261u; ev->termination = 3; }
    uint32_t bounce;
    int stopped = 0;
    for (bounce = 0; bounce < S->max_bounces; ++bounce) {
        float o[3] = {P.org.x, P.org.y, P.org.z}, d[3] = {P.dir.x, P.dir.y, P.dir.z}, t;
        const int has_mis = P.has_mis; /* pending_light_mis.take() (bounce_body.rs:372) */
        const float mis_phase = P.mis_phase;
        P.has_mis = 0;
        int32_t f = scene_intersect(S, o, d, &t);
        if (f < 0) { /* handle_miss (bounce_body/miss.rs:33-54) */
            const float w = has_mis ? balance_heuristic(mis_phase, environment_nee_pdf(S, P.dir)) : 1.0f;
            if (S->hdr) { /* sample_environment_channels' HdrMap arm (rig.rs:756-765) */
                float rgb[3];
                hdr_rgb(S->hdr, P.dir, rgb);
                for (int c = 0; c < NCH; ++c)
                    P.radiance[c] = fmaf(st_intensity(P.st[c]) * w, ixo_rgb_to_spectral(rgb, P.lambdas[c]),
                                         P.radiance[c]);
            } else {
                dir_light L = direction_lighting(&S->env, P.dir, P.observer);
                for (int c = 0; c < NCH; ++c) {
                    float env = light_radiance(&L, spec_power(&S->env, P.lambdas[c]), S->env.exposure);
                    P.radiance[c] = fmaf(st_intensity(P.st[c]) * w, env, P.radiance[c]);
                }
            }
            P.escaped = 1;
            if (ev) { ev->termination = 0; ev->bounces = bounce + 1; }
            stopped = 1;
            break;
        }
        if (S->mat->scattering_sigma_s > 0.0f && P.inside) { /* run_bounce, bounce_body.rs:410-416 */
            const int r = scatter_step(S, &P, bounce, seed, t, ev);
            if (r == 0) continue;
            if (r == 3) {
                if (ev) { ev->termination = 4; ev->bounces = bounce + 1; }
                stopped = 1;
                break;
            }
        }
        int stop = 0;
        facet_event(S, &P, bounce, seed, t, f, ev, &stop);
        if (stop) {
            if (ev) { ev->termination = 2; ev->bounces = bounce + 1; }
            stopped = 1;
            break;
        }
    }
    if (!stopped && ev) ev->bounces = S->max_bounces;
    /* PathState::finish (bounce_body.rs:724-764) */
    if (P.escaped)
        for (int c = 0; c < NCH; ++c) P.radiance[c] += P.split[c];
    path_finish_xyz(S, &P, xyz);
}

void ixo_trace_ray(const ixg_scene* sc, const float origin[3], const float dir[3], uint32_t seed,
                   float hero_rand, float xyz[3], ixg_path_events* ev) {
    scene_t S;
    if (!scene_build(sc, &S)) { xyz[0] = xyz[1] = xyz[2] = NAN; return; }
    trace_core(&S, vfromr(origin), vfromr(dir), seed, hero_rand, xyz, ev);
}

static void sample_core(const scene_t* S, const ixg_scene* sc, uint32_t pixel, uint32_t sample,
                        float xyz[3], ixg_path_events* ev) { /* cpu_frame.rs:45-80 */
    uint32_t seed;
    float jx, jy, hero;
    ixo_sample_draws(pixel, sample, &seed, &jx, &jy, &hero);
    uint32_t x = pixel % sc->width, y = pixel / sc->width;
    float o[3], d[3];
    ixg_camera cam = sc->camera;
    cam.width = ABI_F32((float)sc->width);
    cam.height = ABI_F32((float)sc->height);
    ixo_generate_ray(&cam, (float)x, (float)y, jx, jy, o, d);
    trace_core(S, vfromr(o), vfromr(d), seed, hero, xyz, ev);
}

void ixo_trace_sample(const ixg_scene* sc, uint32_t pixel, uint32_t sample, float xyz[3],
                      ixg_path_events* ev) {
    scene_t S;
    if (!scene_build(sc, &S)) { xyz[0] = xyz[1] = xyz[2] = NAN; return; }
    sample_core(&S, sc, pixel, sample, xyz, ev);
}

typedef struct {
    const ixg_scene* sc;
    const scene_t* S;
    uint32_t sample_offset, spp;
    double* accum;
    int t, T;
    uint64_t first, count;
    float* out_xyz;
    ixg_path_events* events;
} job_t;

static void* accumulate_worker(void* arg) { /* hybrid.rs:337-392: thread t owns t, t+T, ... */
    job_t* J = (job_t*)arg;
    uint32_t npix = J->sc->width * J->sc->height;
    for (uint32_t p = (uint32_t)J->t; p < npix; p += (uint32_t)J->T) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (uint32_t s = 0; s < J->spp; ++s) {
            float xyz[3];
            sample_core(J->S, J->sc, p, J->sample_offset + s, xyz, NULL);
            if (isfinite(xyz[0]) && isfinite(xyz[1]) && isfinite(xyz[2])) { /* sampling.rs:237 */
                s0 += (double)xyz[0];
                s1 += (double)xyz[1];
                s2 += (double)xyz[2];
            }
        }
        J->accum[3 * (size_t)p] += s0;
        J->accum[3 * (size_t)p + 1] += s1;
        J->accum[3 * (size_t)p + 2] += s2;
    }
    return NULL;
}

void ixo_accumulate(const ixg_scene* sc, uint32_t sample_offset, uint32_t spp, double* accum,
                    int threads) {
    scene_t S;
    if (!scene_build(sc, &S)) return;
    if (threads < 1) threads = 1;
    pthread_t th[256];
    job_t jobs[256];
    if (threads > 256) threads = 256;
    for (int t = 0; t < threads; ++t) {
        jobs[t] = (job_t){sc, &S, sample_offset, spp, accum, t, threads, 0, 0, NULL, NULL};
        pthread_create(&th[t], NULL, accumulate_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
}

static void* paths_worker(void* arg) {
    job_t* J = (job_t*)arg;
    for (uint64_t i = J->first + (uint64_t)J->t; i < J->first + J->count; i += (uint64_t)J->T) {
        uint32_t pixel = (uint32_t)(i / J->spp), s = (uint32_t)(i % J->spp);
        float xyz[3];
        ixg_path_events ev;
        sample_core(J->S, J->sc, pixel, J->sample_offset + s, xyz, J->events ? &ev : NULL);
        uint64_t k = i - J->first;
        if (J->out_xyz) { J->out_xyz[3 * k] = xyz[0]; J->out_xyz[3 * k + 1] = xyz[1]; J->out_xyz[3 * k + 2] = xyz[2]; }
        if (J->events) J->events[k] = ev;
    }
    return NULL;
}

void ixo_trace_paths(const ixg_scene* sc, uint32_t sample_offset, uint32_t spp, uint64_t first,
                     uint64_t count, float* out_xyz, ixg_path_events* events, int threads) {
    scene_t S;
    if (!scene_build(sc, &S)) return;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    job_t jobs[256];
    for (int t = 0; t < threads; ++t) {
        jobs[t] = (job_t){sc, &S, sample_offset, spp, NULL, t, threads, first, count, out_xyz, events};
        pthread_create(&th[t], NULL, paths_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
}

/* ================================================================== */
/* Metrics: evaluate_optical_image (color/metrics/)                    */
/* ================================================================== */
enum { FATE_ENTRY_BLOCKED = 0, FATE_LEAKED = 1, FATE_UPWARD = 2, FATE_ABSORBED = 3 };
enum { REASON_NONE = 0, REASON_SHADOW = 1, REASON_ILLUM = 2, REASON_SIDE = 3, REASON_NOHIT = 4,
       REASON_BUDGET = 5, REASON_ENTRY = 6, REASON_INVALID = 7 };

typedef struct {
    const ixg_facet_plane* planes;
    uint32_t n;
    slab_t slab;
    float nd, nf, nc;
    env_t env;          /* scoring environment (observer-free, default head-shadow cosines) */
    float spec550;
    float threshold;
    float head_cos;
} mctx_t;

typedef struct {
    int fate;
    v3 dir;
    float transmittance, exit_cos;
    int32_t facet;
    uint32_t bounces;
    /* detail */
    int32_t exit_facet;
    uint32_t d_bounces;
    uint32_t reason;
} fate_t;

static float fresnel_T(float n1, float n2, float ci, float ct) { /* ray_trace.rs:77-89 */
    float ds = fmaf(n2, ct, n1 * ci);
    float dp = fmaf(n2, ci, n1 * ct);
    if (fabsf(ds) < 1e-6f || fabsf(dp) < 1e-6f) return 0.0f;
    float a = fmaf(n2, -ct, n1 * ci) / ds;
    float b = fmaf(n2, -ci, n1 * ct) / dp;
    float rs = a * a, rp = b * b;
    return clampf(1.0f - midpointf(rs, rp), 0.0f, 1.0f);
}

static void trace_entry_m(const mctx_t* M, v3 point, v3 dir, v3 n, float cos_i, float index,
                          fate_t* F) { /* ray_trace.rs:256-425, polished facets */
    memset(F, 0, sizeof *F);
    F->exit_facet = -1;
    F->reason = REASON_BUDGET;
    float sin2_t = (1.0f / (index * index)) * fmaf(cos_i, -cos_i, 1.0f);
    if (sin2_t > 1.0f) { F->fate = FATE_ENTRY_BLOCKED; F->reason = REASON_ENTRY; return; }
    float cos_t = sqrtf(1.0f - sin2_t);
    float inv = 1.0f / index;
    v3 cur = vnormalize(vadd(vmul(dir, inv), vmul(n, fmaf(inv, cos_i, -cos_t))));
    float entry_T = fresnel_T(1.0f, index, cos_i, cos_t);
    v3 hp = vadd(point, vmul(cur, 1e-4f));
    float sin_crit = 1.0f / index;
    int leaked = 0, upward = 0;
    v3 e_dir = mk(0, 0, 0);
    float e_T = 0, e_cos = 0;
    int32_t e_facet = 0;
    uint32_t e_b = 0;
    float path_len = 0.0f;
    for (uint32_t b = 0; b < 10; ++b) {
        F->d_bounces = b;
        float o[3] = {hp.x, hp.y, hp.z}, d[3] = {cur.x, cur.y, cur.z}, t;
        int32_t f = slab_scan(&M->slab, M->planes, M->n, o, d, &t);
        if (f < 0) { F->reason = REASON_NOHIT; break; }
        path_len += t;
        v3 np = vadd(hp, vmul(cur, t));
        v3 n_out = vfrom(M->planes[f].normal);
        float ct = clampf(vdot(cur, n_out), 0.0f, 1.0f);
        float st = sqrtf(maxf(fmaf(ct, -ct, 1.0f), 0.0f));
        if (st < sin_crit) {
            float s2o = (index * index) * fmaf(ct, -ct, 1.0f);
            if (s2o <= 1.0f) {
                float co = sqrtf(1.0f - s2o);
                v3 od = vnormalize(vadd(vmul(cur, index), vmul(n_out, fmaf(index, -ct, co))));
                /* Escape::record (ray_trace.rs:182-208) */
                F->exit_facet = f;
                F->d_bounces = b;
                F->reason = REASON_SIDE;
                if (n_out.y < -0.05f) {
                    F->reason = REASON_NONE;
                    leaked = 1;
                } else if (od.y > 0.05f) {
                    F->reason = REASON_NONE;
                    upward = 1;
                    e_dir = od;
                    e_T = fresnel_T(index, 1.0f, ct, co);
                    e_cos = co;
                    e_facet = f;
                    e_b = b;
                }
                break;
            }
        }
        cur = vnormalize(vsub(cur, vmul(n_out, 2.0f * ct)));
        hp = vadd(np, vmul(cur, 1e-4f));
        F->d_bounces = b + 1;
    }
    (void)path_len;
    if (leaked) F->fate = FATE_LEAKED;
    else if (upward) {
        F->fate = FATE_UPWARD;
        F->dir = e_dir;
        F->transmittance = entry_T * e_T;
        F->exit_cos = e_cos;
        F->facet = e_facet;
        F->bounces = e_b;
    } else F->fate = FATE_ABSORBED;
}

static float scoring_radiance(const mctx_t* M, v3 dir) { /* lighting.rs:80-100 */
    dir_light L = direction_lighting(&M->env, dir, mk(0, 0, 0));
    return light_radiance(&L, M->spec550, 1.0f);
}

enum { CLS_NONE = -1, CLS_WINDOWED = 0, CLS_EXTINCT = 1, CLS_RETURNED = 2, CLS_OTHER = 3 };

/* classify_ray (classify.rs:109-197); returns class or CLS_NONE on a miss. */
static int classify_m(const mctx_t* M, v3 org, v3 dir, v3 observer, int want_fire, int* has_fire,
                      float* angle, float* weight, int32_t* entry_facet, fate_t* detail) {
    float o[3] = {org.x, org.y, org.z}, d[3] = {dir.x, dir.y, dir.z}, t;
    int32_t f = slab_scan(&M->slab, M->planes, M->n, o, d, &t);
    if (f < 0) return CLS_NONE;
    *entry_facet = f;
    v3 hp = vadd(org, vmul(dir, t));
    v3 n = vfrom(M->planes[f].normal);
    float cos_i = clampf(vdot(vneg(dir), n), 0.0f, 1.0f);
    fate_t F;
    trace_entry_m(M, hp, dir, n, cos_i, M->nd, &F);
    *has_fire = 0;
    if (detail) *detail = F;
    switch (F.fate) {
    case FATE_ENTRY_BLOCKED: return CLS_OTHER;
    case FATE_LEAKED: return CLS_WINDOWED;
    case FATE_ABSORBED: return CLS_EXTINCT;
    default: break;
    }
    /* returned_or_reason (visibility.rs:55-67) */
    if (vdot(F.dir, observer) > M->head_cos) {
        if (detail) detail->reason = REASON_SHADOW;
        return CLS_EXTINCT;
    }
    if (!(scoring_radiance(M, F.dir) > M->threshold)) {
        if (detail) detail->reason = REASON_ILLUM;
        return CLS_EXTINCT;
    }
    if (want_fire) { /* fire_sample (classify.rs:199-256) */
        fate_t A, B;
        trace_entry_m(M, hp, dir, n, cos_i, M->nf, &A);
        trace_entry_m(M, hp, dir, n, cos_i, M->nc, &B);
        if (A.fate == FATE_UPWARD && B.fate == FATE_UPWARD && A.facet == B.facet &&
            A.bounces == B.bounces) {
            float cs = clampf(vdot(A.dir, B.dir), -1.0f, 1.0f);
            *angle = acosf(cs) * 57.2957795130823208767981548141051703f;
            float energy = F.transmittance * A.transmittance * B.transmittance;
            float radiance = maxf(A.exit_cos, 0.0f) * maxf(B.exit_cos, 0.0f);
            *weight = energy * radiance;
            *has_fire = 1;
        }
    }
    return CLS_RETURNED;
}

static void offset_m(uint32_t s, float* jx, float* jy) { /* image.rs:21-36 */
    if (s == 0) { *jx = 0.5f; *jy = 0.5f; return; }
    double radical = (double)bit_reverse(s) / 4294967296.0;
    uint32_t n = s;
    double scale = 1.0 / 3.0, y = 0.0;
    while (n > 0) {
        y += (double)(n % 3) * scale;
        n /= 3;
        scale /= 3.0;
    }
    *jx = (float)radical;
    *jy = (float)y;
}

static float head_shadow_cone_cos(int studio_rig, float deg) { /* visibility.rs:22-40 */
    if (studio_rig || isnan(deg) || deg == 16.0f) return 0.96f;
    if (deg <= 0.0f) return 2.0f;
    float offset = acosf(0.96f) - 16.0f * (F32_PI / 180.0f);
    return cosf(clampf(deg, 1.0f, 89.0f) * (F32_PI / 180.0f) + offset);
}

static int cmp_total(const void* a, const void* b) {
    float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}


static int mctx_build(const ixg_facet_plane* planes, uint32_t n, const ixg_dispersion* disp,
                      const ixg_environment* env, mctx_t* M) { /* prepared.rs:64-94, lighting.rs:133-190 */
    M->planes = planes;
    M->n = n;
    slab_fill(&M->slab, planes, n);
    M->nd = maxf(ixo_dispersion_evaluate(disp, 589.3f), 1.1f);
    M->nf = maxf(ixo_dispersion_evaluate(disp, 486.1f), 1.001f);
    M->nc = maxf(ixo_dispersion_evaluate(disp, 656.3f), 1.001f);
    /* the scoring rig: exposure 1, observer-free, default head-shadow cosines */
    ixg_environment scoring = *env;
    scoring.exposure = 1.0f;
    scoring.head_shadow_deg = 16.0f;
    scoring.backdrop = 0.0f; /* the tray is not part of the scoring radiance (lighting.rs:148-152) */
    if (!env_build(&scoring, &M->env)) return 0;
    M->spec550 = spec_power(&M->env, 550.0f);
    float s[96];
    for (int i = 0; i < 96; ++i) { /* hemisphere_direction (lighting.rs:115-121) */
        float y = ((float)i + 0.5f) / 96.0f;
        float ring = sqrtf(maxf(fmaf(y, -y, 1.0f), 0.0f));
        float a = (float)i * 2.3999632f;
        float sa = sinf(a), ca = cosf(a);
        s[i] = scoring_radiance(M, mk(ring * ca, y, ring * sa));
    }
    qsort(s, 96, sizeof(float), cmp_total);
    int uniform = M->env.model != MODEL_STUDIO;
    float t = s[48] * (uniform ? 0.5f : 4.0f);
    M->threshold = maxf(t, 1.17549435e-38f);
    M->head_cos = head_shadow_cone_cos(M->env.model == MODEL_STUDIO, env->head_shadow_deg);
    return 1;
}

static void camera_ray_m(const ixg_camera* c, float sx, float sy, float pw, float ph, v3* o, v3* d) {
    ixg_camera cam = *c;
    cam.width = ABI_F32(pw);
    cam.height = ABI_F32(ph);
    float oo[3], dd[3];
    ixo_generate_ray(&cam, sx, sy, 0.0f, 0.0f, oo, dd);
    *o = vfromr(oo);
    *d = vfromr(dd);
}

static float fractions2(const ixg_optical_pixel* p) {
    uint32_t h = p->hits > 1 ? p->hits : 1;
    return (float)p->classes[2] / (float)h;
}

int ixo_metric_image(const ixg_metric_request* r, ixg_optical_pixel* pixels, ixg_metric_totals* tot) {
    if (r->samples == 0 || r->samples > 64 || (uint64_t)r->width * r->height == 0 ||
        (uint64_t)r->width * r->height > 1048576ull)
        return IXG_RESOURCE_LIMIT;
    if (r->want_scintillation && !(r->has_pose && r->pose[2] > 0.0f)) return IXG_INVALID;
    mctx_t M;
    if (!mctx_build(r->planes, r->n_planes, &r->material->dispersion, &r->environment, &M))
        return IXG_DECLINED;
    uint32_t pw = r->projection[0] ? r->projection[0] : r->width;
    uint32_t ph = r->projection[1] ? r->projection[1] : r->height;
    uint32_t ox = r->has_tile_origin ? r->tile_origin[0] : 0, oy = r->has_tile_origin ? r->tile_origin[1] : 0;
    float sc0 = r->has_tile_origin ? 1.0f : (float)pw / (float)r->width;
    float sc1 = r->has_tile_origin ? 1.0f : (float)ph / (float)r->height;
    int zero_disp = M.nf == M.nc;
    ixg_camera tcams[5];
    if (r->want_scintillation) { /* temporal_cameras (image_scint.rs:35-48) */
        static const float offs[5] = {-3.0f, -1.5f, 0.0f, 1.5f, 3.0f};
        for (int i = 0; i < 5; ++i) {
            ixo_camera_new(r->pose[0] + offs[i] * (F32_PI / 180.0f), r->pose[1], r->pose[2], 42.0f,
                           pw, ph, 1, &tcams[i]);
            tcams[i].fov_tan = r->camera.fov_tan;
        }
    }
    ixg_optical_pixel total;
    memset(&total, 0, sizeof total);
    for (uint32_t y = 0; y < r->height; ++y)
        for (uint32_t x = 0; x < r->width; ++x) {
            ixg_optical_pixel p;
            memset(&p, 0, sizeof p);
            p.centre_entry_facet = -1;
            p.centre_exit_facet = -1;
            int have_first = 0;
            int32_t first = -1;
            for (uint32_t s = 0; s < r->samples; ++s) { /* PixelSampler::sample_pixel */
                float jx, jy;
                offset_m(s, &jx, &jy);
                v3 o, d;
                camera_ray_m(&r->camera, fmaf((float)x + jx, sc0, (float)ox - 0.5f),
                             fmaf((float)y + jy, sc1, (float)oy - 0.5f), (float)pw, (float)ph, &o, &d);
                int hf;
                float ang = 0, w = 0;
                int32_t ef = -1;
                fate_t det;
                int cls = classify_m(&M, o, d, vneg(d), r->want_fire, &hf, &ang, &w, &ef, &det);
                if (cls == CLS_NONE) continue;
                p.hits++;
                if (s == 0) {
                    p.centre_entry_facet = ef;
                    p.centre_exit_facet = det.exit_facet;
                    p.centre_bounces = det.d_bounces;
                    p.centre_reason = det.reason;
                }
                if (have_first) p.mixed_facets |= (uint32_t)(ef != first);
                else { first = ef; have_first = 1; }
                if (cls == CLS_RETURNED && r->want_fire) {
                    if (hf) {
                        if (isfinite(ang) && isfinite(w)) {
                            p.fire_sum_deg = ABI_F32(fmaf(zero_disp ? 0.0f : ang, w, p.fire_sum_deg));
                            p.fire_qualified++;
                        } else p.fire_rejected++;
                    } else p.fire_rejected++;
                }
                p.classes[cls]++;
            }
            if (p.hits > 0 && r->want_scintillation) { /* TemporalSampler::apply */
                int ret[5];
                for (int i = 0; i < 5; ++i) {
                    v3 o, d;
                    camera_ray_m(&tcams[i], fmaf((float)x + 0.5f, sc0, (float)ox - 0.5f),
                                 fmaf((float)y + 0.5f, sc1, (float)oy - 0.5f), (float)pw, (float)ph, &o, &d);
                    int hf;
                    float a, w;
                    int32_t ef;
                    int cls = classify_m(&M, o, d, vneg(d), 0, &hf, &a, &w, &ef, NULL);
                    ret[i] = cls == CLS_NONE ? -1 : (cls == CLS_RETURNED);
                }
                if (ret[2] >= 0) {
                    uint8_t n = (uint8_t)(ret[2] == 1);
                    for (int i = 0; i < 5; ++i)
                        if (i != 2 && ret[i] == 1) n++;
                    p.scint_returned = n;
                    p.scint_poses = 5;
                }
            }
            total.hits += p.hits;
            for (int c = 0; c < 4; ++c) total.classes[c] += p.classes[c];
            total.fire_sum_deg += p.fire_sum_deg;
            total.fire_qualified += p.fire_qualified;
            total.fire_rejected += p.fire_rejected;
            if (pixels) pixels[(size_t)y * r->width + x] = p;
        }
    if (tot) {
        memset(tot, 0, sizeof *tot);
        tot->total = total;
        uint32_t h = total.hits > 1 ? total.hits : 1;
        tot->windowing_pct = ABI_F32((float)total.classes[0] / (float)h * 100.0f);
        tot->extinction_pct = ABI_F32((float)total.classes[1] / (float)h * 100.0f);
        tot->brilliance_pct = ABI_F32((float)total.classes[2] / (float)h * 100.0f);
        tot->fire_index = ABI_F32(maxf(total.fire_sum_deg / (float)h * 275.0f, 0.1f));
        /* image_scintillation_pct (image_scint.rs:261-287) needs every pixel */
        tot->scintillation_pct = NAN;
    }
    return IXG_OK;
}

/* image_scintillation_pct over a finished pixel array (image_scint.rs:261-287). */
float ixo_image_scintillation_pct(const ixg_optical_pixel* px, uint32_t n) {
    float sum = 0.0f, sum_sq = 0.0f;
    uint32_t count = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (px[i].hits > 0) {
            float f = fractions2(&px[i]);
            sum += f;
            sum_sq = fmaf(f, f, sum_sq);
            count++;
        }
    float var_sum = 0.0f;
    uint32_t poses = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (px[i].scint_poses) {
            float fr = (float)px[i].scint_returned / (float)px[i].scint_poses;
            var_sum += fmaf(fr, -fr, fr);
            poses++;
        }
    float spatial = 0.0f;
    if (count) { /* spatial_scintillation_pct (scintillation.rs:146-171) */
        float mean = sum / (float)count;
        if (mean > 1e-4f) {
            float var = maxf(fmaf(mean, -mean, sum_sq / (float)count), 0.0f);
            float cv = sqrtf(var) / mean;
            spatial = clampf(cv / (1.0f + cv) * 100.0f, 0.0f, 100.0f);
        }
    }
    float temporal = poses ? clampf(((var_sum / (float)poses) / 0.24f) * 100.0f, 0.0f, 100.0f) : 0.0f;
    return clampf(fmaf(0.4f, temporal, 0.6f * spatial), 0.0f, 100.0f); /* scintillation.rs:194 */
}

/* spatial_scintillation_pct (scintillation.rs:146-171) */
static float spatial_pct(float sum, float sum_sq, uint32_t count) {
    if (count == 0) return 0.0f;
    float mean = sum / (float)count;
    if (mean <= 1e-4f) return 0.0f;
    float var = maxf(fmaf(mean, -mean, sum_sq / (float)count), 0.0f);
    float cv = sqrtf(var) / mean;
    return clampf(cv / (1.0f + cv) * 100.0f, 0.0f, 100.0f);
}

/* window_spatial (image_scint.rs:150-173): the 5x5 window of map pixels
 * around (x, y), clipped at the frame, hit pixels only, row-major. */
static float window_spatial_m(const ixg_optical_pixel* px, uint32_t w, uint32_t h, uint32_t x, uint32_t y) {
    float sum = 0.0f, sum_sq = 0.0f;
    uint32_t count = 0;
    uint32_t y0 = y >= 2 ? y - 2 : 0, y1 = y + 3 < h ? y + 3 : h;
    uint32_t x0 = x >= 2 ? x - 2 : 0, x1 = x + 3 < w ? x + 3 : w;
    for (uint32_t wy = y0; wy < y1; ++wy)
        for (uint32_t wx = x0; wx < x1; ++wx) {
            const ixg_optical_pixel* q = &px[(size_t)wy * w + wx];
            if (q->hits > 0) {
                float f = fractions2(q);
                sum += f;
                sum_sq = fmaf(f, f, sum_sq);
                count++;
            }
        }
    return spatial_pct(sum, sum_sq, count);
}

/* scintillation_spatial (image_scint.rs:175-199): 0 where hits == 0 and
 * everywhere when width * height is not the pixel count. */
void ixo_scintillation_spatial(const ixg_optical_pixel* px, uint32_t n, uint32_t w, uint32_t h, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = 0.0f;
    if ((uint64_t)w * h != n) return;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            if (px[(size_t)y * w + x].hits > 0) out[(size_t)y * w + x] = window_spatial_m(px, w, h, x, y);
}

/* scintillation_spatial_at (image_scint.rs:201-232): the window laid out as
 * its own small frame, read by the same window function. */
float ixo_scintillation_spatial_at(const ixg_optical_pixel* px, uint32_t n, uint32_t w, uint32_t h, uint32_t i) {
    if ((uint64_t)w * h != n || i >= n) return 0.0f;
    uint32_t col = i % w, row = i / w;
    uint32_t x0 = col >= 2 ? col - 2 : 0, x1 = col + 3 < w ? col + 3 : w;
    uint32_t y0 = row >= 2 ? row - 2 : 0, y1 = row + 3 < h ? row + 3 : h;
    ixg_optical_pixel local[25];
    uint32_t fw = x1 - x0, fh = y1 - y0, k = 0;
    for (uint32_t wy = y0; wy < y1; ++wy)
        for (uint32_t wx = x0; wx < x1; ++wx) local[k++] = px[(size_t)wy * w + wx];
    if (local[(row - y0) * fw + (col - x0)].hits == 0) return 0.0f;
    return window_spatial_m(local, fw, fh, col - x0, row - y0);
}

/* OpticalImage::temporal_variance (image_types.rs:105-112): 1 and the
 * Bernoulli variance of the five poses, or 0 (None). */
static int temporal_variance_m(const ixg_optical_pixel* px, uint32_t n, uint32_t i, float* v) {
    if (i >= n || px[i].scint_poses == 0) return 0;
    float fr = (float)px[i].scint_returned / (float)px[i].scint_poses;
    *v = fmaf(fr, -fr, fr);
    return 1;
}

/* combine_scintillation_pct (scintillation.rs:194) of the spatial term and
 * temporal_scintillation_pct (scintillation.rs:173-180) of one pixel. */
static float index_of(float spatial, float variance) {
    float temporal = clampf(((variance / 1.0f) / 0.24f) * 100.0f, 0.0f, 100.0f);
    return clampf(fmaf(0.4f, temporal, 0.6f * spatial), 0.0f, 100.0f);
}

/* scintillation_index (image_scint.rs:234-247): 1 and the index of pixel i
 * from spatial[] (0 past its end), or 0 (None). */
int ixo_scintillation_index(const ixg_optical_pixel* px, uint32_t n, const float* spatial, uint32_t n_spatial,
                            uint32_t i, float* out) {
    float v;
    if (!temporal_variance_m(px, n, i, &v)) return 0;
    *out = index_of(i < n_spatial ? spatial[i] : 0.0f, v);
    return 1;
}

/* scintillation_index for every pixel: out[i] = the index, NaN for None. */
void ixo_scintillation_indices(const ixg_optical_pixel* px, uint32_t n, const float* spatial, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        float v;
        out[i] = ixo_scintillation_index(px, n, spatial, n, i, &v) ? v : NAN;
    }
}

/* scintillation_index_at (image_scint.rs:249-259). */
int ixo_scintillation_index_at(const ixg_optical_pixel* px, uint32_t n, uint32_t w, uint32_t h, uint32_t i,
                               float* out) {
    float v;
    if (!temporal_variance_m(px, n, i, &v)) return 0;
    *out = index_of(ixo_scintillation_spatial_at(px, n, w, h, i), v);
    return 1;
}

/* offset (image.rs:21-36) and temporal_cameras (image_scint.rs:35-48). */
void ixo_image_offset(uint32_t s, float* jx, float* jy) { offset_m(s, jx, jy); }

void ixo_temporal_cameras(const float pose[3], float fov_tan, uint32_t width, uint32_t height, ixg_camera out[5]) {
    static const float offs[5] = {-3.0f, -1.5f, 0.0f, 1.5f, 3.0f}; /* SCINT_TEMPORAL_YAW_OFFSETS_DEG */
    for (int i = 0; i < 5; ++i) {
        ixo_camera_new(pose[0] + offs[i] * (F32_PI / 180.0f), pose[1], pose[2], 42.0f, width, height, 1, &out[i]);
        out[i].fov_tan = ABI_F32(fov_tan);
    }
}

/* classify_ray (classify.rs:109-197) of one ray against an explicit
 * observer direction: the class (0 windowed, 1 extinct, 2 returned,
 * 3 entry-blocked / invalid, -1 no hit). */
int ixo_classify_ray(const ixg_facet_plane* planes, uint32_t n, const ixg_dispersion* disp,
                     const ixg_environment* env, const float origin[3], const float dir[3], const float observer[3]) {
    mctx_t M;
    if (!mctx_build(planes, n, disp, env, &M)) return -2;
    int hf;
    float a = 0, w = 0;
    int32_t ef;
    return classify_m(&M, vfromr(origin), vfromr(dir), vfromr(observer), 0, &hf, &a, &w, &ef, NULL);
}

/* ================================================================== */
/* Pose metrics: evaluate_gem_optical_metrics (color/metrics/evaluate.rs) */
/* ================================================================== */

/* ---- stone width (render_setup/stone_width.rs:27-42, geometry/stone_metrics/) ----
 * Plain f64 restatement: every triple of (deduplicated, pow2-normalised)
 * planes plus the six blank-box planes, no pair pruning (the reference's
 * pruning is proven not to change the visited triples). */
typedef struct { double x, y, z; } d3;
static inline d3 dmk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 dsub(d3 a, d3 b) { return dmk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 dadd(d3 a, d3 b) { return dmk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 dscale(d3 a, double s) { return dmk(a.x * s, a.y * s, a.z * s); }
static inline double ddot(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline d3 dcross(d3 a, d3 b) {
    return dmk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline double dabsmax(d3 v) { return fmax(fmax(fabs(v.x), fabs(v.y)), fabs(v.z)); }

enum { MW_MAX_PLANES = 512 };
#define MW_BLANK 64.0          /* meet_solver/mod.rs:218 */
#define MW_EPS_FEAS 1e-5       /* meet_solver/mod.rs:223 */
#define MW_MIN_DET 1e-6        /* meet_solver/mod.rs:235 */
#define MW_DEDUP 1e-6          /* stone_metrics/mod.rs:112 */
#define MW_EPS_FACE 2e-5       /* stone_metrics/mod.rs:109 */

static int normals_coincide_d(d3 a, d3 b) { /* cuts/asc_schedule.rs:335-339 */
    double la = sqrt(ddot(a, a)), lb = sqrt(ddot(b, b));
    double ra = 1.0 / la, rb = 1.0 / lb;
    d3 ah = (isfinite(ra) && ra > 0.0) ? dscale(a, ra) : dmk(0, 0, 0);
    d3 bh = (isfinite(rb) && rb > 0.0) ? dscale(b, rb) : dmk(0, 0, 0);
    return 1.0 - ddot(ah, bh) <= 2e-7;
}

static double pow2_scale_norm(double s) { /* stone_metrics/mod.rs:152-174 */
    if (!(isfinite(s) && s > 0.0)) return 1.0;
    uint64_t raw;
    memcpy(&raw, &s, 8);
    int64_t lift = 0;
    uint64_t bits = raw;
    if ((raw >> 52) == 0) {
        double l = s * 18446744073709551616.0;
        memcpy(&bits, &l, 8);
        lift = 64;
    }
    int64_t biased = (int64_t)(bits >> 52);
    uint64_t mb = (bits & ((1ull << 52) - 1)) | (1023ull << 52);
    double mant;
    memcpy(&mant, &mb, 8);
    int64_t e = biased - 1023 - lift;
    if (mant >= 1.41421356237309504880) e += 1;
    int64_t be = e + 1023;
    if (be < 1) be = 1;
    if (be > 2046) be = 2046;
    uint64_t ob = (uint64_t)be << 52;
    double out;
    memcpy(&out, &ob, 8);
    return out;
}

/* The triple solve of simd/triple_solve.rs (cofactor inverse times rhs). */
static double solve_triple(d3 a, d3 b, d3 c, double ma, double mb, double mc, d3* v) {
    double t0x = b.y * c.z - b.z * c.y, t0y = c.y * a.z - c.z * a.y, t0z = a.y * b.z - a.z * b.y;
    double det = (a.x * t0x + b.x * t0y) + c.x * t0z;
    double t2x = b.x * c.y - b.y * c.x, t2y = c.x * a.y - c.y * a.x, t2z = a.x * b.y - a.y * b.x;
    double t1x = b.z * c.x - b.x * c.z, t1y = c.z * a.x - c.x * a.z, t1z = a.z * b.x - a.x * b.z;
    double inv = 1.0 / det;
    v->x = ((t0x * inv) * ma + (t0y * inv) * mb) + (t0z * inv) * mc;
    v->y = ((t1x * inv) * ma + (t1y * inv) * mb) + (t1z * inv) * mc;
    v->z = ((t2x * inv) * ma + (t2y * inv) * mb) + (t2z * inv) * mc;
    return det;
}

static int cmp_angle(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/* face_area (stone_metrics/mesh.rs:139-203): the face polygon's vertices in
 * angle order about their centroid, shoelace-summed. */
static double face_area_d(d3 n, double m, const d3* V, int nv) {
    d3* on = (d3*)malloc(sizeof(d3) * (size_t)nv);
    double* ang = (double*)malloc(sizeof(double) * 4 * (size_t)nv);
    int k = 0;
    for (int i = 0; i < nv; ++i)
        if (fabs(ddot(n, V[i]) - m) <= MW_EPS_FACE) on[k++] = V[i];
    double area = 0.0;
    if (k >= 3) {
        d3 seed;
        if (fabs(n.x) <= fabs(n.y) && fabs(n.x) <= fabs(n.z)) seed = dmk(1, 0, 0);
        else if (fabs(n.y) <= fabs(n.z)) seed = dmk(0, 1, 0);
        else seed = dmk(0, 0, 1);
        d3 bu = dsub(seed, dscale(n, ddot(n, seed)));
        bu = dscale(bu, 1.0 / sqrt(ddot(bu, bu)));
        d3 bv = dcross(n, bu);
        d3 c = dmk(0, 0, 0);
        for (int i = 0; i < k; ++i) c = dadd(c, on[i]);
        c = dscale(c, 1.0 / (double)k);
        for (int i = 0; i < k; ++i) {
            d3 d = dsub(on[i], c);
            ang[4 * i] = atan2(ddot(bv, d), ddot(bu, d));
            ang[4 * i + 1] = on[i].x;
            ang[4 * i + 2] = on[i].y;
            ang[4 * i + 3] = on[i].z;
        }
        qsort(ang, (size_t)k, 4 * sizeof(double), cmp_angle);
        d3 s = dmk(0, 0, 0);
        for (int i = 0; i < k; ++i) {
            int j = (i + 1) % k;
            d3 a = dsub(dmk(ang[4 * i + 1], ang[4 * i + 2], ang[4 * i + 3]), c);
            d3 b = dsub(dmk(ang[4 * j + 1], ang[4 * j + 2], ang[4 * j + 3]), c);
            s = dadd(s, dcross(a, b));
        }
        area = 0.5 * fabs(ddot(n, s));
    }
    free(on);
    free(ang);
    return area;
}

/* measure_model_width: the short axis of the feasible vertices' x/z bounding
 * box (measure.rs:79-150).  Returns 0 when the arrangement does not measure
 * (no planes, unbounded, fewer than 4 vertices, non-positive volume). */
/* The vertex set behind measure_solid (measure.rs:79-150): the planes
 * deduplicated and power-of-two normalised into N / Mo (the first *real
 * entries, then the six blank faces), every feasible well-conditioned triple
 * vertex in *V (normalised units; the caller frees it).  Returns 1 when the
 * planes close a finite solid of positive volume. */
static int solid_vertices(const ixg_facet_plane* planes, uint32_t n, d3* N, double* Mo, int* real_out,
                          double* scale_out, d3** V_out, int* nv_out) {
    *V_out = NULL;
    *nv_out = 0;
    if (n == 0 || n > MW_MAX_PLANES) return 0;
    int p = 0;
    for (uint32_t i = 0; i < n; ++i) { /* dedup_planes (vertices.rs:67-79) */
        d3 ni = dmk((double)planes[i].normal[0], (double)planes[i].normal[1], (double)planes[i].normal[2]);
        double mi = -(double)planes[i].d;
        int dup = 0;
        for (int j = 0; j < p && !dup; ++j) dup = normals_coincide_d(ni, N[j]) && fabs(mi - Mo[j]) < 1e-9;
        if (!dup) { N[p] = ni; Mo[p] = mi; ++p; }
    }
    double amax = 0.0; /* max_abs_offset (measure.rs:22-33), NaN-propagating */
    for (int i = 0; i < p; ++i) amax = (isnan(amax) || isnan(Mo[i])) ? NAN : fmax(amax, fabs(Mo[i]));
    double scale = pow2_scale_norm(amax);
    for (int i = 0; i < p; ++i) Mo[i] /= scale;
    const int real = p;
    static const double bx[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int i = 0; i < 6; ++i) { N[p] = dmk(bx[i][0], bx[i][1], bx[i][2]); Mo[p] = MW_BLANK; ++p; }
    int cap = 1024, nv = 0;
    d3* V = (d3*)malloc(sizeof(d3) * (size_t)cap);
    int ok = 1;
    for (int a = 0; a < p && ok; ++a)
        for (int b = a + 1; b < p && ok; ++b)
            for (int c = b + 1; c < p; ++c) { /* flush_solid_batch (vertices.rs:96-139) */
                d3 v;
                double det = solve_triple(N[a], N[b], N[c], Mo[a], Mo[b], Mo[c], &v);
                double da = fabs(det);
                if (isnan(da) || da < MW_MIN_DET) continue;
                if (!(isfinite(v.x) && isfinite(v.y) && isfinite(v.z))) continue;
                if (dabsmax(v) > MW_BLANK + 1.0) continue;
                int viol = 0;
                for (int q = 0; q < p && !viol; ++q) viol = ddot(N[q], v) - Mo[q] > MW_EPS_FEAS;
                if (viol) continue;
                if (dabsmax(v) > MW_BLANK - 1.0) { ok = 0; break; }
                int dup = 0; /* VertexAccumulator::insert_if_new */
                for (int q = 0; q < nv && !dup; ++q) dup = dabsmax(dsub(V[q], v)) < MW_DEDUP;
                if (dup) continue;
                if (nv == cap) { cap *= 2; V = (d3*)realloc(V, sizeof(d3) * (size_t)cap); }
                V[nv++] = v;
            }
    if (ok && nv < 4) ok = 0;
    if (ok) {
        double vol = 0.0;
        for (int i = 0; i < real; ++i) vol += Mo[i] * face_area_d(N[i], Mo[i], V, nv) / 3.0;
        if (!(isfinite(vol) && vol > 0.0)) ok = 0;
    }
    *real_out = real;
    *scale_out = scale;
    *V_out = V;
    *nv_out = nv;
    return ok;
}

int ixo_model_width(const ixg_facet_plane* planes, uint32_t n, double* width) {
    d3 N[MW_MAX_PLANES + 6];
    double Mo[MW_MAX_PLANES + 6];
    int real, nv;
    double scale;
    d3* V;
    int ok = solid_vertices(planes, n, N, Mo, &real, &scale, &V, &nv);
    if (ok) {
        double x0 = INFINITY, x1 = -INFINITY, z0 = INFINITY, z1 = -INFINITY;
        for (int i = 0; i < nv; ++i) {
            x0 = fmin(x0, V[i].x); x1 = fmax(x1, V[i].x);
            z0 = fmin(z0, V[i].z); z1 = fmax(z1, V[i].z);
        }
        double dx = x1 - x0, dz = z1 - z0;
        *width = (dx <= dz ? dx : dz) * scale;
    }
    free(V);
    return ok;
}

/* ---- .asc schedules and the solids they cut ---- */

/* raw_asc_planes + dedup_planes (cuts/asc_schedule.rs:152-200, 264-322):
 * one plane per index (one per index-less tier) in schedule order, then the
 * first plane of every near-duplicate group.  Tier t has counts[t] indices,
 * consecutive in idx[]. */
uint32_t ixo_asc_planes(const double* angle_deg, const double* mast, const uint32_t* counts, uint32_t n_tiers,
                        const double* idx, int32_t gear_teeth, double gear_reference_angle,
                        ixg_facet_plane* out, uint32_t cap) {
    uint32_t teeth_abs = gear_teeth < 0 ? 0u - (uint32_t)gear_teeth : (uint32_t)gear_teeth;
    const float teeth = (float)(teeth_abs > 1u ? teeth_abs : 1u);
    const float ref = (float)gear_reference_angle;
    uint32_t raw_n = 0;
    for (uint32_t t = 0; t < n_tiers; ++t) raw_n += counts[t] ? counts[t] : 1u;
    ixg_facet_plane* raw = (ixg_facet_plane*)malloc(sizeof(ixg_facet_plane) * (raw_n ? raw_n : 1u));
    uint32_t r = 0;
    for (uint32_t t = 0; t < n_tiers; ++t) {
        const int crown = angle_deg[t] > 0.0 || (angle_deg[t] == 0.0 && !signbit(angle_deg[t])); /* plane.rs:25 */
        const float theta = (float)fabs(angle_deg[t]) * TO_RAD;
        const float st = sin_cr(theta), ct = cos_cr(theta);
        const float d = -(float)fabs(mast[t]);
        if (counts[t] == 0) {
            plane_new(mk(0.0f, crown ? ct : -ct, st), d, &raw[r++]);
            continue;
        }
        for (uint32_t k = 0; k < counts[t]; ++k) {
            const float phi = 2.0f * F32_PI * ((float)*idx++ + ref) / teeth; /* round_brilliant.rs:64-66 */
            plane_new(mk(st * cos_cr(phi), crown ? ct : -ct, st * sin_cr(phi)), d, &raw[r++]);
        }
    }
    uint32_t kept = 0;
    for (uint32_t i = 0; i < raw_n; ++i) {
        int dup = 0;
        for (uint32_t k = 0; k < kept && !dup; ++k) {
            const ixg_facet_plane* a = &out[k];
            const float sc = fmaxf(fmaxf(fabsf(a->d), fabsf(raw[i].d)), 1.0f);
            dup = normals_coincide_d(dmk(a->normal[0], a->normal[1], a->normal[2]),
                                     dmk(raw[i].normal[0], raw[i].normal[1], raw[i].normal[2])) &&
                  fabsf(a->d - raw[i].d) <= (1.0f / 2048.0f) * sc;
        }
        if (!dup) {
            if (kept == cap) { free(raw); return 0; }
            out[kept++] = raw[i];
        }
    }
    free(raw);
    return kept;
}

typedef struct { double x, z; int id; } outline_pt;

static int cmp_outline(const void* a, const void* b) {
    const outline_pt* p = (const outline_pt*)a;
    const outline_pt* q = (const outline_pt*)b;
    if (p->x != q->x) return p->x < q->x ? -1 : 1;
    if (p->z != q->z) return p->z < q->z ? -1 : 1;
    return (p->id > q->id) - (p->id < q->id);
}

static double outline_turn(outline_pt o, outline_pt a, outline_pt b) { /* brep.rs:573-575 */
    return fma(a.z - o.z, -(b.x - o.x), (a.x - o.x) * (b.z - o.z));
}

/* Incident-plane lists compared lexicographically (the solid's vertex order). */
static const uint8_t* g_incident;
static int g_incident_stride;
static int cmp_incident(const void* a, const void* b) {
    const uint8_t* p = g_incident + (size_t)(*(const int*)a) * (size_t)g_incident_stride;
    const uint8_t* q = g_incident + (size_t)(*(const int*)b) * (size_t)g_incident_stride;
    for (int k = 0; k < g_incident_stride; ++k) {
        /* the list holding the lower plane index first sorts first */
        if (p[k] != q[k]) return p[k] ? -1 : 1;
    }
    return 0;
}

/* The solid's girdle outline (GemPolyhedron::girdle_outline, brep.rs:544-570:
 * the x/z convex hull of the vertices, points within 1e-6 of the bounding
 * radius merged keeping the first vertex in incident-set order) and the
 * length/width ratio of asc_designs.rs:63-88 over it, in f32.  *untouched is
 * the number of planes that carry no face.  Returns 0 when the planes do not
 * close a solid. */
int ixo_solid_outline(const ixg_facet_plane* planes, uint32_t n, double* lw, uint32_t* untouched) {
    d3 N[MW_MAX_PLANES + 6];
    double Mo[MW_MAX_PLANES + 6];
    int real, nv;
    double scale;
    d3* V;
    if (!solid_vertices(planes, n, N, Mo, &real, &scale, &V, &nv)) { free(V); return 0; }
    uint32_t bare = 0;
    for (int i = 0; i < real; ++i) bare += !(face_area_d(N[i], Mo[i], V, nv) > 0.0);
    *untouched = bare;
    /* vertices in incident-set order */
    uint8_t* inc = (uint8_t*)calloc((size_t)nv * (size_t)real, 1);
    int* order = (int*)malloc(sizeof(int) * (size_t)nv);
    for (int v = 0; v < nv; ++v) {
        order[v] = v;
        for (int q = 0; q < real; ++q) inc[(size_t)v * real + q] = fabs(ddot(N[q], V[v]) - Mo[q]) <= MW_EPS_FACE;
    }
    g_incident = inc;
    g_incident_stride = real;
    qsort(order, (size_t)nv, sizeof(int), cmp_incident);
    float* P = (float*)malloc(sizeof(float) * 3 * (size_t)nv); /* the solid's f32 vertices */
    double radius = 0.0;
    for (int k = 0; k < nv; ++k) {
        const d3 v = V[order[k]];
        P[3 * k] = (float)(v.x * scale);
        P[3 * k + 1] = (float)(v.y * scale);
        P[3 * k + 2] = (float)(v.z * scale);
        const double r2 = (double)P[3 * k] * (double)P[3 * k] + (double)P[3 * k + 1] * (double)P[3 * k + 1] +
                          (double)P[3 * k + 2] * (double)P[3 * k + 2];
        radius = fmax(radius, sqrt(r2));
    }
    const double eps = 1e-6 * (double)(float)radius;
    outline_pt* pts = (outline_pt*)malloc(sizeof(outline_pt) * (size_t)nv);
    for (int k = 0; k < nv; ++k) { pts[k].x = (double)P[3 * k]; pts[k].z = (double)P[3 * k + 2]; pts[k].id = k; }
    qsort(pts, (size_t)nv, sizeof(outline_pt), cmp_outline);
    int m = 0;
    for (int k = 0; k < nv; ++k) {
        if (m > 0 && fabs(pts[k].x - pts[m - 1].x) <= eps && fabs(pts[k].z - pts[m - 1].z) <= eps) continue;
        pts[m++] = pts[k];
    }
    outline_pt* hull = (outline_pt*)malloc(sizeof(outline_pt) * (size_t)(2 * m + 2));
    int h = 0;
    for (int k = 0; k < m; ++k) { /* lower chain */
        while (h >= 2 && outline_turn(hull[h - 2], hull[h - 1], pts[k]) <= 0.0) --h;
        hull[h++] = pts[k];
    }
    --h;
    const int base = h;
    for (int k = m - 1; k >= 0; --k) { /* upper chain */
        while (h - base >= 2 && outline_turn(hull[h - 2], hull[h - 1], pts[k]) <= 0.0) --h;
        hull[h++] = pts[k];
    }
    --h;
    int bi = 0, bj = 0;
    float best = 0.0f;
    for (int i = 0; i < h; ++i)
        for (int j = i + 1; j < h; ++j) {
            const float* a = &P[3 * hull[i].id];
            const float* b = &P[3 * hull[j].id];
            const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
            const float d = sqrtf((dx * dx + dy * dy) + dz * dz);
            if (d > best) { best = d; bi = i; bj = j; }
        }
    const float* a = &P[3 * hull[bi].id];
    const float* b = &P[3 * hull[bj].id];
    v3 dir = vnormalize(mk(b[0] - a[0], b[1] - a[1], b[2] - a[2]));
    const v3 perp = mk(-dir.z, 0.0f, dir.x);
    float lo = FLT_MAX, hi = -FLT_MAX;
    for (int i = 0; i < h; ++i) {
        const float* q = &P[3 * hull[i].id];
        const float proj = vdot(mk(q[0], q[1], q[2]), perp);
        lo = fminf(lo, proj);
        hi = fmaxf(hi, proj);
    }
    *lw = (double)(best / (hi - lo));
    free(hull); free(pts); free(P); free(order); free(inc); free(V);
    return 1;
}

/* girdle_half_width (metrics/fan.rs:49-60). */
float ixo_girdle_half_width(const ixg_facet_plane* planes, uint32_t n) {
    double w;
    if (ixo_model_width(planes, n, &w) && isfinite(w) && w > 0.0) return (float)(w * 0.5);
    return 1.0f;
}

/* FanGeometry::origin (fan.rs:92-94). */
static v3 fan_origin(v3 f, v3 r, v3 u, float standoff, float lateral, float gu, float gv) {
    return vadd(vadd(vmul(vneg(f), standoff), vmul(r, gu * lateral)), vmul(u, gv * lateral));
}

/* evaluate_pose (evaluate.rs:327-439 with accumulate.rs:112-173 and
 * scintillation.rs:104-198).  out = {brilliance, fire, scintillation,
 * windowing, extinction}; want_all = 0 is the Fire-less pass (fire 25,
 * scintillation 75). */
int ixo_pose_metrics(const ixg_facet_plane* planes, uint32_t n, const ixg_material* mat,
                     const ixg_environment* env, float yaw, float pitch, int want_all, float out[5]) {
    if (n == 0) { /* placeholder_metrics (evaluate.rs:259-267) */
        out[0] = 85.0f; out[1] = 25.0f; out[2] = 75.0f; out[3] = 5.0f; out[4] = 5.0f;
        return IXG_OK;
    }
    mctx_t M;
    if (!mctx_build(planes, n, &mat->dispersion, env, &M)) return IXG_DECLINED;
    const float hw = ixo_girdle_half_width(planes, n);
    const float standoff = 2.5f * hw;
    const float lateral = 0.95f * hw / sqrtf(0.70f);
    ixg_camera cam;
    ixo_camera_new(yaw, pitch, 1.0f, 42.0f, 1, 1, 1, &cam);
    const v3 F = vfrom(cam.forward), R = vfrom(cam.right), U = vfrom(cam.up);
    v3 TF[5], TR[5], TU[5];
    static const float offs[5] = {-3.0f, -1.5f, 0.0f, 1.5f, 3.0f};
    for (int k = 0; k < 5; ++k) {
        ixg_camera c;
        ixo_camera_new(yaw + offs[k] * (F32_PI / 180.0f), pitch, 1.0f, 42.0f, 1, 1, 1, &c);
        TF[k] = vfrom(c.forward); TR[k] = vfrom(c.right); TU[k] = vfrom(c.up);
    }
    static const float ap[5][2] = {{0.0f, 0.0f}, {0.08f, 0.0f}, {-0.08f, 0.0f}, {0.0f, 0.08f}, {0.0f, -0.08f}};
    uint32_t total = 0, windowed = 0, extinct = 0, returned = 0, cells = 0;
    float fire = 0.0f, cf = 0.0f, cf2 = 0.0f, tv = 0.0f;
    for (int ix = 0; ix < 18; ++ix)
        for (int iz = 0; iz < 18; ++iz) {
            const float gu = fmaf(((float)ix + 0.5f) / 18.0f, 2.0f, -1.0f);
            const float gv = fmaf(((float)iz + 0.5f) / 18.0f, 2.0f, -1.0f);
            if (fmaf(gv, gv, gu * gu) > 0.70f) continue;
            uint32_t ct = 0, cr = 0;
            for (int s = 0; s < 5; ++s) { /* sample_cell (evaluate.rs:380-439) */
                v3 d = vnormalize(vadd(vadd(F, vmul(R, ap[s][0])), vmul(U, ap[s][1])));
                v3 o = fan_origin(F, R, U, standoff, lateral, gu, gv);
                int hf = 0;
                float ang = 0.0f, w = 0.0f;
                int32_t ef;
                int cls = classify_m(&M, o, d, vneg(F), want_all, &hf, &ang, &w, &ef, NULL);
                if (cls == CLS_NONE) continue;
                total++;
                ct++;
                if (cls == CLS_WINDOWED) windowed++;
                else if (cls == CLS_EXTINCT) extinct++;
                else if (cls == CLS_RETURNED) {
                    returned++;
                    cr++;
                    if (hf) fire = fmaf(ang, w, fire);
                }
            }
            if (ct > 0) {
                float frac = (float)cr / (float)ct;
                cf += frac;
                cf2 = fmaf(frac, frac, cf2);
                cells++;
                if (want_all) { /* cell_temporal_variance (scintillation.rs:104-128) */
                    uint32_t k5 = 0;
                    for (int k = 0; k < 5; ++k) {
                        v3 o = fan_origin(TF[k], TR[k], TU[k], standoff, lateral, gu, gv);
                        int hf;
                        float a, w;
                        int32_t ef;
                        if (classify_m(&M, o, TF[k], vneg(TF[k]), 0, &hf, &a, &w, &ef, NULL) == CLS_RETURNED) k5++;
                    }
                    float pr = (float)k5 / 5.0f;
                    tv += fmaf(pr, -pr, pr);
                }
            }
        }
    const float nt = (float)(total > 1 ? total : 1);
    out[3] = clampf((float)windowed / nt * 100.0f, 0.0f, 100.0f);
    out[4] = clampf((float)extinct / nt * 100.0f, 0.0f, 100.0f);
    out[0] = clampf((float)returned / nt * 100.0f, 0.0f, 100.0f);
    if (!want_all) {
        out[1] = 25.0f;
        out[2] = 75.0f;
        return IXG_OK;
    }
    out[1] = maxf(fire / nt * 275.0f, 0.1f);
    float spatial = 0.0f, temporal = 0.0f;
    if (cells) {
        float mean = cf / (float)cells;
        if (mean > 1e-4f) {
            float var = maxf(fmaf(mean, -mean, cf2 / (float)cells), 0.0f);
            float cv = sqrtf(var) / mean;
            spatial = clampf(cv / (1.0f + cv) * 100.0f, 0.0f, 100.0f);
        }
        temporal = clampf(((tv / (float)cells) / 0.24f) * 100.0f, 0.0f, 100.0f);
    }
    out[2] = clampf(fmaf(0.4f, temporal, 0.6f * spatial), 0.0f, 100.0f);
    return IXG_OK;
}

#ifndef IXO_NO_PROBE
/* ------------------------------------------------------------------ */
/* Per-function probes: the CPU side of the reference's tier-2 checks  */
/* (renderer/gpu/ulp.rs and the *_check.rs modules under renderer/gpu/), */
/* each case the oracle's own function for ixg_probe's id, on the same  */
/* records (layouts with the ids in include/ixg.h).                     */
/* ------------------------------------------------------------------ */
static uint32_t f_bits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static float bits_f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }
static stokes st_in(const float* p) { stokes s = {p[0], p[1], p[2], p[3]}; return s; }
static void st_out(float* p, stokes s) { p[0] = s.i; p[1] = s.q; p[2] = s.u; p[3] = s.v; }
static void v_out(float* p, v3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
static void c_out(float* p, cplx c) { p[0] = c.re; p[1] = c.im; }

/* compute_bounce_refraction_geometry's frame for an incident wave normal k on
 * a normal facing the incident side (geometry.rs:306-425). */
static uframe probe_frame(v3 k, v3 normal, v3 c) {
    float cos_i = clampf(vdot(vneg(k), normal), 0.0f, 1.0f);
    float sin_i = sqrtf(maxf(fmaf(cos_i, -cos_i, 1.0f), 0.0f));
    return uframe_build(k, normal, c, cos_i, sin_i);
}

static const uint8_t kProbeIn[IXG_PROBE_COUNT] = {2, 2, 1, 1, 7, 6, 8, 8, 7, 8, 19, 6, 8, 9, 3, 11, 13,
                                                   15, 6, 5, 8, 26, 6, 1, 1, 4, 5, 5, 8, 8};
static const uint8_t kProbeOut[IXG_PROBE_COUNT] = {12, 6, 3, 1, 1, 4, 5, 6, 4, 24, 8, 3, 3, 1, 2, 18, 13,
                                                    5, 3, 3, 3, 3, 2, 1, 1, 2, 3, 3, 3, 1};

/* the scene's biaxial indicatrix at lambda (materials/optics.rs:16-27) */
static biax_t probe_biax(const scene_t* S, float lambda) {
    float nb = ixo_dispersion_evaluate(&S->mat->dispersion, lambda), na = nb - S->mat->biaxial_delta_beta_alpha;
    return biax_from_gamma_axis(na, nb, na + S->mat->birefringence_delta, S->c_raw);
}

int ixo_probe_shape(int32_t fn, uint32_t* in_floats, uint32_t* out_floats) {
    if (fn < 0 || fn >= IXG_PROBE_COUNT) return 0;
    if (in_floats) *in_floats = kProbeIn[fn];
    if (out_floats) *out_floats = kProbeOut[fn];
    return 1;
}

int ixo_probe(const ixg_scene* sc, int32_t fn, const float* in, uint32_t n, float* out) {
    if (fn < 0 || fn >= IXG_PROBE_COUNT) return 0;
    scene_t S;
    if (!scene_build(sc, &S)) return 0;
    for (uint32_t r = 0; r < n; ++r) {
        const float* a = in + (size_t)r * kProbeIn[fn];
        float* o = out + (size_t)r * kProbeOut[fn];
        switch (fn) {
        case IXG_PROBE_SAMPLE_DRAWS: {
            uint32_t seed;
            float jx, jy, hero, lam[NCH];
            ixo_sample_draws(f_bits(a[0]), f_bits(a[1]), &seed, &jx, &jy, &hero);
            ixo_wrapped_hero_wavelengths(hero, lam);
            o[0] = jx; o[1] = jy; o[2] = hero; o[3] = bits_f(seed);
            for (int c = 0; c < NCH; ++c) o[4 + c] = lam[c];
            break;
        }
        case IXG_PROBE_CAMERA_RAY: { /* sample_core's camera half (cpu_frame.rs:45-80) */
            uint32_t seed, pixel = f_bits(a[0]);
            float jx, jy, hero, org[3], dir[3];
            ixo_sample_draws(pixel, f_bits(a[1]), &seed, &jx, &jy, &hero);
            ixg_camera cam = sc->camera;
            cam.width = ABI_F32((float)sc->width);
            cam.height = ABI_F32((float)sc->height);
            ixo_generate_ray(&cam, (float)(pixel % sc->width), (float)(pixel / sc->width), jx, jy, org, dir);
            for (int j = 0; j < 3; ++j) { o[j] = org[j]; o[3 + j] = dir[j]; }
            break;
        }
        case IXG_PROBE_CIE_CMF:
            ixo_cie_cmf(a[0], o);
            break;
        case IXG_PROBE_ILLUMINANT:
            o[0] = spec_power(&S.env, a[0]);
            break;
        case IXG_PROBE_ENVIRONMENT:
            o[0] = env_channel(&S, vfromr(a), a[6], vfromr(a + 3));
            break;
        case IXG_PROBE_FRAME_ROTATION: { /* rotate_stokes_to_plane_of_incidence (absorption.rs:69-87) */
            mat4 R = mueller_frame_rotation(atan2_cr(a[1], a[0]));
            st_out(o, st_apply(&R, st_in(a + 2)));
            break;
        }
        case IXG_PROBE_FRESNEL_REFLECTION: {
            float rs, rp;
            fresnel_rs_rp(a[0], a[1], a[2], a[3], &rs, &rp);
            mat4 M = mueller_fresnel_reflection(rs, rp);
            st_out(o, st_apply(&M, st_in(a + 4)));
            o[4] = clampf(0.5f * fmaf(rp, rp, rs * rs), R_UNPOL_PDF_MIN, R_UNPOL_PDF_MAX);
            break;
        }
        case IXG_PROBE_FRESNEL_TRANSMISSION: { /* exit_split.rs:69-101 */
            float num = 2.0f * a[0] * a[2];
            float t_s = num / fmaf(a[1], a[3], a[0] * a[2]);
            float t_p = num / fmaf(a[0], a[3], a[1] * a[2]);
            mat4 T = mueller_fresnel_transmission(a[0], a[1], a[2], a[3], t_s, t_p);
            st_out(o, st_apply(&T, st_in(a + 4)));
            fresnel_rs_rp(a[0], a[1], a[2], a[3], &o[4], &o[5]);
            break;
        }
        case IXG_PROBE_TIR_RETARDATION: {
            mat4 M = mueller_tir_retardation(tir_phase_delta(a[0], a[1], a[2]));
            st_out(o, st_apply(&M, st_in(a + 3)));
            break;
        }
        case IXG_PROBE_CHANNEL_OPTICS: /* trace_core's per-channel n(lambda) and spectral_absorption */
            for (int c = 0; c < NCH; ++c) {
                o[c] = ixo_dispersion_evaluate(&sc->material->dispersion, a[c]);
                o[NCH + c] = band_sum(sc->material->bands_o, sc->material->n_bands_o, a[c]);
                o[2 * NCH + c] = band_sum(sc->material->bands_e, sc->material->n_bands_e, a[c]);
            }
            break;
        case IXG_PROBE_PLEOCHROIC_ALPHA: { /* apply_absorption's per-channel alpha */
            v3 k = vfromr(a);
            v3 ea = ordinary_eigen(k, S.c_raw);
            v3 eb = vnormalize_or_zero(vcross(k, ea));
            for (int c = 0; c < NCH; ++c)
                o[c] = midpointf(tensor_qf(&S, a[3 + c], a[3 + NCH + c], ea), tensor_qf(&S, a[3 + c], a[3 + NCH + c], eb));
            break;
        }
        case IXG_PROBE_ORDINARY_EIGEN:
            v_out(o, ordinary_eigen(vfromr(a), vfromr(a + 3)));
            break;
        case IXG_PROBE_EXTRAORDINARY_POYNTING:
            v_out(o, extraordinary_poynting_dir(vfromr(a), vfromr(a + 3), a[6], a[7]));
            break;
        case IXG_PROBE_THETA_C: {
            path_t P;
            memset(&P, 0, sizeof P);
            P.lambdas[0] = a[0];
            v3 k = vfromr(a + 2), normal = vfromr(a + 5);
            float cos_i = clampf(vdot(vneg(k), normal), 0.0f, 1.0f);
            o[0] = theta_c_for_bounce(&S, &P, normal, k, cos_i, a[8] != 0.0f, a[1]);
            break;
        }
        case IXG_PROBE_UNIAXIAL_INDICES:
            o[0] = e_index_at(&S, a[0], a[1]);
            o[1] = effective_extraordinary_index(a[1], o[0], a[2]);
            break;
        case IXG_PROBE_ENTRY_SOLVE: {
            v3 c = vfromr(a + 2);
            uframe F = probe_frame(vfromr(a + 5), vfromr(a + 8), c);
            entry_inc I = entry_incidence_frame(1.0f, &F);
            entry_sol sol[2];
            entry_solve_pair(&I, 1.0f, a[0], a[1], c, &F, &sol[0], &sol[1]);
            for (int p = 0; p < 2; ++p) {
                c_out(o + 8 * p, sol[p].r_s);
                c_out(o + 8 * p + 2, sol[p].r_p);
                c_out(o + 8 * p + 4, sol[p].t_o);
                c_out(o + 8 * p + 6, sol[p].t_e);
            }
            o[16] = sol[0].flux_o;
            o[17] = sol[0].flux_e;
            break;
        }
        case IXG_PROBE_INTERNAL_SOLVE:
        case IXG_PROBE_EXIT_TRANSMISSION: {
            v3 c = vfromr(a + 3);
            uframe F = probe_frame(vfromr(a + 6), vfromr(a + 9), c);
            internal_sol sol = internal_solve(a[0], a[1], a[2], c, &F, a[12] != 0.0f);
            if (fn == IXG_PROBE_EXIT_TRANSMISSION) {
                float iu;
                st_out(o, uniaxial_exit_transmission(&sol, a[14], a[13], &iu));
                o[4] = iu;
                break;
            }
            c_out(o, sol.r_o);
            c_out(o + 2, sol.r_e);
            c_out(o + 4, sol.t_s);
            c_out(o + 6, sol.t_p);
            o[8] = sol.flux_ro; o[9] = sol.flux_re; o[10] = sol.flux_ts; o[11] = sol.flux_tp; o[12] = sol.flux_inc;
            break;
        }
        case IXG_PROBE_HG_DIRECTION:
            v_out(o, hg_direction(a[0], a[1], a[2], vfromr(a + 3)));
            break;
        case IXG_PROBE_COSINE_HEMISPHERE:
            v_out(o, cosine_weighted_hemisphere(a[0], a[1], vfromr(a + 2)));
            break;
        case IXG_PROBE_SHADING_NORMAL:
            v_out(o, shading_normal_near_edge(sc->planes, sc->n_planes, vfromr(a), (int32_t)f_bits(a[3]), vfromr(a + 4), a[7]));
            break;
        case IXG_PROBE_FINISH: {
            path_t P;
            memset(&P, 0, sizeof P);
            uint32_t lo = f_bits(a[0]), hi = f_bits(a[1]);
            for (int c = 0; c < NCH; ++c) {
                P.compat[c] = (uint8_t)((c < 4 ? lo >> (8 * c) : hi >> (8 * (c - 4))) & 0xFFu);
                P.path_pdf[c] = a[2 + c];
                P.lambdas[c] = a[2 + NCH + c];
                P.radiance[c] = a[2 + 2 * NCH + c];
            }
            path_finish_xyz(&S, &P, o);
            break;
        }
        case IXG_PROBE_INTERSECT: {
            float t = 0.0f;
            int32_t f = scene_intersect(&S, a, a + 3, &t);
            o[0] = bits_f((uint32_t)f);
            o[1] = f >= 0 ? t : 0.0f;
            break;
        }
        case IXG_PROBE_EXP_LANE:
            o[0] = ixo_exp_lane(a[0]);
            break;
        case IXG_PROBE_EXP: /* e^x in double precision, rounded to float */
            o[0] = (float)exp((double)a[0]);
            break;
        case IXG_PROBE_BIAX_WAVE_INDICES: {
            biax_t B = probe_biax(&S, a[0]);
            biax_wave_indices(&B, vfromr(a + 1), &o[0], &o[1]);
            break;
        }
        case IXG_PROBE_BIAX_EIGEN: {
            biax_t B = probe_biax(&S, a[0]);
            v_out(o, biax_eigen_d(&B, vfromr(a + 1), a[4] != 0.0f));
            break;
        }
        case IXG_PROBE_BIAX_MODE_POYNTING: {
            biax_t B = probe_biax(&S, a[0]);
            v_out(o, biax_mode_poynting(&B, vfromr(a + 1), a[4] != 0.0f));
            break;
        }
        case IXG_PROBE_BIAX_RESOLVE_ENTRY: {
            biax_t B = probe_biax(&S, a[0]);
            v3 k = vfromr(a + 1), normal = vfromr(a + 4), w;
            float cos_i = clampf(vdot(vneg(k), normal), 0.0f, 1.0f);
            (void)biax_resolve_entry(&B, k, normal, cos_i, B.nb, a[7] != 0.0f, &w);
            v_out(o, w);
            break;
        }
        case IXG_PROBE_BIAX_ASSIGNED_ALPHA: {
            biax_t B = probe_biax(&S, a[0]);
            o[0] = biax_tensor_qf(biax_assigned_locals(&S, &B, vfromr(a + 1), a[4] != 0.0f), a[5], a[6], a[7]);
            break;
        }
        default:
            break;
        }
    }
    return 1;
}
#endif
