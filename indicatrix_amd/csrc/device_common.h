// device_common.h -- device arithmetic shared by the film and metric kernels.
//
// Everything here follows the reference's f32 evaluation order (glam's
// left-associated dot, multiply-then-add, explicit FMA only where the
// reference writes mul_add); callers compile with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ixg.h"
#include "exp_ff.h"
#include "transport_params.h"

#define IX_TABLE_QUAL __constant__ static const
#include "spectral_tables.h"

namespace ixg {

#define NCH 8
#define DIRECTION_MATCH_COS_TOL (1.0f - 1e-6f)
#define R_UNPOL_PDF_MIN 1e-4f
#define R_UNPOL_PDF_MAX (1.0f - 1e-4f)
#define R_UNPOL_SELECT_MIN 0.02f
#define R_UNPOL_SELECT_MAX 0.98f
#define FRESNEL_BRANCH_STREAM 0x9E3779B1u
#define FROSTED_DIR_U_STREAM 0x27D4EB2Fu
#define DISTANCE_SAMPLE_STREAM 0xA24BAED4u
#define PHASE_DIR_U_STREAM 0x9FB21C65u
#define PHASE_DIR_V_STREAM 0x1CE4E5B9u
#define FROSTED_DIR_V_STREAM 0x165667B1u
#define RUSSIAN_ROULETTE_STREAM 0x517CC1B7u

// ---------------------------------------------------------------------------
// scalar / vector helpers (glam order: left-associated dot, no contraction)
// ---------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
// The correctly rounded f32 quotient a / b for operands whose quotient and
// residuals stay clear of the overflow and underflow ranges (indices of
// refraction, cosines, Fresnel terms, unit-vector lengths, path lengths): the
// reciprocal, one Newton step and two residual corrections of the device's
// IEEE division, without the operand pre-scaling and special-case fix-up that
// such operands never need, so the bits are the same.  b == 0 gives NaN (IEEE:
// +-inf), and a numerator of -0 gives +0; callers pass neither where it
// matters.
__device__ __forceinline__ float div_rn(float a, float b) {
    float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = a * y;
    float r = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}
// The correctly rounded f32 square root for an operand that is zero or a
// normal number: the hardware root (within one ULP) moved to the neighbour
// the residual's sign selects, as the device's IEEE square root does after
// its denormal-range scaling, which these operands never need.  Zero, +inf
// and NaN pass through unchanged.
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float lo = __int_as_float(__float_as_int(s) - 1);
    const float hi = __int_as_float(__float_as_int(s) + 1);
    const float r_lo = __builtin_fmaf(-lo, s, x);
    const float r_hi = __builtin_fmaf(-hi, s, x);
    const float t = r_lo <= 0.0f ? lo : s;
    return r_hi > 0.0f ? hi : t;
}
__device__ __forceinline__ V3 normalize(V3 a) { return a * div_rn(1.0f, sqrt_rn(dot(a, a))); }
// A zero vector gives the zero vector (1 / 0 is NaN here, not inf: both fail
// the finiteness test).
__device__ __forceinline__ V3 normalize_or_zero(V3 a) {
    float r = div_rn(1.0f, sqrt_rn(dot(a, a)));
    return (__builtin_isfinite(r) && r > 0.0f) ? a * r : v3(0.0f, 0.0f, 0.0f);
}
// Rust f32::clamp (NaN propagates)
__device__ __forceinline__ float clampf(float v, float lo, float hi) {
    v = v < lo ? lo : v;
    return v > hi ? hi : v;
}
__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// powi with LLVM's binary-exponentiation order
template <int N>
__device__ __forceinline__ float powi(float x) {
    float res = 1.0f;
    bool have = false;
    float sq = x;
    int n = N;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        if (n & 1) {
            res = have ? res * sq : sq;
            have = true;
        }
        n >>= 1;
        if (n) sq = sq * sq;
    }
    return res;
}

__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ float unit_draw(uint32_t seed, uint32_t bounce, uint32_t stream) {
    return div_rn((float)hash_u32(seed ^ hash_u32(bounce ^ stream)), 4294967295.0f);
}
// radical_inverse (sampling.rs): digit k of n in base B weighted by the
// digit scale 1/B divided by B k more times, each step correctly rounded --
// the scales are the same f32 sequence for every n, so they are a table
// (folded by the compiler, whose f32 division is the correctly rounded one):
// one fma per digit instead of a division.
template <uint32_t B, int N>
struct DigitScales {
    float v[N];
    constexpr DigitScales() : v() {
        float x = 1.0f / (float)B;
        for (int i = 0; i < N; ++i) {
            v[i] = x;
            x = x / (float)B;
        }
    }
};
__constant__ const DigitScales<3, 21> kDigitScale3{};  // 3^21 > 2^32: every u32 index
__constant__ const DigitScales<5, 14> kDigitScale5{};  // 5^14 > 2^32
template <uint32_t B>
__device__ __forceinline__ float radical_inverse(uint32_t n) {
    static_assert(B == 3 || B == 5, "digit-scale tables exist for bases 3 and 5");
    const float* scale = B == 3 ? kDigitScale3.v : kDigitScale5.v;
    float val = 0.0f;
    for (int k = 0; n > 0; ++k) {
        const uint32_t digit = n % B;
        val = fma_((float)digit, scale[k], val);
        n /= B;
    }
    return val;
}
__device__ __forceinline__ float ld_base2(uint32_t n) {
    return (float)__builtin_bitreverse32(n) / 4294967296.0f;
}
__device__ __forceinline__ float cp_rotate(float x, float off) {
    float s = x + off;
    return s - floorf(s);
}

// ---------------------------------------------------------------------------
// spectra
// ---------------------------------------------------------------------------
// DispersionModel::evaluate (optics/dispersion.rs:170-210), lambda >= 380 nm
__device__ __forceinline__ float dispersion(int kind, const float* c, float lambda) {
    float um = lambda * 1e-3f;
    float l2 = um * um;
    if (kind == IXG_DISP_SELLMEIER3) {
        float n2 = 1.0f;
        n2 += div_rn(c[0] * l2, l2 - c[3]);
        n2 += div_rn(c[1] * l2, l2 - c[4]);
        n2 += div_rn(c[2] * l2, l2 - c[5]);
        return sqrt_rn(fmaxf(n2, 1.0f));
    } else if (kind == IXG_DISP_SELLMEIER1) {
        float n2 = 1.0f + div_rn(c[0] * l2, l2 - c[1]);
        return sqrt_rn(fmaxf(n2, 1.0f));
    }
    float l4 = l2 * l2;
    return fmaxf((c[0] + div_rn(c[1], l2)) + div_rn(c[2], l4), 1.0f);
}

__device__ __forceinline__ void cie_cmf(float lambda, float& x, float& y, float& z) {
    const float end = 380.0f + (float)(IX_CMF_ROWS - 1) * 5.0f;
    if (!(lambda >= 380.0f && lambda <= end)) {
        x = y = z = 0.0f;
        return;
    }
    float pos = div_rn(lambda - 380.0f, 5.0f);
    uint32_t i0 = (uint32_t)floorf(pos);
    i0 = i0 > IX_CMF_ROWS - 2 ? IX_CMF_ROWS - 2 : i0;
    float t = pos - (float)i0;
    x = ix_cmf_table[i0][0] + (ix_cmf_table[i0 + 1][0] - ix_cmf_table[i0][0]) * t;
    y = ix_cmf_table[i0][1] + (ix_cmf_table[i0 + 1][1] - ix_cmf_table[i0][1]) * t;
    z = ix_cmf_table[i0][2] + (ix_cmf_table[i0 + 1][2] - ix_cmf_table[i0][2]) * t;
}

__device__ __forceinline__ float d65(float lambda) {
    float c = clampf(lambda, 380.0f, fma_(10.0f, (float)(IX_D65_ROWS - 1), 380.0f));
    float pos = div_rn(c - 380.0f, 10.0f);
    uint32_t i0 = (uint32_t)floorf(pos);
    i0 = i0 > IX_D65_ROWS - 2 ? IX_D65_ROWS - 2 : i0;
    float frac = pos - (float)i0;
    float v0 = ix_d65_table[i0], v1 = ix_d65_table[i0 + 1];
    return div_rn(fma_(frac, v1 - v0, v0), 100.0f);
}

// e^x correctly rounded to f32 (to within the rare ties exp_ff.h states),
// in single precision: what glibc's expf returns for the arguments here (the
// device library's f32 expf is up to one ULP away).  Used where the value
// feeds every channel of a path: the Planck illuminant and the absorption
// band sums.
__device__ __forceinline__ float exp_rn(float x) { return ixg_exp::exp_ff(x); }

// fmodf for finite x and finite y != 0, exactly (the remainder is always an
// f32), in single precision: the largest y 2^k not above the running
// remainder is subtracted until it drops below |y|; each subtraction is
// exact (Sterbenz: the operands lie within a factor 2).  Infinite or NaN x
// gives NaN.
__device__ __forceinline__ float fmod_exact(float x, float y) {
    if (!(fabsf(x) < __builtin_inff())) return x - x;
    const float ay = fabsf(y);
    float r = fabsf(x);
    while (r >= ay) {
        float t = ay;
        while (t <= 0.5f * r) t *= 2.0f;
        r -= t;
    }
    return __builtin_copysignf(r, x);
}

// f32::midpoint: the f32 nearest (a + b) / 2.  For operands whose sum is
// neither subnormal nor beyond the f32 range this is 0.5 * (a + b): halving
// is exact there, and the f32 sum of two f32 values rounds as the exact sum.
__device__ __forceinline__ float midpoint(float a, float b) { return 0.5f * (a + b); }

__device__ __forceinline__ float spec_power(const EnvParams& E, float lambda) {
    if (E.d65) return d65(lambda);
    float e = exp_rn(fminf(div_rn(14388000.0f, lambda * E.bb_t_k), 80.0f));
    float denom = fmaxf(e - 1.0f, 1e-6f);
    float ratio = div_rn(E.bb_denom_560, denom);
    return clampf(powi<5>(div_rn(560.0f, lambda)) * ratio, 0.01f, 20.0f);
}

__device__ __forceinline__ float smoothstep(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * fma_(-2.0f, t, 3.0f);
}

// The lit models' geometric factor for a unit direction d (rig.rs:230-320,
// 461-489): the iso hemisphere, the light tent (walls, black cards, key
// softbox, spark) and the daylight dome (sky gradient, sun aureole), each
// under the observer's head shadow and with the lit backdrop tray replacing
// its own ground; DaylightSun is the dome plus the 0.27 degree sun disc
// (environment/sun.rs), which neither the head shadow nor the tray touch.
__device__ __forceinline__ float lit_factor_unit(const EnvParams& E, V3 d, V3 observer) {
    const float horizon = smoothstep(-0.05f, 0.05f, d.y);
    const float vis = 1.0f - smoothstep(E.shadow_outer, E.shadow_inner, dot(d, observer));
    if (E.model == ENV_ISO_HEMISPHERE)
        return E.tray > 0.0f ? fma_((1.0f - horizon) * vis, E.tray, horizon * vis) : horizon * vis;
    float above, own_ground;
    if (E.model == ENV_LIGHT_TENT) {
        float walls = fma_(0.08f, fmaxf(d.y, 0.0f), 0.14f) * E.tent_walls;
        if (E.tent_flat > 0.0f) walls = fma_(0.18f * E.tent_walls, E.tent_flat, walls * (1.0f - E.tent_flat));
        float card = 0.0f;
        card = fmaxf(card, smoothstep(0.898794f, 0.9612617f, dot(d, v3(E.ring[4][0], E.ring[4][1], E.ring[4][2]))));
        card = fmaxf(card, smoothstep(0.898794f, 0.9612617f, dot(d, v3(E.ring[8][0], E.ring[8][1], E.ring[8][2]))));
        card = fmaxf(card, smoothstep(0.898794f, 0.9612617f, dot(d, v3(E.ring[12][0], E.ring[12][1], E.ring[12][2]))));
        walls *= fma_(card * E.tent_cards, -0.9f, 1.0f);
        const float key = smoothstep(0.7660444f, 0.9396926f, dot(d, v3(E.key[0], E.key[1], E.key[2]))) *
                          (1.4f * E.spot_mult);
        const float spark = smoothstep(0.9961947f, 0.9993908f, dot(d, v3(E.fill[0], E.fill[1], E.fill[2]))) *
                            (5.0f * E.spot_mult) * E.tent_spark;
        above = ((walls + key) + spark) * (horizon * vis);
        own_ground = E.tent_ground;
    } else {  // daylight dome (and the sky of DaylightSun)
        const float sun_dot = dot(d, v3(E.key[0], E.key[1], E.key[2]));
        const float sky = fma_(0.08f, 1.0f - fmaxf(d.y, 0.0f), 0.10f);
        const float glow = fmaxf(sun_dot, 0.0f);
        const float glow2 = glow * glow;
        const float glow4 = glow2 * glow2;
        const float aureole = (glow4 * glow4) * 0.30f;
        above = (sky + aureole) * (horizon * vis);
        own_ground = 0.04f;
        if (E.model == ENV_DAYLIGHT_SUN) {
            const float ground = E.tray > 0.0f ? ((1.0f - horizon) * vis) * E.tray : own_ground * (1.0f - horizon);
            return (above + ground) + (sun_dot >= 0.9999889f ? E.sun : 0.0f);  // SUN_DISC_COS
        }
    }
    const float ground = E.tray > 0.0f ? ((1.0f - horizon) * vis) * E.tray : own_ground * (1.0f - horizon);
    return above + ground;
}

// EnvironmentMap::radiance_rgb (renderer/env_map.rs:303-417): the direction's
// equirectangular (u, v), u wrapping and v clamped, bilinear over the four
// neighbouring texels; the angles correctly rounded.
__device__ __forceinline__ float rem_euclid1(float x, float m) {
    const float r = fmod_exact(x, m);
    return r < 0.0f ? r + fabsf(m) : r;
}
__device__ __forceinline__ void hdr_rgb(const EnvParams& E, V3 dir, float out[3]) {
    const V3 d = normalize_or_zero(dir);
    const float theta = (float)acos((double)clampf(d.y, -1.0f, 1.0f));
    const float phi = (float)atan2((double)d.x, (double)d.z);
    const float pi = 3.14159265358979323846f;
    const float v = div_rn(theta, pi);
    const float u = rem_euclid1(div_rn(phi, 2.0f * pi), 1.0f);
    const float* px = E.hdr_rgb;
    if (E.hdr_w == 1 && E.hdr_h == 1) {
        out[0] = px[0]; out[1] = px[1]; out[2] = px[2];
        return;
    }
    float fx = fma_(rem_euclid1(u, 1.0f), (float)E.hdr_w, -0.5f);
    float fy = fma_(clampf(v, 0.0f, 1.0f), (float)E.hdr_h, -0.5f);
    if (!(fx == fx)) fx = 0.0f;  // a non-finite direction reads a valid texel
    if (!(fy == fy)) fy = 0.0f;
    const float x0 = floorf(fx), y0 = floorf(fy);
    const float tx = fx - x0, ty = fy - y0;
    const int64_t w = E.hdr_w, h = E.hdr_h;
    const int64_t xc = (int64_t)x0, yc = (int64_t)y0;
    const int64_t x0i = ((xc % w) + w) % w, x1i = (((xc + 1) % w) + w) % w;
    const int64_t y0i = yc < 0 ? 0 : (yc > h - 1 ? h - 1 : yc);
    const int64_t y1i = yc + 1 < 0 ? 0 : (yc + 1 > h - 1 ? h - 1 : yc + 1);
    const float* p00 = px + (y0i * w + x0i) * 3;
    const float* p10 = px + (y0i * w + x1i) * 3;
    const float* p01 = px + (y1i * w + x0i) * 3;
    const float* p11 = px + (y1i * w + x1i) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float top = fma_(p10[c], tx, p00[c] * (1.0f - tx));
        const float bottom = fma_(p11[c], tx, p01[c] * (1.0f - tx));
        out[c] = fma_(bottom, ty, top * (1.0f - ty));
    }
}
// rgb_to_spectral_radiance (renderer/env_map_spectrum.rs:84-125): the grey
// part on three broad bumps, the chroma on three narrow ones.
__device__ __forceinline__ float asym_gauss(float x, float mu, float lo, float hi) {
    const float t = div_rn(x - mu, x < mu ? lo : hi);
    return exp_rn(-0.5f * t * t);
}
__device__ __forceinline__ float rgb_to_spectral(const float rgb[3], float lambda) {
    const float r = fmaxf(rgb[0], 0.0f), g = fmaxf(rgb[1], 0.0f), b = fmaxf(rgb[2], 0.0f);
    const float w = fminf(fminf(r, g), b);
    const float cr = r - w, cg = g - w, cb = b - w;
    const float n_r = fma_(0.82398987f, w, fma_(-0.30506778f, w, 0.26391557f * w));
    const float n_g = fma_(-0.2971578f, w, fma_(1.2528949f, w, -0.4611476f * w));
    const float n_b = fma_(0.078837544f, w, fma_(-0.122806296f, w, 1.2317826f * w));
    const float neutral = fma_(n_r, asym_gauss(lambda, 615.0f, 45.0f, 65.0f),
                               fma_(n_g, asym_gauss(lambda, 545.0f, 45.0f, 45.0f),
                                    n_b * asym_gauss(lambda, 465.0f, 40.0f, 45.0f)));
    const float k_r = fma_(1.0995291f, cr, fma_(0.08711245f, cg, 0.01771266f * cb));
    const float k_g = fma_(0.007455747f, cr, fma_(1.3345385f, cg, 0.014020192f * cb));
    const float k_b = fma_(0.02487475f, cr, fma_(0.058899768f, cg, 1.2609048f * cb));
    const float saturated = fma_(k_r, asym_gauss(lambda, 635.0f, 28.0f, 39.2f),
                                 fma_(k_g, asym_gauss(lambda, 540.0f, 28.0f, 28.0f),
                                      k_b * asym_gauss(lambda, 450.0f, 25.2f, 28.0f)));
    return fmaxf(neutral + saturated, 0.0f);
}

// The ASET contrast view (rig.rs:146-176, 205-222, 439-459): green up to 45
// degrees of exit elevation, red from 45 to 75, blue above, each lit by one
// 20 nm FWHM band at its own gain; below the horizon a neutral D65 tray when
// one is set.  The radiance depends on the wavelength itself, not on an
// illuminant power, so it is its own per-channel evaluation.  d is a unit
// direction; no head shadow.
__device__ __forceinline__ float aset_band(float lambda, float centre) {
    const float z = div_rn(lambda - centre, 8.493218f);
    return exp_rn(-0.5f * z * z);
}
__device__ __forceinline__ float aset_radiance(const EnvParams& E, V3 d, float lambda, float exposure) {
    const float horizon = smoothstep(-0.05f, 0.05f, d.y);
    const float s45 = smoothstep(0.6871068f, 0.7271068f, d.y);
    const float s75 = smoothstep(0.9459258f, 0.9859258f, d.y);
    const float red = (s45 * (1.0f - s75)) * horizon;
    const float green = (1.0f - s45) * horizon;
    const float blue = s75 * horizon;
    const float tray_ground = E.tray > 0.0f ? (1.0f - horizon) * E.tray : 0.0f;
    const float r = (red * aset_band(lambda, 610.0f)) * 8.0f;
    const float g = (green * aset_band(lambda, 540.0f)) * 5.0f;
    const float b = (blue * aset_band(lambda, 460.0f)) * 16.0f;
    const float zones = ((r + g) + b) * exposure;
    return tray_ground > 0.0f ? fma_(tray_ground * d65(lambda), exposure, zones) : zones;
}

// Wavelength-independent lighting of one exit direction (rig.rs:379-405).
// The reference appends the lit spots to a list (key, fill, then the ring in
// order) and folds them into each channel in that order; here every spot has
// a fixed slot (0 key, 1 fill, 2 + i ring light i) and a presence bit, so
// all indices are compile-time constants and the list lives in registers.
// Folding the present slots in slot order is the list's order.
struct DirLight {
    float bg;       // studio: backdrop term; lit models: the scaled factor
    float lights[18];
    uint32_t mask;  // bit k: slot k is lit
    float horizon, ground;  // studio with a tray: fma(horizon, rig, ground * sp)
};

__device__ __forceinline__ void direction_lighting(const EnvParams& E, V3 dir, V3 observer,
                                                   DirLight& L) {
    V3 d = normalize(dir);
    L.mask = 0u;
    if (E.model != ENV_STUDIO) {
        L.bg = lit_factor_unit(E, d, observer);
        return;
    }
    L.bg = fmaxf(fma_(0.012f, fma_(d.y, 0.5f, 0.5f), 0.015f), 0.005f) * E.exposure;
    float key_dot = fmaxf(dot(d, v3(E.key[0], E.key[1], E.key[2])), 0.0f);
    if (key_dot > 0.0f) {
        L.lights[0] = powi<28>(key_dot) * 12.0f * E.spot_mult * E.exposure;
        L.mask |= 1u;
    }
    float fill_dot = fmaxf(dot(d, v3(E.fill[0], E.fill[1], E.fill[2])), 0.0f);
    if (fill_dot > 0.0f) {
        L.lights[1] = powi<18>(fill_dot) * 4.5f * E.exposure;
        L.mask |= 2u;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        float rd = fmaxf(dot(d, v3(E.ring[i][0], E.ring[i][1], E.ring[i][2])), 0.0f);
        if (rd > 0.96f) {
            float spark = (rd - 0.96f) / 0.04f;
            L.lights[2 + i] = powi<6>(spark) * 22.0f * E.spot_mult * E.exposure;
            L.mask |= 4u << i;
        }
    }
    if (E.tray > 0.0f) {  // the tray under the rig (rig.rs:358-364)
        L.horizon = smoothstep(-0.05f, 0.05f, d.y);
        L.ground = ((1.0f - L.horizon) * E.tray) * E.exposure;
    }
}

__device__ __forceinline__ float light_radiance(const EnvParams& E, const DirLight& L, float sp) {
    if (E.model != ENV_STUDIO) return L.bg * (sp * E.exposure);
    float r = L.bg * sp;
#pragma unroll
    for (int i = 0; i < 18; ++i)
        if (L.mask & (1u << i)) r = fma_(L.lights[i], sp, r);
    if (E.tray > 0.0f) r = fma_(L.horizon, r, L.ground * sp);
    return r;
}

// The lit models need no light list: evaluate them without the DirLight
// scratch so the common configs stay in registers.
__device__ __forceinline__ float lit_factor(const EnvParams& E, V3 dir, V3 observer) {
    return lit_factor_unit(E, normalize(dir), observer);
}

// ---------------------------------------------------------------------------
// intersection against the LDS facet set
// ---------------------------------------------------------------------------
// The reference scan (RT/intersect.rs:101-132, simd/slab.rs:147-160): one
// correctly rounded division per plane, t_near the largest entering t and
// t_far the smallest exiting t, ties to the lowest index.
//
// The quotient is div_rn: the scan's operands are unit-normal dot products
// and offsets, |b| <= 1 and, for every plane whose quotient is used,
// |b| > 1e-7, with |a| far from the overflow and underflow ranges.  A plane
// whose |b| <= 1e-7 gets an unused value.

// Written without branches: every plane's quotient is formed and the running
// bests are updated by selects, and a parallel plane with the origin outside
// it (the reference's early miss) is remembered and reported after the loop,
// which returns the same result because the early return has no side effects.
// Straight-line code keeps the wave's lanes (different rays) in lock step.
// One plane of the scan, folded into the running bests.
struct ScanState {
    float t_near, t_far;
    int near_idx, far_idx;
    bool outside_parallel;
};
__device__ __forceinline__ void scan_plane(ScanState& R, float4 p, int i, V3 o, V3 d) {
    const float denom = (p.x * d.x + p.y * d.y) + p.z * d.z;
    const float side = p.w + ((p.x * o.x + p.y * o.y) + p.z * o.z);
    const float t = div_rn(-side, denom);
    const bool valid = fabsf(denom) > 1e-7f;
    const bool up_near = valid && denom < 0.0f && t > R.t_near;
    const bool up_far = valid && !(denom < 0.0f) && t < R.t_far;
    R.t_near = up_near ? t : R.t_near;
    R.near_idx = up_near ? i : R.near_idx;
    R.t_far = up_far ? t : R.t_far;
    R.far_idx = up_far ? i : R.far_idx;
    R.outside_parallel = R.outside_parallel || (!valid && side > 0.0f);
}

// UNROLL planes per iteration, their loads issued together (the facet set is
// wave-uniform: from LDS one broadcast read each, from memory one scalar load
// each), then folded in plane order -- the same updates in the same order.
template <int UNROLL = 1>
__device__ __forceinline__ int intersect(const float4* __restrict__ pl, uint32_t n, V3 o, V3 d, float& t_out) {
    ScanState R{-1e30f, 1e30f, -1, -1, false};
    uint32_t i = 0;
    if (UNROLL > 1) {
        for (; i + UNROLL <= n; i += UNROLL) {
            float4 q[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) q[u] = pl[i + u];
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) scan_plane(R, q[u], (int)(i + u), o, d);
        }
    }
    for (; i < n; ++i) scan_plane(R, pl[i], (int)i, o, d);
    const float t_near = R.t_near, t_far = R.t_far;
    const int near_idx = R.near_idx, far_idx = R.far_idx;
    const bool outside_parallel = R.outside_parallel;
    if (outside_parallel) return -1;
    if (t_near > t_far) return -1;
    if (t_near > 1e-4f) {
        t_out = t_near;
        return near_idx < 0 ? 0 : near_idx;
    }
    if (t_far > 1e-4f) {
        t_out = t_far;
        return far_idx < 0 ? 0 : far_idx;
    }
    return -1;
}

// exp_f32x8's lane polynomial (simd/exp_poly.rs:37-60): Cody-Waite range
// reduction by ln 2 in two parts, a degree-5 minimax polynomial, and the
// exponent built in the float's bits.  Few-ULP, and exactly 1 at +-0.
__device__ __forceinline__ float exp_lane(float x) {
    const float LOG2E = 1.44269504088896340736f;
    const float C1 = __uint_as_float(0x3F318000u);
    const float C2 = -2.1219444e-4f;
    const float HI = 88.02875f, LO = -87.33654f;
    float upper = x < HI ? x : HI;
    x = upper > LO ? upper : LO;
    const float n = floorf(fma_(x, LOG2E, 0.5f));
    x = fma_(n, -C1, x);
    x = fma_(n, -C2, x);
    const float z = x * x;
    float p = 1.9875691e-4f;
    p = fma_(p, x, 1.3981999e-3f);
    p = fma_(p, x, 8.333452e-3f);
    p = fma_(p, x, 4.1665796e-2f);
    p = fma_(p, x, 1.6666665e-1f);
    p = fma_(p, x, 5.0000003e-1f);
    p = fma_(p, z, x) + 1.0f;
    return p * __uint_as_float((uint32_t)(((int32_t)n + 127) << 23));
}

// AbsorptionBand::evaluate (optics/absorption.rs:108-138).
__device__ __forceinline__ float band_eval(const AbsBand& b, float lambda) {
    float t;
    if (b.shape == 0) {
        t = (lambda - b.center_nm) / b.width;
    } else {
        const float delta = b.center_nm - lambda;
        t = (1.0e7f * delta / (lambda * b.center_nm)) / b.width;
    }
    return b.peak * exp_rn(-0.5f * t * t);
}

}  // namespace ixg
