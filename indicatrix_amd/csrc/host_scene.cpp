// host_scene.cpp -- scene-constant math evaluated once per call on the host.
#include "host_scene.h"

#include <cmath>
#include <cstring>

#define IX_TABLE_QUAL static const
#include "spectral_tables.h"

namespace ixg {

bool preset_info(int32_t preset, PresetInfo* o) {
    // environment/mod.rs:321-410 (params) and :519-590 (model, D65, white balance)
    switch (preset) {
    case IXG_PRESET_INCANDESCENT: *o = {ENV_STUDIO, 0, 3200.0f, 1.2f, 1}; return true;
    case IXG_PRESET_RING_LIGHTS: *o = {ENV_STUDIO, 0, 5000.0f, 1.6f, 1}; return true;
    case IXG_PRESET_DARK_SPOTLIGHT: *o = {ENV_STUDIO, 0, 6000.0f, 2.4f, 1}; return true;
    case IXG_PRESET_ILLUMINANT_A: *o = {ENV_STUDIO, 0, 2856.0f, 1.2f, 1}; return true;
    case IXG_PRESET_DAYLIGHT: *o = {ENV_STUDIO, 1, 6500.0f, 1.0f, 0}; return true;
    case IXG_PRESET_ISO_HEMISPHERE: *o = {ENV_ISO_HEMISPHERE, 1, 6500.0f, 1.0f, 0}; return true;
    case IXG_PRESET_LIGHT_TENT: *o = {ENV_LIGHT_TENT, 0, 5000.0f, 1.0f, 1}; return true;
    case IXG_PRESET_WHITE_TRAY: *o = {ENV_LIGHT_TENT, 1, 5000.0f, 0.0f, 0, {5.0f, 0.0f, 0.0f, 0.8f, 1.0f}}; return true;
    case IXG_PRESET_SHOP_LIGHTS: *o = {ENV_LIGHT_TENT, 1, 5000.0f, 1.0f, 0, {3.0f, 0.0f, 1.0f, 0.05f, 0.0f}}; return true;
    case IXG_PRESET_WINDOW_DAYLIGHT: *o = {ENV_LIGHT_TENT, 1, 5000.0f, 1.0f, 0, {0.4f, 0.0f, 0.0f, 0.03f, 0.0f}}; return true;
    case IXG_PRESET_DAYLIGHT_DOME: *o = {ENV_DAYLIGHT_DOME, 1, 6500.0f, 1.0f, 0}; return true;
    case IXG_PRESET_DAYLIGHT_SUN: *o = {ENV_DAYLIGHT_SUN, 1, 6500.0f, 1.0f, 0}; return true;
    case IXG_PRESET_ASET: *o = {ENV_ASET, 1, 6500.0f, 1.0f, 0}; return true;
    default: return false;
    }
}

void cie_cmf_host(float lambda, float out[3]) {
    const float end = 380.0f + (float)(IX_CMF_ROWS - 1) * 5.0f;
    if (!(lambda >= 380.0f && lambda <= end)) {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    float pos = (lambda - 380.0f) / 5.0f;
    size_t i0 = (size_t)std::floor(pos);
    if (i0 > IX_CMF_ROWS - 2) i0 = IX_CMF_ROWS - 2;
    float t = pos - (float)i0;
    for (int c = 0; c < 3; ++c)
        out[c] = ix_cmf_table[i0][c] + (ix_cmf_table[i0 + 1][c] - ix_cmf_table[i0][c]) * t;
}

float d65_host(float lambda) {
    float c = hclamp(lambda, 380.0f, std::fma(10.0f, (float)(IX_D65_ROWS - 1), 380.0f));
    float pos = (c - 380.0f) / 10.0f;
    size_t i0 = (size_t)std::floor(pos);
    if (i0 > IX_D65_ROWS - 2) i0 = IX_D65_ROWS - 2;
    float frac = pos - (float)i0;
    float v0 = ix_d65_table[i0], v1 = ix_d65_table[i0 + 1];
    return std::fma(frac, v1 - v0, v0) / 100.0f;
}

float dispersion_host(const ixg_dispersion& d, float lambda) {
    if (lambda < 380.0f) lambda = 380.0f;
    float um = lambda * 1e-3f;
    float l2 = um * um;
    const float* c = d.coeff;
    if (d.kind == IXG_DISP_SELLMEIER3) {
        float n2 = 1.0f;
        n2 += (c[0] * l2) / (l2 - c[3]);
        n2 += (c[1] * l2) / (l2 - c[4]);
        n2 += (c[2] * l2) / (l2 - c[5]);
        return std::sqrt(std::fmax(n2, 1.0f));
    }
    if (d.kind == IXG_DISP_SELLMEIER1) {
        float n2 = 1.0f + (c[0] * l2) / (l2 - c[1]);
        return std::sqrt(std::fmax(n2, 1.0f));
    }
    float l4 = l2 * l2;
    return std::fmax((c[0] + (c[1] / l2)) + (c[2] / l4), 1.0f);
}

namespace {
struct BB {
    float t_k, denom_560;
};
BB bb_norm(float temp_k) {
    BB n;
    n.t_k = std::fmax(temp_k, 1000.0f);
    float e = std::exp(std::fmin(14388000.0f / (560.0f * n.t_k), 80.0f));
    n.denom_560 = std::fmax(e - 1.0f, 1e-6f);
    return n;
}
float powi5(float x) {
    float x2 = x * x;
    float x4 = x2 * x2;
    return x * x4;
}
const float kXyzToLms[3][3] = {
    {0.8951f, 0.2664f, -0.1614f}, {-0.7502f, 1.7135f, 0.0367f}, {0.0389f, -0.0685f, 1.0296f}};
HV3 to_lms(HV3 a) {
    const float(*m)[3] = kXyzToLms;
    return h3(std::fma(m[0][0], a.x, std::fma(m[0][1], a.y, m[0][2] * a.z)),
              std::fma(m[1][0], a.x, std::fma(m[1][1], a.y, m[1][2] * a.z)),
              std::fma(m[2][0], a.x, std::fma(m[2][1], a.y, m[2][2] * a.z)));
}
}  // namespace

float blackbody(float lambda, float temp_k) {
    BB n = bb_norm(temp_k);
    float e = std::exp(std::fmin(14388000.0f / (lambda * n.t_k), 80.0f));
    float denom = std::fmax(e - 1.0f, 1e-6f);
    float ratio = n.denom_560 / denom;
    return hclamp(powi5(560.0f / lambda) * ratio, 0.01f, 20.0f);
}

void white_balance(float temp_k, float out[3]) {
    HV3 w = h3(0, 0, 0);
    for (int step = 0; step <= 400; ++step) {
        float lambda = 380.0f + (float)step;
        float cmf[3];
        cie_cmf_host(lambda, cmf);
        float b = blackbody(lambda, temp_k);
        w = h3(w.x + cmf[0] * b, w.y + cmf[1] * b, w.z + cmf[2] * b);
    }
    float ty = std::fmax(w.y, 1e-6f);
    const float X = 0.3127f, Y = 0.3290f;
    HV3 target = h3((X / Y) * ty, ty, ((1.0f - X - Y) / Y) * ty);
    HV3 ls = to_lms(w), lt = to_lms(target);
    out[0] = ls.x > 1e-6f ? lt.x / ls.x : 1.0f;
    out[1] = ls.y > 1e-6f ? lt.y / ls.y : 1.0f;
    out[2] = ls.z > 1e-6f ? lt.z / ls.z : 1.0f;
}

bool resolve_environment(const ixg_environment& env, EnvParams* E) {
    PresetInfo pi;
    if (!preset_info(env.preset, &pi)) return false;
    std::memset(E, 0, sizeof *E);
    E->model = pi.model;
    E->d65 = pi.d65;
    BB n = bb_norm(pi.temp_k);
    E->bb_t_k = n.t_k;
    E->bb_denom_560 = n.denom_560;
    E->exposure = env.exposure;
    E->spot_mult = pi.spot_mult;
    E->tray = env.backdrop;
    E->tent_walls = pi.tent[0];
    E->tent_cards = pi.tent[1];
    E->tent_spark = pi.tent[2];
    E->tent_ground = pi.tent[3];
    E->tent_flat = pi.tent[4];
    // head_shadow_cosines (environment/mod.rs:124-137)
    float deg = env.head_shadow_deg;
    if (std::isnan(deg) || deg == 16.0f) {
        E->shadow_outer = 0.9510565f;
        E->shadow_inner = 0.9702957f;
    } else if (deg <= 0.0f) {
        E->shadow_outer = 2.0f;
        E->shadow_inner = 3.0f;
    } else {
        deg = hclamp(deg, 3.0f, 88.0f);
        E->shadow_outer = std::cos((deg + 2.0f) * (kPi / 180.0f));
        E->shadow_inner = std::cos((deg - 2.0f) * (kPi / 180.0f));
    }
    // StudioRig::new (optics/studio_rig.rs:47-80)
    float yaw = env.light_yaw, pitch = env.light_pitch;
    float cp = std::cos(pitch), sp = std::sin(pitch), cy = std::cos(yaw), sy = std::sin(yaw);
    HV3 key = hnormalize(h3(cp * sy, sp, cp * cy));
    float fyaw = std::fma(kPi, 0.78f, yaw);
    float fpitch = hclamp(pitch * 0.65f, 0.15f, 1.2f);
    HV3 fill = hnormalize(h3(std::cos(fpitch) * std::sin(fyaw), std::sin(fpitch),
                             std::cos(fpitch) * std::cos(fyaw)));
    E->key[0] = key.x; E->key[1] = key.y; E->key[2] = key.z;
    // sun_radiance_factor (environment/sun.rs): SUN_RADIANCE * horizon_blend(key)
    if (pi.model == ENV_DAYLIGHT_SUN) {
        float t = hclamp((key.y - -0.05f) / (0.05f - -0.05f), 0.0f, 1.0f);
        E->sun = 40000.0f * (t * t * std::fma(-2.0f, t, 3.0f));
    }
    E->fill[0] = fill.x; E->fill[1] = fill.y; E->fill[2] = fill.z;
    for (int i = 0; i < 16; ++i) {
        float a = std::fma((float)i, kPi * 2.0f / 16.0f, yaw);
        HV3 r = hnormalize(h3(std::sin(a) * 0.75f, sp * 0.8f, std::cos(a) * 0.75f));
        E->ring[i][0] = r.x; E->ring[i][1] = r.y; E->ring[i][2] = r.z;
    }
    if (pi.white_balance) {
        white_balance(pi.temp_k, E->wb);
    } else {
        E->wb[0] = E->wb[1] = E->wb[2] = 1.0f;
    }
    return true;
}

}  // namespace ixg

// ---------------------------------------------------------------------------
// exported host helpers
// ---------------------------------------------------------------------------
using namespace ixg;

extern "C" void ixg_camera_new(float yaw, float pitch, float distance, float fov_deg, uint32_t width,
                               uint32_t height, uint32_t spp, ixg_camera* out) {
    // Camera::new (optics/raytracer/camera.rs:108-147)
    float cp = std::cos(pitch), sp = std::sin(pitch), cy = std::cos(yaw), sy = std::sin(yaw);
    HV3 origin = h3(distance * cp * sy, distance * sp, distance * cp * cy);
    HV3 forward = hnormalize(h3(-origin.x, -origin.y, -origin.z));
    HV3 world_up = std::fabs(cp) < 1e-4f ? h3(0, 0, -1) : (cp < 0.0f ? h3(0, -1, 0) : h3(0, 1, 0));
    HV3 right = hnormalize(hcross(forward, world_up));
    HV3 up = hnormalize(hcross(right, forward));
    out->origin[0] = origin.x; out->origin[1] = origin.y; out->origin[2] = origin.z;
    out->fov_tan = std::tan(fov_deg * (kPi / 180.0f) * 0.5f);
    out->forward[0] = forward.x; out->forward[1] = forward.y; out->forward[2] = forward.z;
    out->width = (float)width;
    out->right[0] = right.x; out->right[1] = right.y; out->right[2] = right.z;
    out->height = (float)height;
    out->up[0] = up.x; out->up[1] = up.y; out->up[2] = up.z;
    out->num_samples = spp;
}

namespace {
void plane(HV3 n, float d, ixg_facet_plane* p) {
    HV3 u = hnormalize(n);
    p->normal[0] = u.x;
    p->normal[1] = u.y;
    p->normal[2] = u.z;
    p->d = d;
}
const float kStar[8] = {6, 18, 30, 42, 54, 66, 78, 90};
const float kMain[8] = {0, 12, 24, 36, 48, 60, 72, 84};
const float kBreak[16] = {95, 1, 11, 13, 23, 25, 35, 37, 47, 49, 59, 61, 71, 73, 83, 85};
const float kToRad = kPi / 180.0f;

// Correctly rounded f32 sine / cosine (double evaluation, one rounding): the
// cut geometry the reference's pinned metrics were produced with has
// sin(42.5 deg) rounded correctly, where glibc's sinf is one ULP high.
float sin_cr(float x) { return (float)std::sin((double)x); }
float cos_cr(float x) { return (float)std::cos((double)x); }

uint32_t tier(ixg_facet_plane* out, uint32_t at, float ang, float sy, const float* idx, int count,
              float d) {
    for (int i = 0; i < count; ++i) {
        // index_to_azimuth (cuts/round_brilliant.rs:61-63), 96-tooth wheel
        float phi = 2.0f * kPi * (idx[i] + 0.0f) / 96.0f;
        plane(h3(sin_cr(ang) * cos_cr(phi), sy * cos_cr(ang), sin_cr(ang) * sin_cr(phi)), d, &out[at++]);
    }
    return at;
}

// The 74 half-spaces of a round brilliant (57 facets + 16 girdle + culet),
// in StandardGemCuts::standard_round_brilliant's push order.
uint32_t brilliant(float table_d, float crown_deg, float crown_d, float pav_deg, float pav_d,
                   ixg_facet_plane* out) {
    uint32_t n = 0;
    plane(h3(0, 1, 0), table_d, &out[n++]);
    n = tier(out, n, 15.0f * kToRad, 1.0f, kStar, 8, -0.45f);
    n = tier(out, n, crown_deg * kToRad, 1.0f, kMain, 8, crown_d);
    n = tier(out, n, 41.0f * kToRad, 1.0f, kBreak, 16, -0.67f);
    for (int i = 0; i < 16; ++i) {
        float phi = 2.0f * kPi * (float)i / 16.0f;
        plane(h3(cos_cr(phi), 0.0f, sin_cr(phi)), -1.0f, &out[n++]);
    }
    n = tier(out, n, pav_deg * kToRad, -1.0f, kMain, 8, pav_d);
    n = tier(out, n, 42.5f * kToRad, -1.0f, kBreak, 16, -0.68f);
    plane(h3(0, -1, 0), -0.88f, &out[n++]);
    return n;
}
}  // namespace

extern "C" uint32_t ixg_cut_standard_round_brilliant(ixg_facet_plane* out) {
    return brilliant(-0.32f, 34.5f, -0.59f, 41.0f, -0.67f, out);
}

extern "C" uint32_t ixg_cut_round_brilliant(float crown_deg, float pavilion_deg, float table_ratio,
                                            ixg_facet_plane* out) {
    // Crown and pavilion mains keep the standard cut's girdle heights
    // (+0.028660 / -0.018470 at radius 1); the table sits where it meets the
    // crown mains at radius table_ratio.
    double ca = (double)crown_deg * M_PI / 180.0, pa = (double)pavilion_deg * M_PI / 180.0;
    double crown_d = std::sin(ca) + std::cos(ca) * 0.028660;
    double pav_d = std::sin(pa) + std::cos(pa) * 0.018470;
    double table_y = (crown_d - std::sin(ca) * (double)table_ratio) / std::cos(ca);
    return brilliant((float)-table_y, crown_deg, (float)-crown_d, pavilion_deg, (float)-pav_d, out);
}
