// host_scene.h -- host-side resolution of an ixg_scene into FilmParams.
//
// Every scene-constant quantity (camera basis, studio light rig, Planck
// normalisation, von Kries scale, head-shadow cone) is evaluated once here in
// FP32 with glibc libm -- the same operations the reference performs per
// trace on the CPU -- so the kernel only does per-path work.
#pragma once
#include <cmath>
#include <cstdint>

#include "../../include/ixg.h"
#include "transport_params.h"

namespace ixg {

struct HV3 {
    float x, y, z;
};
inline HV3 h3(float x, float y, float z) { return HV3{x, y, z}; }
inline float hdot(HV3 a, HV3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline HV3 hscale(HV3 a, float s) { return h3(a.x * s, a.y * s, a.z * s); }
inline HV3 hcross(HV3 a, HV3 b) {
    return h3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline HV3 hnormalize(HV3 a) { return hscale(a, 1.0f / std::sqrt(hdot(a, a))); }
inline float hclamp(float v, float lo, float hi) {
    v = v < lo ? lo : v;
    return v > hi ? hi : v;
}

constexpr float kPi = 3.14159265358979323846f;

// LightingPreset params/model/spectrum (environment/mod.rs:321-410, 519-612).
struct PresetInfo {
    int model;  // EnvModel
    int d65;
    float temp_k;
    float spot_mult;
    int white_balance;
    // TentParams {walls, cards, spark, ground, flat} (environment/mod.rs:185-215);
    // the light tent's own values for every other preset
    float tent[5] = {1.0f, 1.0f, 1.0f, 0.02f, 0.0f};
};
bool preset_info(int32_t preset, PresetInfo* out);

// Resolves the environment; false for presets the device path declines.
bool resolve_environment(const ixg_environment& env, EnvParams* out);

// Planck fit normalised at 560 nm (environment/spectral.rs:101-116).
float blackbody(float lambda_nm, float temp_k);
// compute_illuminant_white_balance (color.rs:268-305).
void white_balance(float temp_k, float out[3]);
// D65 relative power (environment/spectral.rs:49-76, >= 380 nm), host copy.
float d65_host(float lambda_nm);
// DispersionModel::evaluate (optics/dispersion.rs:170-210, >= 380 nm), host copy.
float dispersion_host(const ixg_dispersion& d, float lambda_nm);
// CIE 1931 CMF table lookup (color/cie1931.rs:184-209), host copy.
void cie_cmf_host(float lambda_nm, float out[3]);

}  // namespace ixg
