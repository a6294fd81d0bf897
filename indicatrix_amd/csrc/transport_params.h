// transport_params.h -- the resolved, device-ready form of one film call.
//
// The host turns an ixg_scene into this POD once per call (glibc libm for
// every scene-constant transcendental: camera basis, light rig, white
// balance), and the kernel reads it as a by-value kernel argument.
#pragma once
#include <stdint.h>

namespace ixg {

// LightingModel::gpu_id (environment/mod.rs:281-290)
enum EnvModel : int32_t {
    ENV_STUDIO = 0,
    ENV_ISO_HEMISPHERE = 1,
    ENV_LIGHT_TENT = 2,
    ENV_DAYLIGHT_DOME = 3,
    ENV_ASET = 4,
    ENV_DAYLIGHT_SUN = 5,
    ENV_HDR = 6  // EnvironmentSource::HdrMap
};

struct EnvParams {
    int32_t model;       // EnvModel
    int32_t d65;         // 1: tabulated D65 illuminant, 0: Planck fit
    float bb_t_k;        // BlackbodyNorm::t_k
    float bb_denom_560;  // BlackbodyNorm::denom_560
    float exposure;
    float spot_mult;
    float shadow_outer;  // head-shadow cosines [outer, inner]
    float shadow_inner;
    float key[3];
    float fill[3];
    float ring[16][3];
    float wb[3];         // von Kries LMS scale (1,1,1 for D65 presets)
    float tray;          // lit backdrop radiance at exposure 1 (0 = the model's own ground)
    float tent_walls, tent_cards, tent_spark, tent_ground, tent_flat;  // TentParams
    float sun;           // DaylightSun: the disc's radiance factor (0 below the horizon)
    uint32_t hdr_w, hdr_h;  // ENV_HDR: the panorama and its device texels
    const float* hdr_rgb;
};

struct AbsBand {
    float center_nm, width, peak;
    int32_t shape;  // 0 Gaussian in wavelength, 1 in wavenumber
};

// Absorption of an isotropic-by-symmetry material: the o / e band lists, the
// optic axis as given (eigen-directions) and the AbsorptionTensor3 axes.
struct AbsorbParams {
    int32_t n_o, n_e;
    float path_scale;
    float c_raw[3];
    float a1[3], a2[3], cn[3];
    AbsBand o[8];
    AbsBand e[8];
    int32_t n_b;   // trichroic biaxial stones: the beta bands
    AbsBand b[8];
};

// A uniaxial material (RayMaterialContext::is_anisotropic): the ordinary
// curve is FilmParams::disp, the extraordinary one e_disp when given, else
// n_o + biref.
struct AnisoParams {
    int32_t on;
    float biref;
    int32_t has_e;
    int32_t e_kind;
    float e_disp[6];
    float c_axis[3];  // as given
    // a biaxial stone (biaxial_delta_beta_alpha is Some): n_beta = the
    // dispersion curve, n_alpha = n_beta - dba, n_gamma = n_alpha + biref;
    // the principal frame bx / by / bz (gamma along the normalised axis)
    int32_t biax;
    int32_t trichroic;  // AbsorptionTensor::beta_ray is Some
    float dba;
    float bx[3], by[3], bz[3];
};

struct FilmParams {
    // GpuCameraParams fields
    float cam_origin[3];
    float fov_tan;
    float forward[3];
    float width_f;
    float right[3];
    float height_f;
    float up[3];
    uint32_t width;
    uint32_t height;
    uint32_t spp;
    uint32_t sample_offset;
    uint32_t n_planes;
    uint32_t max_bounces;
    int32_t disp_kind;
    float disp[6];
    float surface_glare;  // clamped to [0,1]
    float edge_radius;    // GemMaterial::edge_rounding_radius (<= 0: sharp edges)
    int32_t frosted;      // 1: some facet is frosted (bitmask after the planes)
    float sigma_s;        // GemMaterial::scattering_sigma_s (0: no volume scattering)
    float scat_g;         // GemMaterial::scattering_g
    uint64_t first_path;  // first (pixel*spp + sample) index of this launch
    uint64_t n_paths;     // paths in this launch
    int32_t absorb;       // 1: the material has absorption bands
    EnvParams env;
    AbsorbParams abs;
    AnisoParams an;
};

// Event-count slots accumulated by the counting kernel.
enum EventSlot : int {
    EV_PATHS = 0,
    EV_BOUNCES = 1,      // loop iterations (one slab scan each)
    EV_TIR = 2,
    EV_REFLECT = 3,
    EV_REFRACT = 4,
    EV_PROBE = 5,        // exit-split probe slab scans
    EV_ESCAPED = 6,
    EV_ROTATIONS = 7,    // Stokes frame rotations
    EV_SCANS = 8,        // bounce slab scans actually run (convex-exit shortcut elides the rest)
    EV_PROBE_SCANS = 9,  // exit-probe slab scans actually run
    EV_NUM = 10
};

// One metric image (a design under one camera pose).
struct MetricImage {
    float origin[3];
    float fov_tan;
    float forward[3];
    float pad0;
    float right[3];
    float pad1;
    float up[3];
    float pad2;
    float t_origin[5][3];  // temporal cameras (same fov_tan)
    float t_forward[5][3];
    float t_right[5][3];
    float t_up[5][3];
    uint32_t plane_offset;
    uint32_t n_planes;
    float nd, nf, nc;
    int32_t zero_dispersion;
};

// What every image of one metric call shares.
struct MetricShared {
    EnvParams env;        // scoring rig: exposure 1, observer-free, default cosines
    float spec550;        // illuminant power at the 550 nm probe wavelength
    float threshold;      // ExitLighting::threshold
    float head_cos;       // head-shadow cone cosine
    uint32_t width, height, samples;
    uint32_t proj_w, proj_h;
    uint32_t origin_x, origin_y;
    float scale_x, scale_y;
    int32_t want_fire, want_scintillation;
    uint32_t n_images;
};

// One design of a fan-grid pose evaluation (PreparedEval, prepared.rs:64-94):
// its facet set, the three line indices and the fan scale (fan.rs:71-95).
struct PoseDesign {
    uint32_t plane_offset, n_planes;
    float nd, nf, nc;
    float standoff, lateral;
    uint32_t pad;
};

// One (design, pose) job: the camera frame at the pose and at the five
// temporal yaw offsets (scintillation.rs:19,47-60), rows forward/right/up.
struct PoseJob {
    float basis[6][3][3];  // [0] = the pose, [1..5] = yaw offsets -3 .. +3 deg
    uint32_t design;
    uint32_t pad[3];
};

constexpr int kPoseMaxCells = 256;  // disc cells of the 18 x 18 grid (180 used)

}  // namespace ixg
