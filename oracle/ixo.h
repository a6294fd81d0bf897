/*
 * ixo.h -- the CPU oracle: a plain-C FP32 restatement of the reference's
 * spectral tracer, used ONLY as a checker (tests, smoke, the bench's CPU
 * baseline leg).  Never linked into libixg.so.
 *
 * Struct layouts come from the public ABI header so both sides read the same
 * scene; every function below restates the reference file:line it cites.
 */
#ifndef IXO_H
#define IXO_H

#include <stdint.h>
#include "../include/ixg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The element type of the ABI structs' float arrays.  The FP64 twin
 * (ixo64.c) compiles ixo.c with its arithmetic in double; the scenes it reads
 * keep f32 fields, which it reads through this type. */
typedef float ixo_abi_f32;

/* RNG and stratified draws (RT/sampling.rs). */
uint32_t ixo_hash_u32(uint32_t x);
float ixo_low_discrepancy_base2(uint32_t n);
float ixo_radical_inverse_base(uint32_t n, uint32_t base);
void ixo_sample_draws(uint32_t pixel, uint32_t sample, uint32_t* seed, float* jx, float* jy,
                      float* hero);
void ixo_wrapped_hero_wavelengths(float hero_rand, float out[8]);

/* Optical primitives. */
float ixo_dispersion_evaluate(const ixg_dispersion* d, float lambda_nm);
void ixo_cie_cmf(float lambda_nm, float out[3]);
float ixo_d65(float lambda_nm);
float ixo_blackbody(float lambda_nm, float temp_k);
void ixo_white_balance(float temp_k, float out[3]);
float ixo_exp_lane(float x);
/* exp_f32x8 (IX/simd/exp_poly.rs:122): ixo_exp_lane on eight lanes (AVX2). */
void ixo_exp_lane8(const float x[8], float out[8]);
float ixo_band_evaluate(const ixg_band* b, float lambda_nm);
void ixo_uniaxial_energy(int kind, float n_inc, float n_o, float n_e, const float c[3], const float k[3],
                         const float normal[3], int pol, float out[2]);
void ixo_uniaxial_entry_r(float n_o, float n_e, const float c[3], const float k[3], const float normal[3], int pol,
                          float out[4]);

/* The biaxial indicatrix (optics/birefringence/biaxial.rs), for the tests.
 * axes = the alpha, beta, gamma principal directions as three columns
 * (axes[0..2] alpha, [3..5] beta, [6..8] gamma). */
void ixo_biax_from_gamma_axis(const float gamma_axis[3], float axes_out[9]);
void ixo_biax_wave_indices(float na, float nb, float ng, const float axes[9], const float k[3], float out[2]);
void ixo_biax_eigen_d(float na, float nb, float ng, const float axes[9], const float k[3], int want_slow,
                      float out[3]);
void ixo_biax_d_to_e(float na, float nb, float ng, const float axes[9], const float d[3], float out[3]);
void ixo_biax_mode_poynting(float na, float nb, float ng, const float axes[9], const float k[3], int want_slow,
                            float out[3]);
float ixo_biax_resolve_entry(float na, float nb, float ng, const float axes[9], const float incident[3],
                             const float normal[3], float cos_i, float n_seed, int want_slow, float dir_out[3]);
void ixo_poynting_direction(const float k[3], const float e[3], float out[3]);

/* 1 (default): next-event estimation at exterior frosted events (DaylightSun's disc). */
extern int ixo_enable_nee;
/* sun_cone_direction (environment/sun.rs): a direction in the sun disc around key. */
void ixo_sun_cone_direction(const float key[3], float u0, float u1, float out[3]);

/* sample_studio_environment (RT/environment/mod.rs:653): observer = 0. */
float ixo_sample_studio_environment(const float dir[3], float lambda_nm, int32_t preset,
                                    float exposure, float light_yaw, float light_pitch);

/* shading_normal_near_edge (RT/intersect.rs:196): the edge-rounded shading normal. */
void ixo_shading_normal_near_edge(const ixg_facet_plane* planes, uint32_t n, const float hit[3], int32_t facet,
                                  const float hn[3], float radius, float out[3]);

/* cosine_weighted_hemisphere (RT/scattering/frosted.rs) about unit n. */
void ixo_cosine_hemisphere(float u1, float u2, const float n[3], float out[3]);

/* sample_henyey_greenstein_direction (RT/scattering/hg.rs:636) about unit fwd. */
void ixo_hg_direction(float u1, float u2, float g, const float fwd[3], float out[3]);

/* rgb_to_spectral_radiance (IX/renderer/env_map_spectrum.rs:84). */
float ixo_rgb_to_spectral(const float rgb[3], float lambda_nm);

/* Camera (RT/camera.rs). */
void ixo_camera_new(float yaw, float pitch, float distance, float fov_deg, uint32_t width,
                    uint32_t height, uint32_t spp, ixg_camera* out);
void ixo_generate_ray(const ixg_camera* cam, float sx, float sy, float jx, float jy,
                      float origin[3], float dir[3]);

/* Geometry. */
uint32_t ixo_cut_standard_round_brilliant(ixg_facet_plane* out);
uint32_t ixo_cut_round_brilliant(float crown_deg, float pavilion_deg, float table_ratio,
                                 ixg_facet_plane* out);
/* intersect_polyhedron_soa (RT/intersect.rs:101): returns facet index or -1. */
int32_t ixo_intersect(const ixg_facet_plane* planes, uint32_t n, const float origin[3],
                      const float dir[3], float* t_out);
/* The facet scan the transport uses: the 8-wide slab scan (IX/simd/slab.rs:381)
 * in the AVX2 build, ixo_intersect in the scalar ones; bit-identical to
 * ixo_intersect.  Single-threaded (one static SoA table). */
int32_t ixo_slab_scan(const ixg_facet_plane* planes, uint32_t n, const float o[3], const float d[3], float* t_out);

/* trace_spectral_ray (RT/transport/mod.rs:118) with an explicit ray; events may be NULL. */
void ixo_trace_ray(const ixg_scene* scene, const float origin[3], const float dir[3],
                   uint32_t rng_seed, float hero_rand, float xyz[3], ixg_path_events* ev);

/* One (pixel, absolute sample) path of a frame (IX/renderer/cpu_frame.rs:45). */
void ixo_trace_sample(const ixg_scene* scene, uint32_t pixel, uint32_t sample, float xyz[3],
                      ixg_path_events* ev);

/* cpu_accumulate (IX/renderer/gpu/hybrid.rs:305) with FP64 per-pixel sums,
 * interleaved pixel partition over `threads` threads.  ADDS into accum. */
void ixo_accumulate(const ixg_scene* scene, uint32_t sample_offset, uint32_t spp,
                    double* accum_xyz, int threads);
/* Per-path raw XYZ + events in (pixel, sample) order for paths [first, first+count). */
void ixo_trace_paths(const ixg_scene* scene, uint32_t sample_offset, uint32_t spp,
                     uint64_t first, uint64_t count, float* out_xyz, ixg_path_events* events,
                     int threads);

/* evaluate_optical_image (IX/color/metrics/image.rs:187); totals->scintillation_pct
 * is left NaN (use ixo_image_scintillation_pct over the pixels). */
int ixo_metric_image(const ixg_metric_request* req, ixg_optical_pixel* pixels,
                     ixg_metric_totals* totals);
float ixo_image_scintillation_pct(const ixg_optical_pixel* px, uint32_t n);
/* scintillation_spatial / _at (IX/color/metrics/image_scint.rs:175-232) and
 * scintillation_index / _at (:234-259); the index functions return 0 for
 * None (no temporal term) and 1 with *out set. */
void ixo_scintillation_spatial(const ixg_optical_pixel* px, uint32_t n, uint32_t w, uint32_t h, float* out);
float ixo_scintillation_spatial_at(const ixg_optical_pixel* px, uint32_t n, uint32_t w, uint32_t h, uint32_t i);
int ixo_scintillation_index(const ixg_optical_pixel* px, uint32_t n, const float* spatial, uint32_t n_spatial,
                            uint32_t i, float* out);
int ixo_scintillation_index_at(const ixg_optical_pixel* px, uint32_t n, uint32_t w, uint32_t h, uint32_t i,
                               float* out);
void ixo_scintillation_indices(const ixg_optical_pixel* px, uint32_t n, const float* spatial, float* out);
/* offset (IX/color/metrics/image.rs:21-36): the sub-pixel offset of sample s. */
void ixo_image_offset(uint32_t s, float* jx, float* jy);
/* temporal_cameras (IX/color/metrics/image_scint.rs:35-48). */
void ixo_temporal_cameras(const float pose[3], float fov_tan, uint32_t width, uint32_t height, ixg_camera out[5]);
/* classify_ray (IX/color/metrics/classify.rs:109-197) with an explicit
 * observer: class 0..3, -1 no hit, -2 no lighting model. */
int ixo_classify_ray(const ixg_facet_plane* planes, uint32_t n, const ixg_dispersion* disp,
                     const ixg_environment* env, const float origin[3], const float dir[3], const float observer[3]);

/* measure_model_width (IX/render_setup/stone_width.rs:27): 1 and the width
 * when the planes bound a finite solid, else 0. */
int ixo_model_width(const ixg_facet_plane* planes, uint32_t n, double* width);
/* girdle_half_width (IX/color/metrics/fan.rs:49): width / 2, or 1.0. */
float ixo_girdle_half_width(const ixg_facet_plane* planes, uint32_t n);
/* StandardGemCuts::from_asc_schedule on plain tier data (IX/geometry/cuts/
 * asc_schedule.rs:20,152-200,264-322): tier t has counts[t] index positions,
 * consecutive in idx[].  Returns the plane count (0 if above cap). */
uint32_t ixo_asc_planes(const double* angle_deg, const double* mast, const uint32_t* counts, uint32_t n_tiers,
                        const double* idx, int32_t gear_teeth, double gear_reference_angle,
                        ixg_facet_plane* out, uint32_t cap);
/* The girdle outline's length/width ratio (IX/geometry/brep.rs:544 with
 * tests/optics_geometry_tests/asc_designs.rs:63-88) and the number of planes
 * that carry no face; 0 when the planes do not close a solid. */
int ixo_solid_outline(const ixg_facet_plane* planes, uint32_t n, double* lw, uint32_t* untouched);
/* evaluate_gem_optical_metrics[_geom_core] (IX/color/metrics/evaluate.rs:158,305)
 * at one pose: out = {brilliance, fire, scintillation, windowing, extinction}. */
int ixo_pose_metrics(const ixg_facet_plane* planes, uint32_t n, const ixg_material* material,
                     const ixg_environment* env, float yaw, float pitch, int want_all,
                     float out[5]);

/* Per-function probes (the CPU side of ixg_probe): the oracle's function for
 * probe id fn on n records, layouts as in include/ixg.h.  Returns 0 for an
 * unknown id or a scene the oracle cannot build. */
int ixo_probe_shape(int32_t fn, uint32_t* in_floats, uint32_t* out_floats);
int ixo_probe(const ixg_scene* scene, int32_t fn, const float* in, uint32_t n, float* out);

#ifdef __cplusplus
}
#endif

#endif
