/*
 * ixo64.c -- the FP64 twin of the CPU oracle (test infrastructure only):
 * ixo.c compiled once more with every internal `float` a `double` and the
 * f32 libm calls their double counterparts, so the same estimator, on the
 * same counter-based sample streams and in the same operation order, runs in
 * double precision.  It is the yardstick that GPU and FP32-oracle errors are
 * stated against; it pins nothing by itself (its source is the FP32 oracle's).
 *
 * The ABI structs of include/ixg.h keep their float fields (the header is
 * read before the substitution, and ixo.c reads those fields through the
 * ixo_abi_f32 type), so the twin consumes the very same scenes.  The
 * f32-specific bit work of the reference -- exp_f32x8's exponent assembly,
 * the 2^-32 draw scale landing on 1.0 -- becomes the plain double function;
 * every constant keeps its f32 value.  Exported names gain the ixo64_ prefix.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ixg.h"
#include "ixo.h"

#define IXO_SCALAR_ONLY 1
#define IXO_NO_PROBE 1
#define IXO_FP64_TWIN 1
#define float double
#define fmaf fma
#define sqrtf sqrt
#define fabsf fabs
#define cosf cos
#define sinf sin
#define tanf tan
#define acosf acos
#define atanf atan
#define atan2f atan2
#define floorf floor
#define expf exp
#define logf log
#define powf pow
#define fmodf fmod
#define fmaxf fmax
#define fminf fmin
#define copysignf copysign

#define ixo_accumulate ixo64_accumulate
#define ixo_asc_planes ixo64_asc_planes
#define ixo_band_evaluate ixo64_band_evaluate
#define ixo_biax_from_gamma_axis ixo64_biax_from_gamma_axis
#define ixo_biax_wave_indices ixo64_biax_wave_indices
#define ixo_biax_eigen_d ixo64_biax_eigen_d
#define ixo_biax_d_to_e ixo64_biax_d_to_e
#define ixo_biax_mode_poynting ixo64_biax_mode_poynting
#define ixo_biax_resolve_entry ixo64_biax_resolve_entry
#define ixo_poynting_direction ixo64_poynting_direction
#define ixo_blackbody ixo64_blackbody
#define ixo_enable_nee ixo64_enable_nee
#define ixo_sun_cone_direction ixo64_sun_cone_direction
#define ixo_camera_new ixo64_camera_new
#define ixo_cie_cmf ixo64_cie_cmf
#define ixo_classify_ray ixo64_classify_ray
#define ixo_cosine_hemisphere ixo64_cosine_hemisphere
#define ixo_cut_round_brilliant ixo64_cut_round_brilliant
#define ixo_cut_standard_round_brilliant ixo64_cut_standard_round_brilliant
#define ixo_d65 ixo64_d65
#define ixo_dispersion_evaluate ixo64_dispersion_evaluate
#define ixo_exp_lane ixo64_exp_lane
#define ixo_exp_lane8 ixo64_exp_lane8
#define ixo_generate_ray ixo64_generate_ray
#define ixo_girdle_half_width ixo64_girdle_half_width
#define ixo_hash_u32 ixo64_hash_u32
#define ixo_hg_direction ixo64_hg_direction
#define ixo_image_offset ixo64_image_offset
#define ixo_image_scintillation_pct ixo64_image_scintillation_pct
#define ixo_intersect ixo64_intersect
#define ixo_low_discrepancy_base2 ixo64_low_discrepancy_base2
#define ixo_metric_image ixo64_metric_image
#define ixo_model_width ixo64_model_width
#define ixo_pose_metrics ixo64_pose_metrics
#define ixo_radical_inverse_base ixo64_radical_inverse_base
#define ixo_rgb_to_spectral ixo64_rgb_to_spectral
#define ixo_sample_draws ixo64_sample_draws
#define ixo_sample_studio_environment ixo64_sample_studio_environment
#define ixo_scintillation_index ixo64_scintillation_index
#define ixo_scintillation_index_at ixo64_scintillation_index_at
#define ixo_scintillation_indices ixo64_scintillation_indices
#define ixo_scintillation_spatial ixo64_scintillation_spatial
#define ixo_scintillation_spatial_at ixo64_scintillation_spatial_at
#define ixo_slab_scan ixo64_slab_scan
#define ixo_shading_normal_near_edge ixo64_shading_normal_near_edge
#define ixo_solid_outline ixo64_solid_outline
#define ixo_temporal_cameras ixo64_temporal_cameras
#define ixo_trace_paths ixo64_trace_paths
#define ixo_trace_ray ixo64_trace_ray
#define ixo_trace_sample ixo64_trace_sample
#define ixo_uniaxial_energy ixo64_uniaxial_energy
#define ixo_uniaxial_entry_r ixo64_uniaxial_entry_r
#define ixo_white_balance ixo64_white_balance
#define ixo_wrapped_hero_wavelengths ixo64_wrapped_hero_wavelengths

#include "ixo.c"
