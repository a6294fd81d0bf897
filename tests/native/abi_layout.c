/* abi_layout.c -- prints the size and every field offset of each struct the
 * C ABI (include/ixg.h) declares, as one JSON object, so tests can hold the
 * ctypes mirror (indicatrix_amd/abi.py) and the Rust binding of the
 * integration guide to the compiler's own layout. */
#include <stddef.h>
#include <stdio.h>

#include "../../include/ixg.h"

static int first_struct = 1, first_field = 1;

static void begin(const char* name, size_t size) {
    printf("%s\"%s\": {\"size\": %zu, \"fields\": {", first_struct ? "" : ", ", name, size);
    first_struct = 0;
    first_field = 1;
}
static void field(const char* name, size_t off, size_t size) {
    printf("%s\"%s\": [%zu, %zu]", first_field ? "" : ", ", name, off, size);
    first_field = 0;
}
static void end(void) { printf("}}"); }

#define S(T) begin(#T, sizeof(T))
#define F(T, f) field(#f, offsetof(T, f), sizeof(((T*)0)->f))

int main(void) {
    printf("{");
    S(ixg_facet_plane); F(ixg_facet_plane, normal); F(ixg_facet_plane, d); end();
    S(ixg_camera); F(ixg_camera, origin); F(ixg_camera, fov_tan); F(ixg_camera, forward); F(ixg_camera, width);
    F(ixg_camera, right); F(ixg_camera, height); F(ixg_camera, up); F(ixg_camera, num_samples); end();
    S(ixg_dispersion); F(ixg_dispersion, kind); F(ixg_dispersion, coeff); end();
    S(ixg_band); F(ixg_band, center_nm); F(ixg_band, width); F(ixg_band, peak); F(ixg_band, shape); end();
    S(ixg_material); F(ixg_material, dispersion); F(ixg_material, birefringence_delta); F(ixg_material, c_axis);
    F(ixg_material, cubic); F(ixg_material, has_e_dispersion); F(ixg_material, e_dispersion);
    F(ixg_material, n_bands_o); F(ixg_material, n_bands_e); F(ixg_material, bands_o); F(ixg_material, bands_e);
    F(ixg_material, absorption_path_scale); F(ixg_material, edge_rounding_radius);
    F(ixg_material, scattering_sigma_s); F(ixg_material, biaxial); F(ixg_material, scattering_g);
    F(ixg_material, biaxial_delta_beta_alpha); F(ixg_material, has_beta_bands); F(ixg_material, n_bands_beta);
    F(ixg_material, bands_beta); end();
    S(ixg_environment); F(ixg_environment, preset); F(ixg_environment, exposure); F(ixg_environment, light_yaw);
    F(ixg_environment, light_pitch); F(ixg_environment, backdrop); F(ixg_environment, surface_glare);
    F(ixg_environment, head_shadow_deg); end();
    S(ixg_hdr_map); F(ixg_hdr_map, width); F(ixg_hdr_map, height); F(ixg_hdr_map, rgb); end();
    S(ixg_scene); F(ixg_scene, camera); F(ixg_scene, width); F(ixg_scene, height); F(ixg_scene, planes);
    F(ixg_scene, n_planes); F(ixg_scene, material); F(ixg_scene, max_bounces); F(ixg_scene, environment);
    F(ixg_scene, facet_finishes); F(ixg_scene, hdr_map); end();
    S(ixg_opts); F(ixg_opts, device); F(ixg_opts, chunk_paths); end();
    S(ixg_path_events); F(ixg_path_events, bounces); F(ixg_path_events, termination);
    F(ixg_path_events, facet_hash); F(ixg_path_events, n_tir); F(ixg_path_events, n_reflect);
    F(ixg_path_events, n_refract); F(ixg_path_events, n_probe); end();
    S(ixg_batch_item); F(ixg_batch_item, scene); F(ixg_batch_item, first_sample); F(ixg_batch_item, samples);
    F(ixg_batch_item, out); end();
    S(ixg_optical_pixel); F(ixg_optical_pixel, hits); F(ixg_optical_pixel, classes);
    F(ixg_optical_pixel, fire_sum_deg); F(ixg_optical_pixel, fire_qualified); F(ixg_optical_pixel, fire_rejected);
    F(ixg_optical_pixel, centre_entry_facet); F(ixg_optical_pixel, centre_exit_facet);
    F(ixg_optical_pixel, centre_bounces); F(ixg_optical_pixel, centre_reason); F(ixg_optical_pixel, mixed_facets);
    F(ixg_optical_pixel, scint_returned); F(ixg_optical_pixel, scint_poses); F(ixg_optical_pixel, pad); end();
    S(ixg_metric_request); F(ixg_metric_request, camera); F(ixg_metric_request, width);
    F(ixg_metric_request, height); F(ixg_metric_request, samples); F(ixg_metric_request, projection);
    F(ixg_metric_request, has_tile_origin); F(ixg_metric_request, tile_origin); F(ixg_metric_request, planes);
    F(ixg_metric_request, n_planes); F(ixg_metric_request, material); F(ixg_metric_request, environment);
    F(ixg_metric_request, want_fire); F(ixg_metric_request, want_scintillation); F(ixg_metric_request, has_pose);
    F(ixg_metric_request, pose); end();
    S(ixg_metric_totals); F(ixg_metric_totals, total); F(ixg_metric_totals, brilliance_pct);
    F(ixg_metric_totals, windowing_pct); F(ixg_metric_totals, extinction_pct); F(ixg_metric_totals, fire_index);
    F(ixg_metric_totals, scintillation_pct); end();
    S(ixg_design); F(ixg_design, plane_offset); F(ixg_design, n_planes); F(ixg_design, dispersion); end();
    S(ixg_asc_error); F(ixg_asc_error, kind); F(ixg_asc_error, line); F(ixg_asc_error, value);
    F(ixg_asc_error, teeth); F(ixg_asc_error, field_count); F(ixg_asc_error, field); F(ixg_asc_error, token);
    F(ixg_asc_error, message); end();
    S(ixg_asc_info); F(ixg_asc_info, gear_teeth); F(ixg_asc_info, symmetry_order); F(ixg_asc_info, mirror);
    F(ixg_asc_info, line_ending); F(ixg_asc_info, gear_reference_angle); F(ixg_asc_info, refractive_index);
    F(ixg_asc_info, n_tiers); F(ixg_asc_info, n_headers); F(ixg_asc_info, n_footnotes);
    F(ixg_asc_info, n_warnings); F(ixg_asc_info, gear_line); F(ixg_asc_info, facet_plane_count); end();
    S(ixg_asc_tier); F(ixg_asc_tier, angle_deg); F(ixg_asc_tier, mast); F(ixg_asc_tier, n_indices);
    F(ixg_asc_tier, n_index_names); end();
    S(ixg_radiance_error); F(ixg_radiance_error, kind); F(ixg_radiance_error, encoding);
    F(ixg_radiance_error, width); F(ixg_radiance_error, height); F(ixg_radiance_error, raw_len);
    F(ixg_radiance_error, first_sample); F(ixg_radiance_error, samples); F(ixg_radiance_error, start);
    F(ixg_radiance_error, end); F(ixg_radiance_error, cumulative); F(ixg_radiance_error, range_len);
    F(ixg_radiance_error, pad); F(ixg_radiance_error, expected_bytes); F(ixg_radiance_error, got_bytes);
    F(ixg_radiance_error, message); end();
    printf("}\n");
    return 0;
}
