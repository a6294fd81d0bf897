/* sanitize_oracle.c -- drives the C oracle (oracle/ixo.c) under
 * AddressSanitizer and UndefinedBehaviorSanitizer over small scenes that
 * reach every transport arm (isotropic and uniaxial pleochroic stones,
 * absorption bands, frosted facets, Henyey-Greenstein scattering, rounded
 * edges, the studio rig, the light tent, the daylight sun, an HDR
 * panorama), the metric image with Fire and scintillation and its per-pixel
 * maps, the fan-grid pose metrics, the stone width and the .asc plane
 * builder.  Exits 0 when every call returned; any sanitizer report aborts. */
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "../../oracle/ixo.h"

static ixg_material diamond(void) {
    ixg_material m;
    memset(&m, 0, sizeof m);
    m.dispersion.kind = IXG_DISP_SELLMEIER3;
    const float c[6] = {4.3356f, 0.3306f, 0.0f, 0.011236f, 0.030625f, 1.0f};
    memcpy(m.dispersion.coeff, c, sizeof c);
    m.cubic = 1;
    m.absorption_path_scale = 1.0f;
    return m;
}

static double frame(ixg_scene* sc, int spp) {
    double acc[12 * 12 * 3];
    memset(acc, 0, sizeof acc);
    ixo_accumulate(sc, 0, (uint32_t)spp, acc, 2);
    double s = 0.0;
    for (int i = 0; i < 12 * 12 * 3; ++i) s += acc[i];
    return s;
}

int main(void) {
    ixg_facet_plane planes[IXG_MAX_PLANES];
    const uint32_t n = ixo_cut_standard_round_brilliant(planes);
    ixg_material mat = diamond();
    ixg_scene sc;
    memset(&sc, 0, sizeof sc);
    ixo_camera_new(0.3f, 0.45f, 5.0f, 18.0f, 12, 12, 4, &sc.camera);
    sc.width = sc.height = 12;
    sc.planes = planes;
    sc.n_planes = n;
    sc.material = &mat;
    sc.max_bounces = 16;
    sc.environment.exposure = 1.0f;
    sc.environment.surface_glare = 0.85f;
    sc.environment.head_shadow_deg = 16.0f;
    double total = 0.0;
    const int presets[] = {IXG_PRESET_ISO_HEMISPHERE, IXG_PRESET_RING_LIGHTS, IXG_PRESET_LIGHT_TENT,
                           IXG_PRESET_DAYLIGHT_SUN, IXG_PRESET_ASET, IXG_PRESET_WHITE_TRAY};
    for (size_t p = 0; p < sizeof presets / sizeof presets[0]; ++p) {
        sc.environment.preset = presets[p];
        total += frame(&sc, 2);
    }
    sc.environment.preset = IXG_PRESET_ISO_HEMISPHERE;
    /* absorbing, rounded, scattering */
    mat.n_bands_o = 2;
    mat.bands_o[0] = (ixg_band){520.0f, 30.0f, 0.8f, 0};
    mat.bands_o[1] = (ixg_band){20000.0f, 1500.0f, 0.4f, 1};
    mat.edge_rounding_radius = 0.03f;
    mat.scattering_sigma_s = 1.5f;
    mat.scattering_g = 0.6f;
    total += frame(&sc, 2);
    /* frosted facets */
    uint8_t finishes[IXG_MAX_PLANES];
    for (uint32_t i = 0; i < n; ++i) finishes[i] = (uint8_t)(i % 3 == 0);
    mat.scattering_sigma_s = 0.0f;
    sc.facet_finishes = finishes;
    total += frame(&sc, 2);
    sc.facet_finishes = NULL;
    /* uniaxial pleochroic */
    ixg_material uni = diamond();
    uni.cubic = 0;
    uni.birefringence_delta = 0.009f;
    uni.c_axis[0] = 0.6f;
    uni.c_axis[1] = 0.64f;
    uni.c_axis[2] = 0.48f;
    uni.n_bands_e = 1;
    uni.bands_e[0] = (ixg_band){600.0f, 40.0f, 1.2f, 0};
    sc.material = &uni;
    total += frame(&sc, 2);
    /* biaxial trichroic, with a rounded edge and a panorama */
    ixg_material bi = uni;
    bi.biaxial = 1;
    bi.biaxial_delta_beta_alpha = 0.004f;
    bi.has_beta_bands = 1;
    bi.n_bands_beta = 1;
    bi.bands_beta[0] = (ixg_band){450.0f, 30.0f, 0.9f, 0};
    sc.material = &bi;
    total += frame(&sc, 2);
    bi.edge_rounding_radius = 0.03f;
    total += frame(&sc, 2);
    sc.material = &mat;
    /* HDR panorama */
    float rgb[8 * 4 * 3];
    for (int i = 0; i < 8 * 4 * 3; ++i) rgb[i] = (float)(i % 7) * 0.3f;
    ixg_hdr_map hdr = {8, 4, rgb};
    sc.hdr_map = &hdr;
    total += frame(&sc, 2);
    sc.hdr_map = NULL;

    /* metric image, maps and pose metrics */
    ixg_material d = diamond();
    ixg_metric_request r;
    memset(&r, 0, sizeof r);
    const float pose[3] = {0.0f, 1.1f, 2.4f};
    ixo_camera_new(pose[0], pose[1], pose[2], 42.0f, 10, 10, 6, &r.camera);
    r.width = r.height = 10;
    r.samples = 6;
    r.planes = planes;
    r.n_planes = n;
    r.material = &d;
    r.environment.preset = IXG_PRESET_RING_LIGHTS;
    r.environment.exposure = 1.0f;
    r.environment.surface_glare = 0.85f;
    r.environment.head_shadow_deg = 16.0f;
    r.want_fire = 1;
    r.want_scintillation = 1;
    r.has_pose = 1;
    memcpy(r.pose, pose, sizeof pose);
    ixg_optical_pixel px[100];
    ixg_metric_totals tot;
    ixo_metric_image(&r, px, &tot);
    float spatial[100], index[100], one = 0.0f;
    ixo_scintillation_spatial(px, 100, 10, 10, spatial);
    ixo_scintillation_indices(px, 100, spatial, index);
    ixo_scintillation_index_at(px, 100, 10, 10, 55, &one);
    float m5[5];
    ixo_pose_metrics(planes, n, &d, &r.environment, 0.0f, 0.45f, 1, m5);

    /* widths, outlines and .asc planes */
    double w = 0.0, lw = 0.0;
    uint32_t untouched = 0;
    ixo_model_width(planes, n, &w);
    ixo_model_width(planes, 3, &w);
    ixo_solid_outline(planes, n, &lw, &untouched);
    const double angle[3] = {41.0, 90.0, -43.0}, mast[3] = {0.5, 1.0, 0.0};
    const uint32_t counts[3] = {4, 2, 3};
    const double idx[9] = {0, 24, 48, 72, 12, 60, 0, 32, 64};
    ixg_facet_plane ap[32];
    ixo_asc_planes(angle, mast, counts, 3, idx, 96, 0.0, ap, 32);
    ixo_asc_planes(angle, mast, counts, 3, idx, 96, 0.0, ap, 1);
    printf("{\"frames_sum\": %.6g, \"brilliance\": %.6g, \"pose_brilliance\": %.6g, \"width\": %.9g}\n", total,
           tot.brilliance_pct, m5[0], w);
    return isfinite(total) ? 0 : 1;
}
