/*
 * ixg.h -- C ABI of the MI355X spectral-transport library (libixg.so).
 *
 * Plain C, POD structs, no exceptions across the boundary.  Every entry point
 * returns an ixg_status; a non-OK status leaves caller-owned output buffers
 * untouched (the "written only on Done" rule of the reference's GPU backend).
 *
 * Each declaration cites the reference interface it replaces.  Paths are
 * relative to the reference checkout; IX/ = crates/indicatrix/src/.
 */
#ifndef IXG_H
#define IXG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes.  GpuAccumulate {Done, Declined, Cancelled} plus the        */
/* GpuFrameError kinds (IX/renderer/gpu_backend/mod.rs:307-326,             */
/* IX/renderer/gpu/frame/mod.rs:383-418).                                    */
/* ------------------------------------------------------------------------ */
typedef enum ixg_status {
    IXG_OK = 0,            /* Done: results were ADDED into the caller's buffer */
    IXG_DECLINED = 1,      /* scene not routable to the GPU (caller falls back)  */
    IXG_CANCELLED = 2,     /* cancel flag observed between chunks                */
    IXG_DEVICE_LOST = 3,   /* HIP error / injected fault; context marked lost    */
    IXG_INVALID = 4,       /* malformed arguments (sizes, null pointers)         */
    IXG_RESOURCE_LIMIT = 5 /* request above a documented limit                   */
} ixg_status;

/* Why ixg_route declines a scene (scene_routes_to_gpu,
 * IX/renderer/gpu_backend/mod.rs:287-298). */
typedef enum ixg_route_reason {
    IXG_ROUTE_OK = 0,
    IXG_ROUTE_UNSUPPORTED_MATERIAL = 1,    /* zoned, scattering anisotropic stones      */
    IXG_ROUTE_UNSUPPORTED_ENVIRONMENT = 2, /* UV lamps; frosted/scatter under sun or HDR */
    IXG_ROUTE_DEVICE_LIMITS = 3,           /* too many planes for the LDS facet set     */
    IXG_ROUTE_NO_DEVICE = 4
} ixg_route_reason;

#define IXG_MAX_PLANES 256   /* LDS facet-set capacity (16 B per plane)        */
#define IXG_MAX_BANDS 8      /* absorption bands per ray (o or e), as the ref  */
#define IXG_NUM_CHANNELS 8   /* hero-wavelength comb width                     */

/* GpuFacetPlane: n.x + d <= 0 inside, unit n, 16 B (IX/geometry/plane.rs:31-53). */
typedef struct ixg_facet_plane {
    float normal[3];
    float d;
} ixg_facet_plane;

/* GpuCameraParams, byte for byte (IX/renderer/buffers/camera.rs:72-88). */
typedef struct ixg_camera {
    float origin[3];
    float fov_tan;
    float forward[3];
    float width;
    float right[3];
    float height;
    float up[3];
    uint32_t num_samples;
} ixg_camera;

/* DispersionModel (IX/optics/dispersion.rs:1-25). */
typedef enum ixg_dispersion_kind {
    IXG_DISP_SELLMEIER1 = 0, /* coeff = {b1, c1}                       */
    IXG_DISP_SELLMEIER3 = 1, /* coeff = {b0, b1, b2, c0, c1, c2}        */
    IXG_DISP_CAUCHY = 2      /* coeff = {a, b, c}                       */
} ixg_dispersion_kind;

typedef struct ixg_dispersion {
    int32_t kind;
    float coeff[6];
} ixg_dispersion;

/* AbsorptionBand (IX/optics/absorption.rs). shape 0 = Gaussian in wavelength,
 * 1 = Gaussian in energy (width is then sigma in cm^-1). */
typedef struct ixg_band {
    float center_nm;
    float width;
    float peak;
    int32_t shape;
} ixg_band;

/* The GemMaterial fields the transport reads (IX/optics/materials/mod.rs:116-190). */
typedef struct ixg_material {
    ixg_dispersion dispersion;            /* ordinary / isotropic curve            */
    float birefringence_delta;            /* n_e - n_o when no e-curve is given    */
    float c_axis[3];                      /* unit optic axis                       */
    int32_t cubic;                        /* 1 = CrystalSystem::Cubic              */
    int32_t has_e_dispersion;             /* uniaxial_extraordinary_dispersion     */
    ixg_dispersion e_dispersion;
    int32_t n_bands_o, n_bands_e;
    ixg_band bands_o[IXG_MAX_BANDS];
    ixg_band bands_e[IXG_MAX_BANDS];
    float absorption_path_scale;
    float edge_rounding_radius;           /* > 0: rounded meet edges (shading)     */
    float scattering_sigma_s;             /* > 0: Henyey-Greenstein volume scatter */
    int32_t biaxial;                      /* 1: biaxial_delta_beta_alpha is Some   */
    float scattering_g;                   /* Henyey-Greenstein asymmetry           */
    /* GemMaterial::biaxial_delta_beta_alpha (materials/mod.rs:146) when biaxial:
     * n_beta = dispersion(lambda), n_alpha = n_beta - this, n_gamma = n_alpha +
     * birefringence_delta, gamma along c_axis (materials/optics.rs:16-27). */
    float biaxial_delta_beta_alpha;
    /* AbsorptionTensor::beta_ray (optics/absorption.rs:337-356): 1 = Some (a
     * trichroic tensor alpha/beta/gamma = o/beta/e bands), 0 = None. Read only
     * for a biaxial material (absorption/wavelength_cache.rs:38-51). */
    int32_t has_beta_bands;
    int32_t n_bands_beta;
    ixg_band bands_beta[IXG_MAX_BANDS];
} ixg_material;

/* LightingPreset ids, the reference's persisted ids
 * (IX/optics/raytracer/environment/mod.rs:470-515). */
typedef enum ixg_preset {
    IXG_PRESET_LIGHT_TENT = 0,
    IXG_PRESET_ISO_HEMISPHERE = 1,
    IXG_PRESET_WHITE_TRAY = 2,
    IXG_PRESET_SHOP_LIGHTS = 3,
    IXG_PRESET_WINDOW_DAYLIGHT = 4,
    IXG_PRESET_DAYLIGHT_DOME = 5,
    IXG_PRESET_DAYLIGHT_SUN = 6,
    IXG_PRESET_DAYLIGHT = 7,
    IXG_PRESET_INCANDESCENT = 8,
    IXG_PRESET_ILLUMINANT_A = 9,
    IXG_PRESET_RING_LIGHTS = 10,
    IXG_PRESET_DARK_SPOTLIGHT = 11,
    IXG_PRESET_ASET = 12
} ixg_preset;

/* EnvironmentSource::Studio (IX/optics/raytracer/environment/source.rs:20-45). */
typedef struct ixg_environment {
    int32_t preset;
    float exposure;
    float light_yaw;
    float light_pitch;
    float backdrop;        /* lit tray radiance at exposure 1, >= 0; 0 = as lit */
    float surface_glare;   /* clamped to [0,1]; 1 = off                   */
    float head_shadow_deg; /* 16 = default, <= 0 = off                    */
} ixg_environment;

/* EnvironmentSource::HdrMap (IX/renderer/env_map.rs): an equirectangular
 * panorama of linear RGB texels, rgb[(y*width + x)*3 + c], row 0 at the +Y
 * pole, u = atan2(x, z) / 2 pi.  Seen without exposure, head shadow or white
 * balance; each channel's radiance is rgb_to_spectral_radiance
 * (IX/renderer/env_map_spectrum.rs:84) of the bilinear texel. */
typedef struct ixg_hdr_map {
    uint32_t width, height;
    const float* rgb;
} ixg_hdr_map;
#define IXG_MAX_HDR_TEXELS (16384u * 8192u)

/* GpuSceneRef (IX/renderer/gpu_backend/mod.rs:217-237). */
typedef struct ixg_scene {
    ixg_camera camera;
    uint32_t width, height;
    const ixg_facet_plane* planes;
    uint32_t n_planes;
    const ixg_material* material;
    uint32_t max_bounces;
    ixg_environment environment;
    /* FacetFinish per plane (GpuSceneRef::facet_finishes, IX/renderer/gpu_backend/
     * mod.rs:217-237): 0 polished, 1 frosted; NULL = every facet polished. */
    const uint8_t* facet_finishes;
    /* The environment is this panorama instead of `environment`'s preset
     * when non-NULL (EnvironmentSource::HdrMap). */
    const ixg_hdr_map* hdr_map;
} ixg_scene;

typedef struct ixg_opts {
    int32_t device;            /* HIP device ordinal                          */
    uint32_t chunk_paths;      /* paths per launch (0 = library default)      */
} ixg_opts;

typedef struct ixg_ctx ixg_ctx;

/* Per-path integer side channel (bit-exact parity with the CPU tracer):
 * the bounce loop's stop point and reason (RT/transport/mod.rs:68-93), a
 * hash of the facet sequence and per-event counts used by the FLOP model. */
typedef struct ixg_path_events {
    uint32_t bounces;       /* loop iterations executed                    */
    uint32_t termination;   /* 0 escaped, 2 russian roulette, 3 hit cap    */
    uint32_t facet_hash;    /* FNV-1a over (facet index, event kind)       */
    uint16_t n_tir, n_reflect, n_refract, n_probe;
} ixg_path_events;

/* ---- lifecycle (GpuBackend::new / disabled / is_lost / adapter_label,
 *      IX/renderer/gpu_backend/backend.rs:125-165) ---- */
ixg_status ixg_create(const ixg_opts* opts, ixg_ctx** out);
void ixg_destroy(ixg_ctx* ctx);
int ixg_is_lost(const ixg_ctx* ctx);
ixg_status ixg_adapter_label(const ixg_ctx* ctx, char* buf, size_t n);
ixg_status ixg_last_error(const ixg_ctx* ctx, char* buf, size_t n);
/* try_recover (IX/renderer/gpu_backend/backend.rs:704): clears the lost flag. */
ixg_status ixg_try_recover(ixg_ctx* ctx);
/* fail_on_turn test seam (IX/renderer/gpu_backend/backend.rs:81-85): the
 * n-th chunk from now reports IXG_DEVICE_LOST without touching the device. */
ixg_status ixg_inject_fault(ixg_ctx* ctx, int32_t fail_on_chunk);

/* scene_routes_to_gpu (IX/renderer/gpu_backend/mod.rs:287-298). */
ixg_route_reason ixg_route(const ixg_scene* scene);

/* ---- film ----
 * GpuBackend::try_accumulate_cancellable (backend.rs:226): traces absolute
 * sample indices [sample_offset, sample_offset+spp) for every pixel and ADDS
 * the per-pixel XYZ sums (FP64, non-finite samples dropped but counted) into
 * accum_xyz[w*h*3] only on IXG_OK.  cancel may be NULL; it is one byte,
 * non-zero = cancel, the in-memory form of Rust's AtomicBool. */
ixg_status ixg_accumulate(ixg_ctx* ctx, const ixg_scene* scene, uint32_t sample_offset,
                          uint32_t spp, double* accum_xyz, const volatile uint8_t* cancel);
/* One picture of a batch (GpuBatchItem, IX/renderer/gpu_backend/mod.rs:246-259):
 * its scene, absolute sample range and its own FP64 output (w*h*3), which
 * receives exactly what ixg_accumulate would ADD for the scene alone. */
typedef struct ixg_batch_item {
    const ixg_scene* scene;
    uint32_t first_sample;
    uint32_t samples;
    double* out;
} ixg_batch_item;
/* Called once per item, in input order: IXG_OK (Done), IXG_DECLINED,
 * IXG_CANCELLED, or IXG_INVALID for a malformed item. */
typedef void (*ixg_batch_done_fn)(void* user, uint32_t index, ixg_status outcome);
/* GpuBackend::try_accumulate_batch_cancellable (IX/renderer/gpu_backend/
 * backend.rs:451): traces the items as one stream on the library's stream,
 * the next picture's chunks queued before the previous picture is read back,
 * and reports each through on_done.  Each Done item is bit-identical to
 * ixg_accumulate on it alone; a declined or cancelled item leaves its out
 * untouched.  An unsupported material or environment declines only its own
 * item; a device loss declines every item not yet reported; the cancel flag
 * is polled once per chunk and, once seen, reports the first unfinished item
 * and all later ones IXG_CANCELLED (items already fully queued may still
 * finish Done).  Returns IXG_INVALID (nothing traced or reported) for bad
 * arguments, else IXG_OK once every item has been reported. */
ixg_status ixg_accumulate_batch(ixg_ctx* ctx, const ixg_batch_item* items, uint32_t n,
                                const volatile uint8_t* cancel, ixg_batch_done_fn on_done, void* user);
/* One chunk of a dispatch lane: WorkerLane::render_chunk (indicatrix-dispatch/
 * src/lane.rs:113-142) as the coordinator's own lane runs it (coordinator/
 * lanes/mod.rs:536-570 via stream_emit/tracer.rs:328 trace_range).  Traces
 * [first_sample, first_sample + samples) of the scene in sub-batches sized
 * toward 100 ms each (ixg_next_subbatch), adding each sub-batch only once it
 * completed, and stops between (or inside, discarding that sub-batch)
 * sub-batches once *cancel is non-zero.  sum_xyz (w*h*3 f32, the
 * ChunkResult::sum Vec3 layout) is overwritten with the summed XYZ of the
 * traced prefix [first_sample, first_sample + *done) -- always a prefix, never
 * a gap.  *rate (may be NULL) = samples per second over the sub-batches after
 * the first (the steady state without the scene upload), 0 with fewer than
 * two; seed_subbatch = the first sub-batch (0 = 1, seeded_batch_size without
 * a remembered size) and *subbatch_out (may be NULL) = the converged size to
 * seed the next chunk of the same scene with.  Returns IXG_OK when the whole
 * range was traced, IXG_CANCELLED when the flag ended it, or the failing call's
 * status (ChunkResult::partial: the prefix before the failure is kept). */
ixg_status ixg_render_chunk(ixg_ctx* ctx, const ixg_scene* scene, uint32_t first_sample, uint32_t samples,
                            uint32_t seed_subbatch, float* sum_xyz, uint32_t* done, double* rate,
                            uint32_t* subbatch_out, const volatile uint8_t* cancel);
/* stream_emit/sizing.rs next_batch_size: the next sub-batch after one of
 * prev samples took elapsed_s seconds -- prev scaled toward 100 ms, at most
 * x4 per step and never above 2048 or below 1 (an unmeasurable time grows by
 * x4). */
uint32_t ixg_next_subbatch(uint32_t prev, double elapsed_s);
/* Same, with accum_xyz a device pointer (inputs resident in HBM), on the
 * library's stream; no host copy of the result. */
ixg_status ixg_accumulate_device(ixg_ctx* ctx, const ixg_scene* scene, uint32_t sample_offset,
                                 uint32_t spp, double* d_accum_xyz);
/* ixg_accumulate_device without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The film
 * runs on a fixed scratch area that the library keeps in device memory, one
 * per GPU, for the life of the process (about 18.1 MiB: the path counter,
 * the facet set, IXG_NOALLOC_MAX_PATHS path records of 12 B and a panorama
 * slot of IXG_NOALLOC_MAX_HDR_TEXELS texels).  The frame is traced in chunks
 * of whole pixels that fit the path records, so any frame size works; spp
 * above IXG_NOALLOC_MAX_PATHS or a larger panorama is IXG_RESOURCE_LIMIT.
 * Results equal ixg_accumulate_device's bit for bit.  Every call on a GPU
 * shares that GPU's scratch: calls that overlap on one GPU race and must be
 * serialised by the caller. */
#define IXG_NOALLOC_MAX_PATHS (1u << 20)
#define IXG_NOALLOC_MAX_HDR_TEXELS (1024u * 512u)
ixg_status ixg_accumulate_device_noalloc(ixg_ctx* ctx, const ixg_scene* scene, uint32_t sample_offset,
                                         uint32_t spp, double* d_accum_xyz);

/* Per-path diagnostics for the parity tiers: out_xyz[paths*3] (f32, raw
 * per-sample XYZ) and events[paths] for paths = w*h*spp in (pixel, sample)
 * order.  Either output may be NULL. */
ixg_status ixg_trace_paths(ixg_ctx* ctx, const ixg_scene* scene, uint32_t sample_offset,
                           uint32_t spp, float* out_xyz, ixg_path_events* events);
/* Diagnostic: the per-path XYZ that the film kernel ixg_accumulate runs
 * (the persistent, feature-specialised instantiation the timed path
 * launches) writes before its per-pixel reduction, in (pixel, sample)
 * order, out_xyz[3 * (pixel * spp + s)]; at most 2^26 paths.  The same
 * paths as ixg_trace_paths, from the production kernel. */
ixg_status ixg_film_samples(ixg_ctx* ctx, const ixg_scene* scene, uint32_t sample_offset, uint32_t spp,
                            float* out_xyz);
/* trace_spectral_ray (RT/transport/mod.rs:118) for explicit rays, the entry
 * the reference's golden tables pin: rays[8*i..] = {origin xyz, dir xyz,
 * hero_rand, rng_seed bits as float}.  The scene's camera and frame size are
 * ignored.  out_xyz[n*3] f32; events[n] may be NULL. */
ixg_status ixg_trace_rays(ixg_ctx* ctx, const ixg_scene* scene, const float* rays, uint32_t n,
                          float* out_xyz, ixg_path_events* events);
/* Event totals of one film call (FLOP model input), summed on the device:
 * {paths, loop iterations, TIR, partial reflections, refractions, exit-split
 * probes, escapes, Stokes frame rotations, bounce slab scans run, probe slab
 * scans run}. */
#define IXG_EVENT_SLOTS 10
ixg_status ixg_count_events(ixg_ctx* ctx, const ixg_scene* scene, uint32_t sample_offset,
                            uint32_t spp, uint64_t totals[IXG_EVENT_SLOTS]);
/* Average duration (ms) of the last film call's transport kernel launches,
 * measured with HIP events on the library's stream. */
double ixg_last_kernel_ms(const ixg_ctx* ctx);
/* The HIP stream (a hipStream_t) every call on this context launches on, so
 * a caller can order its own work and timing events against the library's.
 * No counterpart in the reference, whose GPU queue is internal to its
 * renderer; it is for hosts that share the device with other HIP work. */
void* ixg_stream(ixg_ctx* ctx);
uint32_t ixg_last_kernel_launches(const ixg_ctx* ctx);

/* ---- per-function probes ----
 * The reference's per-function GPU checks (renderer/gpu/ulp.rs with the
 * *_check.rs modules under renderer/gpu/ and renderer/gpu/transport_check/;
 * labels and budgets in crates/indicatrix/docs/gpu_harness_last_run.json:
 * 24-73): ixg_probe evaluates the device function `fn` on n input records
 * and writes n output records, with the scene supplying the camera,
 * material, environment and facet set the function reads (the scene must
 * route; the material's absorption bands are read only when it has some).
 * Diagnostic: the film and metric paths never call it.  Records are floats;
 * u32 / i32 fields carry their bits.  in -> out per record:
 *   SAMPLE_DRAWS      pixel u32, sample u32 -> jx, jy, hero, seed u32, lambda[8]
 *   CAMERA_RAY        pixel u32, sample u32 -> origin[3], dir[3]
 *   CIE_CMF           lambda -> x, y, z
 *   ILLUMINANT        lambda -> the environment's illuminant power
 *   ENVIRONMENT       dir[3], observer[3], lambda -> radiance (exposure applied)
 *   FRAME_ROTATION    cos_psi, sin_psi, stokes[4] -> stokes[4]
 *   FRESNEL_REFLECTION n1, n2, cos_i, cos_t, stokes[4] -> stokes[4], clamped R_unpol
 *   FRESNEL_TRANSMISSION n1, n2, cos_i, cos_t, stokes[4] -> stokes[4], r_s, r_p
 *   TIR_RETARDATION   n1, cos_i, sin_i, stokes[4] -> stokes[4]
 *   CHANNEL_OPTICS    lambda[8] -> n_o[8], alpha_o[8], alpha_e[8]
 *   PLEOCHROIC_ALPHA  k[3], alpha_o[8], alpha_e[8] -> alpha[8]
 *   ORDINARY_EIGEN    k[3], c[3] -> o[3]
 *   EXTRAORDINARY_POYNTING w[3], c[3], n_o, n_e -> s[3]
 *   THETA_C           lambda_hero, n_o, k[3], normal[3], inside -> theta
 *   UNIAXIAL_INDICES  lambda, n_o, theta -> n_e, n_eff
 *   ENTRY_SOLVE       n_o, n_e, c[3], k[3], normal[3] -> {r_s, r_p, t_o, t_e} (re, im)
 *                     for s then p incidence, flux_o, flux_e
 *   INTERNAL_SOLVE    n_inc, n_o, n_e, c[3], k[3], normal[3], ordinary ->
 *                     r_o, r_e, t_s, t_p (re, im), flux_ro, flux_re, flux_ts, flux_tp, flux_inc
 *   EXIT_TRANSMISSION INTERNAL_SOLVE's 13, inc_i, r_branch -> stokes[4], i_unit
 *   HG_DIRECTION      u1, u2, g, fwd[3] -> dir[3]
 *   COSINE_HEMISPHERE u1, u2, n[3] -> dir[3]
 *   SHADING_NORMAL    hit[3], facet i32, normal[3], radius -> normal[3]
 *   FINISH            compat lo u32, compat hi u32, pdf[8], lambda[8], radiance[8] -> xyz[3]
 *   INTERSECT         origin[3], dir[3] -> facet i32 (-1 = miss), t
 *   EXP_LANE          x -> exp_f32x8's lane value
 *   EXP               x -> e^x correctly rounded
 * The biaxial probes read the scene's biaxial material: lambda gives
 * n_beta = dispersion(lambda), n_alpha = n_beta - delta_beta_alpha,
 * n_gamma = n_alpha + birefringence_delta, gamma along the optic axis.
 *   BIAX_WAVE_INDICES lambda, k[3] -> n_slow, n_fast
 *   BIAX_EIGEN        lambda, k[3], want_slow -> D[3]
 *   BIAX_MODE_POYNTING lambda, k[3], want_slow -> S[3]
 *   BIAX_RESOLVE_ENTRY lambda, k[3], normal[3] (facing k), want_slow -> wave normal[3]
 *   BIAX_ASSIGNED_ALPHA lambda, k[3], is_extra, alpha_o, alpha_beta, alpha_e -> alpha */
typedef enum ixg_probe_fn {
    IXG_PROBE_SAMPLE_DRAWS = 0,
    IXG_PROBE_CAMERA_RAY,
    IXG_PROBE_CIE_CMF,
    IXG_PROBE_ILLUMINANT,
    IXG_PROBE_ENVIRONMENT,
    IXG_PROBE_FRAME_ROTATION,
    IXG_PROBE_FRESNEL_REFLECTION,
    IXG_PROBE_FRESNEL_TRANSMISSION,
    IXG_PROBE_TIR_RETARDATION,
    IXG_PROBE_CHANNEL_OPTICS,
    IXG_PROBE_PLEOCHROIC_ALPHA,
    IXG_PROBE_ORDINARY_EIGEN,
    IXG_PROBE_EXTRAORDINARY_POYNTING,
    IXG_PROBE_THETA_C,
    IXG_PROBE_UNIAXIAL_INDICES,
    IXG_PROBE_ENTRY_SOLVE,
    IXG_PROBE_INTERNAL_SOLVE,
    IXG_PROBE_EXIT_TRANSMISSION,
    IXG_PROBE_HG_DIRECTION,
    IXG_PROBE_COSINE_HEMISPHERE,
    IXG_PROBE_SHADING_NORMAL,
    IXG_PROBE_FINISH,
    IXG_PROBE_INTERSECT,
    IXG_PROBE_EXP_LANE,
    IXG_PROBE_EXP,
    IXG_PROBE_BIAX_WAVE_INDICES,
    IXG_PROBE_BIAX_EIGEN,
    IXG_PROBE_BIAX_MODE_POYNTING,
    IXG_PROBE_BIAX_RESOLVE_ENTRY,
    IXG_PROBE_BIAX_ASSIGNED_ALPHA,
    IXG_PROBE_COUNT
} ixg_probe_fn;
/* Floats per input and output record of fn; 0 for an unknown id. */
int ixg_probe_shape(int32_t fn, uint32_t* in_floats, uint32_t* out_floats);
ixg_status ixg_probe(ixg_ctx* ctx, const ixg_scene* scene, int32_t fn, const float* in, uint32_t n, float* out);

/* ---- metrics ----
 * OpticalPixel (IX/color/metrics/image_types.rs:59-75): per-pixel class
 * counts {windowed, extinct, returned, entry-blocked/invalid} over the
 * pixel's deterministic sub-pixel samples, the weighted Fire sum, and the
 * temporal-scintillation pose counts. */
typedef struct ixg_optical_pixel {
    uint32_t hits;
    uint32_t classes[4];
    float fire_sum_deg;
    uint32_t fire_qualified;
    uint32_t fire_rejected;
    int32_t centre_entry_facet; /* -1 = none */
    int32_t centre_exit_facet;  /* -1 = none */
    uint32_t centre_bounces;
    uint32_t centre_reason;     /* ExtinctionReason ordinal */
    uint32_t mixed_facets;
    uint8_t scint_returned;
    uint8_t scint_poses;
    uint8_t pad[2];
} ixg_optical_pixel;

/* OpticalImageRequest (IX/color/metrics/image_types.rs:89-164). */
typedef struct ixg_metric_request {
    ixg_camera camera;             /* origin/forward/right/up/fov_tan are read     */
    uint32_t width, height;        /* <= 1048576 pixels                            */
    uint32_t samples;              /* 1..64                                        */
    uint32_t projection[2];        /* 0,0 = width,height                           */
    int32_t has_tile_origin;
    uint32_t tile_origin[2];
    const ixg_facet_plane* planes;
    uint32_t n_planes;
    const ixg_material* material;
    ixg_environment environment;
    int32_t want_fire;
    int32_t want_scintillation;
    int32_t has_pose;
    float pose[3];                 /* yaw, pitch, distance (needed for scintillation) */
} ixg_metric_request;

/* Image totals plus the image-domain metrics: shares of the visible area
 * (OpticalImage::total.fractions), the Fire index over all traced rays
 * (accumulate.rs:140-152 convention, x275, floor 0.1) and the image
 * scintillation (image_scint.rs:261-287). */
typedef struct ixg_metric_totals {
    ixg_optical_pixel total;
    float brilliance_pct, windowing_pct, extinction_pct, fire_index, scintillation_pct;
} ixg_metric_totals;

/* evaluate_optical_image (IX/color/metrics/image.rs:187-262) on the GPU.
 * pixels[width*height] may be NULL.  Errors: IXG_RESOURCE_LIMIT (pixels,
 * samples), IXG_INVALID (non-unit planes, bad camera, scintillation without
 * a pose), IXG_DECLINED (environment without a device lighting model). */
ixg_status ixg_metric_image(ixg_ctx* ctx, const ixg_metric_request* req, ixg_optical_pixel* pixels,
                            ixg_metric_totals* totals);

/* The per-pixel scintillation maps of a finished image (the pixels an
 * ixg_metric_image call returned): scintillation_spatial
 * (IX/color/metrics/image_scint.rs:175-199) into spatial[n_pixels] -- the
 * squashed CV of the brilliance fraction over the 5x5 window of hit pixels,
 * 0..100, 0 where hits == 0 and everywhere when width*height != n_pixels --
 * and scintillation_index (image_scint.rs:234-247) into index[n_pixels] --
 * 0.6 spatial + 0.4 of the five-pose Bernoulli variance over 0.24, NaN where
 * the temporal term was not computed (None).  Host buffers; either output
 * may be NULL.  Computed on the device, equal to the CPU pass bit for bit. */
ixg_status ixg_scintillation_maps(ixg_ctx* ctx, const ixg_optical_pixel* pixels, uint32_t n_pixels, uint32_t width,
                                  uint32_t height, float* spatial, float* index);
/* The same on device buffers (no copies), synchronised before it returns. */
ixg_status ixg_scintillation_maps_device(ixg_ctx* ctx, const ixg_optical_pixel* d_pixels, uint32_t n_pixels,
                                         uint32_t width, uint32_t height, float* d_spatial, float* d_index);
/* The hover read-outs on host pixels, equal to the full maps bit for bit:
 * scintillation_spatial_at (image_scint.rs:201-232) and
 * scintillation_index_at (image_scint.rs:249-259), which returns 0 (None)
 * or 1 with *out set. */
float ixg_scintillation_spatial_at(const ixg_optical_pixel* pixels, uint32_t n_pixels, uint32_t width,
                                   uint32_t height, uint32_t i);
int ixg_scintillation_index_at(const ixg_optical_pixel* pixels, uint32_t n_pixels, uint32_t width, uint32_t height,
                               uint32_t i, float* out);

/* One design of a batched metric sweep: its facet set inside a shared plane
 * table and its dispersion curve. */
typedef struct ixg_design {
    uint32_t plane_offset, n_planes;
    ixg_dispersion dispersion;
} ixg_design;

/* Batched tilt sweep (MET/sweep.rs, MET/pose_batch.rs:73-199 generalised to
 * images): every (design, tilt) image of width x height at `samples`, with
 * Fire and scintillation, in ONE device pass; camera yaw 0 / 180 deg by the
 * tilt's sign, pitch 90 - |tilt|, at `distance`, fov 42 deg.
 * out_metrics[(d*n_tilts + t)*5] = {brilliance, fire, scintillation,
 * windowing, extinction}. */
ixg_status ixg_metric_sweep(ixg_ctx* ctx, const ixg_facet_plane* planes, uint32_t total_planes,
                            const ixg_design* designs, uint32_t n_designs, const float* tilts_deg,
                            uint32_t n_tilts, uint32_t width, uint32_t height, uint32_t samples,
                            float distance, const ixg_environment* env, float* out_metrics);

/* How ixg_metric_sweep launches (no reference counterpart: the reference
 * evaluates metrics on the CPU).  use_graph = 1 captures the sweep's kernels
 * in a HIP graph that later calls of the same shape replay; block = 0 lets
 * the autotuner time the 2-D grid and each persistent work-queue shape once
 * per image size and keep the fastest, block = -1 forces the 2-D grid, and
 * block > 0 with tile forces one queue shape (lanes per block, pixels per
 * queue item: 64/64, 64/256, 128/128, 128/512, 256/256, 256/1024), and
 * block = -64, -128 or -256 forces the flow kernel (every lane advancing one
 * slab scan of its pixel's traces per step) with that many lanes per block
 * and tile pixels per item (0 = the whole image); the autotuner also times
 * the flow shapes.  Defaults: use_graph 1, block 0. */
ixg_status ixg_set_metric_mode(ixg_ctx* ctx, int use_graph, int block, int tile);
/* The shape the last sweep ran with (block -1 = 2-D grid, -64/-128/-256 =
 * the flow kernel) and whether it was replayed from a graph. */
ixg_status ixg_metric_tuning(ixg_ctx* ctx, int* block, int* tile, uint32_t* grid_blocks, int* graphed);
/* Device time (ms) of the last ixg_metric_sweep's launches -- the pixel
 * kernel (or its graph) and the per-image reductions, autotuning excluded --
 * measured with HIP events on the library's stream (no reference
 * counterpart). */
double ixg_metric_last_ms(ixg_ctx* ctx);

/* measure_model_width (IX/render_setup/stone_width.rs:27-42): the short axis
 * of the solid's x/z vertex bounding box (IX/geometry/stone_metrics/
 * measure.rs:63-150), in f64 on the host.  IXG_DECLINED when the planes do
 * not bound a finite solid of positive volume (the reference's None). */
ixg_status ixg_model_width(const ixg_facet_plane* planes, uint32_t n_planes, double* width);

/* evaluate_gem_optical_metrics (IX/color/metrics/evaluate.rs:158) and its
 * Fire-less twin evaluate_gem_optical_metrics_geom_core (evaluate.rs:305),
 * batched: every (design, pose) of the 18 x 18 analytical ray fan in one
 * device pass, each design's fan scaled to its measured girdle half-width
 * (fan.rs:49-95).  poses[2p] = yaw, poses[2p+1] = pitch (radians).
 * out_metrics[(d*n_poses + p)*5] = {brilliance, fire, scintillation,
 * windowing, extinction}; want_all = 0 skips Fire and scintillation and
 * reports 25 / 75 for them (placeholder_metrics, evaluate.rs:259); a design
 * with no planes reports the placeholder metrics. */
ixg_status ixg_pose_metrics(ixg_ctx* ctx, const ixg_facet_plane* planes, uint32_t total_planes,
                            const ixg_design* designs, uint32_t n_designs, const float* poses,
                            uint32_t n_poses, const ixg_environment* env, int want_all, float* out_metrics);

/* sweep_axes (IX/color/metrics/pose_batch.rs:199-217) generalised to many
 * designs -- evaluate_full_axis_profile_at_azimuth (profile.rs:200) for each
 * (design, azimuth): the 181 poses of axis_poses (pose_batch.rs:118) in one
 * device pass, laid out by tilt_slot (pose_batch.rs:133).
 * out_curves[((d*n_azimuths + a)*3 + k)*181 + t], k = brilliance,
 * extinction, windowing; t = tilt + 90 deg. */
ixg_status ixg_axis_profiles(ixg_ctx* ctx, const ixg_facet_plane* planes, uint32_t total_planes,
                             const ixg_design* designs, uint32_t n_designs, const float* azimuths_deg,
                             uint32_t n_azimuths, const ixg_environment* env, float* out_curves);

/* ---- metric surrogate (new work; the reference has no surrogate) ----
 * A BF16 MLP 16 -> 4 x 256 (GELU) -> 3, FP32 master weights and Adam,
 * trained and evaluated with MFMA kernels.  x, t, y are device pointers:
 * x [batch x 16], t and y [batch x 3], f32, batch a multiple of 64 and at
 * most the max_batch given at creation.  Parameters are exchanged in the
 * padded FP32 layout of ixg_mlp_param_count() floats:
 *   W0 [256 x 64] | W1..W3 [256 x 256] | W4 [64 x 256] | b0..b3 [256] | b4 [64]
 * (row-major, out x in; input columns >= 16 and output rows >= 3 are zero). */
typedef struct ixg_mlp ixg_mlp;
uint32_t ixg_mlp_param_count(void);
ixg_status ixg_mlp_create(ixg_ctx* ctx, uint32_t max_batch, uint64_t seed, ixg_mlp** out);
void ixg_mlp_destroy(ixg_mlp* mlp);
ixg_status ixg_mlp_set_params(ixg_mlp* mlp, const float* host, size_t count);
ixg_status ixg_mlp_get_params(ixg_mlp* mlp, float* host, size_t count);
/* The optimiser state, for a checkpoint that resumes training exactly: the
 * Adam moments m, v (count = ixg_mlp_param_count() each, FP32) and the step
 * count behind the bias corrections.  set_params resets them to zero. */
ixg_status ixg_mlp_get_state(ixg_mlp* mlp, float* m, float* v, size_t count, uint64_t* step);
ixg_status ixg_mlp_set_state(ixg_mlp* mlp, const float* m, const float* v, size_t count, uint64_t step);
ixg_status ixg_mlp_infer(ixg_mlp* mlp, const float* d_x, uint32_t batch, float* d_y);
/* Loss (mean squared error over batch x 3) and its gradient w.r.t. every
 * parameter (host, padded layout), without updating. */
ixg_status ixg_mlp_gradients(ixg_mlp* mlp, const float* d_x, const float* d_t, uint32_t batch,
                             float* host_grad, size_t count, float* loss);
/* One Adam step (beta 0.9/0.999, eps 1e-8) at learning rate lr; loss may be NULL. */
ixg_status ixg_mlp_train_step(ixg_mlp* mlp, const float* d_x, const float* d_t, uint32_t batch,
                              float lr, float* loss);
ixg_status ixg_mlp_sync(ixg_mlp* mlp);

/* ---- GemCAD .asc designs (indicatrix-formats/src/asc, IX/geometry/cuts/asc_schedule.rs) ----
 * A parsed design is an opaque ixg_asc; its header, tiers and text fields are
 * read with the accessors below and its facet planes with ixg_asc_planes. */
#define IXG_ASC_MAX_GEAR_TEETH 720 /* AscParseError::MAX_GEAR_TEETH (asc/error.rs) */
typedef struct ixg_asc ixg_asc;

/* AscParseError's variants (asc/error.rs), plus text that is not UTF-8. */
typedef enum ixg_asc_error_kind {
    IXG_ASC_ERR_NONE = 0,
    IXG_ASC_ERR_EMPTY_INPUT = 1,
    IXG_ASC_ERR_GEAR_LINE_MISSING_FIELDS = 2,
    IXG_ASC_ERR_GEAR_TEETH_NOT_NUMERIC = 3,
    IXG_ASC_ERR_GEAR_REFERENCE_ANGLE_NOT_NUMERIC = 4,
    IXG_ASC_ERR_SYMMETRY_LINE_MISSING_FIELDS = 5,
    IXG_ASC_ERR_SYMMETRY_ORDER_NOT_NUMERIC = 6,
    IXG_ASC_ERR_REFRACTIVE_INDEX_MISSING = 7,
    IXG_ASC_ERR_REFRACTIVE_INDEX_NOT_NUMERIC = 8,
    IXG_ASC_ERR_NOT_WHOLE_NUMBER = 9,
    IXG_ASC_ERR_GEAR_TEETH_ZERO = 10,
    IXG_ASC_ERR_GEAR_TEETH_TOO_LARGE = 11,
    IXG_ASC_ERR_MIRROR_FLAG_INVALID = 12,
    IXG_ASC_ERR_MISSING_GEAR_LINE = 13,
    IXG_ASC_ERR_MISSING_SYMMETRY_LINE = 14,
    IXG_ASC_ERR_MISSING_REFRACTIVE_INDEX_LINE = 15,
    IXG_ASC_ERR_NO_TIER_RECORDS = 16,
    IXG_ASC_ERR_TIER_RECORD_TOO_SHORT = 17,
    IXG_ASC_ERR_ANGLE_NOT_NUMERIC = 18,
    IXG_ASC_ERR_MAST_NOT_NUMERIC = 19,
    IXG_ASC_ERR_NON_FINITE_VALUE = 20,
    IXG_ASC_ERR_SYMMETRY_ORDER_ZERO = 21,
    IXG_ASC_ERR_REFRACTIVE_INDEX_OUT_OF_RANGE = 22,
    IXG_ASC_ERR_INVALID_UTF8 = 23
} ixg_asc_error_kind;

/* A parse failure: the variant, its 1-based line (0 for whole-file errors),
 * the variant's fields and the reference's Display wording in message. */
typedef struct ixg_asc_error {
    int32_t kind;          /* ixg_asc_error_kind */
    uint32_t line;
    double value;          /* NotWholeNumber / NonFiniteValue / RefractiveIndexOutOfRange */
    int32_t teeth;         /* GearTeethTooLarge */
    uint32_t field_count;  /* TierRecordTooShort */
    char field[32];        /* e.g. "angle", "gear tooth count" */
    char token[128];
    char message[512];
} ixg_asc_error;

typedef enum ixg_asc_line_ending { IXG_ASC_LF = 0, IXG_ASC_CRLF = 1 } ixg_asc_line_ending;

/* AscSchedule's scalar fields (asc/schedule.rs). */
typedef struct ixg_asc_info {
    int32_t gear_teeth;            /* signed: a negative count is a handedness convention */
    uint32_t symmetry_order;
    int32_t mirror;
    int32_t line_ending;           /* ixg_asc_line_ending */
    double gear_reference_angle;
    double refractive_index;
    uint32_t n_tiers, n_headers, n_footnotes, n_warnings;
    uint32_t gear_line;            /* 1-based line of the gear record kept (parse_asc_with_gear_line) */
    uint32_t facet_plane_count;    /* AscSchedule::facet_plane_count: planes before dedup */
} ixg_asc_info;

/* One `a` record (AscTier, asc/tier.rs); a culet is angle_deg == -0.0. */
typedef struct ixg_asc_tier {
    double angle_deg;
    double mast;
    uint32_t n_indices;
    uint32_t n_index_names;
} ixg_asc_tier;

typedef enum ixg_asc_text_field {
    IXG_ASC_TEXT_VERSION = 0,   /* gemcad_version                 */
    IXG_ASC_TEXT_HEADER = 1,    /* headers[i]                     */
    IXG_ASC_TEXT_FOOTNOTE = 2,  /* footnotes[i]                   */
    IXG_ASC_TEXT_WARNING = 3,   /* warnings[i]                    */
    IXG_ASC_TEXT_TIER_NAME = 4, /* tiers[i].name ("c/d" joined)   */
    IXG_ASC_TEXT_TIER_NOTES = 5 /* tiers[i].notes                 */
} ixg_asc_text_field;

/* parse_asc / parse_asc_with_gear_line (asc/parse.rs:54,68): UTF-8 text, a
 * leading BOM ignored.  IXG_INVALID with *err filled on a parse error. */
ixg_status ixg_asc_parse_text(const char* text, size_t len, ixg_asc** out, ixg_asc_error* err);
/* parse_asc_bytes (asc/decode.rs): raw file bytes, UTF-8 or Windows-1252,
 * lone CR line ends read as LF. */
ixg_status ixg_asc_parse_bytes(const uint8_t* bytes, size_t len, ixg_asc** out, ixg_asc_error* err);
void ixg_asc_free(ixg_asc* schedule);
ixg_status ixg_asc_info_get(const ixg_asc* schedule, ixg_asc_info* info);
/* Tier i; indices (may be NULL) receives its n_indices index positions. */
ixg_status ixg_asc_tier_get(const ixg_asc* schedule, uint32_t i, ixg_asc_tier* tier, double* indices,
                            uint32_t cap);
/* A text field (ixg_asc_text_field) as UTF-8, NUL-terminated in buf[cap];
 * returns its full byte length (call with cap 0 to size the buffer). */
size_t ixg_asc_text(const ixg_asc* schedule, int32_t field, uint32_t i, char* buf, size_t cap);
/* AscTier::index_names[k] of tier: the position it binds to, and the name. */
size_t ixg_asc_index_name(const ixg_asc* schedule, uint32_t tier, uint32_t k, uint32_t* position,
                          char* buf, size_t cap);
/* StandardGemCuts::from_asc_schedule (IX/geometry/cuts/asc_schedule.rs:20) on
 * a parsed design: its deduplicated facet planes, in schedule order. *n_out
 * is always the plane count; IXG_RESOURCE_LIMIT when cap is below it. */
ixg_status ixg_asc_planes(const ixg_asc* schedule, ixg_facet_plane* out, uint32_t cap, uint32_t* n_out);
/* The same on plain tier data (tiers[t].n_indices positions each, in order,
 * from indices[]; n_index_names is ignored), with the two bookkeeping views
 * of asc_schedule.rs:29,60: raw_to_plane[facet_plane_count] (may be NULL)
 * maps every raw plane to its output plane (asc_plane_provenance), and
 * tier_ends[n_tiers] (may be NULL) is the number of output planes up to the
 * end of each tier (asc_tier_plane_boundaries at gear_reference_angle).
 * out may be NULL to count only. */
ixg_status ixg_asc_schedule_planes(const ixg_asc_tier* tiers, uint32_t n_tiers, const double* indices,
                                   uint32_t n_indices, int32_t gear_teeth, double gear_reference_angle,
                                   ixg_facet_plane* out, uint32_t cap, uint32_t* n_out,
                                   uint32_t* raw_to_plane, uint32_t* tier_ends);

/* ---- host helpers (same f32 arithmetic as the reference, glibc libm) ---- */
/* Camera::new (RT/camera.rs:108). */
void ixg_camera_new(float yaw, float pitch, float distance, float fov_deg, uint32_t width,
                    uint32_t height, uint32_t spp, ixg_camera* out);
/* StandardGemCuts::standard_round_brilliant (IX/geometry/cuts/round_brilliant.rs:71). */
uint32_t ixg_cut_standard_round_brilliant(ixg_facet_plane* out /* >= 74 */);
/* 57-facet round brilliant from crown/pavilion angles (deg) and table ratio. */
uint32_t ixg_cut_round_brilliant(float crown_deg, float pavilion_deg, float table_ratio,
                                 ixg_facet_plane* out /* >= 74 */);

/* ---- radiance payloads: the body of a FRAME / PREVIEW message ----
 * indicatrix-net/src/radiance: width*height pixels of three little-endian f32
 * (XYZ running sums, 12 B each, radiance/mod.rs:43); a compressed encoding
 * first byte-shuffles the 4n bytes into 4 planes (radiance/shuffle.rs). */
#define IXG_MAX_FRAME_LEN (512u * 1024u * 1024u) /* framing/mod.rs:26 */
typedef enum ixg_payload_encoding { /* PayloadEncoding, messages/encoding.rs:43 */
    IXG_PAYLOAD_RAW = 0,
    IXG_PAYLOAD_SHUFFLE_ZSTD = 1,
    IXG_PAYLOAD_SHUFFLE_LZ4 = 2
} ixg_payload_encoding;
/* PayloadEncoding::ShuffleZstd { level } as one int32: the family in bits
 * 0-7, the encoder's u8 level in bits 8-15 (zero for the other arms). */
#define IXG_PAYLOAD_ZSTD(level) ((int32_t)IXG_PAYLOAD_SHUFFLE_ZSTD | (int32_t)(((level) & 0xFF) << 8))
#define IXG_PAYLOAD_ZSTD_LEVEL_OF(encoding) (((encoding) >> 8) & 0xFF)
typedef enum ixg_radiance_error_kind { /* RadianceError, radiance/mod.rs:66-119 */
    IXG_RAD_OK = 0,
    IXG_RAD_LENGTH_MISMATCH = 1,
    IXG_RAD_MISALIGNED = 2,
    IXG_RAD_TOO_LARGE = 3,
    IXG_RAD_RAW_LEN_MISMATCH = 4,
    IXG_RAD_UNSUPPORTED_ENCODING = 5,
    IXG_RAD_DECOMPRESS_FAILED = 6,
    IXG_RAD_SHORT_OUTPUT = 7,
    IXG_RAD_FRAME_OUTSIDE_RANGE = 8,
    IXG_RAD_SAMPLE_COUNT_OVERRUN = 9,
    IXG_RAD_ZERO_SAMPLE_FRAME = 10
} ixg_radiance_error_kind;
/* The variant's fields where it has them; message = its Display text. */
typedef struct ixg_radiance_error {
    int32_t kind;
    int32_t encoding;
    uint32_t width, height;
    uint32_t raw_len;
    uint32_t first_sample, samples, start, end;
    uint32_t cumulative, range_len;
    uint32_t pad;
    uint64_t expected_bytes, got_bytes;
    char message[256];
} ixg_radiance_error;
/* shuffle_into / unshuffle_into (radiance/shuffle.rs:35,60) on device
 * buffers: n_bytes a multiple of 4 (IXG_INVALID otherwise, where the
 * reference panics), at most IXG_MAX_FRAME_LEN, src and dst distinct. */
ixg_status ixg_radiance_shuffle_device(ixg_ctx* ctx, const void* d_src, uint64_t n_bytes, void* d_dst);
ixg_status ixg_radiance_unshuffle_device(ixg_ctx* ctx, const void* d_src, uint64_t n_bytes, void* d_dst);
/* The summing arms of PayloadDecoder::decode_and_add (radiance/payload.rs:225-250):
 * adds the payload (shuffled = 0: raw Vec3 bytes, 4-byte aligned; 1: the
 * planes, add_unshuffled_valid, shuffle.rs:122) onto d_acc[pixels*3] f32,
 * skipping every pixel that fails is_valid_sample, and writes the number
 * skipped to *d_dropped (device u32). */
ixg_status ixg_radiance_add_device(ixg_ctx* ctx, const void* d_payload, int32_t shuffled, uint64_t pixels,
                                   float* d_acc, uint32_t* d_dropped);
/* PayloadDecoder::prepare's header checks (radiance/payload.rs:282-320) in
 * their order (encoding: IXG_PAYLOAD_RAW, IXG_PAYLOAD_ZSTD(level) or
 * IXG_PAYLOAD_SHUFFLE_LZ4): TooLarge, RawLenMismatch, then for Raw LengthMismatch of the
 * wire bytes, for a compressed encoding UnsupportedEncoding unless
 * compression != 0 (this library carries no zstd / LZ4 codec: a caller that
 * decompresses itself passes 1 and then hands the planes to
 * ixg_radiance_add_device).  IXG_OK or IXG_INVALID with err filled. */
ixg_status ixg_radiance_check_payload(int32_t encoding, uint32_t raw_len, uint64_t wire_len, uint32_t width,
                                      uint32_t height, int32_t compression, ixg_radiance_error* err);
/* Accumulator::apply_frame's admission checks (client/accumulate/mod.rs:414-465):
 * ZeroSampleFrame, then with a requested range [start, end) FrameOutsideRange
 * (frame_within_range, overflow-safe) and SampleCountOverrun of
 * samples_done + samples (saturating) against end - start. */
ixg_status ixg_radiance_check_frame(uint32_t first_sample, uint32_t samples, int32_t has_range, uint32_t start,
                                    uint32_t end, uint32_t samples_done, ixg_radiance_error* err);

#ifdef __cplusplus
}
#endif
#endif /* IXG_H */
