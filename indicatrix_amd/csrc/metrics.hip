// metrics.hip -- brilliance / fire / scintillation images on CDNA4 (gfx950).
//
// The reference evaluates these on the CPU only (color/metrics/): per pixel,
// up to 64 deterministic sub-pixel rays are classified by a single-wavelength
// scalar-Fresnel trace at the d line (<= 10 internal bounces), the returned
// ones replayed at the F and C lines for Fire, and the pixel centre re-traced
// from five yaw-offset cameras for temporal scintillation.
//
// Device layout: either a 2-D grid (blockIdx.y = image = design x camera
// pose, blockIdx.x = a 256-pixel tile of it) or a persistent work queue over
// (image, tile) items; the image's facet set is staged in LDS.  One lane owns one pixel and writes its OpticalPixel; a
// second kernel reduces each image in raster order (one lane per image), so
// every f32 sum is formed in the reference's order.
//
// Reference map (paths relative to crates/indicatrix/src/color/metrics/):
//   offset / sample_pixel      image.rs:21-175
//   classify_ray / fire_sample classify.rs:109-256
//   trace_entry / interior     ray_trace.rs:77-425
//   returned_or_reason         visibility.rs:55-67, lighting.rs:208
//   temporal sampler           image_scint.rs:35-130
//   image scintillation        image_scint.rs:261-287, scintillation.rs:146-198
#include "device_common.h"

namespace ixg {

constexpr int kMBlock = 256;

// Occupancy target of the work-queue and pose kernels (waves per SIMD; 0 =
// the compiler's choice).  Measured on the 512-design sweep (64x64 @ 32, 21
// tilts): 4 waves (no spills, 97-122 VGPRs) 2.94 s, 5 waves 2.59 s, 6 waves
// 2.44 s, 7 waves (72 VGPRs, 80 B/lane spilled) 2.37-2.43 s -- the scan is
// latency-bound on its LDS reads, so occupancy beats spill-free code.  Reading
// the facet set with wave-uniform scalar loads in place of LDS, and loading 4
// planes per scan step, were both slower (2.74-3.28 s).
#ifndef IXG_METRIC_WAVES_PER_EU
#define IXG_METRIC_WAVES_PER_EU 7
#endif
#if IXG_METRIC_WAVES_PER_EU > 0
#define IXG_METRIC_WAVES_ATTR __attribute__((amdgpu_waves_per_eu(IXG_METRIC_WAVES_PER_EU)))
#else
#define IXG_METRIC_WAVES_ATTR
#endif

enum { FATE_ENTRY_BLOCKED = 0, FATE_LEAKED = 1, FATE_UPWARD = 2, FATE_ABSORBED = 3 };
enum { REASON_NONE = 0, REASON_SHADOW = 1, REASON_ILLUM = 2, REASON_SIDE = 3, REASON_NOHIT = 4,
       REASON_BUDGET = 5, REASON_ENTRY = 6 };
enum { CLS_NONE = -1, CLS_WINDOWED = 0, CLS_EXTINCT = 1, CLS_RETURNED = 2, CLS_OTHER = 3 };

struct Fate {
    int fate;
    V3 dir;
    float transmittance, exit_cos;
    int facet;
    uint32_t bounces;
    int exit_facet;
    uint32_t d_bounces;
    uint32_t reason;
};

__device__ __forceinline__ float fresnel_T(float n1, float n2, float ci, float ct) {
    float ds = fma_(n2, ct, n1 * ci);
    float dp = fma_(n2, ci, n1 * ct);
    if (fabsf(ds) < 1e-6f || fabsf(dp) < 1e-6f) return 0.0f;
    float a = fma_(n2, -ct, n1 * ci) / ds;
    float b = fma_(n2, -ci, n1 * ct) / dp;
    return clampf(1.0f - 0.5f * (a * a + b * b), 0.0f, 1.0f);
}

// trace_entry for a polished entry facet (ray_trace.rs:256-425).
__device__ __forceinline__ void trace_entry(const float4* __restrict__ pl, uint32_t n, V3 point,
                                            V3 dir, V3 nrm, float cos_i, float index, Fate& F) {
    F.fate = FATE_ABSORBED;
    F.exit_facet = -1;
    F.d_bounces = 0;
    F.reason = REASON_BUDGET;
    F.facet = 0;
    F.bounces = 0;
    F.transmittance = 0.0f;
    F.exit_cos = 0.0f;
    F.dir = v3(0.0f, 0.0f, 0.0f);
    const float sin2_t = (1.0f / (index * index)) * fma_(cos_i, -cos_i, 1.0f);
    if (sin2_t > 1.0f) {
        F.fate = FATE_ENTRY_BLOCKED;
        F.reason = REASON_ENTRY;
        return;
    }
    const float cos_t = sqrt_rn(1.0f - sin2_t);
    const float inv = 1.0f / index;
    V3 cur = normalize(dir * inv + nrm * fma_(inv, cos_i, -cos_t));
    const float entry_T = fresnel_T(1.0f, index, cos_i, cos_t);
    V3 hp = point + cur * 1e-4f;
    const float sin_crit = 1.0f / index;
    const float idx2 = index * index;
    for (uint32_t b = 0; b < 10; ++b) {
        F.d_bounces = b;
        float t;
        int f = intersect(pl, n, hp, cur, t);
        if (f < 0) {
            F.reason = REASON_NOHIT;
            return;
        }
        V3 np = hp + cur * t;
        float4 pf = pl[f];
        V3 n_out = v3(pf.x, pf.y, pf.z);
        float ct = clampf(dot(cur, n_out), 0.0f, 1.0f);
        float one_m = fma_(ct, -ct, 1.0f);
        float st = sqrt_rn(fmaxf(one_m, 0.0f));
        if (st < sin_crit) {
            float s2o = idx2 * one_m;
            if (s2o <= 1.0f) {
                float co = sqrt_rn(1.0f - s2o);
                V3 od = normalize(cur * index + n_out * fma_(index, -ct, co));
                F.exit_facet = f;
                F.reason = REASON_SIDE;
                if (n_out.y < -0.05f) {
                    F.reason = REASON_NONE;
                    F.fate = FATE_LEAKED;
                } else if (od.y > 0.05f) {
                    F.reason = REASON_NONE;
                    F.fate = FATE_UPWARD;
                    F.dir = od;
                    F.transmittance = entry_T * fresnel_T(index, 1.0f, ct, co);
                    F.exit_cos = co;
                    F.facet = f;
                    F.bounces = b;
                }
                return;
            }
        }
        cur = normalize(cur - n_out * (2.0f * ct));
        hp = np + cur * 1e-4f;
        F.d_bounces = b + 1;
    }
}

__device__ __forceinline__ float scoring_radiance(const MetricShared& S, V3 dir) {
    // ASET reads its probe wavelength (spec550 holds 550) in place of a power
    if (S.env.model == ENV_ASET) return aset_radiance(S.env, normalize(dir), S.spec550, 1.0f);
    if (S.env.model != ENV_STUDIO) return lit_factor(S.env, dir, v3(0.0f, 0.0f, 0.0f)) * (S.spec550 * 1.0f);
    DirLight L;
    direction_lighting(S.env, dir, v3(0.0f, 0.0f, 0.0f), L);
    return light_radiance(S.env, L, S.spec550);
}

struct Classified {
    int cls;
    int has_fire;
    float angle, weight;
    int entry_facet;
};

// classify_ray (classify.rs:109-197) + fire_sample (classify.rs:199-256).
// `observer` is the unit direction towards the eye the head-shadow cone is
// centred on: -dir for a perspective pixel, -forward for the fan.
__device__ __forceinline__ void classify_ray(const MetricShared& S, const float4* __restrict__ pl, uint32_t n_planes,
                                             float nd, float nf, float nc, V3 org, V3 dir, V3 observer,
                                             bool want_fire, Classified& C, Fate& F) {
    C.cls = CLS_NONE;
    C.has_fire = 0;
    C.entry_facet = -1;
    float t;
    int f = intersect(pl, n_planes, org, dir, t);
    if (f < 0) return;
    C.entry_facet = f;
    V3 hp = org + dir * t;
    float4 pf = pl[f];
    V3 nrm = v3(pf.x, pf.y, pf.z);
    float cos_i = clampf(dot(neg(dir), nrm), 0.0f, 1.0f);
    trace_entry(pl, n_planes, hp, dir, nrm, cos_i, nd, F);
    if (F.fate == FATE_ENTRY_BLOCKED) { C.cls = CLS_OTHER; return; }
    if (F.fate == FATE_LEAKED) { C.cls = CLS_WINDOWED; return; }
    if (F.fate == FATE_ABSORBED) { C.cls = CLS_EXTINCT; return; }
    if (dot(F.dir, observer) > S.head_cos) {
        F.reason = REASON_SHADOW;
        C.cls = CLS_EXTINCT;
        return;
    }
    if (!(scoring_radiance(S, F.dir) > S.threshold)) {
        F.reason = REASON_ILLUM;
        C.cls = CLS_EXTINCT;
        return;
    }
    C.cls = CLS_RETURNED;
    if (want_fire) {
        Fate A, B;
        trace_entry(pl, n_planes, hp, dir, nrm, cos_i, nf, A);
        trace_entry(pl, n_planes, hp, dir, nrm, cos_i, nc, B);
        if (A.fate == FATE_UPWARD && B.fate == FATE_UPWARD && A.facet == B.facet &&
            A.bounces == B.bounces) {
            float cs = clampf(dot(A.dir, B.dir), -1.0f, 1.0f);
            C.angle = acosf(cs) * 57.2957795130823208767981548141051703f;
            float energy = F.transmittance * A.transmittance * B.transmittance;
            float radiance = fmaxf(A.exit_cos, 0.0f) * fmaxf(B.exit_cos, 0.0f);
            C.weight = energy * radiance;
            C.has_fire = 1;
        }
    }
}

__device__ __forceinline__ void classify(const MetricShared& S, const MetricImage& I,
                                         const float4* __restrict__ pl, V3 org, V3 dir,
                                         bool want_fire, Classified& C, Fate& F) {
    classify_ray(S, pl, I.n_planes, I.nd, I.nf, I.nc, org, dir, neg(dir), want_fire, C, F);
}

// Camera::generate_ray with zero jitter (camera.rs:149-167).
__device__ __forceinline__ void cam_ray(const float* fwd, const float* right, const float* up,
                                        float fov_tan, float sx, float sy, float pw, float ph,
                                        V3& dir) {
    float aspect = pw / ph;
    float u = ((sx + 0.0f) / pw - 0.5f) * 2.0f * aspect * fov_tan;
    float v = (0.5f - (sy + 0.0f) / ph) * 2.0f * fov_tan;
    dir = normalize((v3(fwd[0], fwd[1], fwd[2]) + v3(right[0], right[1], right[2]) * u) +
                    v3(up[0], up[1], up[2]) * v);
}

// image.rs:21-36: sample 0 is the pixel centre, then the base-2 and base-3
// radical inverses evaluated in double precision and rounded to f32.  A
// request has at most 64 samples, so the offsets are a table the compiler
// evaluates (its constant evaluation rounds as IEEE double arithmetic does):
// the kernels read f32 values and compute nothing in double.
struct SampleOffsets {
    float jx[64], jy[64];
    constexpr SampleOffsets() : jx(), jy() {
        jx[0] = 0.5f;
        jy[0] = 0.5f;
        for (uint32_t s = 1; s < 64; ++s) {
            uint32_t rev = 0;
            for (int b = 0; b < 32; ++b) rev |= ((s >> b) & 1u) << (31 - b);
            jx[s] = (float)((double)rev / 4294967296.0);
            uint32_t n = s;
            double scale = 1.0 / 3.0, y = 0.0;
            while (n > 0) {
                y += (double)(n % 3) * scale;
                n /= 3;
                scale /= 3.0;
            }
            jy[s] = (float)y;
        }
    }
};
__constant__ const SampleOffsets kSampleOffsets{};

__device__ __forceinline__ void sample_offset(uint32_t s, float& jx, float& jy) {
    const uint32_t i = s < 64u ? s : 63u;  // requests hold at most 64 samples
    jx = kSampleOffsets.jx[i];
    jy = kSampleOffsets.jy[i];
}

// One pixel p of image I (sample_pixel, image.rs:110-175, plus the temporal
// sampler): the OpticalPixel the reference's evaluator produces for it.
__device__ __forceinline__ ixg_optical_pixel metric_pixel(const MetricShared& S, const MetricImage& I,
                                                          const float4* __restrict__ sp, uint32_t p) {
    const uint32_t x = p % S.width, y = p / S.width;
    const float pw = (float)S.proj_w, ph = (float)S.proj_h;
    const V3 org = v3(I.origin[0], I.origin[1], I.origin[2]);

    ixg_optical_pixel px;
    px.hits = 0;
    uint32_t n_cls[4] = {0u, 0u, 0u, 0u};  // summed by selects: no dynamic index
    px.fire_sum_deg = 0.0f;
    px.fire_qualified = px.fire_rejected = 0;
    px.centre_entry_facet = -1;
    px.centre_exit_facet = -1;
    px.centre_bounces = 0;
    px.centre_reason = 0;
    px.mixed_facets = 0;
    px.scint_returned = 0;
    px.scint_poses = 0;
    px.pad[0] = px.pad[1] = 0;
    int first = -1;
    for (uint32_t s = 0; s < S.samples; ++s) {
        float jx, jy;
        sample_offset(s, jx, jy);
        float sx = fma_((float)x + jx, S.scale_x, (float)S.origin_x - 0.5f);
        float sy = fma_((float)y + jy, S.scale_y, (float)S.origin_y - 0.5f);
        V3 dir;
        cam_ray(I.forward, I.right, I.up, I.fov_tan, sx, sy, pw, ph, dir);
        Classified C;
        Fate det;
        classify(S, I, sp, org, dir, S.want_fire != 0, C, det);
        if (C.cls == CLS_NONE) continue;
        px.hits++;
        if (s == 0) {
            px.centre_entry_facet = C.entry_facet;
            px.centre_exit_facet = det.exit_facet;
            px.centre_bounces = det.d_bounces;
            px.centre_reason = det.reason;
        }
        if (first >= 0) px.mixed_facets |= (uint32_t)(C.entry_facet != first);
        else first = C.entry_facet;
        if (C.cls == CLS_RETURNED && S.want_fire) {
            if (C.has_fire && __builtin_isfinite(C.angle) && __builtin_isfinite(C.weight)) {
                px.fire_sum_deg = fma_(I.zero_dispersion ? 0.0f : C.angle, C.weight, px.fire_sum_deg);
                px.fire_qualified++;
            } else {
                px.fire_rejected++;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) n_cls[k] += C.cls == k ? 1u : 0u;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) px.classes[k] = n_cls[k];
    if (px.hits > 0 && S.want_scintillation) {
        // TemporalSampler::apply (image_scint.rs:95-118): the centre ray from
        // five yaw-offset cameras; pose 2 (no offset) must hit for any count.
        float sx = fma_((float)x + 0.5f, S.scale_x, (float)S.origin_x - 0.5f);
        float sy = fma_((float)y + 0.5f, S.scale_y, (float)S.origin_y - 0.5f);
        int ret[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            V3 dir;
            cam_ray(I.t_forward[i], I.t_right[i], I.t_up[i], I.fov_tan, sx, sy, pw, ph, dir);
            Classified C;
            Fate unused;
            classify(S, I, sp, v3(I.t_origin[i][0], I.t_origin[i][1], I.t_origin[i][2]), dir, false, C,
                     unused);
            ret[i] = C.cls == CLS_NONE ? -1 : (C.cls == CLS_RETURNED ? 1 : 0);
        }
        if (ret[2] >= 0) {
            uint32_t nret = ret[2] == 1 ? 1u : 0u;
#pragma unroll
            for (int i = 0; i < 5; ++i)
                if (i != 2 && ret[i] == 1) ++nret;
            px.scint_returned = (uint8_t)nret;
            px.scint_poses = 5;
        }
    }
    return px;
}

// 2-D grid: blockIdx.y = image, blockIdx.x = a kMBlock-pixel tile.
__global__ __launch_bounds__(kMBlock) void metric_pixels_kernel(const MetricShared S,
                                                                const MetricImage* __restrict__ images,
                                                                const float4* __restrict__ planes,
                                                                ixg_optical_pixel* __restrict__ out) {
    __shared__ float4 sp[IXG_MAX_PLANES];
    const uint32_t img = blockIdx.y;
    const MetricImage& I = images[img];
    for (uint32_t i = threadIdx.x; i < I.n_planes; i += blockDim.x) sp[i] = planes[I.plane_offset + i];
    __syncthreads();
    const uint32_t npx = S.width * S.height;
    const uint32_t p = blockIdx.x * kMBlock + threadIdx.x;
    if (p >= npx) return;
    out[(size_t)img * npx + p] = metric_pixel(S, I, sp, p);
}

// Persistent work queue over (image, tile) items, image-major: a resident
// block takes the next item from a global counter, restages the facet set in
// LDS only when the image changes (consecutive items mostly share a design),
// and runs the tile's pixels one per lane, TILE / BLOCK passes per item.  The
// item count bounds the loop, so every block leaves once the queue is empty.

template <int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) IXG_METRIC_WAVES_ATTR void metric_queue_kernel(const MetricShared S,
                                                             const MetricImage* __restrict__ images,
                                                             const float4* __restrict__ planes,
                                                             ixg_optical_pixel* __restrict__ out,
                                                             uint32_t* __restrict__ counter, uint32_t n_items) {
    __shared__ float4 sp[IXG_MAX_PLANES];
    __shared__ uint32_t s_item;
    const uint32_t npx = S.width * S.height;
    const uint32_t tiles = (npx + TILE - 1) / TILE;
    uint32_t cur = 0xFFFFFFFFu;
    while (true) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        // uniform across the block: scalar loads of the image record, in SGPRs
        const uint32_t item = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_item);
        if (item >= n_items) break;
        const uint32_t img = item / tiles, tile = item % tiles;
        const MetricImage& I = images[img];
        if (img != cur) {
            const MetricImage& J = images[img];
            for (uint32_t i = threadIdx.x; i < J.n_planes; i += BLOCK) sp[i] = planes[J.plane_offset + i];
            cur = img;
        }
        __syncthreads();
#pragma unroll 1
        for (uint32_t k = 0; k < TILE; k += BLOCK) {
            const uint32_t p = tile * TILE + k + threadIdx.x;
            if (p < npx) out[(size_t)img * npx + p] = metric_pixel(S, I, sp, p);
        }
        __syncthreads();  // s_item and sp are rewritten by the next item
    }
}

// ---------------------------------------------------------------------------
// Flow kernel: the same pixels with the lanes kept together in the slab scan.
//
// metric_pixel nests three loops per pixel (samples; <= 10 bounces per trace;
// the F / C replays of a returned ray and the five temporal rays), and across
// a wave those trip counts differ, so most lanes sit masked while the longest
// trace runs.  Here a lane runs its pixel as a chain of traces -- per sample
// the camera ray's entry scan, the d-line trace and, for a returned ray, the
// F and C replays; then the five temporal camera rays -- and each iteration
// of the loop advances every lane by exactly one scan of its current trace.
// A lane whose pixel is complete writes it and takes the next pixel of the
// block's item (one LDS atomic per wave).  Every lane computes its pixel with
// the operations of metric_pixel in the same order, so the records are the
// same bits.
// ---------------------------------------------------------------------------
enum FlowTask : uint32_t {
    FT_CAM = 0,     // the camera ray of sample s: its entry scan
    FT_DLINE = 1,   // the sample's d-line trace
    FT_FLINE = 2,   // the F-line replay of a returned ray
    FT_CLINE = 3,   // the C-line replay
    FT_TCAM = 4,    // temporal camera ti: its entry scan
    FT_TDLINE = 5,  // its d-line trace
};

// One trace_entry in progress: what its bounce loop carries (the index and
// its square are re-read from the image for the line being traced).
struct TraceRun {
    V3 hp, cur;
    float sin_crit, entry_T;
    uint32_t b;  // bounce; also the fate's d_bounces, and its bounces at an upward exit
};

// The fate of a flow trace: Fate less what the trace itself holds (the
// bounce counts are TraceRun::b), its small integers packed in one word:
// bits 0-1 fate, 2-10 exit facet + 1, 11-18 exit facet of an upward exit
// (Fate::facet), 19-21 reason.
struct FlowFate {
    uint32_t bits;
    V3 dir;
    float transmittance, exit_cos;
};
__device__ __forceinline__ uint32_t ff_fate(uint32_t b) { return b & 3u; }
__device__ __forceinline__ int32_t ff_exit(uint32_t b) { return (int32_t)((b >> 2) & 511u) - 1; }
__device__ __forceinline__ uint32_t ff_facet(uint32_t b) { return (b >> 11) & 255u; }
__device__ __forceinline__ uint32_t ff_reason(uint32_t b) { return (b >> 19) & 7u; }
__device__ __forceinline__ uint32_t ff_pack(uint32_t fate, int32_t exit_facet, uint32_t facet, uint32_t reason) {
    return fate | ((uint32_t)(exit_facet + 1) << 2) | (facet << 11) | (reason << 19);
}
__device__ __forceinline__ uint32_t ff_with_reason(uint32_t b, uint32_t reason) {
    return (b & ~(7u << 19)) | (reason << 19);
}

// trace_entry up to its bounce loop; false when the trace ends without a
// scan (entry blocked).
__device__ __forceinline__ bool trace_begin(V3 point, V3 dir, V3 nrm, float cos_i, float index, FlowFate& F,
                                            TraceRun& R) {
    F.bits = ff_pack(FATE_ABSORBED, -1, 0u, REASON_BUDGET);
    F.transmittance = 0.0f;
    F.exit_cos = 0.0f;
    F.dir = v3(0.0f, 0.0f, 0.0f);
    R.b = 0;
    const float sin2_t = (1.0f / (index * index)) * fma_(cos_i, -cos_i, 1.0f);
    if (sin2_t > 1.0f) {
        F.bits = ff_pack(FATE_ENTRY_BLOCKED, -1, 0u, REASON_ENTRY);
        return false;
    }
    const float cos_t = sqrt_rn(1.0f - sin2_t);
    const float inv = 1.0f / index;
    R.cur = normalize(dir * inv + nrm * fma_(inv, cos_i, -cos_t));
    R.entry_T = fresnel_T(1.0f, index, cos_i, cos_t);
    R.hp = point + R.cur * 1e-4f;
    R.sin_crit = 1.0f / index;
    return true;
}

// One iteration of trace_entry's bounce loop after its scan gave (f, t);
// false when the trace has ended.
__device__ __forceinline__ bool trace_step(const float4* __restrict__ pl, int f, float t, float index, FlowFate& F,
                                           TraceRun& R) {
    if (f < 0) {
        F.bits = ff_with_reason(F.bits, REASON_NOHIT);
        return false;
    }
    const V3 np = R.hp + R.cur * t;
    const float4 pf = pl[f];
    const V3 n_out = v3(pf.x, pf.y, pf.z);
    const float ct = clampf(dot(R.cur, n_out), 0.0f, 1.0f);
    const float one_m = fma_(ct, -ct, 1.0f);
    const float st = sqrt_rn(fmaxf(one_m, 0.0f));
    if (st < R.sin_crit) {
        const float s2o = (index * index) * one_m;
        if (s2o <= 1.0f) {
            const float co = sqrt_rn(1.0f - s2o);
            const V3 od = normalize(R.cur * index + n_out * fma_(index, -ct, co));
            if (n_out.y < -0.05f) {
                F.bits = ff_pack(FATE_LEAKED, f, 0u, REASON_NONE);
            } else if (od.y > 0.05f) {
                F.bits = ff_pack(FATE_UPWARD, f, (uint32_t)f, REASON_NONE);
                F.dir = od;
                F.transmittance = R.entry_T * fresnel_T(index, 1.0f, ct, co);
                F.exit_cos = co;
            } else {
                F.bits = ff_pack(ff_fate(F.bits), f, 0u, REASON_SIDE);
            }
            return false;
        }
    }
    R.cur = normalize(R.cur - n_out * (2.0f * ct));
    R.hp = np + R.cur * 1e-4f;
    ++R.b;
    return R.b < 10u;
}

// classify_ray's verdict on a finished d-line trace (classify.rs:140-183).
__device__ __forceinline__ int classify_fate(const MetricShared& S, FlowFate& F, V3 observer) {
    const uint32_t fate = ff_fate(F.bits);
    if (fate == FATE_ENTRY_BLOCKED) return CLS_OTHER;
    if (fate == FATE_LEAKED) return CLS_WINDOWED;
    if (fate == FATE_ABSORBED) return CLS_EXTINCT;
    if (dot(F.dir, observer) > S.head_cos) {
        F.bits = ff_with_reason(F.bits, REASON_SHADOW);
        return CLS_EXTINCT;
    }
    if (!(scoring_radiance(S, F.dir) > S.threshold)) {
        F.bits = ff_with_reason(F.bits, REASON_ILLUM);
        return CLS_EXTINCT;
    }
    return CLS_RETURNED;
}

// What one lane carries between scans.  The per-pixel counts (each <= 64
// samples) are 8-bit fields:
//   counts  classes[0..3]            (byte k = class k)
//   tally   hits | fire_qualified << 8 | fire_rejected << 16 | mixed << 24
//   centre  entry + 1 (bits 0-8) | exit + 1 (9-17) | bounces (18-21) | reason (22-24)
//   step    sample (0-7) | task (8-10) | temporal camera (12-14) |
//           temporal hits (16-20) | temporal returns (24-28)
//   facets  first entry facet + 1 (0-8) | this sample's entry facet (16-23)
//   a_bits  the F-line fate's fate (0-1) | facet (2-9) | bounces (10-13)
struct FlowLane {
    uint32_t p;
    uint32_t step, counts, tally, centre, facets;
    float fire_sum;
    V3 hp0, dir0;   // the traced ray's entry point and direction (its facet is in `facets`)
    float d_T;      // d-line transmittance (Fire energy)
    uint32_t a_bits;
    V3 a_dir;
    float a_T, a_exit;
    TraceRun R;     // the running trace; R.hp / R.cur also hold a camera ray
    FlowFate F;     // its fate
};
__device__ __forceinline__ uint32_t fl_sample(const FlowLane& L) { return L.step & 255u; }
__device__ __forceinline__ uint32_t fl_task(const FlowLane& L) { return (L.step >> 8) & 7u; }
__device__ __forceinline__ uint32_t fl_ti(const FlowLane& L) { return (L.step >> 12) & 7u; }
__device__ __forceinline__ void fl_set_task(FlowLane& L, uint32_t task) { L.step = (L.step & ~(7u << 8)) | (task << 8); }
__device__ __forceinline__ int32_t fl_entry(const FlowLane& L) { return (int32_t)((L.facets >> 16) & 255u); }

// The index of the line a trace runs at: F and C for the replays, else d.
__device__ __forceinline__ float fl_index(const MetricImage& I, uint32_t task) {
    return task == FT_FLINE ? I.nf : (task == FT_CLINE ? I.nc : I.nd);
}

// The next camera ray of the lane's pixel: sample `sample` of the pixel's
// sub-pixel pattern, or (temporal) the pixel centre from temporal camera ti
// (Camera::generate_ray with zero jitter, camera.rs:149-167).
__device__ __forceinline__ void flow_camera(const MetricShared& S, const MetricImage& I, FlowLane& L, bool temporal) {
    const uint32_t x = L.p % S.width, y = L.p / S.width;
    float jx = 0.5f, jy = 0.5f;
    if (!temporal) sample_offset(fl_sample(L), jx, jy);
    const float sx = fma_((float)x + jx, S.scale_x, (float)S.origin_x - 0.5f);
    const float sy = fma_((float)y + jy, S.scale_y, (float)S.origin_y - 0.5f);
    const uint32_t i = fl_ti(L);
    const float* fwd = temporal ? I.t_forward[i] : I.forward;
    const float* right = temporal ? I.t_right[i] : I.right;
    const float* up = temporal ? I.t_up[i] : I.up;
    const float* org = temporal ? I.t_origin[i] : I.origin;
    cam_ray(fwd, right, up, I.fov_tan, sx, sy, (float)S.proj_w, (float)S.proj_h, L.R.cur);
    L.R.hp = v3(org[0], org[1], org[2]);
    fl_set_task(L, temporal ? FT_TCAM : FT_CAM);
}

__device__ __forceinline__ void flow_pixel_start(const MetricShared& S, const MetricImage& I, FlowLane& L,
                                                 uint32_t p) {
    L.p = p;
    L.step = 0u;
    L.counts = L.tally = L.centre = L.facets = 0u;  // centre 0: entry -1, exit -1, bounces 0, reason 0
    L.fire_sum = 0.0f;
    flow_camera(S, I, L, false);
}

// The pixel's record, as metric_pixel returns it.
__device__ __forceinline__ ixg_optical_pixel flow_record(const FlowLane& L) {
    ixg_optical_pixel px;
    px.hits = L.tally & 255u;
    px.classes[0] = L.counts & 255u;
    px.classes[1] = (L.counts >> 8) & 255u;
    px.classes[2] = (L.counts >> 16) & 255u;
    px.classes[3] = L.counts >> 24;
    px.fire_sum_deg = L.fire_sum;
    px.fire_qualified = (L.tally >> 8) & 255u;
    px.fire_rejected = (L.tally >> 16) & 255u;
    px.centre_entry_facet = (int32_t)(L.centre & 511u) - 1;
    px.centre_exit_facet = (int32_t)((L.centre >> 9) & 511u) - 1;
    px.centre_bounces = (L.centre >> 18) & 15u;
    px.centre_reason = (L.centre >> 22) & 7u;
    px.mixed_facets = L.tally >> 24;
    px.scint_returned = 0;
    px.scint_poses = 0;
    const uint32_t t_hit = (L.step >> 16) & 31u, t_ret = (L.step >> 24) & 31u;
    if (t_hit & 4u) {  // pose 2 (no offset) must meet the stone for any count
        px.scint_returned = (uint8_t)__popc(t_ret);
        px.scint_poses = 5;
    }
    px.pad[0] = px.pad[1] = 0;
    return px;
}

// The camera ray (L.R.hp, L.R.cur) met facet f at t: the entry of
// classify_ray (classify.rs:118-131).
__device__ __forceinline__ void flow_entry(FlowLane& L, int f, float t) {
    L.facets = (L.facets & 0xFFFFu) | ((uint32_t)f << 16);
    L.dir0 = L.R.cur;
    L.hp0 = L.R.hp + L.R.cur * t;
}

// The entry facet's normal and the cosine of incidence (classify.rs:125-126).
__device__ __forceinline__ void flow_entry_geometry(const float4* __restrict__ sp, const FlowLane& L, V3& nrm,
                                                    float& cos_i) {
    const float4 pf = sp[fl_entry(L)];
    nrm = v3(pf.x, pf.y, pf.z);
    cos_i = clampf(dot(neg(L.dir0), nrm), 0.0f, 1.0f);
}

// The F-line fate (in L.F, bounces in L.R.b) kept while the C line runs.
__device__ __forceinline__ void flow_keep_fline(FlowLane& L) {
    L.a_bits = ff_fate(L.F.bits) | (ff_facet(L.F.bits) << 2) | (L.R.b << 10);
    L.a_dir = L.F.dir;
    L.a_T = L.F.transmittance;
    L.a_exit = L.F.exit_cos;
}

// fire_sample's tail (classify.rs:226-256) with the C-line fate in L.F.
__device__ __forceinline__ bool flow_fire(const FlowLane& L, float& angle, float& weight) {
    const uint32_t a_fate = L.a_bits & 3u, a_facet = (L.a_bits >> 2) & 255u, a_bounces = L.a_bits >> 10;
    if (a_fate == FATE_UPWARD && ff_fate(L.F.bits) == FATE_UPWARD && a_facet == ff_facet(L.F.bits) &&
        a_bounces == L.R.b) {
        const float cs = clampf(dot(L.a_dir, L.F.dir), -1.0f, 1.0f);
        angle = acosf(cs) * 57.2957795130823208767981548141051703f;
        const float energy = L.d_T * L.a_T * L.F.transmittance;
        const float radiance = fmaxf(L.a_exit, 0.0f) * fmaxf(L.F.exit_cos, 0.0f);
        weight = energy * radiance;
        return true;
    }
    return false;
}

// Advance the lane by the scan result (f, t) of its current trace or camera
// ray; false once its pixel is complete.  Each heavy step (classification,
// trace set-up, camera ray) has one call site, so the inlined code stays
// small.
__device__ __forceinline__ bool flow_advance(const MetricShared& S, const MetricImage& I,
                                             const float4* __restrict__ sp, FlowLane& L, int f, float t) {
    const uint32_t task = fl_task(L);
    const bool temporal = task >= FT_TCAM;
    bool cam_miss = false;
    if (task == FT_CAM || task == FT_TCAM) {
        if (f < 0) {
            cam_miss = true;
        } else {
            flow_entry(L, f, t);
            V3 nrm;
            float cos_i;
            flow_entry_geometry(sp, L, nrm, cos_i);
            fl_set_task(L, task + 1);  // FT_DLINE / FT_TDLINE
            if (trace_begin(L.hp0, L.dir0, nrm, cos_i, I.nd, L.F, L.R)) return true;
            // entry blocked: the d-line trace has ended
        }
    } else if (trace_step(sp, f, t, fl_index(I, task), L.F, L.R)) {
        return true;
    }
    // a camera ray missed, or a trace ended with its fate in L.F
    int cls = CLS_NONE;
    const bool dline = task == FT_CAM || task == FT_DLINE || task == FT_TCAM || task == FT_TDLINE;
    if (dline && !cam_miss) cls = classify_fate(S, L.F, neg(L.dir0));
    if (temporal) {
        const uint32_t ti = fl_ti(L);
        L.step |= cam_miss ? 0u : (1u << (16 + ti));
        L.step |= cls == CLS_RETURNED ? (1u << (24 + ti)) : 0u;
        if (ti + 1u < 5u) {
            L.step += 1u << 12;
            flow_camera(S, I, L, true);
            return true;
        }
        return false;
    }
    bool has_fire = false;
    float angle = 0.0f, weight = 0.0f;
    int line = 0;  // the replay to set up next: 1 F, 2 C
    if (dline) {
        if (cls != CLS_NONE && fl_sample(L) == 0) {  // the centre sample's record
            L.centre = (uint32_t)(fl_entry(L) + 1) | ((uint32_t)(ff_exit(L.F.bits) + 1) << 9) | (L.R.b << 18) |
                       (ff_reason(L.F.bits) << 22);
        }
        L.d_T = L.F.transmittance;
        if (cls == CLS_RETURNED && S.want_fire) line = 1;
    } else if (task == FT_FLINE) {
        flow_keep_fline(L);
        line = 2;
    } else {  // FT_CLINE
        cls = CLS_RETURNED;
        has_fire = flow_fire(L, angle, weight);
    }
    while (line) {
        V3 nrm;
        float cos_i;
        flow_entry_geometry(sp, L, nrm, cos_i);
        fl_set_task(L, line == 1 ? FT_FLINE : FT_CLINE);
        if (trace_begin(L.hp0, L.dir0, nrm, cos_i, line == 1 ? I.nf : I.nc, L.F, L.R)) return true;
        // entry blocked at this index: the replay has ended at once
        if (line == 1) {
            flow_keep_fline(L);
            line = 2;
        } else {
            cls = CLS_RETURNED;
            has_fire = flow_fire(L, angle, weight);
            line = 0;
        }
    }
    // the sample's tail (metric_pixel's sample loop body after classify)
    if (cls != CLS_NONE) {
        L.tally += 1u;  // hits
        const uint32_t entry = (uint32_t)fl_entry(L);
        const uint32_t first = L.facets & 511u;
        if (first > 0u) L.tally |= (uint32_t)(entry + 1u != first) << 24;
        else L.facets |= entry + 1u;
        if (cls == CLS_RETURNED && S.want_fire) {
            if (has_fire && __builtin_isfinite(angle) && __builtin_isfinite(weight)) {
                L.fire_sum = fma_(I.zero_dispersion ? 0.0f : angle, weight, L.fire_sum);
                L.tally += 1u << 8;
            } else {
                L.tally += 1u << 16;
            }
        }
        L.counts += 1u << (8 * cls);
    }
    const uint32_t s = fl_sample(L) + 1u;
    L.step = (L.step & ~255u) | s;
    const bool more_samples = s < S.samples;
    if (more_samples || ((L.tally & 255u) > 0u && S.want_scintillation)) {
        flow_camera(S, I, L, !more_samples);  // ti is 0
        return true;
    }
    return false;
}

// Occupancy target of the flow kernel (waves per SIMD).  Its lanes carry a
// whole pixel's trace state (packed: 128 VGPRs, nothing spilled).
#ifndef IXG_METRIC_FLOW_WAVES_PER_EU
#define IXG_METRIC_FLOW_WAVES_PER_EU 4
#endif

// Persistent over (image, tile) items as metric_queue_kernel; inside an item
// the lanes take pixels one at a time and flow through their traces.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(IXG_METRIC_FLOW_WAVES_PER_EU))) void
metric_flow_kernel(const MetricShared S, const MetricImage* __restrict__ images, const float4* __restrict__ planes,
                   ixg_optical_pixel* __restrict__ out, uint32_t* __restrict__ counter, uint32_t n_items,
                   uint32_t tile) {
    __shared__ float4 sp[IXG_MAX_PLANES];
    __shared__ uint32_t s_item, s_next;
    const uint32_t npx = S.width * S.height;
    const uint32_t tiles = (npx + tile - 1) / tile;
    const uint32_t lane = __lane_id();
    uint32_t cur = 0xFFFFFFFFu;
    while (true) {
        if (threadIdx.x == 0) {
            s_item = atomicAdd(counter, 1u);
            s_next = 0u;
        }
        __syncthreads();
        const uint32_t item = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_item);
        if (item >= n_items) break;
        const uint32_t img = item / tiles;
        const uint32_t first_px = (item % tiles) * tile;
        const uint32_t n_px = min(tile, npx - first_px);
        const MetricImage& I = images[img];
        if (img != cur) {
            for (uint32_t i = threadIdx.x; i < I.n_planes; i += BLOCK) sp[i] = planes[I.plane_offset + i];
            cur = img;
        }
        __syncthreads();
        ixg_optical_pixel* dst = out + (size_t)img * npx;
        FlowLane L;
        bool busy = false, dry = false;
        while (true) {
            if (!dry) {
                // lanes without a pixel take the next ones: one LDS atomic per wave
                const uint64_t need = __ballot(!busy);
                if (need) {
                    const uint32_t leader = (uint32_t)(__ffsll((unsigned long long)need) - 1);
                    const uint32_t cnt = (uint32_t)__popcll(need);
                    uint32_t base = 0;
                    if (lane == leader) base = atomicAdd(&s_next, cnt);
                    base = (uint32_t)__shfl((int)base, (int)leader);
                    if (!busy) {
                        const uint32_t k = base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
                        if (k < n_px) {
                            flow_pixel_start(S, I, L, first_px + k);
                            busy = true;
                        }
                    }
                    dry = base + cnt >= n_px;
                }
            }
            if (__ballot(busy) == 0ull) break;
            if (busy) {
                float t = 0.0f;
                const int f = intersect(sp, I.n_planes, L.R.hp, L.R.cur, t);
                busy = flow_advance(S, I, sp, L, f, t);
                if (!busy) dst[L.p] = flow_record(L);
            }
        }
        __syncthreads();  // s_item, s_next and sp are rewritten by the next item
    }
}

__global__ void metric_zero_counter_kernel(uint32_t* c) { *c = 0u; }

// Per-image totals and metrics in raster order (one lane per image).
__global__ __launch_bounds__(64) void metric_reduce_kernel(const MetricShared S,
                                                           const ixg_optical_pixel* __restrict__ px,
                                                           ixg_metric_totals* __restrict__ totals) {
    const uint32_t img = blockIdx.x * 64 + threadIdx.x;
    if (img >= S.n_images) return;
    const uint32_t npx = S.width * S.height;
    const ixg_optical_pixel* P = px + (size_t)img * npx;
    ixg_metric_totals T;
    T.total.hits = 0;
    T.total.classes[0] = T.total.classes[1] = T.total.classes[2] = T.total.classes[3] = 0;
    T.total.fire_sum_deg = 0.0f;
    T.total.fire_qualified = T.total.fire_rejected = 0;
    T.total.centre_entry_facet = T.total.centre_exit_facet = -1;
    T.total.centre_bounces = T.total.centre_reason = T.total.mixed_facets = 0;
    T.total.scint_returned = T.total.scint_poses = 0;
    T.total.pad[0] = T.total.pad[1] = 0;
    float sum = 0.0f, sum_sq = 0.0f, var_sum = 0.0f;
    uint32_t count = 0, poses = 0;
    for (uint32_t i = 0; i < npx; ++i) {
        const ixg_optical_pixel q = P[i];
        T.total.hits += q.hits;
        for (int c = 0; c < 4; ++c) T.total.classes[c] += q.classes[c];
        T.total.fire_sum_deg += q.fire_sum_deg;
        T.total.fire_qualified += q.fire_qualified;
        T.total.fire_rejected += q.fire_rejected;
        if (q.hits > 0) {
            float f = (float)q.classes[2] / (float)(q.hits > 1 ? q.hits : 1);
            sum += f;
            sum_sq = fma_(f, f, sum_sq);
            ++count;
        }
    }
    for (uint32_t i = 0; i < npx; ++i) {
        const ixg_optical_pixel q = P[i];
        if (q.scint_poses) {
            float fr = (float)q.scint_returned / (float)q.scint_poses;
            var_sum += fma_(fr, -fr, fr);
            ++poses;
        }
    }
    float spatial = 0.0f;
    if (count) {
        float mean = sum / (float)count;
        if (mean > 1e-4f) {
            float var = fmaxf(fma_(mean, -mean, sum_sq / (float)count), 0.0f);
            float cv = sqrt_rn(var) / mean;
            spatial = clampf(cv / (1.0f + cv) * 100.0f, 0.0f, 100.0f);
        }
    }
    float temporal = poses ? clampf(((var_sum / (float)poses) / 0.24f) * 100.0f, 0.0f, 100.0f) : 0.0f;
    const float h = (float)(T.total.hits > 1 ? T.total.hits : 1);
    T.windowing_pct = clampf((float)T.total.classes[0] / h * 100.0f, 0.0f, 100.0f);
    T.extinction_pct = clampf((float)T.total.classes[1] / h * 100.0f, 0.0f, 100.0f);
    T.brilliance_pct = clampf((float)T.total.classes[2] / h * 100.0f, 0.0f, 100.0f);
    T.fire_index = fmaxf(T.total.fire_sum_deg / h * 275.0f, 0.1f);
    T.scintillation_pct = clampf(fma_(0.4f, temporal, 0.6f * spatial), 0.0f, 100.0f);
    totals[img] = T;
}

// The scoring radiance at explicit directions (the ambient-level probes).
__global__ void scoring_probe_kernel(const MetricShared S, const float* __restrict__ dirs, uint32_t n,
                                     float* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = scoring_radiance(S, v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}

// ---------------------------------------------------------------------------
// Fan-grid pose metrics (evaluate.rs:327-439): one block per (design, pose)
// job.  The five aperture rays of every disc cell (and, on the full pass, the
// cell's five temporal yaw-offset rays) are classified one per lane into LDS;
// lane 0 then folds them in the reference's (cell, sample) order, so the Fire
// fma chain, the cell-fraction sums and the temporal variances are formed
// exactly as the serial grid loop forms them.
// ---------------------------------------------------------------------------
constexpr int kPoseBlock = 256;
enum { CLS_RETURNED_FIRE = 4 };

__device__ __forceinline__ V3 row(const float (*b)[3], int k) { return v3(b[k][0], b[k][1], b[k][2]); }

// FanGeometry::origin (fan.rs:92-94)
__device__ __forceinline__ V3 fan_origin(V3 f, V3 r, V3 up, float standoff, float lateral, float u, float v) {
    return (neg(f) * standoff + r * (u * lateral)) + up * (v * lateral);
}

__global__ __launch_bounds__(kPoseBlock) IXG_METRIC_WAVES_ATTR void pose_kernel(const MetricShared S, const PoseDesign* __restrict__ designs,
                                                          const PoseJob* __restrict__ jobs,
                                                          const float4* __restrict__ planes,
                                                          const float2* __restrict__ cells, uint32_t n_cells,
                                                          int want_all, float* __restrict__ out) {
    __shared__ float4 sp[IXG_MAX_PLANES];
    __shared__ int8_t s_cls[kPoseMaxCells * 5];
    __shared__ float s_ang[kPoseMaxCells * 5];
    __shared__ float s_w[kPoseMaxCells * 5];
    __shared__ uint8_t s_ret[kPoseMaxCells * 5];
    const PoseJob& J = jobs[blockIdx.x];
    const PoseDesign D = designs[J.design];
    float* o = out + (size_t)blockIdx.x * 5;
    if (D.n_planes == 0) {  // placeholder_metrics (evaluate.rs:259-267)
        if (threadIdx.x == 0) {
            o[0] = 85.0f;
            o[1] = 25.0f;
            o[2] = 75.0f;
            o[3] = 5.0f;
            o[4] = 5.0f;
        }
        return;
    }
    for (uint32_t i = threadIdx.x; i < D.n_planes; i += kPoseBlock) sp[i] = planes[D.plane_offset + i];
    __syncthreads();
    const uint32_t n_main = n_cells * 5;
    const uint32_t n_items = want_all ? 2 * n_main : n_main;
    for (uint32_t i = threadIdx.x; i < n_items; i += kPoseBlock) {
        if (i < n_main) {
            // classify_aperture_sample (classify.rs:91-106), sub-aperture s
            const uint32_t c = i / 5, s = i % 5;
            const float2 uv = cells[c];
            const V3 f = row(J.basis[0], 0), r = row(J.basis[0], 1), up = row(J.basis[0], 2);
            const float sr = s == 1 ? 0.08f : (s == 2 ? -0.08f : 0.0f);
            const float su = s == 3 ? 0.08f : (s == 4 ? -0.08f : 0.0f);
            const V3 dir = normalize((f + r * sr) + up * su);
            const V3 org = fan_origin(f, r, up, D.standoff, D.lateral, uv.x, uv.y);
            Classified C;
            Fate unused;
            classify_ray(S, sp, D.n_planes, D.nd, D.nf, D.nc, org, dir, neg(f), want_all != 0, C, unused);
            int8_t code = (int8_t)C.cls;
            if (C.cls == CLS_RETURNED && C.has_fire) {
                code = CLS_RETURNED_FIRE;
                s_ang[i] = C.angle;
                s_w[i] = C.weight;
            }
            s_cls[i] = code;
        } else {
            // cell_returned_at_yaw_offset (scintillation.rs:79-101)
            const uint32_t j = i - n_main, c = j / 5, k = j % 5;
            const float2 uv = cells[c];
            const V3 f = row(J.basis[1 + k], 0), r = row(J.basis[1 + k], 1), up = row(J.basis[1 + k], 2);
            const V3 org = fan_origin(f, r, up, D.standoff, D.lateral, uv.x, uv.y);
            Classified C;
            Fate unused;
            classify_ray(S, sp, D.n_planes, D.nd, D.nf, D.nc, org, f, neg(f), false, C, unused);
            s_ret[j] = (uint8_t)(C.cls == CLS_RETURNED);
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    // run_grid / sample_cell accumulation (evaluate.rs:345-439)
    uint32_t total = 0, windowed = 0, extinct = 0, returned = 0, counted = 0;
    float fire = 0.0f, cf = 0.0f, cf2 = 0.0f, tv = 0.0f;
    for (uint32_t c = 0; c < n_cells; ++c) {
        uint32_t ct = 0, cr = 0;
        for (uint32_t s = 0; s < 5; ++s) {
            const uint32_t k = c * 5 + s;
            const int code = s_cls[k];
            if (code < 0) continue;
            ++total;
            ++ct;
            if (code == CLS_WINDOWED) ++windowed;
            else if (code == CLS_EXTINCT) ++extinct;
            else if (code == CLS_RETURNED || code == CLS_RETURNED_FIRE) {
                ++returned;
                ++cr;
                if (code == CLS_RETURNED_FIRE) fire = fma_(s_ang[k], s_w[k], fire);
            }
        }
        if (ct > 0) {
            const float frac = (float)cr / (float)ct;
            cf += frac;
            cf2 = fma_(frac, frac, cf2);
            ++counted;
            if (want_all) {  // cell_temporal_variance (scintillation.rs:110-128)
                uint32_t n5 = 0;
                for (uint32_t k = 0; k < 5; ++k) n5 += s_ret[c * 5 + k];
                const float p = (float)n5 / 5.0f;
                tv += fma_(p, -p, p);
            }
        }
    }
    // finish_metrics (accumulate.rs:112-173, scintillation.rs:146-198)
    const float nt = (float)(total > 1 ? total : 1);
    o[3] = clampf((float)windowed / nt * 100.0f, 0.0f, 100.0f);
    o[4] = clampf((float)extinct / nt * 100.0f, 0.0f, 100.0f);
    o[0] = clampf((float)returned / nt * 100.0f, 0.0f, 100.0f);
    if (!want_all) {
        o[1] = 25.0f;
        o[2] = 75.0f;
        return;
    }
    o[1] = fmaxf(fire / nt * 275.0f, 0.1f);
    float spatial = 0.0f, temporal = 0.0f;
    if (counted) {
        const float mean = cf / (float)counted;
        if (mean > 1e-4f) {
            const float var = fmaxf(fma_(mean, -mean, cf2 / (float)counted), 0.0f);
            const float cv = sqrt_rn(var) / mean;
            spatial = clampf(cv / (1.0f + cv) * 100.0f, 0.0f, 100.0f);
        }
        temporal = clampf(((tv / (float)counted) / 0.24f) * 100.0f, 0.0f, 100.0f);
    }
    o[2] = clampf(fma_(0.4f, temporal, 0.6f * spatial), 0.0f, 100.0f);
}

}  // namespace ixg

extern "C" hipError_t ixg_launch_scoring_probe(const ixg::MetricShared* S, const float* d_dirs, uint32_t n,
                                               float* d_out, hipStream_t stream) {
    hipLaunchKernelGGL(ixg::scoring_probe_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, *S, d_dirs, n,
                       d_out);
    return hipGetLastError();
}

extern "C" hipError_t ixg_launch_metric_pixels(const ixg::MetricShared* S, const ixg::MetricImage* d_images,
                                               const float4* d_planes, ixg_optical_pixel* d_out,
                                               hipStream_t stream) {
    uint32_t npx = S->width * S->height;
    dim3 grid((npx + ixg::kMBlock - 1) / ixg::kMBlock, S->n_images);
    hipLaunchKernelGGL(ixg::metric_pixels_kernel, grid, dim3(ixg::kMBlock), 0, stream, *S, d_images,
                       d_planes, d_out);
    return hipGetLastError();
}

// Queue shapes the autotuner chooses among: (block lanes, pixels per item).
extern "C" int ixg_metric_queue_shapes(int* block, int* tile, int max_n) {
    static const int shapes[][2] = {{64, 64}, {64, 256}, {128, 128}, {128, 512}, {256, 256}, {256, 1024}};
    const int n = (int)(sizeof shapes / sizeof shapes[0]);
    for (int i = 0; i < n && i < max_n; ++i) {
        block[i] = shapes[i][0];
        tile[i] = shapes[i][1];
    }
    return n;
}

extern "C" hipError_t ixg_launch_metric_queue(const ixg::MetricShared* S, const ixg::MetricImage* d_images,
                                              const float4* d_planes, ixg_optical_pixel* d_out, uint32_t* d_counter,
                                              int block, int tile, uint32_t grid_blocks, hipStream_t stream) {
    const uint32_t npx = S->width * S->height;
    const uint32_t n_items = S->n_images * ((npx + tile - 1) / tile);
    hipLaunchKernelGGL(ixg::metric_zero_counter_kernel, dim3(1), dim3(1), 0, stream, d_counter);
#define IXG_QUEUE(B, T)                                                                                        \
    if (block == B && tile == T) {                                                                             \
        hipLaunchKernelGGL((ixg::metric_queue_kernel<B, T>), dim3(grid_blocks), dim3(B), 0, stream, *S, d_images, \
                           d_planes, d_out, d_counter, n_items);                                               \
        return hipGetLastError();                                                                              \
    }
    IXG_QUEUE(64, 64)
    IXG_QUEUE(64, 256)
    IXG_QUEUE(128, 128)
    IXG_QUEUE(128, 512)
    IXG_QUEUE(256, 256)
    IXG_QUEUE(256, 1024)
#undef IXG_QUEUE
    return hipErrorInvalidValue;
}

// The flow kernel: block lanes, pixels per item.
extern "C" hipError_t ixg_launch_metric_flow(const ixg::MetricShared* S, const ixg::MetricImage* d_images,
                                             const float4* d_planes, ixg_optical_pixel* d_out, uint32_t* d_counter,
                                             int block, int tile, uint32_t grid_blocks, hipStream_t stream) {
    const uint32_t npx = S->width * S->height;
    if (tile <= 0) return hipErrorInvalidValue;
    const uint32_t n_items = S->n_images * ((npx + (uint32_t)tile - 1) / (uint32_t)tile);
    hipLaunchKernelGGL(ixg::metric_zero_counter_kernel, dim3(1), dim3(1), 0, stream, d_counter);
#define IXG_FLOW(B)                                                                                          \
    if (block == B) {                                                                                        \
        hipLaunchKernelGGL((ixg::metric_flow_kernel<B>), dim3(grid_blocks), dim3(B), 0, stream, *S, d_images, \
                           d_planes, d_out, d_counter, n_items, (uint32_t)tile);                             \
        return hipGetLastError();                                                                            \
    }
    IXG_FLOW(64)
    IXG_FLOW(128)
    IXG_FLOW(256)
#undef IXG_FLOW
    return hipErrorInvalidValue;
}

extern "C" hipError_t ixg_metric_flow_occupancy(int block, int* per_cu) {
    if (block == 64) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ixg::metric_flow_kernel<64>, 64, 0);
    if (block == 128) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ixg::metric_flow_kernel<128>, 128, 0);
    if (block == 256) return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ixg::metric_flow_kernel<256>, 256, 0);
    return hipErrorInvalidValue;
}

// Resident blocks per CU of a queue shape (the grid is this x CUs x fill).
extern "C" hipError_t ixg_metric_queue_occupancy(int block, int tile, int* per_cu) {
#define IXG_OCC(B, T)                                                                                 \
    if (block == B && tile == T)                                                                      \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ixg::metric_queue_kernel<B, T>, B, 0);
    IXG_OCC(64, 64)
    IXG_OCC(64, 256)
    IXG_OCC(128, 128)
    IXG_OCC(128, 512)
    IXG_OCC(256, 256)
    IXG_OCC(256, 1024)
#undef IXG_OCC
    return hipErrorInvalidValue;
}

extern "C" hipError_t ixg_launch_metric_reduce(const ixg::MetricShared* S, const ixg_optical_pixel* d_px,
                                               ixg_metric_totals* d_totals, hipStream_t stream) {
    dim3 grid((S->n_images + 63) / 64);
    hipLaunchKernelGGL(ixg::metric_reduce_kernel, grid, dim3(64), 0, stream, *S, d_px, d_totals);
    return hipGetLastError();
}

extern "C" hipError_t ixg_launch_pose(const ixg::MetricShared* S, const ixg::PoseDesign* d_designs,
                                      const ixg::PoseJob* d_jobs, uint32_t n_jobs, const float4* d_planes,
                                      const float2* d_cells, uint32_t n_cells, int want_all, float* d_out,
                                      hipStream_t stream) {
    if (n_cells > (uint32_t)ixg::kPoseMaxCells) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ixg::pose_kernel, dim3(n_jobs), dim3(ixg::kPoseBlock), 0, stream, *S, d_designs, d_jobs,
                       d_planes, d_cells, n_cells, want_all, d_out);
    return hipGetLastError();
}
