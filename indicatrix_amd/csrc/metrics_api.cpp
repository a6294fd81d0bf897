// metrics_api.cpp -- host side of the metric images and the batched sweep.
//
// Scene constants are resolved once per call on the host (glibc libm):
// the three line indices, the scoring rig, the head-shadow cone and the
// temporal cameras.  The ambient level that sets the illumination threshold
// (96 golden-angle hemisphere probes, lighting.rs:115-190) is evaluated by
// the device's own radiance function and sorted here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/ixg.h"
#include "host_scene.h"
#include "transport_params.h"

extern "C" hipError_t ixg_launch_metric_pixels(const ixg::MetricShared* S, const ixg::MetricImage* d_images,
                                               const float4* d_planes, ixg_optical_pixel* d_out,
                                               hipStream_t stream);
extern "C" hipError_t ixg_launch_metric_reduce(const ixg::MetricShared* S, const ixg_optical_pixel* d_px,
                                               ixg_metric_totals* d_totals, hipStream_t stream);
extern "C" hipError_t ixg_launch_metric_queue(const ixg::MetricShared* S, const ixg::MetricImage* d_images,
                                              const float4* d_planes, ixg_optical_pixel* d_out, uint32_t* d_counter,
                                              int block, int tile, uint32_t grid_blocks, hipStream_t stream);
extern "C" hipError_t ixg_metric_queue_occupancy(int block, int tile, int* per_cu);
extern "C" hipError_t ixg_launch_metric_flow(const ixg::MetricShared* S, const ixg::MetricImage* d_images,
                                             const float4* d_planes, ixg_optical_pixel* d_out, uint32_t* d_counter,
                                             int block, int tile, uint32_t grid_blocks, hipStream_t stream);
extern "C" hipError_t ixg_metric_flow_occupancy(int block, int* per_cu);
extern "C" int ixg_metric_queue_shapes(int* block, int* tile, int max_n);
extern "C" hipError_t ixg_launch_scoring_probe(const ixg::MetricShared* S, const float* d_dirs, uint32_t n,
                                               float* d_out, hipStream_t stream);

// libixg internals shared with api.cpp
extern "C" hipStream_t ixg_internal_stream(ixg_ctx* ctx);
extern "C" int ixg_internal_device(ixg_ctx* ctx);
extern "C" ixg_status ixg_internal_fail(ixg_ctx* ctx, ixg_status st, const char* what);
extern "C" ixg_status ixg_internal_hip_fail(ixg_ctx* ctx, hipError_t e, const char* where);
extern "C" int ixg_internal_lost(ixg_ctx* ctx);
extern "C" void** ixg_internal_metric_plan(ixg_ctx* ctx, void (*free_fn)(void*));

namespace {

using namespace ixg;

constexpr float kToRad = kPi / 180.0f;

// visibility.rs:22-40
float head_shadow_cone_cos(bool studio_rig, float deg) {
    if (studio_rig || std::isnan(deg) || deg == 16.0f) return 0.96f;
    if (deg <= 0.0f) return 2.0f;
    float offset = std::acos(0.96f) - 16.0f * kToRad;
    return std::cos(hclamp(deg, 1.0f, 89.0f) * kToRad + offset);
}

struct Resolved {
    MetricShared S;
};

ixg_status resolve_shared(ixg_ctx* ctx, const ixg_environment& env, uint32_t width, uint32_t height,
                          uint32_t samples, MetricShared* S) {
    std::memset(S, 0, sizeof *S);
    ixg_environment scoring = env;
    scoring.exposure = 1.0f;
    scoring.head_shadow_deg = 16.0f;
    scoring.backdrop = 0.0f;  // the tray is not part of the scoring radiance (lighting.rs:148-152)
    if (!resolve_environment(scoring, &S->env))
        return ixg_internal_fail(ctx, IXG_DECLINED, "environment has no device lighting model");
    PresetInfo pi;
    preset_info(env.preset, &pi);
    // studio_spec_power (rig.rs:579-585): the probe wavelength itself for ASET
    S->spec550 = pi.model == ENV_ASET ? 550.0f : pi.d65 ? d65_host(550.0f) : blackbody(550.0f, pi.temp_k);
    S->head_cos = head_shadow_cone_cos(pi.model == ENV_STUDIO, env.head_shadow_deg);
    S->width = width;
    S->height = height;
    S->samples = samples;
    // ambient level: the device radiance at 96 hemisphere probes, median
    std::vector<float> dirs(96 * 3);
    for (int i = 0; i < 96; ++i) {
        float y = ((float)i + 0.5f) / 96.0f;
        float ring = std::sqrt(std::fmax(std::fma(y, -y, 1.0f), 0.0f));
        float a = (float)i * 2.3999632f;
        dirs[3 * i] = ring * std::cos(a);
        dirs[3 * i + 1] = y;
        dirs[3 * i + 2] = ring * std::sin(a);
    }
    hipStream_t st = ixg_internal_stream(ctx);
    float *d_dirs = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc(&d_dirs, dirs.size() * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&d_out, 96 * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(d_dirs, dirs.data(), dirs.size() * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = ixg_launch_scoring_probe(S, d_dirs, 96, d_out, st);
    std::vector<float> rad(96);
    if (e == hipSuccess) e = hipMemcpyAsync(rad.data(), d_out, 96 * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_dirs) (void)hipFree(d_dirs);
    if (d_out) (void)hipFree(d_out);
    if (e != hipSuccess) return ixg_internal_hip_fail(ctx, e, "scoring probe");
    std::sort(rad.begin(), rad.end());
    bool uniform = pi.model != ENV_STUDIO;
    S->threshold = std::fmax(rad[48] * (uniform ? 0.5f : 4.0f), 1.17549435e-38f);
    return IXG_OK;
}

void fill_camera(const ixg_camera& c, MetricImage* I) {
    std::memcpy(I->origin, c.origin, sizeof I->origin);
    I->fov_tan = c.fov_tan;
    std::memcpy(I->forward, c.forward, sizeof I->forward);
    std::memcpy(I->right, c.right, sizeof I->right);
    std::memcpy(I->up, c.up, sizeof I->up);
}

void fill_temporal(const float pose[3], float fov_tan, uint32_t pw, uint32_t ph, MetricImage* I) {
    static const float offs[5] = {-3.0f, -1.5f, 0.0f, 1.5f, 3.0f};  // scintillation.rs:19
    for (int i = 0; i < 5; ++i) {
        ixg_camera c;
        ixg_camera_new(pose[0] + offs[i] * kToRad, pose[1], pose[2], 42.0f, pw, ph, 1, &c);
        std::memcpy(I->t_origin[i], c.origin, sizeof I->t_origin[i]);
        std::memcpy(I->t_forward[i], c.forward, sizeof I->t_forward[i]);
        std::memcpy(I->t_right[i], c.right, sizeof I->t_right[i]);
        std::memcpy(I->t_up[i], c.up, sizeof I->t_up[i]);
    }
    (void)fov_tan;
}

void fill_indices(const ixg_dispersion& d, MetricImage* I) {
    I->nd = std::fmax(dispersion_host(d, 589.3f), 1.1f);
    I->nf = std::fmax(dispersion_host(d, 486.1f), 1.001f);
    I->nc = std::fmax(dispersion_host(d, 656.3f), 1.001f);
    I->zero_dispersion = I->nf == I->nc;
}

bool planes_valid(const ixg_facet_plane* p, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        float s = p[i].normal[0] * p[i].normal[0] + p[i].normal[1] * p[i].normal[1] +
                  p[i].normal[2] * p[i].normal[2];
        if (!std::isfinite(p[i].d) || !std::isfinite(s) || std::fabs(s - 1.0f) > 0.01f) return false;
    }
    return true;
}

// Runs pixels + reduction for n images already described in `images`.
ixg_status run_images(ixg_ctx* ctx, const MetricShared& S, const std::vector<MetricImage>& images,
                      const ixg_facet_plane* planes, uint32_t total_planes, ixg_optical_pixel* host_px,
                      std::vector<ixg_metric_totals>& totals) {
    hipStream_t st = ixg_internal_stream(ctx);
    const size_t npx = (size_t)S.width * S.height;
    std::vector<float4> pl(total_planes);
    for (uint32_t i = 0; i < total_planes; ++i)
        pl[i] = make_float4(planes[i].normal[0], planes[i].normal[1], planes[i].normal[2], planes[i].d);
    float4* d_pl = nullptr;
    MetricImage* d_img = nullptr;
    ixg_optical_pixel* d_px = nullptr;
    ixg_metric_totals* d_tot = nullptr;
    hipError_t e = hipMalloc(&d_pl, std::max<size_t>(1, pl.size()) * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&d_img, images.size() * sizeof(MetricImage));
    if (e == hipSuccess) e = hipMalloc(&d_px, images.size() * npx * sizeof(ixg_optical_pixel));
    if (e == hipSuccess) e = hipMalloc(&d_tot, images.size() * sizeof(ixg_metric_totals));
    if (e == hipSuccess) e = hipMemcpyAsync(d_pl, pl.data(), pl.size() * sizeof(float4), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_img, images.data(), images.size() * sizeof(MetricImage), hipMemcpyHostToDevice, st);
    // gridDim.y <= 65535 images per launch
    for (size_t first = 0; e == hipSuccess && first < images.size(); first += 65535) {
        MetricShared Sc = S;
        Sc.n_images = (uint32_t)std::min<size_t>(65535, images.size() - first);
        e = ixg_launch_metric_pixels(&Sc, d_img + first, d_pl, d_px + first * npx, st);
        if (e == hipSuccess) e = ixg_launch_metric_reduce(&Sc, d_px + first * npx, d_tot + first, st);
    }
    totals.resize(images.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(totals.data(), d_tot, images.size() * sizeof(ixg_metric_totals), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && host_px)
        e = hipMemcpyAsync(host_px, d_px, images.size() * npx * sizeof(ixg_optical_pixel), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_pl) (void)hipFree(d_pl);
    if (d_img) (void)hipFree(d_img);
    if (d_px) (void)hipFree(d_px);
    if (d_tot) (void)hipFree(d_tot);
    if (e != hipSuccess) return ixg_internal_hip_fail(ctx, e, "metric images");
    return IXG_OK;
}

// ---------------------------------------------------------------------------
// The sweep plan: device buffers that persist across calls (grown on demand),
// the launch shape chosen by the autotuner for each image size, and the
// captured graph of [counter fill, pixel kernel, reduction] that later calls
// with the same shape replay after copying in their own designs.
// ---------------------------------------------------------------------------
struct Shape {
    int block = 0;  // -1: 2-D grid kernel; > 0: the queue kernel's lanes per block;
                    // -64 / -128 / -256: the flow kernel with that many lanes per block
    int tile = 0;   // pixels per queue item
    uint32_t grid = 0;
};

bool is_flow(int block) { return block == -64 || block == -128 || block == -256; }

// The flow kernel's pixels per item when the shape leaves it open (tile 0):
// whole images while there are at least two items per resident block, else
// images cut into equal parts until there are (at least 64 pixels each).
int flow_tile(int tile, size_t n_images, uint32_t npx, uint32_t grid) {
    if (tile > 0) return tile;
    const size_t want = 2 * (size_t)std::max<uint32_t>(grid, 1);
    const size_t parts = std::max<size_t>(1, (want + n_images - 1) / std::max<size_t>(n_images, 1));
    const size_t t = (npx + parts - 1) / parts;
    return (int)std::min<size_t>(npx, std::max<size_t>(64, t));
}

// Resident blocks per CU of a queue or flow shape.
hipError_t shape_occupancy(int block, int tile, int* per_cu) {
    return is_flow(block) ? ixg_metric_flow_occupancy(-block, per_cu) : ixg_metric_queue_occupancy(block, tile, per_cu);
}

struct SweepPlan {
    int cus = 0;
    int use_graph = 1;
    Shape forced;  // block 0 = autotune
    float4* d_pl = nullptr;
    size_t cap_pl = 0;
    MetricImage* d_img = nullptr;
    size_t cap_img = 0;
    ixg_optical_pixel* d_px = nullptr;
    size_t cap_px = 0;
    ixg_metric_totals* d_tot = nullptr;
    size_t cap_tot = 0;
    uint32_t* d_counter = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint8_t> graph_key;
    std::map<std::array<uint32_t, 4>, Shape> tuned;
    Shape last;
    int last_graphed = 0;
    std::vector<float> tune_ms;  // per candidate, last autotune
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the sweep's launches
    double last_ms = 0.0;                     // their device time, last sweep
};

void drop_graph(SweepPlan* P) {
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    P->exec = nullptr;
    P->graph = nullptr;
    P->graph_key.clear();
}

void plan_free(void* p) {
    SweepPlan* P = static_cast<SweepPlan*>(p);
    drop_graph(P);
    void* bufs[] = {P->d_pl, P->d_img, P->d_px, P->d_tot, P->d_counter};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (P->ev0) (void)hipEventDestroy(P->ev0);
    if (P->ev1) (void)hipEventDestroy(P->ev1);
    delete P;
}

SweepPlan* get_plan(ixg_ctx* ctx) {
    void** slot = ixg_internal_metric_plan(ctx, plan_free);
    if (!*slot) {
        SweepPlan* P = new SweepPlan();
        if (hipDeviceGetAttribute(&P->cus, hipDeviceAttributeMultiprocessorCount, ixg_internal_device(ctx)) !=
            hipSuccess)
            P->cus = 1;
        *slot = P;
    }
    return static_cast<SweepPlan*>(*slot);
}

template <class T>
hipError_t grow(T** buf, size_t* cap, size_t n, SweepPlan* P) {
    if (n <= *cap) return hipSuccess;
    drop_graph(P);  // the graph holds the old addresses
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(buf, n * sizeof(T));
    if (e == hipSuccess) *cap = n;
    return e;
}

// Pixel kernel(s) + reduction for n images already in the plan's buffers.
hipError_t enqueue_sweep(SweepPlan* P, const MetricShared& S, size_t n_images, const Shape& sh, hipStream_t st) {
    const size_t npx = (size_t)S.width * S.height;
    hipError_t e = hipSuccess;
    if (sh.block == -1) {
        for (size_t first = 0; e == hipSuccess && first < n_images; first += 65535) {
            MetricShared Sc = S;
            Sc.n_images = (uint32_t)std::min<size_t>(65535, n_images - first);
            e = ixg_launch_metric_pixels(&Sc, P->d_img + first, P->d_pl, P->d_px + first * npx, st);
        }
    } else {
        MetricShared Sc = S;
        Sc.n_images = (uint32_t)n_images;
        e = is_flow(sh.block)
                ? ixg_launch_metric_flow(&Sc, P->d_img, P->d_pl, P->d_px, P->d_counter, -sh.block,
                                         flow_tile(sh.tile, n_images, (uint32_t)npx, sh.grid), sh.grid, st)
                : ixg_launch_metric_queue(&Sc, P->d_img, P->d_pl, P->d_px, P->d_counter, sh.block, sh.tile, sh.grid,
                                          st);
    }
    for (size_t first = 0; e == hipSuccess && first < n_images; first += 65535) {
        MetricShared Sc = S;
        Sc.n_images = (uint32_t)std::min<size_t>(65535, n_images - first);
        e = ixg_launch_metric_reduce(&Sc, P->d_px + first * npx, P->d_tot + first, st);
    }
    return e;
}

// Times the candidate shapes at full residency on the first images of this
// call and keeps the fastest for this image size and call-size class: the
// flow kernel at 128 and 256 lanes per block (items sized per call) and
// three work-queue shapes.  The sample is about 2^26 sub-pixel rays' worth,
// at least 16 images, and enough images to give every resident flow lane 16
// pixels (or the whole call), so the flow kernel's items are as large as in
// a full call and its measurement is not dominated by the end of its items.
// A one-image pass over every candidate first loads each kernel's code, so
// the timed passes measure only the work.  (The flow kernel at 64 lanes
// wins only on samples cut into many small items; on whole images it ran
// 20-25 % behind, so it is not a candidate.)
hipError_t autotune(SweepPlan* P, const MetricShared& S, size_t n_images, hipStream_t st, Shape* best) {
    const size_t npx = std::max<size_t>(1, (size_t)S.width * S.height);
    const size_t rays_per_image = npx * std::max<uint32_t>(S.samples, 1);
    std::vector<Shape> cands;
    size_t flow_lanes = 0;
    for (int lanes : {128, 256}) {
        int per_cu = 0;
        if (ixg_metric_flow_occupancy(lanes, &per_cu) != hipSuccess || per_cu <= 0) continue;
        cands.push_back(Shape{-lanes, 0, (uint32_t)(per_cu * P->cus)});
        flow_lanes = std::max<size_t>(flow_lanes, (size_t)per_cu * P->cus * lanes);
    }
    const int queue[][2] = {{64, 64}, {128, 128}, {256, 1024}};
    for (const auto& q : queue) {
        int per_cu = 0;
        if (ixg_metric_queue_occupancy(q[0], q[1], &per_cu) != hipSuccess || per_cu <= 0) continue;
        cands.push_back(Shape{q[0], q[1], (uint32_t)(per_cu * P->cus)});
    }
    const size_t n = std::min<size_t>(
        n_images, std::max<size_t>({16, ((size_t)1 << 26) / rays_per_image, (16 * flow_lanes + npx - 1) / npx}));
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    float best_ms = 1e30f;
    P->tune_ms.assign(cands.size(), -1.0f);
    for (size_t c = 0; e == hipSuccess && c < cands.size(); ++c) e = enqueue_sweep(P, S, 1, cands[c], st);
    for (size_t c = 0; e == hipSuccess && c < cands.size(); ++c) {
        e = hipEventRecord(a, st);
        if (e == hipSuccess) e = enqueue_sweep(P, S, n, cands[c], st);
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e == hipSuccess) {
            P->tune_ms[c] = ms;
            if (ms < best_ms) {
                best_ms = ms;
                *best = cands[c];
            }
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    return e;
}

ixg_status run_sweep(ixg_ctx* ctx, const MetricShared& S, const std::vector<MetricImage>& images,
                     const ixg_facet_plane* planes, uint32_t total_planes, std::vector<ixg_metric_totals>& totals) {
    hipStream_t st = ixg_internal_stream(ctx);
    SweepPlan* P = get_plan(ctx);
    const size_t npx = (size_t)S.width * S.height;
    const size_t n = images.size();
    std::vector<float4> pl(total_planes);
    for (uint32_t i = 0; i < total_planes; ++i)
        pl[i] = make_float4(planes[i].normal[0], planes[i].normal[1], planes[i].normal[2], planes[i].d);
    hipError_t e = grow(&P->d_pl, &P->cap_pl, std::max<size_t>(1, pl.size()), P);
    if (e == hipSuccess) e = grow(&P->d_img, &P->cap_img, n, P);
    if (e == hipSuccess) e = grow(&P->d_px, &P->cap_px, n * npx, P);
    if (e == hipSuccess) e = grow(&P->d_tot, &P->cap_tot, n, P);
    if (e == hipSuccess && !P->d_counter) e = hipMalloc(&P->d_counter, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpyAsync(P->d_pl, pl.data(), pl.size() * sizeof(float4), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(P->d_img, images.data(), n * sizeof(MetricImage), hipMemcpyHostToDevice, st);
    Shape sh;
    if (e == hipSuccess) {
        if (P->forced.block != 0) {
            sh = P->forced;
            if (sh.block != -1) {
                int per_cu = 0;
                e = shape_occupancy(sh.block, sh.tile, &per_cu);
                sh.grid = (uint32_t)std::max(1, per_cu) * (uint32_t)P->cus;
            }
        } else {
            // the image size, and whether the call is large enough for the flow kernel alone
            int per_cu = 0;
            const size_t lanes = ixg_metric_flow_occupancy(128, &per_cu) == hipSuccess
                                     ? (size_t)std::max(per_cu, 1) * P->cus * 128 : 1;
            const uint32_t large = n * S.width * S.height >= 8 * lanes ? 1u : 0u;
            const std::array<uint32_t, 4> key{S.width, S.height, S.samples, large};
            auto it = P->tuned.find(key);
            if (it != P->tuned.end()) {
                sh = it->second;
            } else {
                e = autotune(P, S, n, st, &sh);
                if (e == hipSuccess) P->tuned[key] = sh;
            }
        }
    }
    if (e == hipSuccess && !P->ev0) e = hipEventCreate(&P->ev0);
    if (e == hipSuccess && !P->ev1) e = hipEventCreate(&P->ev1);
    if (e == hipSuccess) e = hipEventRecord(P->ev0, st);
    if (e == hipSuccess) {
        if (P->use_graph) {
            // the graph bakes in kernel arguments: shared constants, image
            // count, shape and buffer addresses
            std::vector<uint8_t> key(sizeof(MetricShared) + sizeof(size_t) + sizeof(Shape) + 4 * sizeof(void*));
            uint8_t* k = key.data();
            std::memcpy(k, &S, sizeof S);
            k += sizeof S;
            std::memcpy(k, &n, sizeof n);
            k += sizeof n;
            std::memcpy(k, &sh, sizeof sh);
            k += sizeof sh;
            const void* ptrs[4] = {P->d_pl, P->d_img, P->d_px, P->d_tot};
            std::memcpy(k, ptrs, sizeof ptrs);
            if (key != P->graph_key || !P->exec) {
                drop_graph(P);
                e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    hipError_t le = enqueue_sweep(P, S, n, sh, st);
                    hipGraph_t g = nullptr;
                    e = hipStreamEndCapture(st, &g);
                    if (e == hipSuccess) e = le;
                    if (e == hipSuccess) {
                        P->graph = g;
                        e = hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0);
                    } else if (g) {
                        (void)hipGraphDestroy(g);
                    }
                }
                if (e == hipSuccess) P->graph_key = key;
            }
            if (e == hipSuccess) e = hipGraphLaunch(P->exec, st);
        } else {
            e = enqueue_sweep(P, S, n, sh, st);
        }
    }
    if (e == hipSuccess) e = hipEventRecord(P->ev1, st);
    totals.resize(n);
    if (e == hipSuccess)
        e = hipMemcpyAsync(totals.data(), P->d_tot, n * sizeof(ixg_metric_totals), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, P->ev0, P->ev1);
    if (e == hipSuccess) P->last_ms = ms;
    if (e != hipSuccess) {
        drop_graph(P);
        return ixg_internal_hip_fail(ctx, e, "metric sweep");
    }
    P->last = sh;
    if (is_flow(sh.block)) P->last.tile = flow_tile(sh.tile, n, (uint32_t)npx, sh.grid);
    P->last_graphed = P->use_graph;
    return IXG_OK;
}

}  // namespace

extern "C" ixg_status ixg_metric_image(ixg_ctx* ctx, const ixg_metric_request* r, ixg_optical_pixel* pixels,
                                       ixg_metric_totals* totals) {
    if (!ctx || !r || !r->planes || !r->material) return IXG_INVALID;
    if (ixg_internal_lost(ctx)) return ixg_internal_fail(ctx, IXG_DEVICE_LOST, "context is lost");
    const uint64_t count = (uint64_t)r->width * r->height;
    if (count == 0 || count > 1048576ull || r->samples == 0 || r->samples > 64)
        return ixg_internal_fail(ctx, IXG_RESOURCE_LIMIT, "pixels or samples out of range");
    if (r->want_scintillation && !(r->has_pose && std::isfinite(r->pose[0]) && std::isfinite(r->pose[1]) &&
                                   std::isfinite(r->pose[2]) && r->pose[2] > 0.0f))
        return ixg_internal_fail(ctx, IXG_INVALID, "scintillation needs a camera pose");
    if (r->n_planes == 0) return ixg_internal_fail(ctx, IXG_INVALID, "empty geometry");
    if (r->n_planes > IXG_MAX_PLANES) return ixg_internal_fail(ctx, IXG_DECLINED, "too many planes");
    if (!planes_valid(r->planes, r->n_planes)) return ixg_internal_fail(ctx, IXG_INVALID, "non-unit plane");
    if (hipSetDevice(ixg_internal_device(ctx)) != hipSuccess) return IXG_DEVICE_LOST;
    MetricShared S;
    ixg_status stt = resolve_shared(ctx, r->environment, r->width, r->height, r->samples, &S);
    if (stt != IXG_OK) return stt;
    S.proj_w = r->projection[0] ? r->projection[0] : r->width;
    S.proj_h = r->projection[1] ? r->projection[1] : r->height;
    S.origin_x = r->has_tile_origin ? r->tile_origin[0] : 0;
    S.origin_y = r->has_tile_origin ? r->tile_origin[1] : 0;
    S.scale_x = r->has_tile_origin ? 1.0f : (float)S.proj_w / (float)r->width;
    S.scale_y = r->has_tile_origin ? 1.0f : (float)S.proj_h / (float)r->height;
    S.want_fire = r->want_fire ? 1 : 0;
    S.want_scintillation = r->want_scintillation ? 1 : 0;
    S.n_images = 1;
    std::vector<MetricImage> images(1);
    MetricImage& I = images[0];
    std::memset(&I, 0, sizeof I);
    fill_camera(r->camera, &I);
    if (r->want_scintillation) fill_temporal(r->pose, r->camera.fov_tan, S.proj_w, S.proj_h, &I);
    I.plane_offset = 0;
    I.n_planes = r->n_planes;
    fill_indices(r->material->dispersion, &I);
    std::vector<ixg_metric_totals> tot;
    stt = run_images(ctx, S, images, r->planes, r->n_planes, pixels, tot);
    if (stt != IXG_OK) return stt;
    if (totals) *totals = tot[0];
    return IXG_OK;
}

extern "C" ixg_status ixg_metric_sweep(ixg_ctx* ctx, const ixg_facet_plane* planes, uint32_t total_planes,
                                       const ixg_design* designs, uint32_t n_designs, const float* tilts_deg,
                                       uint32_t n_tilts, uint32_t width, uint32_t height, uint32_t samples,
                                       float distance, const ixg_environment* env, float* out_metrics) {
    if (!ctx || !planes || !designs || !tilts_deg || !env || !out_metrics) return IXG_INVALID;
    if (ixg_internal_lost(ctx)) return ixg_internal_fail(ctx, IXG_DEVICE_LOST, "context is lost");
    const uint64_t count = (uint64_t)width * height;
    if (count == 0 || count > 1048576ull || samples == 0 || samples > 64 || n_designs == 0 || n_tilts == 0)
        return ixg_internal_fail(ctx, IXG_RESOURCE_LIMIT, "sweep shape out of range");
    if ((uint64_t)n_designs * n_tilts * count * sizeof(ixg_optical_pixel) > (64ull << 30))
        return ixg_internal_fail(ctx, IXG_RESOURCE_LIMIT, "sweep pixel buffer above 64 GiB");
    for (uint32_t d = 0; d < n_designs; ++d) {
        if (designs[d].n_planes == 0 || designs[d].n_planes > IXG_MAX_PLANES ||
            (uint64_t)designs[d].plane_offset + designs[d].n_planes > total_planes)
            return ixg_internal_fail(ctx, IXG_INVALID, "design plane range");
    }
    if (!planes_valid(planes, total_planes)) return ixg_internal_fail(ctx, IXG_INVALID, "non-unit plane");
    if (hipSetDevice(ixg_internal_device(ctx)) != hipSuccess) return IXG_DEVICE_LOST;
    MetricShared S;
    ixg_status stt = resolve_shared(ctx, *env, width, height, samples, &S);
    if (stt != IXG_OK) return stt;
    S.proj_w = width;
    S.proj_h = height;
    S.scale_x = S.scale_y = 1.0f;
    S.want_fire = 1;
    S.want_scintillation = 1;
    std::vector<MetricImage> images((size_t)n_designs * n_tilts);
    // camera per tilt: yaw 0 or 180 deg by sign, pitch 90 - |tilt| (pose_batch.rs:118-131)
    std::vector<MetricImage> per_tilt(n_tilts);
    for (uint32_t t = 0; t < n_tilts; ++t) {
        MetricImage& T = per_tilt[t];
        std::memset(&T, 0, sizeof T);
        float tilt = tilts_deg[t];
        float yaw = tilt < 0.0f ? 180.0f * kToRad : 0.0f;
        float pitch = (90.0f - std::fabs(tilt)) * kToRad;
        ixg_camera c;
        ixg_camera_new(yaw, pitch, distance, 42.0f, width, height, samples, &c);
        fill_camera(c, &T);
        float pose[3] = {yaw, pitch, distance};
        fill_temporal(pose, c.fov_tan, width, height, &T);
    }
    for (uint32_t d = 0; d < n_designs; ++d) {
        MetricImage base;
        std::memset(&base, 0, sizeof base);
        fill_indices(designs[d].dispersion, &base);
        for (uint32_t t = 0; t < n_tilts; ++t) {
            MetricImage I = per_tilt[t];
            I.plane_offset = designs[d].plane_offset;
            I.n_planes = designs[d].n_planes;
            I.nd = base.nd;
            I.nf = base.nf;
            I.nc = base.nc;
            I.zero_dispersion = base.zero_dispersion;
            images[(size_t)d * n_tilts + t] = I;
        }
    }
    std::vector<ixg_metric_totals> tot;
    stt = run_sweep(ctx, S, images, planes, total_planes, tot);
    if (stt != IXG_OK) return stt;
    for (size_t i = 0; i < tot.size(); ++i) {
        out_metrics[5 * i + 0] = tot[i].brilliance_pct;
        out_metrics[5 * i + 1] = tot[i].fire_index;
        out_metrics[5 * i + 2] = tot[i].scintillation_pct;
        out_metrics[5 * i + 3] = tot[i].windowing_pct;
        out_metrics[5 * i + 4] = tot[i].extinction_pct;
    }
    return IXG_OK;
}

extern "C" double ixg_metric_last_ms(ixg_ctx* ctx) {
    if (!ctx) return 0.0;
    return get_plan(ctx)->last_ms;
}

extern "C" ixg_status ixg_set_metric_mode(ixg_ctx* ctx, int use_graph, int block, int tile) {
    if (!ctx) return IXG_INVALID;
    if (block > 0) {
        int per_cu = 0;
        if (ixg_metric_queue_occupancy(block, tile, &per_cu) != hipSuccess)
            return ixg_internal_fail(ctx, IXG_INVALID, "unknown queue shape");
    } else if (is_flow(block)) {
        if (tile < 0) return ixg_internal_fail(ctx, IXG_INVALID, "flow tile must be >= 0 (0 = whole image)");
    } else if (block < -1) {
        return ixg_internal_fail(ctx, IXG_INVALID,
                                 "block must be -1 (grid), 0 (autotune), a queue shape or -64/-128/-256 (flow)");
    }
    SweepPlan* P = get_plan(ctx);
    P->use_graph = use_graph ? 1 : 0;
    P->forced = Shape{block, block != -1 ? tile : 0, 0};
    return IXG_OK;
}

extern "C" ixg_status ixg_metric_tuning(ixg_ctx* ctx, int* block, int* tile, uint32_t* grid_blocks, int* graphed) {
    if (!ctx) return IXG_INVALID;
    SweepPlan* P = get_plan(ctx);
    if (block) *block = P->last.block;
    if (tile) *tile = P->last.tile;
    if (grid_blocks) *grid_blocks = P->last.grid;
    if (graphed) *graphed = P->last_graphed;
    return IXG_OK;
}

// ---------------------------------------------------------------------------
// Fan-grid pose metrics (evaluate_gem_optical_metrics, evaluate.rs:158-439)
// and the tilt-axis sweep built on them (pose_batch.rs:73-199).
// ---------------------------------------------------------------------------
extern "C" void ixg_internal_half_widths(const ixg_facet_plane* planes, const ixg_design* designs, uint32_t n,
                                         float* out);
extern "C" hipError_t ixg_launch_pose(const ixg::MetricShared* S, const ixg::PoseDesign* d_designs,
                                      const ixg::PoseJob* d_jobs, uint32_t n_jobs, const float4* d_planes,
                                      const float2* d_cells, uint32_t n_cells, int want_all, float* d_out,
                                      hipStream_t stream);

namespace {

// run_grid's disc cells in (ix, iz) order (evaluate.rs:351-358, fan.rs:22).
std::vector<float2> fan_cells() {
    std::vector<float2> cells;
    for (int ix = 0; ix < 18; ++ix)
        for (int iz = 0; iz < 18; ++iz) {
            const float u = std::fma(((float)ix + 0.5f) / 18.0f, 2.0f, -1.0f);
            const float v = std::fma(((float)iz + 0.5f) / 18.0f, 2.0f, -1.0f);
            if (std::fma(v, v, u * u) > 0.70f) continue;
            cells.push_back(make_float2(u, v));
        }
    return cells;
}

// camera_view_basis (camera.rs:16-21): Camera::new's frame at distance 1.
void view_basis(float yaw, float pitch, float (*b)[3]) {
    ixg_camera c;
    ixg_camera_new(yaw, pitch, 1.0f, 42.0f, 1, 1, 1, &c);
    std::memcpy(b[0], c.forward, sizeof b[0]);
    std::memcpy(b[1], c.right, sizeof b[1]);
    std::memcpy(b[2], c.up, sizeof b[2]);
}

ixg_status run_poses(ixg_ctx* ctx, const ixg_facet_plane* planes, uint32_t total_planes, const ixg_design* designs,
                     uint32_t n_designs, const float* poses, uint32_t n_poses, const ixg_environment& env,
                     int want_all, float* out) {
    if ((uint64_t)n_designs * n_poses > (1ull << 24))
        return ixg_internal_fail(ctx, IXG_RESOURCE_LIMIT, "more than 2^24 (design, pose) jobs");
    for (uint32_t d = 0; d < n_designs; ++d)
        if (designs[d].n_planes > IXG_MAX_PLANES ||
            (uint64_t)designs[d].plane_offset + designs[d].n_planes > total_planes)
            return ixg_internal_fail(ctx, IXG_INVALID, "design plane range");
    if (!planes_valid(planes, total_planes)) return ixg_internal_fail(ctx, IXG_INVALID, "non-unit plane");
    for (uint32_t p = 0; p < 2 * n_poses; ++p)
        if (!std::isfinite(poses[p])) return ixg_internal_fail(ctx, IXG_INVALID, "non-finite pose");
    if (hipSetDevice(ixg_internal_device(ctx)) != hipSuccess) return IXG_DEVICE_LOST;
    MetricShared S;
    ixg_status stt = resolve_shared(ctx, env, 0, 0, 0, &S);
    if (stt != IXG_OK) return stt;

    // PreparedEval per design (prepared.rs:64-94): indices and fan scale
    std::vector<float> half(n_designs);
    ixg_internal_half_widths(planes, designs, n_designs, half.data());
    std::vector<PoseDesign> pd(n_designs);
    for (uint32_t d = 0; d < n_designs; ++d) {
        MetricImage idx;
        fill_indices(designs[d].dispersion, &idx);
        PoseDesign& P = pd[d];
        std::memset(&P, 0, sizeof P);
        P.plane_offset = designs[d].plane_offset;
        P.n_planes = designs[d].n_planes;
        P.nd = idx.nd;
        P.nf = idx.nf;
        P.nc = idx.nc;
        P.standoff = 2.5f * half[d];                            // fan.rs:28,77
        P.lateral = 0.95f * half[d] / std::sqrt(0.70f);         // fan.rs:25,78
    }
    // the frames of each pose, shared by every design (scintillation.rs:47-60)
    static const float offs[5] = {-3.0f, -1.5f, 0.0f, 1.5f, 3.0f};
    std::vector<PoseJob> per_pose(n_poses);
    for (uint32_t p = 0; p < n_poses; ++p) {
        PoseJob& J = per_pose[p];
        std::memset(&J, 0, sizeof J);
        const float yaw = poses[2 * p], pitch = poses[2 * p + 1];
        view_basis(yaw, pitch, J.basis[0]);
        for (int k = 0; k < 5; ++k) view_basis(yaw + offs[k] * kToRad, pitch, J.basis[1 + k]);
    }
    const size_t n_jobs = (size_t)n_designs * n_poses;
    std::vector<PoseJob> jobs(n_jobs);
    for (uint32_t d = 0; d < n_designs; ++d)
        for (uint32_t p = 0; p < n_poses; ++p) {
            PoseJob J = per_pose[p];
            J.design = d;
            jobs[(size_t)d * n_poses + p] = J;
        }
    const std::vector<float2> cells = fan_cells();
    std::vector<float4> pl(std::max<uint32_t>(1, total_planes));
    for (uint32_t i = 0; i < total_planes; ++i)
        pl[i] = make_float4(planes[i].normal[0], planes[i].normal[1], planes[i].normal[2], planes[i].d);

    hipStream_t st = ixg_internal_stream(ctx);
    float4* d_pl = nullptr;
    PoseDesign* d_des = nullptr;
    PoseJob* d_jobs = nullptr;
    float2* d_cells = nullptr;
    float* d_out = nullptr;
    hipError_t e = hipMalloc(&d_pl, pl.size() * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&d_des, pd.size() * sizeof(PoseDesign));
    if (e == hipSuccess) e = hipMalloc(&d_jobs, n_jobs * sizeof(PoseJob));
    if (e == hipSuccess) e = hipMalloc(&d_cells, cells.size() * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&d_out, n_jobs * 5 * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(d_pl, pl.data(), pl.size() * sizeof(float4), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_des, pd.data(), pd.size() * sizeof(PoseDesign), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_jobs, jobs.data(), n_jobs * sizeof(PoseJob), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_cells, cells.data(), cells.size() * sizeof(float2), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = ixg_launch_pose(&S, d_des, d_jobs, (uint32_t)n_jobs, d_pl, d_cells, (uint32_t)cells.size(), want_all,
                            d_out, st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, n_jobs * 5 * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    void* bufs[] = {d_pl, d_des, d_jobs, d_cells, d_out};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (e != hipSuccess) return ixg_internal_hip_fail(ctx, e, "pose metrics");
    return IXG_OK;
}

// tilt_slot (pose_batch.rs:133-143): where job k of an axis lands in its
// 181-point tilt curve.
uint32_t tilt_slot(uint32_t job) {
    const uint32_t half = 90;
    if (job == 0) return half;
    if (job <= half) return 2 * half - (job - 1);
    return job - 1 - half;
}

}  // namespace

extern "C" ixg_status ixg_pose_metrics(ixg_ctx* ctx, const ixg_facet_plane* planes, uint32_t total_planes,
                                       const ixg_design* designs, uint32_t n_designs, const float* poses,
                                       uint32_t n_poses, const ixg_environment* env, int want_all,
                                       float* out_metrics) {
    if (!ctx || !designs || !poses || !env || !out_metrics || (!planes && total_planes)) return IXG_INVALID;
    if (ixg_internal_lost(ctx)) return IXG_DEVICE_LOST;
    if (n_designs == 0 || n_poses == 0) return IXG_OK;
    return run_poses(ctx, planes, total_planes, designs, n_designs, poses, n_poses, *env, want_all ? 1 : 0,
                     out_metrics);
}

extern "C" ixg_status ixg_axis_profiles(ixg_ctx* ctx, const ixg_facet_plane* planes, uint32_t total_planes,
                                        const ixg_design* designs, uint32_t n_designs, const float* azimuths_deg,
                                        uint32_t n_azimuths, const ixg_environment* env, float* out_curves) {
    if (!ctx || !designs || !azimuths_deg || !env || !out_curves || (!planes && total_planes)) return IXG_INVALID;
    if (ixg_internal_lost(ctx)) return IXG_DEVICE_LOST;
    if (n_designs == 0 || n_azimuths == 0) return IXG_OK;
    constexpr uint32_t kPer = 181;  // EVALUATIONS_PER_AXIS (profile.rs:369)
    // axis_poses (pose_batch.rs:118-131): the table-up pole, then pitch
    // 0..89 deg at the positive azimuth, then at the positive + 180 deg
    std::vector<float> poses((size_t)n_azimuths * kPer * 2);
    for (uint32_t a = 0; a < n_azimuths; ++a) {
        const float pos = azimuths_deg[a] * kToRad;
        const float neg = (azimuths_deg[a] + 180.0f) * kToRad;
        float* P = poses.data() + (size_t)a * kPer * 2;
        P[0] = pos;
        P[1] = 90.0f * kToRad;
        for (uint32_t i = 0; i < 90; ++i) {
            const float pitch = (float)i * kToRad;
            P[2 * (1 + i)] = pos;
            P[2 * (1 + i) + 1] = pitch;
            P[2 * (91 + i)] = neg;
            P[2 * (91 + i) + 1] = pitch;
        }
    }
    const uint32_t n_poses = n_azimuths * kPer;
    std::vector<float> m((size_t)n_designs * n_poses * 5);
    ixg_status st = run_poses(ctx, planes, total_planes, designs, n_designs, poses.data(), n_poses, *env, 0, m.data());
    if (st != IXG_OK) return st;
    // assemble_axis (pose_batch.rs:146-156): [brilliance, extinction, windowing]
    for (uint32_t d = 0; d < n_designs; ++d)
        for (uint32_t a = 0; a < n_azimuths; ++a)
            for (uint32_t j = 0; j < kPer; ++j) {
                const float* r = m.data() + ((size_t)d * n_poses + (size_t)a * kPer + j) * 5;
                float* curves = out_curves + ((size_t)d * n_azimuths + a) * 3 * kPer;
                const uint32_t slot = tilt_slot(j);
                curves[slot] = r[0];
                curves[kPer + slot] = r[4];
                curves[2 * kPer + slot] = r[3];
            }
    return IXG_OK;
}
