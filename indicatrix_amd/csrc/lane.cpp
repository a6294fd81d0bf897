// lane.cpp -- the device as one dispatch lane: a chunk of absolute samples
// traced in adaptive sub-batches, cancellable between them, returning the
// summed radiance of the traced prefix.
//
// WorkerLane::render_chunk (indicatrix-dispatch/src/lane.rs:113-142) as the
// coordinator's own lane runs it (OwnLane, apps/indicatrix-worker/src/
// coordinator/lanes/mod.rs:536-570, through stream_emit/tracer.rs:328
// trace_range): sub-batches sized toward 100 ms each (stream_emit/sizing.rs
// next_batch_size: scaled by target / elapsed, at most x4 per step, at most
// 2048 samples), each one added only once it has completed, so the result is
// always a prefix [first_sample, first_sample + done) of the range.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/ixg.h"

namespace {

constexpr double kTargetSubbatchS = 0.100;  // sizing.rs TARGET_SUBBATCH
constexpr uint32_t kMaxSubbatch = 2048;     // sizing.rs MAX_SUBBATCH

uint32_t clamp_u32(uint64_t v, uint32_t lo, uint32_t hi) {
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(v, lo), hi);
}

}  // namespace

extern "C" uint32_t ixg_next_subbatch(uint32_t prev, double elapsed_s) {
    const uint32_t grown = clamp_u32((uint64_t)prev * 4u, 1u, kMaxSubbatch);
    if (!(elapsed_s > 0.0)) return grown;
    const double scaled = (double)prev * (kTargetSubbatchS / elapsed_s);
    return (uint32_t)std::llround(std::min(std::max(scaled, 1.0), (double)grown));
}

extern "C" ixg_status ixg_render_chunk(ixg_ctx* ctx, const ixg_scene* scene, uint32_t first_sample,
                                       uint32_t samples, uint32_t seed_subbatch, float* sum_xyz, uint32_t* done,
                                       double* rate, uint32_t* subbatch_out, const volatile uint8_t* cancel) {
    if (!ctx || !scene || !sum_xyz || !done) return IXG_INVALID;
    if (scene->width == 0 || scene->height == 0) return IXG_INVALID;
    if ((uint64_t)first_sample + samples > 0xFFFFFFFFull) return IXG_INVALID;
    const size_t n = (size_t)scene->width * scene->height * 3;
    std::vector<double> total(n, 0.0);
    *done = 0;
    if (rate) *rate = 0.0;
    std::memset(sum_xyz, 0, n * sizeof(float));
    uint32_t batch = std::min(std::max(seed_subbatch, 1u), kMaxSubbatch);
    uint32_t produced = 0;
    uint64_t steady_samples = 0;
    double steady_s = 0.0;
    int batches = 0;
    ixg_status st = IXG_OK;
    while (produced < samples) {
        if (cancel && *cancel) {
            st = IXG_CANCELLED;
            break;
        }
        const uint32_t this_batch = std::min(batch, samples - produced);
        const auto t0 = std::chrono::steady_clock::now();
        // ADDs into total only when the whole sub-batch completed
        st = ixg_accumulate(ctx, scene, first_sample + produced, this_batch, total.data(), cancel);
        if (st != IXG_OK) break;
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        produced += this_batch;
        if (batches++ > 0) {  // the first sub-batch carries the scene upload and warm-up
            steady_samples += this_batch;
            steady_s += elapsed;
        }
        batch = ixg_next_subbatch(batch, elapsed);
    }
    for (size_t i = 0; i < n; ++i) sum_xyz[i] = (float)total[i];
    *done = produced;
    if (rate && steady_s > 0.0) *rate = (double)steady_samples / steady_s;
    if (subbatch_out) *subbatch_out = batch;
    return produced == samples ? IXG_OK : st;
}
