"""Throughput on one MI355X: Msamples/s of 8-lambda spectral paths, and
design evaluations/s of the batched metric sweep and the surrogate-screened
Optimize.

One step = one pass of the film hot path over one frame of the workload
(every pixel x spp paths, absolute sample indices advancing each step),
accumulated in FP64 into a device-resident buffer; `value` is the film rate.
The same JSON line carries the uniaxial 1024x1024 @ 256 frame (`film_c2`),
the sweep (512 seeded designs x 21 tilts, 64x64 @ 32 samples), the surrogate
training step (`surrogate_training`) and the screened Optimize (8192
candidates per step through the surrogate, top 64 traced at 128x128 @ 64
samples x 21 tilts), each with the roofline of its dominant kernel: exact
algorithmic FLOPs (indicatrix_amd/flops.py) over that kernel's device time
from HIP events on the library's stream.  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                  [--film-only] [--copy-calibration] [--part c2|sweep|pose|mlp|optimize]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_FP32_PEAK_TFLOPS = 157.3    # FP32 vector ALU (the film and metric kernels)
MI355X_BF16_PEAK_TFLOPS = 2500.0   # dense BF16 MFMA (the surrogate GEMMs)
MI355X_HBM_PEAK_GBS = 8000.0
LIBRARY_CHUNK_PATHS = 1 << 23      # paths per film launch (the library's default chunk)
HOST_CPU_SHARE = 16                # host threads one GPU's share of the box allows
# HBM bytes per launch of each benchmark's dominant kernel: FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 in KiB (the scales a 1 GiB device copy reads), from
# rocprofv3 --pmc runs recorded with their commands under profiles/.
FILM_KERNEL_HBM_BYTES = (667.3 * 2.0 + 13433.8) * 1024           # film C1, 1 048 576 paths
FILM_C2_KERNEL_HBM_BYTES = (1033222.0 * 2.0 + 6310438.3) * 1024   # film C2, 2^23 paths (mostly private memory)
SWEEP_KERNEL_HBM_BYTES = (10527.94 * 2.0 + 2603463.28) * 1024     # the 512-design sweep launch
MLP_STEP_HBM_BYTES = (53473.5 * 2.0 + 56344.8) * 1024            # one training step's four launches


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU work of the bounded CPU-baseline sample")
    ap.add_argument("--film-only", action="store_true",
                    help="measure the film only (profiling runs)")
    ap.add_argument("--optimize-steps", type=int, default=50)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 1 GiB device copy (scales rocprof FETCH/WRITE_SIZE)")
    ap.add_argument("--mlp-steps", type=int, default=200)
    ap.add_argument("--part", choices=["c2", "sweep", "pose", "mlp", "optimize"],
                    help="run one secondary workload alone and print its figures (profiling runs)")
    return ap.parse_args()


SURROGATE_WEIGHTS = os.path.join(ROOT, "indicatrix_amd", "data", "surrogate.npz")


def host_threads():
    """Threads for the CPU baselines: every core this process may run on, up
    to one GPU's share of the host."""
    return max(1, min(len(os.sched_getaffinity(0)), HOST_CPU_SHARE))


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "share_cap": HOST_CPU_SHARE}


def film_roofline(entry, totals, kernel_ms, paths_per_launch, traffic=None):
    """Roofline of film_persistent_kernel from an exact count (flop_counts.json
    entry) scaled per path to the paths one launch processes."""
    from indicatrix_amd import flops
    dev, algo = flops.film_device_flops(entry, totals)
    per_path = dev / entry["paths"]
    launch = per_path * paths_per_launch
    achieved = launch / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "pipe": "fp32 vector ALU", "achieved": round(achieved, 3),
            "peak": MI355X_FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / MI355X_FP32_PEAK_TFLOPS, 4),
            "traffic": traffic, "kernel": "film_persistent_kernel", "kernel_ms": round(kernel_ms, 4),
            "flops_per_launch": round(launch), "flops_per_path": round(per_path, 1),
            "algorithm_flops_per_path": round(algo / entry["paths"], 1),
            "paths_per_launch": paths_per_launch,
            "flops_counted_on": f"{entry['width']}x{entry['height']}, sample indices "
                                f"{entry['sample_offset']}..{entry['sample_offset'] + entry['spp'] - 1}",
            "events": totals}


def copy_calibration():
    """A plain device copy of known size, so the HBM byte counters of a
    profiled run can be scaled (read 1 GiB, write 1 GiB per launch)."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": a.numel() * 4, "launches": 3}), flush=True)


def film_c2_rate(backend, steps):
    """C2: the 121-facet seeded oval in a uniaxial pleochroic material,
    1024x1024 @ 256 spp, depth 24 (one full frame per step)."""
    import torch
    from indicatrix_amd import flops, workloads
    sc = workloads.film_c2_scene()
    spp = workloads.C2["spp"]
    paths = sc.width * sc.height * spp
    accum = torch.zeros(sc.width * sc.height * 3, dtype=torch.float64, device="cuda")
    backend.accumulate_device(sc, 0, 4, accum.data_ptr())   # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    launch_ms = []
    for k in range(steps):
        backend.accumulate_device(sc, 4 + k * spp, spp, accum.data_ptr())
        launch_ms.append(backend.last_kernel_ms())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    entry = flops.counts()["film_c2"]
    if (entry["width"], entry["height"], entry["planes"]) != (sc.width, sc.height, sc.scene.n_planes):
        raise RuntimeError("flop_counts.json film_c2 does not describe this frame")
    totals = backend.count_events(sc, entry["sample_offset"], entry["spp"])
    px_per_launch = min(max(1, LIBRARY_CHUNK_PATHS // spp), sc.width * sc.height)
    roof = film_roofline(entry, totals, sum(launch_ms) / len(launch_ms), px_per_launch * spp,
                         round(FILM_C2_KERNEL_HBM_BYTES))
    return {"workload": "film 1024x1024 @ 256 spp, depth 24, 121-facet seeded oval (tier angles U[30,50]), "
                        "uniaxial n_e - n_o = 0.009, pleochroic Gaussian o/e bands, random optic axis",
            "steps": steps, "ms_per_step": round(dt * 1e3 / steps, 2),
            "value": round(paths * steps / dt / 1e6, 2), "unit": "Msamples/s", "roofline": roof}, sc


def film_c2_cpu_baseline(sc, target_s=8.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    threads = host_threads()
    t0 = time.perf_counter()
    oracle_lib.accumulate(sc, 0, 1, nthreads=threads)
    one = time.perf_counter() - t0
    spp = max(1, min(64, int(target_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    oracle_lib.accumulate(sc, 1, spp, nthreads=threads)
    dt = time.perf_counter() - t0
    paths = sc.width * sc.height * spp
    return {"value": round(paths / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"the C2 frame at {spp} spp (sample indices 1..{spp}): {paths} paths in {dt:.2f} s "
                      f"on {threads} threads"}


def sweep_rate(backend, n_designs=512):
    """C3: seeded designs (32..256 planes, Sellmeier-1 from n_D / Abbe) x 21
    tilts, 64x64 @ 32 samples, one ixg_metric_sweep call."""
    from indicatrix_amd import designs, flops, metrics, tracer
    planes, ds, _ = designs.sweep_batch(n_designs)
    env = tracer.LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0)
    # warm-up and autotune on a call of the same size class (at least 8 pixels per resident lane)
    metrics.metric_sweep(backend, planes, ds[:64], designs.TILTS_C3, 64, 64, 32, 2.4, env)
    t0 = time.perf_counter()
    out = metrics.metric_sweep(backend, planes, ds, designs.TILTS_C3, 64, 64, 32, 2.4, env)
    dt = time.perf_counter() - t0
    dev_ms = metrics.metric_last_ms(backend)
    block, tile, grid, graphed = metrics.metric_tuning(backend)
    result = {"designs": n_designs, "tilts": len(designs.TILTS_C3), "seconds": round(dt, 3),
              "design_evals_per_s": round(n_designs / dt, 2),
              "mean_metrics": [round(float(v), 4) for v in out.mean(axis=0).mean(axis=0)],
              "shape": {"block": block, "tile": tile, "grid_blocks": grid, "graphed": graphed}}
    cnt = flops.counts()["sweep_c3"]
    if n_designs == cnt["n_designs"] and [d[1] for d in ds] == cnt["design_planes"]:
        fl = flops.sweep_flops()
        achieved = fl / (dev_ms * 1e-3) / 1e12
        result["roofline"] = {
            "bound": "valu", "pipe": "fp32 vector ALU", "achieved": round(achieved, 3),
            "peak": MI355X_FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / MI355X_FP32_PEAK_TFLOPS, 4),
            "traffic": round(SWEEP_KERNEL_HBM_BYTES), "kernel": (f"metric_queue_kernel<{block},{tile}>" if block > 0 else
                                        f"metric_flow_kernel<{-block}> ({tile} px per item)" if block < -1 else
                                        "metric_pixels_kernel"),
            "kernel_ms": round(dev_ms, 3), "flops_per_launch": fl,
            "note": "one launch covers the whole sweep; kernel_ms includes the per-image reductions"}
    return result, (planes, ds, env)


def sweep_cpu_baseline(planes, ds, env, n_sample=8):
    """The oracle's evaluate_optical_image on the host's cores for a stated
    subset of the sweep: n_sample designs at evenly spaced ranks of their plane
    count, x 21 tilts.  The rate is the subset's; scaled by the exact FLOP
    counts of the subset and of the whole sweep it gives the whole-sweep
    estimate beside it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from indicatrix_amd import designs, flops
    threads = host_threads()
    order = sorted(range(len(ds)), key=lambda i: (ds[i][1], i))
    pick = [order[int((k + 0.5) * len(ds) / n_sample)] for k in range(n_sample)]
    t0 = time.perf_counter()
    oracle_lib.metric_sweep(planes, [ds[i] for i in pick], designs.TILTS_C3, 64, 64, 32, 2.4, env,
                            nthreads=threads)
    dt = time.perf_counter() - t0
    out = {"value": round(n_sample / dt, 4), "unit": "design evals/s", "cores": threads, "kind": "port",
           "sample": f"{n_sample} designs at evenly spaced plane-count ranks "
                     f"({', '.join(str(ds[i][1]) for i in pick)} planes) x 21 tilts, 64x64 @ 32 samples: "
                     f"{dt:.2f} s on {threads} threads", "host": host_cpu()}
    cnt = flops.counts()["sweep_c3"]
    if len(ds) == cnt["n_designs"]:
        sub = sum(cnt["design_flops"][i] for i in pick)
        out["whole_sweep_estimate"] = round(len(ds) / (dt * cnt["sweep_flops"] / sub), 4)
    return out


AXIS_AZIMUTHS_DEG = [0.0, 45.0, 90.0, 135.0]   # PROFILE_AZIMUTHS_DEG (color/metrics/profile.rs:26)


def pose_sweep_rate(backend, n_designs=64):
    """M10: tilt-axis profiles (profile.rs:200, pose_batch.rs:199) -- the 57-facet
    standard round brilliant in Diamond under RingLights, 4 azimuths x 181 poses
    (724 Fire-less fan-grid evaluations, the reference's full tilt sweep), then
    the same sweep over n_designs seeded designs (32..256 planes) in one call.
    Timed per call: the host's girdle-width measurement is included."""
    from indicatrix_amd import designs, metrics, tracer
    srb = tracer.planes_array(tracer.standard_round_brilliant())
    disp = tracer.GemMaterial.diamond().to_abi().dispersion
    env = tracer.LightingPreset.studio("RingLights", 1.0, 0.85, 0.95)
    one = [(0, len(srb), disp)]
    metrics.axis_profiles(backend, srb, one, AXIS_AZIMUTHS_DEG, env)  # warm
    t0 = time.perf_counter()
    curves = metrics.axis_profiles(backend, srb, one, AXIS_AZIMUTHS_DEG, env)
    dt1 = time.perf_counter() - t0
    planes, ds, _ = designs.sweep_batch(n_designs)
    t0 = time.perf_counter()
    metrics.axis_profiles(backend, planes, ds, AXIS_AZIMUTHS_DEG, env)
    dt2 = time.perf_counter() - t0
    poses = 181 * len(AXIS_AZIMUTHS_DEG)
    return {"srb_axes": {"poses": poses, "seconds": round(dt1, 5), "poses_per_s": round(poses / dt1, 1),
                         "mean_brilliance_pct": round(float(curves[0, :, 0].mean()), 4)},
            "designs_axes": {"designs": n_designs, "poses": n_designs * poses, "seconds": round(dt2, 4),
                             "poses_per_s": round(n_designs * poses / dt2, 1)}}, (srb, disp, env)


def pose_sweep_cpu_baseline(srb, disp, env):
    """The oracle's evaluate_pose over the same 724 poses (Fire-less) on the
    host's cores, one pose per task."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    threads = host_threads()
    poses = [p for az in AXIS_AZIMUTHS_DEG for p in oracle_lib.axis_poses(az)]
    t0 = time.perf_counter()
    oracle_lib.pose_metrics(srb, disp, env, poses, want_all=False, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(len(poses) / dt, 1), "unit": "poses/s", "cores": threads, "kind": "port",
            "sample": f"the SRB tilt sweep, {len(poses)} Fire-less poses in {dt:.3f} s on {threads} threads"}


def mlp_train_rate(backend, steps, batch=4096):
    """C4: surrogate training steps -- BF16 MFMA forward and backward, FP32
    master weights and Adam -- at batch 4096 on a fixed synthetic batch,
    timed with HIP events on the library's stream around `steps` steps
    launched back to back."""
    import torch
    from indicatrix_amd import flops, surrogate
    net = surrogate.SurrogateMLP(backend, max_batch=batch, seed=0)
    g = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn(batch, surrogate.N_IN, device="cuda", generator=g)
    t = torch.randn(batch, surrogate.N_OUT, device="cuda", generator=g)
    torch.cuda.synchronize()
    for _ in range(10):
        net.train_step(x, t, 1e-3, want_loss=False)
    stream = torch.cuda.ExternalStream(backend.stream())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    net.sync()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        net.train_step(x, t, 1e-3, want_loss=False)
    e1.record(stream)
    e1.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / steps
    loss = net.train_step(x, t, 1e-3, want_loss=True)
    net.close()
    fl = flops.mlp_train_step_flops(batch)
    achieved = fl / (ms * 1e-3) / 1e12
    return {"workload": "MLP 16 -> 4 x 256 (GELU) -> 3, BF16 weights/activations, FP32 accumulation, "
                        "master weights and Adam, batch 4096",
            "steps": steps, "ms_per_step": round(ms, 4), "wall_ms_per_step": round(wall * 1e3 / steps, 4),
            "samples_per_s": round(batch / (ms * 1e-3)), "final_loss": round(float(loss), 5),
            "roofline": {"bound": "mfma", "pipe": "bf16 dense MFMA", "achieved": round(achieved, 3),
                         "peak": MI355X_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(achieved / MI355X_BF16_PEAK_TFLOPS, 4), "traffic": round(MLP_STEP_HBM_BYTES),
                         "kernel": "mlp_forward/backward/dw/reduce (one training step)",
                         "kernel_ms": round(ms, 4), "flops_per_launch": fl,
                         "note": "flops = 2 M N K of the forward, backward-data and weight-gradient GEMMs "
                                 "over the logical layer shapes; time = the step's four launches"}}


def screened_rate(backend, steps):
    """C5: the surrogate-screened Optimize loop with the committed surrogate
    (refined on the designs it has verified, 100 Adam steps per step)."""
    import numpy as np
    from indicatrix_amd import optimize, surrogate
    if not os.path.exists(SURROGATE_WEIGHTS):
        return {"skipped": f"{SURROGATE_WEIGHTS} missing"}
    net = surrogate.SurrogateMLP(backend, max_batch=8192 * 21, seed=0)
    net.load(SURROGATE_WEIGHTS)
    opt = optimize.ScreenedOptimizer(backend, net, nd=2.05, abbe=35.0, seed=11)
    recs = [opt.step() for _ in range(steps)]
    net.close()
    screen = sum(r["screen_s"] for r in recs)
    verify = sum(r["verify_s"] for r in recs)
    refine = sum(r["refine_s"] for r in recs)
    total = sum(r["step_s"] for r in recs)
    cand = sum(r["candidates"] for r in recs)
    verified = sum(r["verified"] for r in recs)
    tracer_only = verified / verify   # the same tracer work per verified design
    later = recs[2:] if len(recs) > 2 else recs
    err = np.concatenate([r["rel_err"] for r in later])
    # per (design, tilt), before the tilt average
    err_tilt = np.concatenate([(np.abs(r["pred"] - r["traced"]) / np.maximum(np.abs(r["traced"]), 1e-6))
                               .reshape(-1, 3) for r in later])

    def r5(v):
        return [round(float(x), 5) for x in v]

    return {"steps": steps, "candidates_per_step": recs[0]["candidates"],
            "verified_per_step": recs[0]["verified"], "verify_shape": list(optimize.VERIFY_SHAPE),
            "screen_s_per_step": round(screen / steps, 4), "verify_s_per_step": round(verify / steps, 3),
            "refine_s_per_step": round(refine / steps, 4),
            "design_evals_per_s": round(cand / total, 1),
            "tracer_only_design_evals_per_s": round(tracer_only, 2),
            "gain_vs_tracer_only": round(cand / total / tracer_only, 1),
            "metrics": ["brilliance_pct", "fire_index", "scintillation_pct"],
            "rel_err_median_by_step": [r5(np.median(r["rel_err"], axis=0)) for r in recs],
            "rel_err_median_from_step2": r5(np.median(err, axis=0)),
            "rel_err_mean_from_step2": r5(err.mean(axis=0)),
            "rel_err_p90_from_step2": r5(np.percentile(err, 90, axis=0)),
            "per_tilt_rel_err_median_from_step2": r5(np.median(err_tilt, axis=0)),
            "per_tilt_rel_err_mean_from_step2": r5(err_tilt.mean(axis=0)),
            "objective_rel_err_median_from_step2": round(float(np.median(
                np.concatenate([r["obj_rel_err"] for r in later]))), 5),
            "best_verified_objective": round(recs[-1]["best"], 4)}


def cpu_baseline(scene, target_s):
    """The CPU oracle (a C restatement of the reference tracer) on this host's
    cores, over a bounded sample of the same workload: the same frame at the
    spp that makes about `target_s` seconds of CPU work (sample indices
    0..spp-1), sized from a timed 1-spp pass."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    threads = host_threads()
    t0 = time.perf_counter()
    oracle_lib.accumulate(scene, 0, 1, nthreads=threads)
    one = time.perf_counter() - t0
    spp = max(1, min(4096, int(target_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    oracle_lib.accumulate(scene, 0, spp, nthreads=threads)
    dt = time.perf_counter() - t0
    paths = scene.width * scene.height * spp
    return {"value": round(paths / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads,
            "kind": "port",
            "sample": f"the bench frame at {spp} spp (sample indices 0..{spp - 1}): {paths} paths "
                      f"in {dt:.2f} s on {threads} threads (AVX2 slab scan and exponential, -O3)",
            "host": host_cpu()}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                                   "(the film would shard by sample range across GPUs; only one is measured here)"}))
        sys.exit(2)
    import torch
    from indicatrix_amd import flops, tracer, workloads

    if args.copy_calibration:
        copy_calibration()
        return
    if args.part:
        backend = tracer.GpuBackend()
        if args.part == "c2":
            result = film_c2_rate(backend, 2)[0]
        elif args.part == "sweep":
            result = sweep_rate(backend)[0]
        elif args.part == "pose":
            result = pose_sweep_rate(backend)[0]
        elif args.part == "mlp":
            result = mlp_train_rate(backend, args.mlp_steps)
        else:
            result = screened_rate(backend, args.optimize_steps)
        print(json.dumps({"part": args.part, "result": result}), flush=True)
        backend.close()
        return
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device visible"}))
        sys.exit(1)
    scene = workloads.film_c1_scene()
    spp = workloads.C1["spp"]
    w, h = scene.width, scene.height
    paths_per_step = w * h * spp
    backend = tracer.GpuBackend()
    accum = torch.zeros(w * h * 3, dtype=torch.float64, device="cuda")
    ptr = accum.data_ptr()

    print(f"# bench: {backend.adapter_label()} film {w}x{h} @ {spp} spp, depth "
          f"{scene.scene.max_bounces}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        backend.accumulate_device(scene, i * spp, spp, ptr)
    torch.cuda.synchronize()

    # exact algorithmic FLOPs of this frame (sample indices 0..63) with the
    # device's own event totals for the same paths (untimed)
    entry = flops.counts()["film_c1"]
    if (entry["width"], entry["height"], entry["spp"], entry["planes"]) != (w, h, spp, scene.scene.n_planes):
        raise RuntimeError("flop_counts.json film_c1 does not describe this frame")
    totals = backend.count_events(scene, entry["sample_offset"], entry["spp"])

    kernel_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        backend.accumulate_device(scene, (args.warmup + k) * spp, spp, ptr)
        kernel_ms.append(backend.last_kernel_ms())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0

    ms_per_step = dt * 1e3 / args.steps
    value = paths_per_step * args.steps / dt / 1e6
    avg_kernel_ms = sum(kernel_ms) / len(kernel_ms)
    roofline = film_roofline(entry, totals, avg_kernel_ms, paths_per_step, round(FILM_KERNEL_HBM_BYTES))
    line = {
        "metric": "Msamples/s (8-lambda spectral paths)",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded counter RNG; parametric round-brilliant geometry)",
        "config": {"workload": "film 128x128 @ 64 spp, depth 16, 57-facet round brilliant "
                               "(crown 34.5, pavilion 40.75, table 57%), Diamond Sellmeier-3, "
                               "uniform D65 hemisphere, full Mueller transport",
                   "paths_per_step": paths_per_step, "planes": scene.scene.n_planes},
        "roofline": roofline,
    }
    if not args.film_only:
        print("# bench: film C2", file=sys.stderr, flush=True)
        c2, c2_scene = film_c2_rate(backend, 2)
        line["film_c2"] = c2
        print("# bench: metric sweep", file=sys.stderr, flush=True)
        sweep, sweep_inputs = sweep_rate(backend)
        print("# bench: surrogate training", file=sys.stderr, flush=True)
        line["surrogate_training"] = mlp_train_rate(backend, args.mlp_steps)
        print("# bench: screened optimize", file=sys.stderr, flush=True)
        line["design_evals"] = {"sweep": sweep, "screened_optimize": screened_rate(backend, args.optimize_steps)}
        print("# bench: pose sweep", file=sys.stderr, flush=True)
        pose_sweep, pose_inputs = pose_sweep_rate(backend)
        line["pose_sweep"] = pose_sweep
    if not args.no_cpu_baseline:
        print("# bench: cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(scene, args.cpu_seconds)
        if not args.film_only:
            line["design_evals"]["sweep"]["cpu_baseline"] = sweep_cpu_baseline(*sweep_inputs)
            line["film_c2"]["cpu_baseline"] = film_c2_cpu_baseline(c2_scene)
            line["pose_sweep"]["cpu_baseline"] = pose_sweep_cpu_baseline(*pose_inputs)
    print(json.dumps(line), flush=True)
    backend.close()


if __name__ == "__main__":
    main()
