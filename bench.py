"""Film throughput on one MI355X: Msamples/s of 8-lambda spectral paths.

One step = one pass of the film hot path over one frame of the workload
(every pixel x spp paths, absolute sample indices advancing each step),
accumulated in FP64 into a device-resident buffer.  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_FP32_PEAK_TFLOPS = 157.3
MI355X_HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU work of the bounded CPU-baseline sample")
    return ap.parse_args()


def cpu_baseline(scene, target_s):
    """The CPU oracle (a C restatement of the reference tracer) on this host's
    cores, over a bounded sample of the same workload: the same frame at the
    spp that makes about `target_s` seconds of CPU work (sample indices
    0..spp-1), sized from a timed 1-spp pass."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    threads = min(os.cpu_count() or 1, 16)
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
                      f"in {dt:.2f} s on {threads} threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                                   "(the film shards by sample range across GPUs, see DESIGN.md)"}))
        sys.exit(2)
    import torch
    from indicatrix_amd import flops, tracer, workloads

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

    # algorithmic FLOPs of one step (event counts are exact per path; untimed)
    totals = backend.count_events(scene, 0, spp)
    step_flops = flops.film_flops(totals, scene.scene.n_planes)

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
    achieved = step_flops / (avg_kernel_ms * 1e-3) / 1e12
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
        "roofline": {"bound": "mfma", "pipe": "fp32-valu (non-matrix FP32)",
                     "achieved": round(achieved, 3), "peak": MI355X_FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / MI355X_FP32_PEAK_TFLOPS, 4),
                     "traffic": None,
                     "kernel": "film_persistent_kernel", "kernel_ms": round(avg_kernel_ms, 4),
                     "flops_per_launch": step_flops, "events": totals},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(scene, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    backend.close()


if __name__ == "__main__":
    main()
