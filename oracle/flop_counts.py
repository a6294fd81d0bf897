"""Exact FLOP counts of the benchmark workloads from the oracle's counting
build (oracle/ixo_flops.cpp), written to indicatrix_amd/data/flop_counts.json
for the roofline figures.  Test and measurement infrastructure: it runs the
oracle here, on the CPU; the JSON it writes is data.

Counting rule (ixo_flops.cpp): add, subtract, multiply, divide, square root
and every transcendental 1, fused multiply-add 2; negation, abs, min/max,
compares, conversions and integer work 0.  The counts are of the reference's
algorithm as the oracle restates it, operation for operation, on the exact
inputs named below.

  film C1    the bench frame (128x128, sample indices 0..63), every path
  film C2    the 1024x1024 frame at sample index 0 (1,048,576 of the 268 M
             paths of one 256-spp frame), per-path mean
  sweep C3   every image of the sweep: 512 seeded designs x 21 tilts,
             64x64 @ 32 samples with Fire and scintillation (about 4 CPU-hours
             of counting; --jobs processes)

  python3 oracle/flop_counts.py [--jobs 8]
"""
import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
OUT = os.path.join(ROOT, "indicatrix_amd", "data", "flop_counts.json")

SWEEP = dict(n_designs=512, width=64, height=64, samples=32, distance=2.4, preset="IsoHemisphere")


def sweep_inputs():
    from indicatrix_amd import designs, tracer
    planes, ds, _ = designs.sweep_batch(SWEEP["n_designs"])
    env = tracer.LightingPreset.studio(SWEEP["preset"], 1.0, 0.0, 0.0)
    return planes, ds, env, designs.TILTS_C3


def count_design(i):
    import oracle_lib
    planes, ds, env, tilts = sweep_inputs()
    jobs = oracle_lib.sweep_requests(planes, [ds[i]], tilts, SWEEP["width"], SWEEP["height"], SWEEP["samples"],
                                     SWEEP["distance"], env)
    return [oracle_lib.metric_image_flops(req) for _, _, req in jobs]


def film(scene_fn, spp):
    import oracle_lib
    sc = scene_fn()
    t0 = time.time()
    _, flops = oracle_lib.accumulate_flops(sc, 0, spp)
    paths = sc.width * sc.height * spp
    print(f"  {paths} paths, {flops} FLOPs ({flops / paths:.1f} per path) in {time.time() - t0:.1f} s",
          flush=True)
    return {"sample_offset": 0, "spp": spp, "width": sc.width, "height": sc.height, "paths": paths,
            "planes": sc.scene.n_planes, "flops": flops}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    from indicatrix_amd import workloads
    out = {"rule": "add/sub/mul/div/sqrt/transcendental = 1, fma = 2; neg/abs/min/max/compare/convert = 0",
           "source": "oracle/ixo_flops.cpp over oracle/ixo.c; python3 oracle/flop_counts.py"}
    print("film C1", flush=True)
    out["film_c1"] = film(workloads.film_c1_scene, workloads.C1["spp"])
    print("film C2 (sample index 0)", flush=True)
    out["film_c2"] = film(workloads.film_c2_scene, 1)
    out["film_c2"]["frame_spp"] = workloads.C2["spp"]
    print("sweep C3", flush=True)
    planes, ds, env, tilts = sweep_inputs()
    t0 = time.time()
    with Pool(args.jobs) as pool:
        per = pool.map(count_design, range(len(ds)), chunksize=1)
    tot = [sum(v) for v in per]
    print(f"  {len(ds)} designs x {len(tilts)} tilts in {time.time() - t0:.1f} s: {sum(tot)} FLOPs", flush=True)
    out["sweep_c3"] = dict(SWEEP, tilts=len(tilts), design_planes=[int(d[1]) for d in ds],
                           design_flops=[int(v) for v in tot], sweep_flops=int(sum(tot)))
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
