"""The exact FLOP counts the benchmark's roofline figures use
(indicatrix_amd/data/flop_counts.json, written by oracle/flop_counts.py) are
the counting build's on the benchmark's own inputs, and the device-side
correction and MLP counts in indicatrix_amd/flops.py hold.  CPU only."""
import numpy as np

import oracle_lib
from indicatrix_amd import designs, flops, workloads


def test_film_c1_count_is_the_counting_builds():
    e = flops.counts()["film_c1"]
    sc = workloads.film_c1_scene()
    assert (e["width"], e["height"], e["planes"]) == (sc.width, sc.height, sc.scene.n_planes)
    _, n = oracle_lib.accumulate_flops(sc, e["sample_offset"], e["spp"])
    assert n == e["flops"]


def test_film_c2_count_is_the_counting_builds():
    e = flops.counts()["film_c2"]
    sc = workloads.film_c2_scene()
    assert (e["width"], e["height"], e["planes"]) == (sc.width, sc.height, sc.scene.n_planes)
    _, n = oracle_lib.accumulate_flops(sc, e["sample_offset"], e["spp"])
    assert n == e["flops"]


def test_sweep_counts_cover_the_bench_sweep():
    c = flops.counts()["sweep_c3"]
    planes, ds, _ = designs.sweep_batch(c["n_designs"])
    assert [d[1] for d in ds] == c["design_planes"]
    assert sum(c["design_flops"]) == c["sweep_flops"] == flops.sweep_flops()
    # one design recounted image by image: the smallest, 21 tilts
    i = int(np.argmin(c["design_planes"]))
    from indicatrix_amd import tracer
    env = tracer.LightingPreset.studio(c["preset"], 1.0, 0.0, 0.0)
    jobs = oracle_lib.sweep_requests(planes, [ds[i]], designs.TILTS_C3, c["width"], c["height"], c["samples"],
                                     c["distance"], env)
    assert sum(oracle_lib.metric_image_flops(r) for _, _, r in jobs) == c["design_flops"][i]


def test_device_correction_counts_only_the_skipped_scans():
    e = {"flops": 10_000_000, "planes": 74, "paths": 100}
    none_skipped = {"bounces": 50, "scans": 50, "probe": 7, "probe_scans": 7}
    assert flops.film_device_flops(e, none_skipped) == (10_000_000, 10_000_000)
    # 5 bounces and 2 probes resolved from one plane: 12 x 73 FLOPs saved each
    some = {"bounces": 50, "scans": 45, "probe": 7, "probe_scans": 5}
    assert flops.film_device_flops(e, some)[0] == 10_000_000 - 7 * 12 * 73


def test_mlp_step_count():
    """Per row: forward 2 (16*256 + 3*256*256 + 256*3) = 402,944; backward
    data over the last four layers 2 (3*256*256 + 256*3) = 394,752; weight
    gradients = the forward's 402,944."""
    assert flops.mlp_forward_flops(1) == 402_944
    assert flops.mlp_train_step_flops(1) == 402_944 + 394_752 + 402_944
    assert flops.mlp_train_step_flops(4096) == 4096 * 1_200_640
