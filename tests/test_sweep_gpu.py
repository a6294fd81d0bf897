"""The batched metric sweep (one device pass over designs x tilts) against the
oracle's evaluate_optical_image on every (design, tilt) image."""
import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, designs, metrics, tracer, workloads

pytestmark = pytest.mark.gpu


def test_sweep_matches_per_image_oracle(backend):
    planes, ds, meta = designs.sweep_batch(3, first_seed=40)
    tilts = np.array([-27.0, 0.0, 12.0], np.float32)
    env = tracer.LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0)
    w = h = 24
    samples = 6
    print("sweep 3 designs x 3 tilts", flush=True)
    out = metrics.metric_sweep(backend, planes, ds, tilts, w, h, samples, 2.4, env)
    to_rad = np.float32(np.float32(np.pi) / np.float32(180.0))
    for i, (off, cnt, disp) in enumerate(ds):
        pl = (abi.FacetPlane * cnt).from_buffer_copy(planes[off:off + cnt].tobytes())
        mat = abi.Material()
        mat.dispersion = disp
        mat.cubic = 1
        mat.absorption_path_scale = 1.0
        for j, t in enumerate(tilts):
            yaw = float(np.float32(180.0) * to_rad) if t < 0 else 0.0
            pitch = float(np.float32(np.float32(90.0) - np.float32(abs(t))) * to_rad)
            cam = tracer.camera_new(yaw, pitch, 2.4, 42.0, w, h, samples)
            r = metrics.make_request(cam, w, h, samples, pl, mat, env, True, True, (yaw, pitch, 2.4))
            _, tot, scint = oracle_lib.metric_image(r)
            g = out[i, j]
            assert g[0] == pytest.approx(tot.brilliance_pct, abs=1e-5)
            assert g[3] == pytest.approx(tot.windowing_pct, abs=1e-5)
            assert g[4] == pytest.approx(tot.extinction_pct, abs=1e-5)
            assert g[1] == pytest.approx(tot.fire_index, rel=2e-3, abs=1e-3)
            assert g[2] == pytest.approx(scint, abs=1e-3)


def test_sweep_shapes_and_graph_replays_match_the_eager_grid(backend):
    """Every launch shape computes the same pixels, so the sweep's metrics are
    bit-identical across the 2-D grid, each persistent-queue shape, the
    autotuned choice and graph replays (with other device work and other
    inputs between replays)."""
    env = tracer.LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0)
    tilts = np.array([-21.0, 0.0, 9.0, 30.0], np.float32)
    a_planes, a_ds, _ = designs.sweep_batch(6, first_seed=70)
    b_planes, b_ds, _ = designs.sweep_batch(6, first_seed=90)
    w = h = 20
    spp = 4
    try:
        print("sweep eager grid", flush=True)
        metrics.set_metric_mode(backend, use_graph=False, block=-1)
        ref_a = metrics.metric_sweep(backend, a_planes, a_ds, tilts, w, h, spp, 2.4, env)
        ref_b = metrics.metric_sweep(backend, b_planes, b_ds, tilts, w, h, spp, 2.4, env)
        assert metrics.metric_tuning(backend)[0] == -1
        for block, tile in [(64, 64), (64, 256), (128, 128), (128, 512), (256, 256), (256, 1024),
                            (-64, 0), (-128, 0), (-256, 0), (-128, 100), (-64, 7)]:
            print(f"sweep queue/flow {block}x{tile}", flush=True)
            metrics.set_metric_mode(backend, use_graph=False, block=block, tile=tile)
            got = metrics.metric_sweep(backend, a_planes, a_ds, tilts, w, h, spp, 2.4, env)
            assert np.array_equal(got, ref_a), (block, tile)
        print("sweep autotuned graph replays", flush=True)
        metrics.set_metric_mode(backend, use_graph=True, block=0)
        film = workloads.film_c1_scene(8, 8)
        for k in range(3):
            got_a = metrics.metric_sweep(backend, a_planes, a_ds, tilts, w, h, spp, 2.4, env)
            assert np.array_equal(got_a, ref_a), k
            assert metrics.metric_tuning(backend)[3]
            acc = np.zeros((64, 3))
            assert backend.try_accumulate(film, 0, 2, acc)       # eager work between replays
            got_b = metrics.metric_sweep(backend, b_planes, b_ds, tilts, w, h, spp, 2.4, env)
            assert np.array_equal(got_b, ref_b), k
        block, tile, grid, graphed = metrics.metric_tuning(backend)
        assert graphed and grid > 0 and (block == -1 or (block > 0 and tile >= block) or
                                         (block in (-64, -128, -256) and tile > 0))
    finally:
        metrics.set_metric_mode(backend, use_graph=True, block=0)


def test_flow_kernel_matches_grid_on_studio_rig_and_c3_shape(backend):
    """The flow kernel against the 2-D grid kernel on the lit studio rig
    (light list, head shadow, Fire replays) and on a slice of the C3 sweep
    shape (64x64 at 32 samples): the metrics of every image are bit-identical."""
    tilts = np.array([-30.0, -3.0, 15.0], np.float32)
    planes, ds, _ = designs.sweep_batch(4, first_seed=7)
    cases = [(tracer.LightingPreset.studio("RingLights", 1.0, 0.85, 0.95), 32, 32, 9),
             (tracer.LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0), 64, 64, 32)]
    try:
        for env, w, h, spp in cases:
            print(f"flow vs grid {w}x{h}@{spp}", flush=True)
            metrics.set_metric_mode(backend, use_graph=False, block=-1)
            ref = metrics.metric_sweep(backend, planes, ds, tilts, w, h, spp, 2.4, env)
            for block, tile in [(-128, 0), (-256, 333)]:
                metrics.set_metric_mode(backend, use_graph=False, block=block, tile=tile)
                got = metrics.metric_sweep(backend, planes, ds, tilts, w, h, spp, 2.4, env)
                assert metrics.metric_tuning(backend)[0] == block
                assert np.array_equal(got, ref), (block, tile, w)
    finally:
        metrics.set_metric_mode(backend, use_graph=True, block=0)
