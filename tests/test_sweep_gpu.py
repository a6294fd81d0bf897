"""The batched metric sweep (one device pass over designs x tilts) against the
oracle's evaluate_optical_image on every (design, tilt) image."""
import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, designs, metrics, tracer

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
