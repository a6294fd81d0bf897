"""Frosted facets on the device against the oracle, path for path: the same
integer side channel (bounces, termination, facet and event-kind hash), XYZ
within the transcendental budget.  Both device templates: the studio rig
(per-channel illuminant) and a lit model (one scaled factor)."""
import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import tracer

pytestmark = pytest.mark.gpu

GIRDLE = range(33, 49)


@pytest.mark.parametrize("preset,which", [("RingLights", "girdle"), ("IsoHemisphere", "girdle"),
                                          ("LightTent", "pavilion"), ("RingLights", "all")])
def test_frosted_paths_match_oracle(backend, preset, which):
    if which == "girdle":
        fr = [1 if i in GIRDLE else 0 for i in range(74)]
    elif which == "pavilion":
        fr = [1 if i >= 49 else 0 for i in range(74)]
    else:
        fr = [1] * 74
    cam = tracer.camera_new(0.35, 1.15, 5.0, 18.0, 32, 24, 8)
    env = tracer.LightingPreset.studio(preset, 1.0, 0.85, 0.95)
    sc = tracer.FrameScene(cam, 32, 24, tracer.standard_round_brilliant(), tracer.GemMaterial.diamond(), 12,
                           env, facet_finishes=fr)
    print(f"frosted {which} under {preset} 32x24x8", flush=True)
    g_xyz, g_ev = backend.trace_paths(sc, 0, 8)
    o_xyz, o_ev = oracle_lib.trace_paths(sc, 0, 8)
    same = ((g_ev["bounces"] == o_ev["bounces"]) & (g_ev["termination"] == o_ev["termination"])
            & (g_ev["facet_hash"] == o_ev["facet_hash"]))
    a, b = g_xyz[same].astype(np.float64), o_xyz[same].astype(np.float64)
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-6)
    rel = np.abs(a - b) / scale
    print(f"  flip {1 - same.mean():.2e} q999 {np.quantile(rel, 0.999):.2e}", flush=True)
    assert 1.0 - same.mean() < 2e-3
    assert np.quantile(rel, 0.999) < 1e-4
    sc.set_finishes(None)
    sharp, _ = backend.trace_paths(sc, 0, 8)
    assert not np.array_equal(sharp, g_xyz)


@pytest.mark.parametrize("kind,preset", [("uniaxial", "RingLights"), ("uniaxial", "IsoHemisphere"),
                                         ("biaxial", "RingLights")])
def test_frosted_anisotropic_paths_match_oracle(backend, kind, preset):
    """Frosted girdle and pavilion on uniaxial / biaxial stones: the frosted
    arm from the anisotropic bounce geometry, a 50/50 eigenmode draw on a
    transmission into the stone, the re-coupling draw after a frosted
    internal reflection."""
    from indicatrix_amd import workloads
    m = workloads.c2_material(bands=True) if kind == "uniaxial" else tracer.GemMaterial.alexandrite()
    fr = [1 if (i in GIRDLE or i >= 60) else 0 for i in range(74)]
    cam = tracer.camera_new(0.35, 1.15, 5.0, 18.0, 32, 24, 8)
    env = tracer.LightingPreset.studio(preset, 1.0, 0.85, 0.95)
    sc = tracer.FrameScene(cam, 32, 24, tracer.standard_round_brilliant(), m, 12, env, facet_finishes=fr)
    print(f"frosted {kind} under {preset} 32x24x8", flush=True)
    g_xyz, g_ev = backend.trace_paths(sc, 0, 8)
    o_xyz, o_ev = oracle_lib.trace_paths(sc, 0, 8)
    same = ((g_ev["bounces"] == o_ev["bounces"]) & (g_ev["termination"] == o_ev["termination"])
            & (g_ev["facet_hash"] == o_ev["facet_hash"]))
    a, b = g_xyz[same].astype(np.float64), o_xyz[same].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-6)
    print(f"  flip {1 - same.mean():.2e} q999 {np.quantile(rel, 0.999):.2e} max {rel.max():.2e}", flush=True)
    if kind == "biaxial":
        # no transcendental in the biaxial indicatrix: paths agree to rounding
        assert 1.0 - same.mean() < 3e-3
        assert np.quantile(rel, 0.999) < 1e-5
    else:
        # the uniaxial transport's acos / sin / cos / atan differ from glibc's by
        # an ULP here and there, and from this top view one-ULP differences flip
        # about 2 % of paths in the oracle itself (tests/test_uniaxial_sensitivity.py):
        # bounded at that scale, and the frames held to the z-score criteria
        import zstats
        assert 1.0 - same.mean() < 1e-2
        assert (rel.max(axis=1) > 1e-3).mean() < 1e-2
        g32, _ = backend.trace_paths(sc, 0, 32)
        o32, _ = oracle_lib.trace_paths(sc, 32, 32)
        ok, res = zstats.image_equivalent(o32, g32, 32, 24, 32)
        assert ok, res
    f_xyz = backend.film_samples(sc, 0, 8)   # the timed EXTRA instantiation, path for path
    assert f_xyz.tobytes() == g_xyz.tobytes()
