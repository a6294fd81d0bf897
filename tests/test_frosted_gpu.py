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


@pytest.mark.parametrize("kind", ["isotropic", "uniaxial", "biaxial"])
def test_sun_light_samples_match_oracle(backend, kind):
    """Under DaylightSun an exterior frosted event light-samples the sun disc
    (nee_contribution_frosted_exterior) and the next miss weighs the
    BSDF-sampled continuation by the balance heuristic: the deposits and the
    weights path for path against the oracle."""
    from indicatrix_amd import workloads
    m = {"isotropic": tracer.GemMaterial.diamond, "uniaxial": lambda: workloads.c2_material(bands=True),
         "biaxial": tracer.GemMaterial.alexandrite}[kind]()
    fr = [1 if (i in GIRDLE or i >= 49) else 0 for i in range(74)]
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, 32, 24, 8)
    env = tracer.LightingPreset.studio("DaylightSun", 1.0, 0.4, 0.9)
    sc = tracer.FrameScene(cam, 32, 24, tracer.standard_round_brilliant(), m, 12, env, facet_finishes=fr)
    print(f"sun light samples {kind} 32x24x8", flush=True)
    g_xyz, g_ev = backend.trace_paths(sc, 0, 8)
    o_xyz, o_ev = oracle_lib.trace_paths(sc, 0, 8)
    same = ((g_ev["bounces"] == o_ev["bounces"]) & (g_ev["termination"] == o_ev["termination"])
            & (g_ev["facet_hash"] == o_ev["facet_hash"]))
    a, b = g_xyz[same].astype(np.float64), o_xyz[same].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-6)
    print(f"  flip {1 - same.mean():.2e} q999 {np.quantile(rel, 0.999):.2e} max {rel.max():.2e}", flush=True)
    if kind == "uniaxial":
        assert 1.0 - same.mean() < 1e-2 and (rel.max(axis=1) > 1e-3).mean() < 1e-2
    else:
        assert 1.0 - same.mean() < 3e-3 and np.quantile(rel, 0.999) < 1e-4
    f_xyz = backend.film_samples(sc, 0, 8)   # the timed EXTRA instantiation
    assert f_xyz.tobytes() == g_xyz.tobytes()


def test_scattering_under_the_sun_matches_oracle(backend):
    """Scatters light-sample only a panorama: under DaylightSun a scattering
    stone is the plain estimator, path for path."""
    from test_scattering import scattering_scene
    sc = scattering_scene(1.5, 0.6, preset="DaylightSun", size=32, spp=8)
    g_xyz, g_ev = backend.trace_paths(sc, 0, 8)
    o_xyz, o_ev = oracle_lib.trace_paths(sc, 0, 8)
    same = ((g_ev["bounces"] == o_ev["bounces"]) & (g_ev["termination"] == o_ev["termination"])
            & (g_ev["facet_hash"] == o_ev["facet_hash"]))
    a, b = g_xyz[same].astype(np.float64), o_xyz[same].astype(np.float64)
    rel = np.abs(a - b) / np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-6)
    print(f"scattering under the sun: flip {1 - same.mean():.2e} q999 {np.quantile(rel, 0.999):.2e}", flush=True)
    assert 1.0 - same.mean() < 3e-3 and np.quantile(rel, 0.999) < 1e-4
