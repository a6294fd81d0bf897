"""Biaxial stones on the GPU, through the C ABI, against the CPU oracle.

The biaxial facet event (mode A / mode B indices from each channel's
indicatrix at the hero's resolved wave normals, 50/50 entry and re-coupling
draws, per-channel biaxial walk-off, assigned-mode trichroic absorption) on
the built-in Alexandrite (trichroic bands) and Topaz (one isotropic band):
the integer side channel path for path, XYZ within the device's
transcendental budget (the general indicatrix branch itself has no
transcendental: correctly rounded division and square root only), the timed
persistent instantiation bit-identical to the diagnostic kernel, frame sums,
per-pixel z-scores and the white furnace.
"""
import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, tracer

pytestmark = pytest.mark.gpu


def scene(material, preset="RingLights", size=32, spp=8, bounces=16):
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, size, size, spp)
    env = tracer.LightingPreset.studio(preset, 1.0, 0.85, 0.95)
    return tracer.FrameScene(cam, size, size, tracer.standard_round_brilliant(), material, bounces, env)


MATERIALS = {"alexandrite": tracer.GemMaterial.alexandrite, "topaz": tracer.GemMaterial.topaz}


def _same(g_ev, o_ev):
    return ((g_ev["bounces"] == o_ev["bounces"]) & (g_ev["termination"] == o_ev["termination"])
            & (g_ev["facet_hash"] == o_ev["facet_hash"]))


def _rel(g, o):
    a, b = g.astype(np.float64), o.astype(np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-6)


@pytest.mark.parametrize("preset", ["RingLights", "IsoHemisphere"])
@pytest.mark.parametrize("name", list(MATERIALS))
def test_biaxial_paths_match_oracle(backend, name, preset):
    sc = scene(MATERIALS[name](), preset)
    print(f"biaxial paths {name} {preset} 32x32x8", flush=True)
    g_xyz, g_ev = backend.trace_paths(sc, 0, 8)
    o_xyz, o_ev = oracle_lib.trace_paths(sc, 0, 8)
    same = _same(g_ev, o_ev)
    rel = _rel(g_xyz[same], o_xyz[same])
    print(f"  flip {1.0 - same.mean():.2e} q999 {np.quantile(rel, 0.999):.2e} max {rel.max():.2e}", flush=True)
    assert 1.0 - same.mean() < 2e-3
    for f in ("n_tir", "n_reflect", "n_refract", "n_probe"):
        assert np.array_equal(g_ev[f][same], o_ev[f][same])
    # the indicatrix functions are bit-exact on the device (no transcendental
    # in the general branch); what remains is the frame rotation, TIR
    # retardation, exponentials and lighting (measured: flip 0, max 7e-7)
    assert rel.max() < 1e-5
    assert (g_ev["n_refract"] > 0).mean() > 0.3


@pytest.mark.parametrize("name", list(MATERIALS))
def test_timed_biaxial_kernel_per_path(backend, name):
    """The persistent instantiation the accumulate path launches for a
    biaxial stone writes the diagnostic kernel's XYZ bits path for path, and
    the per-pixel sums are their FP64 sample-order reduction."""
    from test_film_gpu import _reduce_like_the_device
    sc = scene(MATERIALS[name](), "IsoHemisphere", size=48, spp=4)
    print(f"timed biaxial kernel {name} 48x48x4", flush=True)
    f_xyz = backend.film_samples(sc, 16, 4)
    g_xyz, _ = backend.trace_paths(sc, 16, 4)
    assert f_xyz.tobytes() == g_xyz.tobytes()
    acc = np.zeros((48 * 48, 3), np.float64)
    assert backend.try_accumulate(sc, 16, 4, acc)
    assert acc.tobytes() == _reduce_like_the_device(f_xyz, 48 * 48, 4).tobytes()


def test_biaxial_accumulate_matches_oracle_sums(backend):
    sc = scene(tracer.GemMaterial.alexandrite(), "RingLights", size=32, spp=16)
    acc = np.zeros((32 * 32, 3))
    print("accumulate biaxial 32x32x16", flush=True)
    assert backend.try_accumulate(sc, 0, 16, acc)
    ref = oracle_lib.accumulate(sc, 0, 16)
    assert acc.sum() > 0
    assert np.abs(acc.sum(axis=0) - ref.sum(axis=0)).max() / ref.sum() < 2e-3


def test_biaxial_statistical_image_equivalence(backend):
    import zstats
    sc = scene(tracer.GemMaterial.alexandrite(), "IsoHemisphere", size=40, spp=32)
    print("z-score biaxial 40x40x32", flush=True)
    g_xyz, _ = backend.trace_paths(sc, 0, 32)
    o_xyz, _ = oracle_lib.trace_paths(sc, 32, 32)
    ok, res = zstats.image_equivalent(o_xyz, g_xyz, 40, 40, 32)
    assert ok, res


def test_biaxial_white_furnace_on_the_device(backend):
    from test_hdr_env import furnace_target
    m = tracer.GemMaterial("furnace probe", abi.DISP_CAUCHY, [1.5, 0.0, 0.0], birefringence_delta=0.02,
                           cubic=False, c_axis=(0.3, 0.8, 0.52))
    m.biaxial_delta_beta_alpha = 0.008
    sc = scene(m, size=12, spp=96)
    sc.set_hdr_map(np.full((1, 1, 3), 2.5, np.float32))
    acc = np.zeros((12 * 12, 3))
    print("biaxial furnace 12x12x96", flush=True)
    assert backend.try_accumulate(sc, 0, 96, acc)
    mean = acc.sum(axis=0) / (12 * 12 * 96)
    target = furnace_target(2.5)
    assert (np.abs(mean - target) / target).max() < 0.01, (mean, target)
