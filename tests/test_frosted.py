"""Frosted facets (FacetFinish::Frosted, RT/scattering/frosted.rs) in the
oracle and through the route (CPU).

The oracle's cosine-weighted hemisphere is pinned by what the reference's
construction fixes: unit directions on the normal's side, the pole at u1 = 0,
and the cosine-law moments (E[cos] = 2/3, E[cos^2] = 1/2, no azimuthal
bias).  Polished finishes given explicitly equal no finishes bit for bit; a
frosted girdle changes the image.  The device follows the oracle path for
path in tests/test_frosted_gpu.py.  Next-event estimation at frosted events
is not carried, so the route declines a frosted stone under DaylightSun.
Uniaxial and biaxial stones take the same arm from their own bounce geometry
(the eigenmode of a transmission into the stone drawn 50/50,
scattering/frosted.rs:340-390); a non-absorbing frosted stone of either kind
conserves energy in the analytic white furnace.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, tracer

F3 = C.c_float * 3
GIRDLE = range(33, 49)   # the standard brilliant's 16 girdle planes


def hemi(u1, u2, n):
    out = F3()
    oracle_lib.lib().ixo_cosine_hemisphere(u1, u2, F3(*n), out)
    return np.array(out[:], np.float64)


def test_cosine_hemisphere_moments():
    for n in [(0.0, 1.0, 0.0), (1.0, 0.0, 0.0), (0.6, -0.48, 0.64)]:
        nv = np.array(n)
        pole = hemi(0.0, 0.3, n)
        assert np.abs(pole - nv).max() < 1e-6
        m = 96
        cos = []
        az = np.zeros(2)
        for i in range(m):
            for j in range(m):
                d = hemi((i + 0.5) / m, (j + 0.5) / m, n)
                assert abs(np.linalg.norm(d) - 1.0) < 1e-6
                c = d @ nv
                assert c >= -1e-7
                cos.append(c)
                t = d - nv * c
                az += t[:2] if abs(n[2]) < 0.9 else t[1:]
        cos = np.array(cos)
        assert abs(cos.mean() - 2.0 / 3.0) < 2e-3
        assert abs((cos ** 2).mean() - 0.5) < 2e-3
        assert np.abs(az).max() / (m * m) < 2e-3


def _scene(finishes, preset="RingLights", size=24, spp=8, material=None):
    cam = tracer.camera_new(0.35, 1.15, 5.0, 18.0, size, size, spp)
    env = tracer.LightingPreset.studio(preset, 1.0, 0.85, 0.95)
    pl = tracer.standard_round_brilliant()
    return tracer.FrameScene(cam, size, size, pl, material or tracer.GemMaterial.diamond(), 12, env,
                             facet_finishes=finishes)


def test_explicit_polished_equals_no_finishes_and_frosting_changes_the_image():
    a = oracle_lib.accumulate(_scene(None), 0, 8)
    b = oracle_lib.accumulate(_scene([0] * 74), 0, 8)
    assert a.tobytes() == b.tobytes()
    fr = [1 if i in GIRDLE else 0 for i in range(74)]
    c = oracle_lib.accumulate(_scene(fr), 0, 8)
    assert not np.array_equal(a, c)
    assert np.isfinite(c).all() and c.sum() > 0


def test_route_for_frosted_stones():
    fr = [1 if i in GIRDLE else 0 for i in range(74)]
    assert tracer.route(_scene(fr)) == abi.ROUTE_OK
    assert tracer.route(_scene(fr, "DaylightSun")) == abi.ROUTE_UNSUPPORTED_ENVIRONMENT
    assert tracer.route(_scene(None, "DaylightSun")) == abi.ROUTE_OK
    m = tracer.GemMaterial.diamond().to_abi()
    m.cubic = 0
    m.birefringence_delta = 0.009
    assert tracer.route(_scene(fr, material=m)) == abi.ROUTE_OK           # frosted uniaxial
    assert tracer.route(_scene(fr, material=tracer.GemMaterial.alexandrite())) == abi.ROUTE_OK
    assert tracer.route(_scene(fr, "DaylightSun", material=m)) == abi.ROUTE_UNSUPPORTED_ENVIRONMENT
    assert tracer.route(_scene(None, material=m)) == abi.ROUTE_OK


def _aniso(kind):
    from indicatrix_amd import workloads
    return (workloads.c2_material(bands=True) if kind == "uniaxial" else tracer.GemMaterial.alexandrite()).to_abi()


@pytest.mark.parametrize("kind", ["uniaxial", "biaxial"])
def test_frosted_anisotropic_stones(kind):
    fr = [1 if i in GIRDLE else 0 for i in range(74)]
    a = oracle_lib.accumulate(_scene(None, material=_aniso(kind)), 0, 8)
    b = oracle_lib.accumulate(_scene([0] * 74, material=_aniso(kind)), 0, 8)
    assert a.tobytes() == b.tobytes()
    c = oracle_lib.accumulate(_scene(fr, material=_aniso(kind)), 0, 8)
    assert not np.array_equal(a, c) and np.isfinite(c).all() and c.sum() > 0


@pytest.mark.parametrize("kind", ["isotropic", "uniaxial", "biaxial"])
def test_frosted_white_furnace(kind):
    # the frosted arm's weights (cosine-weighted direction, Fresnel-chosen side,
    # pdf by the clamped reflectance) conserve energy: a frosted, non-absorbing
    # stone in a uniform panorama returns its spectrum within the truncation
    from test_hdr_env import furnace_target
    m = tracer.GemMaterial("furnace probe", abi.DISP_CAUCHY, [1.5, 0.0, 0.0],
                           birefringence_delta=0.0 if kind == "isotropic" else 0.02, cubic=kind == "isotropic",
                           c_axis=(0.3, 0.8, 0.52))
    if kind == "biaxial":
        m.biaxial_delta_beta_alpha = 0.008
    fr = [1 if (i in GIRDLE or i >= 49) else 0 for i in range(74)]
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, 12, 12, 96)
    env = tracer.LightingPreset.studio("Daylight", 1.0, 0.4, 0.35)
    sc = tracer.FrameScene(cam, 12, 12, tracer.standard_round_brilliant(), m, 48, env, facet_finishes=fr)
    sc.set_hdr_map(np.full((1, 1, 3), 2.5, np.float32))
    acc = oracle_lib.accumulate(sc, 0, 96)
    mean = acc.sum(axis=0) / (12 * 12 * 96)
    target = furnace_target(2.5)
    assert (np.abs(mean - target) / target).max() < 0.015, (mean, target)
