"""Frosted facets (FacetFinish::Frosted, RT/scattering/frosted.rs) in the
oracle and through the route (CPU).

The oracle's cosine-weighted hemisphere is pinned by what the reference's
construction fixes: unit directions on the normal's side, the pole at u1 = 0,
and the cosine-law moments (E[cos] = 2/3, E[cos^2] = 1/2, no azimuthal
bias).  Polished finishes given explicitly equal no finishes bit for bit; a
frosted girdle changes the image.  The device follows the oracle path for
path in tests/test_frosted_gpu.py.  Next-event estimation at frosted events
is not carried, so the route declines a frosted stone under DaylightSun, and
frosted uniaxial stones, for which the transport has no frosted arm.
"""
import ctypes as C
import math

import numpy as np

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
    assert tracer.route(_scene(fr, material=m)) == abi.ROUTE_UNSUPPORTED_MATERIAL
    assert tracer.route(_scene(None, material=m)) == abi.ROUTE_OK
