"""HDR panorama environments (EnvironmentSource::HdrMap) in the oracle (CPU).

The analytic furnace anchor: a non-absorbing dielectric inside a uniform
environment returns that environment's radiance in every direction, so the
mean pixel equals the panorama spectrum integrated against the CIE
observer (RT/intersect.rs:530-615 runs it with rounded edges; 16 bounces
lose 0.6 %, the bound is 1 %).  rgb_to_spectral_radiance is pinned by its
construction: a grey input uses only the broad bumps and scales linearly, a
negative input is clamped to zero, and it is never negative.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, tracer

F3 = C.c_float * 3


def spectral(rgb, lam):
    return oracle_lib.lib().ixo_rgb_to_spectral(F3(*rgb), lam)


def furnace_target(L0):
    cmf = (C.c_float * 3)()
    t = np.zeros(3)
    for step in range(0, 401):
        lam = 380.0 + step
        oracle_lib.lib().ixo_cie_cmf(lam, cmf)
        t += np.array(cmf[:]) * spectral([L0, L0, L0], lam)
    return t / 106.856


def furnace_scene(radius=0.0, grid=12, spp=96):
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, grid, grid, spp)
    mat = tracer.GemMaterial("furnace probe", abi.DISP_CAUCHY, [1.5, 0.0, 0.0]).to_abi()
    mat.edge_rounding_radius = radius
    env = tracer.LightingPreset.studio("Daylight", 1.0, 0.4, 0.35)
    sc = tracer.FrameScene(cam, grid, grid, tracer.standard_round_brilliant(), mat, 16, env)
    sc.set_hdr_map(np.full((1, 1, 3), 2.5, np.float32))
    return sc


@pytest.mark.parametrize("radius", [0.0, 0.03])
def test_white_furnace(radius):
    sc = furnace_scene(radius)
    acc = oracle_lib.accumulate(sc, 0, 96)
    mean = acc.sum(axis=0) / (12 * 12 * 96)
    target = furnace_target(2.5)
    rel = np.abs(mean - target) / target
    assert rel.max() < 0.01, (mean, target, rel)


def test_rgb_to_spectral_construction():
    lams = np.arange(380.0, 781.0, 5.0)
    g1 = np.array([spectral([1.0, 1.0, 1.0], l) for l in lams])
    g3 = np.array([spectral([3.0, 3.0, 3.0], l) for l in lams])
    assert np.all(g1 > 0.0)
    assert np.allclose(g3, 3.0 * g1, rtol=1e-6)
    assert all(spectral([-1.0, -2.0, -0.5], l) == 0.0 for l in lams)
    red = np.array([spectral([1.0, 0.0, 0.0], l) for l in lams])
    blue = np.array([spectral([0.0, 0.0, 1.0], l) for l in lams])
    assert 600.0 <= lams[np.argmax(red)] <= 660.0
    assert 430.0 <= lams[np.argmax(blue)] <= 470.0
    assert np.all(red >= 0.0) and np.all(blue >= 0.0)


def test_route_for_panoramas():
    sc = furnace_scene()
    assert tracer.route(sc) == abi.ROUTE_OK
    sc.material_abi.scattering_sigma_s = 1.0
    assert tracer.route(sc) == abi.ROUTE_UNSUPPORTED_ENVIRONMENT
    sc.material_abi.scattering_sigma_s = 0.0
    sc.set_finishes([1] * 74)
    assert tracer.route(sc) == abi.ROUTE_UNSUPPORTED_ENVIRONMENT
