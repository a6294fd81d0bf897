"""Henyey-Greenstein volume scattering (RT/scattering/hg.rs) in the oracle
and through the route (CPU).

The direction sampler is pinned by the phase function's closed forms: the
mean cosine of a Henyey-Greenstein lobe is g (1/3 for cos^2 at g = 0), every
direction is a unit vector, the azimuth is unbiased.  The free-flight step
is pinned by its limits: a vanishing scattering coefficient reaches every
facet with every channel's survival ratio exactly 1, so the frame equals the
non-scattering one bit for bit; a dense medium changes the image.  The
device follows the oracle path for path in tests/test_scattering_gpu.py.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, tracer, workloads

F3 = C.c_float * 3


def hg(u1, u2, g, fwd):
    out = F3()
    oracle_lib.lib().ixo_hg_direction(u1, u2, g, F3(*fwd), out)
    return np.array(out[:], np.float64)


@pytest.mark.parametrize("g", [0.0, 0.5, -0.3, 0.85])
def test_hg_sampler_mean_cosine_is_g(g):
    fwd = np.array([0.36, 0.48, 0.8])
    m = 128
    cos, az = [], np.zeros(2)
    for i in range(m):
        for j in range(m):
            d = hg((i + 0.5) / m, (j + 0.5) / m, g, fwd)
            assert abs(np.linalg.norm(d) - 1.0) < 1e-6
            c = d @ fwd
            cos.append(c)
            az += (d - fwd * c)[:2]
    cos = np.array(cos)
    assert abs(cos.mean() - g) < 2e-3, cos.mean()
    if g == 0.0:
        assert abs((cos ** 2).mean() - 1.0 / 3.0) < 2e-3
    assert np.abs(az).max() / (m * m) < 2e-3


def scattering_scene(sigma_s, g=0.6, material=None, preset="RingLights", size=24, spp=8):
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, size, size, spp)
    m = material or tracer.GemMaterial.diamond().to_abi()
    m.scattering_sigma_s = sigma_s
    m.scattering_g = g
    env = tracer.LightingPreset.studio(preset, 1.0, 0.85, 0.95)
    return tracer.FrameScene(cam, size, size, tracer.standard_round_brilliant(), m, 12, env)


def test_vanishing_scattering_is_the_clear_stone_bit_for_bit():
    clear = oracle_lib.accumulate(scattering_scene(0.0), 0, 8)
    faint = oracle_lib.accumulate(scattering_scene(1e-12), 0, 8)
    assert clear.tobytes() == faint.tobytes()


def test_dense_medium_changes_the_image():
    clear = oracle_lib.accumulate(scattering_scene(0.0), 0, 8)
    milky = oracle_lib.accumulate(scattering_scene(3.0), 0, 8)
    assert np.isfinite(milky).all()
    assert not np.array_equal(clear, milky)
    assert milky.sum() > 0


def test_route_for_scattering_materials():
    assert tracer.route(scattering_scene(2.0)) == abi.ROUTE_OK
    assert tracer.route(scattering_scene(2.0, preset="DaylightSun")) == abi.ROUTE_OK   # no sun sample at scatters
    assert tracer.route(scattering_scene(2.0, g=1.0)) == abi.ROUTE_UNSUPPORTED_MATERIAL
    m = tracer.GemMaterial.diamond().to_abi()
    m.cubic = 0
    m.birefringence_delta = 0.009
    assert tracer.route(scattering_scene(2.0, material=m)) == abi.ROUTE_OK          # uniaxial
    assert tracer.route(scattering_scene(2.0, material=tracer.GemMaterial.alexandrite().to_abi())) == abi.ROUTE_OK
    assert tracer.route(scattering_scene(2.0, material=m, preset="DaylightSun")) == abi.ROUTE_OK
    assert tracer.route(workloads.film_absorbing_scene(8, 8)) == abi.ROUTE_OK


def _aniso_material(kind, bands):
    if kind == "uniaxial":
        m = workloads.c2_material(bands=bands)
    else:
        m = tracer.GemMaterial.alexandrite()
        if not bands:
            m.bands_o, m.bands_e, m.bands_beta = [], [], None
    return m.to_abi()


@pytest.mark.parametrize("kind", ["uniaxial", "biaxial"])
def test_vanishing_scattering_in_anisotropic_stones_is_the_clear_stone(kind):
    # try_scatter_step runs for every interior segment of a scattering stone
    # (bounce_body.rs:410-416, hg.rs:435-651), anisotropic ones included:
    # without bands a vanishing coefficient leaves the frame bit for bit
    clear = oracle_lib.accumulate(scattering_scene(0.0, material=_aniso_material(kind, False)), 0, 8)
    faint = oracle_lib.accumulate(scattering_scene(1e-12, material=_aniso_material(kind, False)), 0, 8)
    assert clear.tobytes() == faint.tobytes()


@pytest.mark.parametrize("kind", ["uniaxial", "biaxial"])
def test_anisotropic_scattering_white_furnace(kind):
    # a non-absorbing scattering stone still returns the uniform panorama's
    # spectrum: scattering only redistributes (albedo 1)
    from test_hdr_env import furnace_target
    m = _aniso_material(kind, False)
    m.dispersion.kind = abi.DISP_CAUCHY
    m.dispersion.coeff[0], m.dispersion.coeff[1], m.dispersion.coeff[2] = 1.5, 0.0, 0.0
    m.scattering_sigma_s = 0.4
    m.scattering_g = 0.5
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, 12, 12, 96)
    env = tracer.LightingPreset.studio("Daylight", 1.0, 0.4, 0.35)
    sc = tracer.FrameScene(cam, 12, 12, tracer.standard_round_brilliant(), m, 48, env)
    sc.set_hdr_map(np.full((1, 1, 3), 2.5, np.float32))
    acc = oracle_lib.accumulate(sc, 0, 96)
    mean = acc.sum(axis=0) / (12 * 12 * 96)
    target = furnace_target(2.5)
    assert (np.abs(mean - target) / target).max() < 0.015, (mean, target)


@pytest.mark.parametrize("kind", ["uniaxial", "biaxial"])
def test_anisotropic_scattering_absorbs_through_the_assigned_mode(kind):
    # with bands, the scatter step's coefficients are the assigned mode's
    # (absorption.rs:224-249): the frame differs from the clear stone's and
    # from the band-free stone's
    base = oracle_lib.accumulate(scattering_scene(0.8, material=_aniso_material(kind, False)), 0, 8)
    banded = oracle_lib.accumulate(scattering_scene(0.8, material=_aniso_material(kind, True)), 0, 8)
    assert np.all(np.isfinite(banded)) and banded.sum() < base.sum()
