"""The DaylightSun and ASET lighting models in the oracle (CPU).

Pinned by the reference's own environment tests (RT/environment/tests.rs):
outside the 0.27 degree disc the sun preset is the dome bit for bit, inside
it adds SUN_RADIANCE times the illuminant (within 1e-4), and a key below the
horizon has no sun (:495-560); the ASET view lights green up to 45 degrees of
elevation, red to 75 and blue above, each a 20 nm FWHM band peaking at its
centre, and is black below the horizon (:1310-1380).  The route accepts both
presets and still declines the UV lamps, as scene_routes_to_gpu does.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi

SUN, DOME, ASET = abi.PRESETS["DaylightSun"], abi.PRESETS["DaylightDome"], abi.PRESETS["Aset"]


def sample(preset, d, lam, exposure=1.0, yaw=0.4, pitch=0.35):
    f = oracle_lib.lib().ixo_sample_studio_environment
    return np.float32(f((C.c_float * 3)(*[float(x) for x in d]), lam, preset, exposure, yaw, pitch))


def key_dir(yaw, pitch):
    """StudioRig::new's key direction (optics/studio_rig.rs:47-80), f32."""
    cp, sp = np.float32(math.cos(np.float32(pitch))), np.float32(math.sin(np.float32(pitch)))
    cy, sy = np.float32(math.cos(np.float32(yaw))), np.float32(math.sin(np.float32(yaw)))
    v = np.array([cp * sy, sp, cp * cy], np.float32)
    return v / np.float32(np.sqrt(np.float32(v @ v)))


def tilted_from(k, degrees):
    """k turned by `degrees` about an axis perpendicular to it."""
    k = np.asarray(k, np.float64)
    a = np.array([0.0, 1.0, 0.0]) if abs(k[0]) > 0.9 else np.array([1.0, 0.0, 0.0])
    t = a - k * (k @ a)
    t /= np.linalg.norm(t)
    r = math.radians(degrees)
    return (k * math.cos(r) + t * math.sin(r)).astype(np.float32)


def test_sun_disc_adds_the_sun_and_only_inside_the_disc():
    exposure = 1.3
    spd = oracle_lib.lib().ixo_d65(560.0)
    for yaw, pitch in [(0.3, 0.6), (0.85, 1.2566), (-0.5, 1.2)]:
        k = key_dir(yaw, pitch)
        for deg in [0.0, 0.1, 0.2]:
            d = tilted_from(k, deg)
            sun = float(sample(SUN, d, 560.0, exposure, yaw, pitch)) - float(sample(DOME, d, 560.0, exposure, yaw, pitch))
            expected = 40000.0 * (spd * exposure)
            assert abs(sun / expected - 1.0) < 1e-4, (yaw, pitch, deg, sun, expected)
        for deg in [0.3, 0.5, 2.0, 30.0, 120.0]:
            d = tilted_from(k, deg)
            for lam in [450.0, 560.0, 650.0]:
                a = sample(SUN, d, lam, exposure, yaw, pitch)
                b = sample(DOME, d, lam, exposure, yaw, pitch)
                assert a.tobytes() == b.tobytes(), (yaw, pitch, deg, lam)


def test_sun_below_the_horizon_is_off():
    for pitch in [-0.3, -0.06]:
        k = key_dir(0.3, pitch)
        assert sample(SUN, k, 560.0, 1.0, 0.3, pitch).tobytes() == sample(DOME, k, 560.0, 1.0, 0.3, pitch).tobytes()
    # at the horizon itself the disc carries half its radiance (the smoothstep midpoint)
    k = key_dir(0.3, 0.0)
    half = float(sample(SUN, k, 560.0, 1.0, 0.3, 0.0)) - float(sample(DOME, k, 560.0, 1.0, 0.3, 0.0))
    assert abs(half / (20000.0 * oracle_lib.lib().ixo_d65(560.0)) - 1.0) < 1e-4


def aset_at(elevation_deg, lam):
    e = math.radians(elevation_deg)
    return float(sample(ASET, [math.cos(e), math.sin(e), 0.0], lam))


def test_aset_zones_follow_the_elevation():
    RED, GREEN, BLUE = 610.0, 540.0, 460.0
    for elevation, lit, dark in [(20.0, GREEN, [RED, BLUE]), (60.0, RED, [GREEN, BLUE]),
                                 (82.0, BLUE, [RED, GREEN]), (89.0, BLUE, [RED, GREEN])]:
        on = aset_at(elevation, lit)
        assert on > 0.5, (elevation, on)
        for other in dark:
            assert aset_at(elevation, other) < on * 1e-3, (elevation, other)


def test_aset_bands_peak_at_their_centres_with_20nm_fwhm():
    for elevation, centre in [(20.0, 540.0), (60.0, 610.0), (82.0, 460.0)]:
        peak = aset_at(elevation, centre)
        for step in range(1, 40):
            off = step * 0.5
            assert aset_at(elevation, centre + off) < peak
            assert aset_at(elevation, centre - off) < peak
        for side in [-10.0, 10.0]:
            assert abs(aset_at(elevation, centre + side) / peak - 0.5) < 5e-3
        assert aset_at(elevation, centre + 80.0) < peak * 1e-6


def test_aset_is_black_below_the_horizon():
    for elevation in [-90.0, -45.0, -10.0, -6.0]:
        for lam in [460.0, 540.0, 610.0, 550.0]:
            assert aset_at(elevation, lam) == 0.0


@pytest.mark.parametrize("preset,routed", [(6, True), (12, True), (13, False), (14, False)])
def test_route_accepts_sun_and_aset_and_declines_the_uv_lamps(preset, routed):
    from indicatrix_amd import tracer
    from indicatrix_amd._lib import lib
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, 8, 8, 1)
    env = abi.Environment(preset, 1.0, 0.4, 0.35, 0.0, 1.0, 16.0)
    sc = tracer.FrameScene(cam, 8, 8, tracer.standard_round_brilliant(), tracer.GemMaterial.diamond(), 8, env)
    assert (lib().ixg_route(sc.ref()) == 0) is routed
