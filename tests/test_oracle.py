"""Pins the CPU oracle before anything is checked against it.

Every check below is either a golden value held by the reference's own tests
(tests/golden/reference_goldens.json, extracted by the script beside it) or a
closed form / invariant chosen so that a dropped term or wrong index fails.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "reference_goldens.json")))


def f32(x):
    return float(np.float32(x))


def bits(x):
    return int(np.array([x], np.float32).view(np.uint32)[0])


def glam_normalize(v):
    v = np.asarray(v, np.float32)
    d = np.float32(np.float32(v[0] * v[0]) + np.float32(v[1] * v[1])) + np.float32(v[2] * v[2])
    return v * (np.float32(1.0) / np.sqrt(np.float32(d)))


def diamond_material():
    m = abi.Material()
    m.dispersion.kind = abi.DISP_SELLMEIER3
    for i, v in enumerate([4.3356, 0.3306, 0.0, 0.011236, 0.030625, 1.0]):
        m.dispersion.coeff[i] = v
    m.c_axis[1] = 1.0
    m.cubic = 1
    m.absorption_path_scale = 1.0
    return m


def golden_scene(preset, yaw, pitch, max_bounces=12):
    L = oracle_lib.lib()
    planes = (abi.FacetPlane * 80)()
    n = L.ixo_cut_standard_round_brilliant(planes)
    mat = diamond_material()
    sc = abi.Scene()
    sc.planes = C.cast(planes, C.POINTER(abi.FacetPlane))
    sc.n_planes = n
    sc.material = C.pointer(mat)
    sc.max_bounces = max_bounces
    sc.environment = abi.Environment(abi.PRESETS[preset], 1.0, yaw, pitch, 0.0, 1.0, 16.0)
    sc._keep = (planes, mat)
    return sc


def hero_of(seed):
    return f32(np.float32(oracle_lib.lib().ixo_hash_u32(seed)) / np.float32(4294967295.0))


@pytest.mark.parametrize("material", ["diamond", "sapphire"])
@pytest.mark.parametrize("row", range(4))
def test_trace_spectral_ray_matches_reference_golden_bits(row, material):
    """golden_regression.rs:47-130: exact f32 bits of both columns.  Diamond is
    the cubic isotropic transport; Sapphire (uniaxial, n_e - n_o = -0.0081,
    pleochroic energy-domain bands) pins the uniaxial entry / internal solves,
    the eigenmode draws, walk-off, o/e re-coupling and the assigned-mode
    absorption."""
    from indicatrix_amd import tracer
    t = GOLDEN["golden_regression"]["traces"][row]
    yaw = math.pi if t["light_yaw"] == "PI" else t["light_yaw"]
    sc = golden_scene(t["preset"], yaw, t["light_pitch"])
    if material == "sapphire":
        mat = tracer.GemMaterial.sapphire().to_abi()
        sc.material = C.pointer(mat)
        sc._keep = sc._keep + (mat,)
    d = glam_normalize(t["dir"]) if t["normalize"] else np.asarray(t["dir"], np.float32)
    xyz = (C.c_float * 3)()
    oracle_lib.lib().ixo_trace_ray(C.byref(sc), (C.c_float * 3)(*t["origin"]), (C.c_float * 3)(*d),
                                   t["seed"], hero_of(t["seed"]), xyz, None)
    assert [bits(v) for v in xyz] == t[f"{material}_bits"]


def test_studio_environment_matches_reference_golden_bits():
    """golden_regression.rs:180-205: sample_studio_environment, RingLights."""
    e = GOLDEN["golden_regression"]["environment"]
    L = oracle_lib.lib()
    for d, norm, expect in zip(e["dirs"], e["normalize"], e["bits"]):
        v = glam_normalize(d) if norm else np.asarray(d, np.float32)
        got = [bits(L.ixo_sample_studio_environment((C.c_float * 3)(*v), lam, abi.PRESETS["RingLights"],
                                                    1.0, 0.85, 0.95)) for lam in e["lambdas"]]
        assert got == expect


def fibonacci_dirs():
    libm = C.CDLL("libm.so.6")
    libm.cosf.restype = libm.sinf.restype = C.c_float
    libm.cosf.argtypes = libm.sinf.argtypes = [C.c_float]
    golden = np.float32(np.float32(math.pi) * (np.float32(3.0) - np.sqrt(np.float32(5.0))))
    out = []
    for k in range(64):
        kf = np.float32(k) + np.float32(0.5)
        phi = np.float32(kf * golden)
        y = np.float32(np.float32(1.0) - kf * np.float32(0.03125))  # fma exact: product exact
        r = np.sqrt(np.float32(max(1.0 - float(y) * float(y), 0.0)))
        out.append((np.float32(r * np.float32(libm.cosf(phi))), y,
                    np.float32(r * np.float32(libm.sinf(phi)))))
    return out


@pytest.mark.parametrize("row", range(7))
def test_environment_baseline_rows(row):
    """environment/tests.rs:786-1160 BASELINE_BITS, rows Daylight..DaylightDome (the UV
    lamps are CPU-only in the reference and not on this path), <= 2 ULP."""
    eb = GOLDEN["environment_baseline"]
    name = eb["order"][row]
    L = oracle_lib.lib()
    idx, worst = 0, 0
    for d in fibonacci_dirs():
        for lam in eb["lambdas"]:
            v = L.ixo_sample_studio_environment((C.c_float * 3)(*d), lam, abi.PRESETS[name],
                                                eb["exposure"], eb["light_yaw"], eb["light_pitch"])
            worst = max(worst, abs(bits(v) - eb["bits"][row][idx]))
            idx += 1
    assert worst <= eb["max_ulp"]


def fnv1a_xyz(acc32):
    """fnv1a_hash_xyz (renderer/gpu/pin_tests.rs:19-33): FNV-1a over the
    little-endian bytes of every pixel's X, Y, Z f32 bits."""
    h = 0xcbf29ce484222325
    for byte in np.ascontiguousarray(acc32, "<f4").tobytes():
        h = ((h ^ byte) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def cpu_accumulate_f32(scene, sample_offset, spp):
    """cpu_accumulate (renderer/gpu/hybrid.rs:305) in its own arithmetic: per
    pixel, the finite samples summed in f32 in sample order (sampling.rs:237),
    added into a zeroed f32 buffer."""
    xyz, _ = oracle_lib.trace_paths(scene, sample_offset, spp)
    xyz = xyz.reshape(-1, spp, 3)
    acc = np.zeros((xyz.shape[0], 3), np.float32)
    for p in range(xyz.shape[0]):
        s = np.zeros(3, np.float32)
        for k in range(spp):
            if np.all(np.isfinite(xyz[p, k])):
                s = (s + xyz[p, k]).astype(np.float32)
        acc[p] = (acc[p] + s).astype(np.float32)
    return acc


def test_cpu_accumulate_diamond_light_tent_pin():
    """pin_tests.rs:128-141: the 24x16 Diamond round brilliant in the light tent
    (studio(1.0, 0.4, 0.35) over the 0.23 backdrop tray), 8 bounces, camera
    (0.35, 0.28, 5.0, 18 deg), 3 spp at sample offset 0 -- the whole-frame FNV-1a
    hash of the f32 sums.  The reference records that the offset-0 hash is the
    same under glibc and the Windows runtime (pin_tests.rs:116-126); offset 5 is
    not checked for that reason."""
    from indicatrix_amd import tracer
    L = oracle_lib.lib()
    planes = (abi.FacetPlane * 80)()
    n = L.ixo_cut_standard_round_brilliant(planes)
    cam = abi.Camera()
    L.ixo_camera_new(0.35, 0.28, 5.0, 18.0, 24, 16, 3, C.byref(cam))
    env = abi.Environment(abi.PRESETS["LightTent"], 1.0, 0.4, 0.35, 0.23, 1.0, 16.0)
    sc = tracer.FrameScene(cam, 24, 16, (abi.FacetPlane * n)(*planes[:n]), diamond_material(), 8, env)
    assert fnv1a_xyz(cpu_accumulate_f32(sc, 0, 3)) == 0x99c71f3d943d7cdb


def test_sample_draws_hand_computed():
    """sampling.rs:391-438: hand-evaluated radical inverses and the zero seed."""
    L = oracle_lib.lib()
    assert L.ixo_hash_u32(0) == 0
    assert L.ixo_low_discrepancy_base2(1) == 0.5
    assert L.ixo_low_discrepancy_base2(2) == 0.25
    assert abs(L.ixo_radical_inverse_base(1, 3) - 1 / 3) < 1e-6
    assert abs(L.ixo_radical_inverse_base(2, 3) - 2 / 3) < 1e-6
    assert abs(L.ixo_radical_inverse_base(1, 5) - 0.2) < 1e-6
    assert abs(L.ixo_radical_inverse_base(2, 5) - 0.4) < 1e-6
    seed, jx, jy, hero = C.c_uint32(), C.c_float(), C.c_float(), C.c_float()
    for pixel in range(50):
        for s in range(16):
            L.ixo_sample_draws(pixel, s, C.byref(seed), C.byref(jx), C.byref(jy), C.byref(hero))
            assert -0.5 <= jx.value < 0.5 and -0.5 <= jy.value < 0.5 and 0.0 <= hero.value < 1.0
    L.ixo_sample_draws(0, 0, C.byref(seed), C.byref(jx), C.byref(jy), C.byref(hero))
    assert seed.value == 0


def test_hero_comb_wraps_in_band():
    """transport/mod.rs:58-66: hero in slot 0, 50 nm apart mod 400, inside [380, 780)."""
    L = oracle_lib.lib()
    out = (C.c_float * 8)()
    for h in np.linspace(0.0, 0.999, 37):
        L.ixo_wrapped_hero_wavelengths(float(h), out)
        lam = np.array(out[:])
        assert abs(lam[0] - (380.0 + 400.0 * h)) < 1e-3
        assert np.all(lam >= 380.0) and np.all(lam < 780.0 + 1e-3)
        steps = np.mod(np.diff(lam), 400.0)
        assert np.allclose(steps, 50.0, atol=1e-3)


def test_diamond_dispersion_closed_forms():
    """diamond_through_emerald.rs:33-34: n_d = 2.41726, V_d = 55.27."""
    L = oracle_lib.lib()
    d = abi.Dispersion(abi.DISP_SELLMEIER3, (C.c_float * 6)(4.3356, 0.3306, 0.0, 0.011236, 0.030625, 1.0))
    nd = L.ixo_dispersion_evaluate(C.byref(d), 589.3)
    nf = L.ixo_dispersion_evaluate(C.byref(d), 486.1)
    nc = L.ixo_dispersion_evaluate(C.byref(d), 656.3)
    assert abs(nd - 2.41726) < 2e-5
    assert abs((nd - 1.0) / (nf - nc) - 55.27) < 0.05


def test_cmf_y_integral():
    """tests/raytracer_tests/fixtures.rs:54-66: integral of y-bar at 1 nm = 106.856."""
    L = oracle_lib.lib()
    out = (C.c_float * 3)()
    total = 0.0
    for lam in range(380, 781):
        L.ixo_cie_cmf(float(lam), out)
        total += out[1]
    assert abs(total - 106.856) < 0.01


def test_exp_lane_accuracy_and_identity():
    """simd/exp_poly.rs: exp_lane(0) == 1 exactly (absorption-free media are a no-op),
    and a few-ULP polynomial elsewhere."""
    L = oracle_lib.lib()
    assert L.ixo_exp_lane(0.0) == 1.0
    assert L.ixo_exp_lane(-0.0) == 1.0
    for x in np.linspace(-20, 20, 101):
        got = L.ixo_exp_lane(float(np.float32(x)))
        ref = math.exp(float(np.float32(x)))
        assert abs(got - ref) <= 4e-7 * ref


def test_slab_scan_unit_cube_brute_force():
    """RT/intersect.rs:231-260 style: a unit cube's entry/exit match the analytic slab."""
    L = oracle_lib.lib()
    cube = (abi.FacetPlane * 6)()
    for i, (n, d) in enumerate([((1, 0, 0), -1), ((-1, 0, 0), -1), ((0, 1, 0), -1),
                                ((0, -1, 0), -1), ((0, 0, 1), -1), ((0, 0, -1), -1)]):
        cube[i] = abi.FacetPlane((C.c_float * 3)(*n), d)
    rng = np.random.default_rng(5)
    t = C.c_float()
    for _ in range(200):
        o = rng.uniform(-4, 4, 3).astype(np.float32)
        d = rng.normal(size=3).astype(np.float32)
        d /= np.linalg.norm(d)
        f = L.ixo_intersect(cube, 6, (C.c_float * 3)(*o), (C.c_float * 3)(*d), C.byref(t))
        # analytic slab
        tn, tf = -1e30, 1e30
        for a in range(3):
            if abs(d[a]) > 1e-7:
                t1, t2 = (-1 - o[a]) / d[a], (1 - o[a]) / d[a]
                tn, tf = max(tn, min(t1, t2)), min(tf, max(t1, t2))
            elif abs(o[a]) > 1:
                tn, tf = 1, 0
        if tn > tf or tf <= 1e-4:
            assert f < 0
        else:
            expect = tn if tn > 1e-4 else tf
            assert f >= 0 and abs(t.value - expect) < 1e-4


def test_interleaved_partition_is_thread_count_invariant():
    """cpu_frame.rs:84-103: any interleaved pixel partition gives the same sums."""
    L = oracle_lib.lib()
    sc = golden_scene("RingLights", 0.85, 0.95, max_bounces=8)
    cam = abi.Camera()
    L.ixo_camera_new(0.35, 0.28, 5.0, 18.0, 12, 10, 4, C.byref(cam))
    sc.camera = cam
    sc.width, sc.height = 12, 10

    class W:
        width, height = 12, 10

        def ref(self):
            return C.byref(sc)
    a = oracle_lib.accumulate(W(), 0, 4, nthreads=1)
    b = oracle_lib.accumulate(W(), 0, 4, nthreads=3)
    assert np.array_equal(a, b)
    assert a.sum() > 0


# --------------------------------------------------------------------------
# metric images: structural properties the reference's own tests pin
# (color/metrics/image_tests.rs, tests/raytracer_tests/optical_metrics.rs)
# --------------------------------------------------------------------------
def _metric_request(material=None, w=24, h=20, samples=6, want_fire=True, scint=False,
                    tile=None, proj=None, preset="IsoHemisphere"):
    from indicatrix_amd import metrics, tracer
    planes = (abi.FacetPlane * 80)()
    n = oracle_lib.lib().ixo_cut_standard_round_brilliant(planes)
    planes = (abi.FacetPlane * n)(*planes[:n])
    cam = abi.Camera()
    pw, ph = proj or (w, h)
    oracle_lib.lib().ixo_camera_new(0.3, 0.9, 2.4, 42.0, pw, ph, samples, C.byref(cam))
    env = abi.Environment(abi.PRESETS[preset], 1.0, 0.85, 0.95, 0.0, 1.0, 16.0)
    mat = material or diamond_material()
    r = metrics.make_request(cam, w, h, samples, planes, mat, env, want_fire=want_fire,
                             want_scintillation=scint, pose=(0.3, 0.9, 2.4))
    if proj:
        r.projection[0], r.projection[1] = proj
    if tile:
        r.has_tile_origin = 1
        r.tile_origin[0], r.tile_origin[1] = tile
    return r


def test_metric_counts_conserve_hits_and_repeat():
    """image_tests.rs:10-37."""
    r = _metric_request()
    a, ta, _ = oracle_lib.metric_image(r)
    b, tb, _ = oracle_lib.metric_image(r)
    assert np.array_equal(a["classes"].sum(axis=1), a["hits"])
    assert np.array_equal(a.tobytes(), b.tobytes())
    assert np.all(a["hits"] <= r.samples)
    assert np.any(a["hits"] == 0) and ta.total.hits > 0


def test_zero_dispersion_has_no_fire_and_fire_preserves_counts():
    """image_tests.rs:70-99: a flat index gives no fire; fire on/off keeps class counts."""
    flat = diamond_material()
    flat.dispersion = abi.Dispersion(abi.DISP_CAUCHY, (C.c_float * 6)(2.4, 0.0, 0.0, 0, 0, 0))
    px, tot, _ = oracle_lib.metric_image(_metric_request(flat))
    assert tot.total.classes[2] > 0 and np.all(px["fire_sum_deg"] == 0.0)
    base, tb, _ = oracle_lib.metric_image(_metric_request(want_fire=False))
    fire, tf, _ = oracle_lib.metric_image(_metric_request(want_fire=True))
    assert list(tb.total.classes) == list(tf.total.classes)
    assert tf.total.fire_sum_deg > 0.0


def test_row_tiles_match_full_projection():
    """image_tests.rs:102-137: a tile of the projection classifies exactly like those rows
    of the full map (screen coordinates are rounded once per path, so Fire sums agree to
    rounding)."""
    full, _, _ = oracle_lib.metric_image(_metric_request(w=24, h=20))
    tile, _, _ = oracle_lib.metric_image(_metric_request(w=24, h=5, tile=(0, 8), proj=(24, 20)))
    rows = full.reshape(20, 24)[8:13].reshape(-1)
    for f in ("hits", "classes", "centre_entry_facet", "centre_exit_facet", "centre_bounces"):
        assert np.array_equal(tile[f], rows[f]), f
    assert np.allclose(tile["fire_sum_deg"], rows["fire_sum_deg"], rtol=2e-3, atol=1e-5)


def test_centre_sample_hits_the_camera_ray_facet():
    """image_tests.rs:227-262: sample 0 is the undisturbed pixel-centre camera ray."""
    r = _metric_request(w=16, h=12)
    px, _, _ = oracle_lib.metric_image(r)
    L = oracle_lib.lib()
    o, d, t = (C.c_float * 3)(), (C.c_float * 3)(), C.c_float()
    for y in range(12):
        for x in range(16):
            cam = r.camera
            cam.width, cam.height = 16.0, 12.0
            L.ixo_generate_ray(C.byref(cam), float(x), float(y), 0.0, 0.0, o, d)
            f = L.ixo_intersect(r.planes, r.n_planes, o, d, C.byref(t))
            assert px["centre_entry_facet"][y * 16 + x] == (f if f >= 0 else -1)


def test_high_dispersion_has_more_fire():
    """tests/raytracer_tests/optical_metrics.rs / metrics_tests.rs:139: fire orders by dispersion."""
    lo = diamond_material()
    lo.dispersion = abi.Dispersion(abi.DISP_CAUCHY, (C.c_float * 6)(2.40, 0.004, 0.0, 0, 0, 0))
    hi = diamond_material()
    hi.dispersion = abi.Dispersion(abi.DISP_CAUCHY, (C.c_float * 6)(2.40, 0.030, 0.0, 0, 0, 0))
    _, tl, _ = oracle_lib.metric_image(_metric_request(lo, samples=8))
    _, th, _ = oracle_lib.metric_image(_metric_request(hi, samples=8))
    assert th.fire_index > 2.0 * tl.fire_index


# ---------------------------------------------------------------------------
# absorption (optics/absorption.rs, RT/absorption.rs): bands and Beer-Lambert
# ---------------------------------------------------------------------------
def test_band_closed_forms():
    """AbsorptionBand::evaluate: peak at the centre, exp(-1/2) one width away,
    in both the wavelength and the energy (wavenumber) domain."""
    L = oracle_lib.lib()
    b = abi.Band(550.0, 30.0, 1.7, 0)
    assert L.ixo_band_evaluate(C.byref(b), 550.0) == f32(1.7)
    assert L.ixo_band_evaluate(C.byref(b), 580.0) == pytest.approx(1.7 * math.exp(-0.5), rel=2e-7)
    assert L.ixo_band_evaluate(C.byref(b), 520.0) == pytest.approx(1.7 * math.exp(-0.5), rel=2e-7)
    e = abi.Band(600.0, 1500.0, 0.9, 1)
    assert L.ixo_band_evaluate(C.byref(e), 600.0) == f32(0.9)
    lam = 1e7 / (1e7 / 600.0 + 1500.0)   # one sigma in wavenumber
    assert L.ixo_band_evaluate(C.byref(e), lam) == pytest.approx(0.9 * math.exp(-0.5), rel=2e-5)


CUBE = [((1, 0, 0), -1), ((-1, 0, 0), -1), ((0, 1, 0), -1), ((0, -1, 0), -1), ((0, 0, 1), -1),
        ((0, 0, -1), -1)]


def _cube_scene(mat, max_bounces=12):
    planes = (abi.FacetPlane * 6)()
    for i, (n, d) in enumerate(CUBE):
        planes[i] = abi.FacetPlane((C.c_float * 3)(*n), d)
    sc = abi.Scene()
    sc.planes = C.cast(planes, C.POINTER(abi.FacetPlane))
    sc.n_planes = 6
    sc.material = C.pointer(mat)
    sc.max_bounces = max_bounces
    sc.environment = abi.Environment(abi.PRESETS["IsoHemisphere"], 1.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    sc._keep = (planes, mat)
    return sc


def _flat_bands(mat, alpha_o, alpha_e, c_axis=(0.0, 1.0, 0.0), scale=1.0):
    """Bands so wide that every channel's coefficient is exactly the peak."""
    mat.n_bands_o = 1
    mat.n_bands_e = 1
    mat.bands_o[0] = abi.Band(550.0, 1e7, alpha_o, 0)
    mat.bands_e[0] = abi.Band(550.0, 1e7, alpha_e, 0)
    for i in range(3):
        mat.c_axis[i] = c_axis[i]
    mat.absorption_path_scale = scale
    return mat


def _straight_paths(sc, seeds):
    """Y of rays sent straight down the z axis through the cube, keyed by seed,
    for the paths that enter and leave once (no internal reflection)."""
    out = {}
    for seed in seeds:
        xyz = (C.c_float * 3)()
        ev = abi.PathEvents()
        oracle_lib.lib().ixo_trace_ray(C.byref(sc), (C.c_float * 3)(0.1, 0.2, 5.0),
                                       (C.c_float * 3)(0.0, 0.0, -1.0), seed, hero_of(seed), xyz,
                                       C.byref(ev))
        if ev.n_refract == 2 and ev.n_reflect == 0 and ev.n_tir == 0 and ev.termination == 0:
            out[seed] = (xyz[1], ev.facet_hash)
    return out


@pytest.mark.parametrize("c_axis,alpha", [((0.0, 0.0, 1.0), 0.2), ((1.0, 0.0, 0.0), 0.4),
                                          ((0.0, 1.0, 0.0), 0.4)])
def test_beer_lambert_through_a_slab(c_axis, alpha):
    """A ray crossing the 2-unit cube once transmits exp(-alpha L) with
    L = 2 - 1e-4 (the self-hit offset).  With alpha_o = 0.2, alpha_e = 0.6 in a
    cubic material the coefficient is the unpolarised mean of the two eigen-
    directions' quadratic forms: alpha_o along the optic axis, the midpoint
    (0.4) across it."""
    clear = _straight_paths(_cube_scene(diamond_material()), range(64))
    mat = _flat_bands(diamond_material(), 0.2, 0.6, c_axis)
    dark = _straight_paths(_cube_scene(mat), range(64))
    assert len(clear) >= 10 and clear.keys() == dark.keys()
    expect = math.exp(-alpha * (2.0 - 1e-4))
    for s in clear:
        assert dark[s][1] == clear[s][1]                 # same facet sequence
        assert dark[s][0] / clear[s][0] == pytest.approx(expect, rel=2e-6)


def test_absorption_path_scale_squares_the_transmittance():
    clear = _straight_paths(_cube_scene(diamond_material()), range(48))
    one = _straight_paths(_cube_scene(_flat_bands(diamond_material(), 0.3, 0.3)), range(48))
    two = _straight_paths(_cube_scene(_flat_bands(diamond_material(), 0.3, 0.3, scale=2.0)), range(48))
    for s in clear:
        r1 = one[s][0] / clear[s][0]
        r2 = two[s][0] / clear[s][0]
        assert r2 == pytest.approx(r1 * r1, rel=4e-6)


def test_zero_peak_bands_are_exactly_transparent():
    sc0 = _cube_scene(diamond_material())
    sc1 = _cube_scene(_flat_bands(diamond_material(), 0.0, 0.0))
    for seed in range(32):
        a = (C.c_float * 3)()
        b = (C.c_float * 3)()
        d = (C.c_float * 3)(0.3, -0.2, -1.0)
        oracle_lib.lib().ixo_trace_ray(C.byref(sc0), (C.c_float * 3)(0.1, 0.2, 5.0), d, seed, hero_of(seed), a, None)
        oracle_lib.lib().ixo_trace_ray(C.byref(sc1), (C.c_float * 3)(0.1, 0.2, 5.0), d, seed, hero_of(seed), b, None)
        assert [bits(v) for v in a] == [bits(v) for v in b]



# ---------------------------------------------------------------------------
# uniaxial transport (raytracer/uniaxial_fresnel/, refraction/uniaxial_*.rs)
# ---------------------------------------------------------------------------
F3 = C.c_float * 3


def _neff(n_o, n_e, theta):
    return n_o * n_e / math.sqrt((n_e * math.cos(theta)) ** 2 + (n_o * math.sin(theta)) ** 2)


def test_uniaxial_interfaces_conserve_energy():
    """Lekner's boundary solution: R + T = 1 (Poynting flux along the normal)
    at entry for s and p incidence, and inside for an o or e wave meeting the
    surface, total internal reflection included (T = 0, R = 1)."""
    L = oracle_lib.lib()
    rng = np.random.default_rng(8)
    out = (C.c_float * 2)()
    worst, n_tir = 0.0, 0
    for _ in range(400):
        c = rng.normal(size=3)
        c /= np.linalg.norm(c)
        th = rng.uniform(0.0, 1.5)
        k = (math.sin(th), 0.0, -math.cos(th))
        for pol in (0, 1):
            L.ixo_uniaxial_energy(0, 1.0, 1.64, 1.66, F3(*c), F3(*k), F3(0, 0, 1), pol, out)
            worst = max(worst, abs(out[0] + out[1] - 1.0))
            n_inc = 2.0 if pol == 0 else _neff(2.0, 2.15, math.acos(abs(float(np.dot(k, c)))))
            L.ixo_uniaxial_energy(1, n_inc, 2.0, 2.15, F3(*c), F3(*k), F3(0, 0, 1), pol, out)
            worst = max(worst, abs(out[0] + out[1] - 1.0))
            n_tir += out[1] == 0.0
    assert worst < 2e-5, worst
    assert n_tir > 100


def test_uniaxial_entry_isotropic_limit_is_scalar_fresnel():
    """n_e = n_o: r_ss, r_pp equal the scalar r_s, r_p (the crate's sign
    convention) and the cross terms vanish, for any optic axis."""
    L = oracle_lib.lib()
    r = (C.c_float * 4)()
    for th in (0.05, 0.4, 0.9, 1.35):
        k = (math.sin(th), 0.0, -math.cos(th))
        ci = math.cos(th)
        ct = math.sqrt(1.0 - (math.sin(th) / 2.0) ** 2)
        for c in ((0.3, 0.4, 0.866), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)):
            L.ixo_uniaxial_entry_r(2.0, 2.0, F3(*c), F3(*k), F3(0, 0, 1), 0, r)
            assert r[0] == pytest.approx((ci - 2 * ct) / (ci + 2 * ct), abs=2e-6) and abs(r[1]) < 1e-6
            assert abs(r[2]) < 1e-6 and abs(r[3]) < 1e-6
            L.ixo_uniaxial_entry_r(2.0, 2.0, F3(*c), F3(*k), F3(0, 0, 1), 1, r)
            assert r[2] == pytest.approx((2 * ci - ct) / (2 * ci + ct), abs=2e-6) and abs(r[3]) < 1e-6
            assert abs(r[0]) < 1e-6


def test_uniaxial_normal_incidence_axis_in_surface():
    """Normal incidence, optic axis in the surface: light polarised along the
    axis sees n_e, across it n_o, so the unpolarised reflectance is the mean
    of the two normal-incidence reflectances."""
    L = oracle_lib.lib()
    out = (C.c_float * 2)()
    n_o, n_e = 1.544, 1.553
    rs = []
    for pol in (0, 1):
        L.ixo_uniaxial_energy(0, 1.0, n_o, n_e, F3(0.6, 0.8, 0.0), F3(0, 0, -1), F3(0, 0, 1), pol, out)
        rs.append(out[0])
    expect = 0.5 * (((n_o - 1) / (n_o + 1)) ** 2 + ((n_e - 1) / (n_e + 1)) ** 2)
    assert 0.5 * (rs[0] + rs[1]) == pytest.approx(expect, rel=1e-5)


def _c2_sums(material, spp, first=0, w=64, h=64):
    from indicatrix_amd import workloads
    return oracle_lib.accumulate(workloads.film_c2_scene(w, h, spp, material=material), first, spp).sum(axis=0)


def test_uniaxial_film_converges_to_the_isotropic_limit():
    """The uniaxial estimator (Lekner entry / exit, eigenmode draws, walk-off,
    o<->e re-coupling) at a birefringence just above the anisotropy
    threshold renders the same film as the isotropic tracer, within 4 sigma
    of the Monte Carlo noise measured from independent sample blocks."""
    from indicatrix_amd import workloads
    iso = workloads.c2_material(bands=False)
    iso.cubic = True
    iso.birefringence_delta = 0.0
    uni = workloads.c2_material(bands=False, birefringence=2e-4)
    blocks = np.array([_c2_sums(uni, 32, 32 * b) for b in range(8)])
    sigma = blocks.std(axis=0, ddof=1) / math.sqrt(8)
    a = _c2_sums(iso, 256)
    b = blocks.sum(axis=0)
    assert (np.abs(b - a) < 4.0 * sigma * 8 + 1e-9).all(), (a, b, sigma)


def test_pleochroic_bands_absorb_and_depend_on_the_axis():
    from indicatrix_amd import workloads
    clear = _c2_sums(workloads.c2_material(bands=False), 16)
    dark = _c2_sums(workloads.c2_material(), 16)
    assert (dark < clear).all()
    m = workloads.c2_material()
    m.c_axis = (0.0, 0.0, 1.0)
    turned = _c2_sums(m, 16)
    assert np.abs(turned - dark).max() / dark.max() > 1e-3


# ---------------------------------------------------------------------------
# statistical image equivalence (renderer/gpu/z_stats.rs), pinned on the CPU
# ---------------------------------------------------------------------------
def test_z_criteria_accept_independent_samples_and_reject_a_planted_bias():
    import zstats
    from indicatrix_amd import workloads
    w = h = 40
    spp = 32
    sc = workloads.film_c1_scene(w, h)
    a, _ = oracle_lib.trace_paths(sc, 0, spp)
    b, _ = oracle_lib.trace_paths(sc, spp, spp)
    ok, res = zstats.image_equivalent(a, b, w, h, spp)
    assert ok, res
    # a planted mistake -- half the light of every path kept twice over -- is
    # caught: per-pixel tests at 32 spp resolve gross or local errors, while
    # the frame-level sums elsewhere resolve small global ones
    ok_bad, res_bad = zstats.image_equivalent(a, b * 1.5, w, h, spp)
    assert not ok_bad, res_bad
