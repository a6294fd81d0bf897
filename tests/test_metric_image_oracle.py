"""The metric-image oracle pinned: the sub-pixel offsets, the temporal
cameras, the scintillation terms and the per-pixel spatial map and index
(IX/color/metrics/image.rs, image_scint.rs, scintillation.rs), by the
reference's own tests (image_scint_tests.rs, image_tests.rs) and by closed
forms, chosen so a swapped base, a wrong yaw offset, a transposed window or a
mis-weighted blend fails one of them.  CPU only."""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, metrics, tracer

POSE = (0.6, 0.45, 2.4)   # image_scint_tests.rs:12


def f32(x):
    return float(np.float32(x))


def ulps(a, b):
    ia = np.array([a], np.float32).view(np.int32)[0]
    ib = np.array([b], np.float32).view(np.int32)[0]
    return abs(int(ia) - int(ib))


def srb():
    cut = tracer.standard_round_brilliant()
    return (abi.FacetPlane * len(cut))(*cut)


def request(preset="LightTent", w=30, h=19, samples=2, want_fire=False, scint=True, pose=POSE,
            material=None, head_shadow=16.0, cam_pose=POSE):
    planes = srb()
    cam = abi.Camera()
    oracle_lib.lib().ixo_camera_new(cam_pose[0], cam_pose[1], cam_pose[2], 42.0, w, h, 1, C.byref(cam))
    env = abi.Environment(abi.PRESETS[preset], 1.0, 0.85, 0.95, 0.0, 1.0, head_shadow)
    mat = material or tracer.GemMaterial.diamond().to_abi()
    r = metrics.make_request(cam, w, h, samples, planes, mat, env, want_fire=want_fire,
                             want_scintillation=scint, pose=pose)
    r._keep = (planes, mat)
    return r


# ---- offset (image.rs:21-36) -------------------------------------------
def radical(n, base):
    """The exact radical inverse of n in `base`."""
    x, scale = Fraction(0), Fraction(1, base)
    while n:
        x += (n % base) * scale
        n //= base
        scale /= base
    return x


def test_offset_is_the_centre_then_base2_base3_radical_inverses():
    assert oracle_lib.image_offset(0) == (0.5, 0.5)
    # the first few by hand: (1/2, 1/3), (1/4, 2/3), (3/4, 1/9), (1/8, 4/9)
    for s, (x, y) in {1: (0.5, 1 / 3), 2: (0.25, 2 / 3), 3: (0.75, 1 / 9), 4: (0.125, 4 / 9)}.items():
        jx, jy = oracle_lib.image_offset(s)
        assert jx == x and ulps(jy, y) <= 1, (s, jx, jy)
    for s in range(1, 4096):
        jx, jy = oracle_lib.image_offset(s)
        # base 2 is exact in f64 and in f32 for s < 2^24
        assert jx == f32(float(radical(s, 2))), s
        assert ulps(jy, float(radical(s, 3))) <= 1, s
        assert 0.0 <= jx < 1.0 and 0.0 <= jy < 1.0


def test_offsets_are_distinct_over_the_sample_budget():
    pts = {oracle_lib.image_offset(s) for s in range(64)}
    assert len(pts) == 64


# ---- temporal cameras (image_scint.rs:35-48, tests :105-143) ------------
def orbit_frame(yaw, pitch, dist):
    """Camera::new's frame in closed form for cos(pitch) > 0: the origin on
    the orbit sphere, forward towards the centre, right = (cos yaw, 0,
    -sin yaw), up = right x forward."""
    o = np.array([dist * math.cos(pitch) * math.sin(yaw), dist * math.sin(pitch),
                  dist * math.cos(pitch) * math.cos(yaw)])
    f = -o / np.linalg.norm(o)
    r = np.array([math.cos(yaw), 0.0, -math.sin(yaw)])
    u = np.cross(r, f)
    return o, f, r, u


def test_temporal_cameras_are_the_grid_offsets():
    cams = oracle_lib.temporal_cameras(POSE, 0.5, 30, 19)
    assert len(cams) == 5
    for cam, off_deg in zip(cams, (-3.0, -1.5, 0.0, 1.5, 3.0)):
        yaw = POSE[0] + math.radians(off_deg)
        o, f, r, u = orbit_frame(yaw, POSE[1], POSE[2])
        assert np.allclose(list(cam.origin), o, atol=2e-6)
        assert np.allclose(list(cam.forward), f, atol=2e-6)
        assert np.allclose(list(cam.right), r, atol=2e-6)
        assert np.allclose(list(cam.up), u, atol=2e-6)
        assert cam.fov_tan == f32(0.5), "the request's field of view"
    # the base pose is the request camera itself, bit for bit
    base = abi.Camera()
    oracle_lib.lib().ixo_camera_new(POSE[0], POSE[1], POSE[2], 42.0, 30, 19, 1, C.byref(base))
    for fld in ("origin", "forward", "right", "up"):
        assert list(getattr(cams[2], fld)) == list(getattr(base, fld))
    # the offsets turn the view about +Y by the stated angles
    f0 = np.array(list(cams[2].forward))
    for cam, off_deg in zip(cams, (-3.0, -1.5, 0.0, 1.5, 3.0)):
        f = np.array(list(cam.forward))
        yaw_seen = math.degrees(math.atan2(f[0], f[2]) - math.atan2(f0[0], f0[2]))
        assert abs(yaw_seen - off_deg) < 1e-4


# ---- the temporal term on a traced image (image_scint_tests.rs:146-176) --
def test_temporal_variance_is_bernoulli_and_capped():
    px, _, _ = oracle_lib.metric_image(request())
    computed = flickering = 0
    for p in px:
        assert p["scint_poses"] in (0, 5)
        assert p["scint_returned"] <= p["scint_poses"]
        if p["scint_poses"]:
            computed += 1
            assert p["hits"] > 0, "a pixel without hits has no temporal term"
            fr = f32(f32(p["scint_returned"]) / 5.0)
            v = f32(fr - fr * fr)
            assert any(abs(a - v) < 1e-6 for a in (0.0, 0.16, 0.24)), v
            flickering += v > 0.0
    assert computed > 50, computed
    assert flickering > 0, "some pixel must flip within +/-3 degrees"


def test_scintillation_off_leaves_every_other_field_unchanged():
    """image_scint_tests.rs:67-103."""
    off, toff, _ = oracle_lib.metric_image(request(want_fire=True, scint=False))
    on, ton, _ = oracle_lib.metric_image(request(want_fire=True, scint=True))
    assert np.all(off["scint_poses"] == 0) and np.all(off["scint_returned"] == 0)
    for f in ("hits", "classes", "fire_sum_deg", "fire_qualified", "fire_rejected", "centre_entry_facet",
              "centre_exit_facet", "centre_bounces", "centre_reason", "mixed_facets"):
        assert np.array_equal(off[f], on[f]), f
    assert list(toff.total.classes) == list(ton.total.classes)
    assert toff.total.fire_sum_deg == ton.total.fire_sum_deg


def test_scintillation_without_a_pose_is_invalid():
    """image_scint_tests.rs:234-250."""
    r = request(pose=None)
    px = (abi.OpticalPixel * (r.width * r.height))()
    assert oracle_lib.lib().ixo_metric_image(C.byref(r), px, None) == abi.IXG_INVALID
    r = request(pose=(0.6, 0.4, 0.0))
    assert oracle_lib.lib().ixo_metric_image(C.byref(r), px, None) == abi.IXG_INVALID


# ---- the spatial map (image_scint_tests.rs:178-208) ----------------------
def map_from(size, brilliant):
    px = np.zeros(size * size, metrics.PIXEL_DTYPE)
    for y in range(size):
        for x in range(size):
            p = px[y * size + x]
            p["hits"] = 1
            p["classes"][2 if brilliant(x, y) else 1] = 1
    return px


def window_reference(px, w, h, x, y):
    """window_spatial written out in float64 (a closed-form check of the
    5x5 window, not its rounding): CV squashed through cv / (1 + cv)."""
    fr = [p["classes"][2] / max(p["hits"], 1) for p in px]
    vals = [fr[wy * w + wx] for wy in range(max(y - 2, 0), min(y + 3, h))
            for wx in range(max(x - 2, 0), min(x + 3, w)) if px[wy * w + wx]["hits"] > 0]
    mean = sum(vals) / len(vals)
    if mean <= 1e-4:
        return 0.0
    var = max(sum(v * v for v in vals) / len(vals) - mean * mean, 0.0)
    cv = math.sqrt(var) / mean
    return cv / (1 + cv) * 100


def test_spatial_term_zero_on_a_uniform_map_positive_on_a_checkerboard():
    uniform = map_from(9, lambda x, y: True)
    assert np.all(np.abs(oracle_lib.scintillation_spatial(uniform, 9, 9)) < 1e-3)
    checker = map_from(9, lambda x, y: (x + y) % 2 == 0)
    sp = oracle_lib.scintillation_spatial(checker, 9, 9)
    assert np.all((sp >= 0) & (sp <= 100)) and np.all(sp > 10.0), sp
    for i in (0, 8, 40, 72, 80):
        assert np.float32(oracle_lib.scintillation_spatial_at(checker, 9, 9, i)).tobytes() == sp[i].tobytes()
    # against the window written out: an interior pixel sees 13 of 25 bright
    # (cv = sqrt(12/13)), a corner 5 of 9
    for y in range(9):
        for x in range(9):
            assert abs(sp[y * 9 + x] - window_reference(checker, 9, 9, x, y)) < 1e-4
    assert abs(sp[40] - (lambda cv: cv / (1 + cv) * 100)(math.sqrt(12 / 13))) < 1e-4
    # no hits reads zero in both; a size / pixel-count mismatch reads all zero
    hole = map_from(5, lambda x, y: (x + y) % 2 == 0)
    hole[12] = np.zeros(1, metrics.PIXEL_DTYPE)[0]
    assert oracle_lib.scintillation_spatial(hole, 5, 5)[12] == 0.0
    assert oracle_lib.scintillation_spatial_at(hole, 5, 5, 12) == 0.0
    assert np.all(oracle_lib.scintillation_spatial(hole, 6, 5) == 0.0)


def test_spatial_window_is_not_transposed():
    # a map bright in its left columns only: a transposed window (x and y
    # swapped) reads a different neighbourhood at (1, 3) of a 7 x 5 frame
    w, h = 7, 5
    px = map_from(7, lambda x, y: x < 2)[: w * h]
    for y in range(h):
        for x in range(w):
            px[y * w + x]["classes"][:] = 0
            px[y * w + x]["classes"][2 if x < 2 else 1] = 1
    sp = oracle_lib.scintillation_spatial(px, w, h)
    for y in range(h):
        for x in range(w):
            assert abs(sp[y * w + x] - window_reference(px, w, h, x, y)) < 1e-4, (x, y)


# ---- the index (image_scint_tests.rs:210-232) ------------------------------
def test_index_combines_sixty_forty():
    px = map_from(3, lambda x, y: True)
    px[4]["scint_returned"], px[4]["scint_poses"] = 2, 5   # variance 0.24, the cap
    spatial = np.full(9, 50.0, np.float32)
    idx = oracle_lib.scintillation_index(px, spatial, 4)
    assert abs(idx - (0.4 * 100.0 + 0.6 * 50.0)) < 1e-3
    px[4]["scint_returned"] = 5   # never flipping: the spatial share only
    assert abs(oracle_lib.scintillation_index(px, spatial, 4) - 30.0) < 1e-3
    assert oracle_lib.scintillation_index(px, spatial, 3) is None
    assert oracle_lib.scintillation_index(px, spatial, 9) is None
    px[4]["scint_returned"] = 1   # 1 of 5: variance 0.16 -> two thirds of the cap
    assert abs(oracle_lib.scintillation_index(px, spatial, 4) - (0.4 * 100 * 0.16 / 0.24 + 30.0)) < 1e-3
    agg = oracle_lib.lib().ixo_image_scintillation_pct(oracle_lib.pixels_ctypes(px), 9)
    assert 0.0 <= agg <= 100.0
    full = oracle_lib.scintillation_spatial(px, 3, 3)
    for i in range(9):
        a = oracle_lib.scintillation_index_at(px, 3, 3, i)
        b = oracle_lib.scintillation_index(px, full, i)
        assert (a is None and b is None) or np.float32(a).tobytes() == np.float32(b).tobytes()


def test_index_on_a_traced_image_agrees_with_the_aggregate_terms():
    px, _, agg = oracle_lib.metric_image(request())
    sp = oracle_lib.scintillation_spatial(px, 30, 19)
    assert np.all(sp[px["hits"] == 0] == 0.0)
    idx = [oracle_lib.scintillation_index(px, sp, i) for i in range(len(px))]
    assert all((v is None) == (p["scint_poses"] == 0) for v, p in zip(idx, px))
    assert all(0.0 <= v <= 100.0 for v in idx if v is not None)
    assert 0.0 < agg < 100.0


# ---- classification against each pixel's own eye (image_tests.rs:270-333) --
def test_perspective_rays_use_their_own_eye_direction_for_the_observer_cone():
    w = h = 41
    r = request(w=w, h=h, samples=1, scint=False, pose=None, head_shadow=30.0, cam_pose=(0.0, 1.2, 2.4))
    px, _, _ = oracle_lib.metric_image(r)
    planes = list(srb())
    disp = tracer.GemMaterial.diamond().to_abi().dispersion
    cam = r.camera
    o, d = (C.c_float * 3)(), (C.c_float * 3)()
    bucket = {1: 1, 2: 2}
    moved = 0
    for y in range(h):
        for x in range(w):
            cam.width, cam.height = float(w), float(h)
            oracle_lib.lib().ixo_generate_ray(C.byref(cam), float(x), float(y), 0.0, 0.0, o, d)
            own = oracle_lib.classify_ray(planes, disp, r.environment, list(o), list(d), [-v for v in d])
            axis = oracle_lib.classify_ray(planes, disp, r.environment, list(o), list(d),
                                           [-v for v in cam.forward])
            p = px[y * w + x]
            assert p["classes"][1] == int(bucket.get(own, 0) == 1), (x, y)
            assert p["classes"][2] == int(bucket.get(own, 0) == 2), (x, y)
            bo, ba = bucket.get(own, 0), bucket.get(axis, 0)
            moved += bo != ba and bo * ba != 0
    assert moved > 0, "no pixel changed between extinction and brilliance with the cone on the eye"


def test_fire_distribution_is_finite_and_non_negative():
    """image_tests.rs:335-376, for Diamond and a quartz-like low dispersion."""
    quartz = tracer.GemMaterial.isotropic_from_nd_abbe(1.544, 70.0).to_abi()
    for mat in (tracer.GemMaterial.diamond().to_abi(), quartz):
        r = request(preset="IsoHemisphere", w=48, h=48, samples=4, want_fire=True, scint=False, pose=None,
                    material=mat, cam_pose=(0.0, 1.2, 2.4))
        px, _, _ = oracle_lib.metric_image(r)
        hit = px[px["hits"] > 0]
        fire = hit["fire_sum_deg"] / np.maximum(hit["hits"], 1)
        assert len(fire) and np.all(np.isfinite(fire)) and np.all(fire >= 0.0)


# ---- the library's host read-outs against the oracle (no device needed) ----
def test_library_hover_readouts_equal_the_oracle():
    from scint_maps_cases import cases
    for name, px, w, h in cases():
        n = len(px)
        for i in sorted({0, n // 2, n - 1, min(12, n - 1)}):
            a = metrics.scintillation_spatial_at(px, w, h, i)
            b = oracle_lib.scintillation_spatial_at(px, w, h, i)
            assert np.float32(a).tobytes() == np.float32(b).tobytes(), (name, i)
            a = metrics.scintillation_index_at(px, w, h, i)
            b = oracle_lib.scintillation_index_at(px, w, h, i)
            assert (a is None) == (b is None), (name, i)
            if a is not None:
                assert np.float32(a).tobytes() == np.float32(b).tobytes(), (name, i)
    assert metrics.scintillation_index_at(cases()[0][1], 9, 9, 81) is None


def test_all_pixel_indices_equal_the_single_pixel_reads():
    from scint_maps_cases import cases
    for name, px, w, h in cases():
        sp = oracle_lib.scintillation_spatial(px, w, h)
        allv = oracle_lib.scintillation_indices(px, sp)
        for i in range(len(px)):
            v = oracle_lib.scintillation_index(px, sp, i)
            assert (np.isnan(allv[i]) and v is None) or np.float32(v).tobytes() == allv[i].tobytes(), (name, i)
