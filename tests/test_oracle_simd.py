"""The oracle's 8-wide slab scan and exponential (the reference's AVX2
kernels, IX/simd/slab.rs:193 and IX/simd/exp_poly.rs:71) against the scalar
functions they replace, bit for bit, as the reference's own tests hold them
(slab.rs:401 slab_scan_matches_scalar_bitwise, exp_poly.rs:142 and :186), and
the AVX2 build's frames against the scalar build's.  The FLOP-counting build
(oracle/ixo_flops.cpp) is held to the counting rule on the slab scan and to
the FP32 oracle's results.  CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, metrics, tracer, workloads

f3 = C.c_float * 3


def plane(n, d):
    return abi.FacetPlane(f3(*[float(v) for v in n]), float(d))


def bits(x):
    return np.array(x, np.float32).view(np.uint32)


def exp8(L, lanes):
    x = (C.c_float * 8)(*lanes)
    out = (C.c_float * 8)()
    L.ixo_exp_lane8(x, out)
    return np.array(out[:], np.float32)


def test_exp8_lane_identical_and_accurate():
    """exp_poly.rs:142: the sweep x = 0.37 k - 87 up to 10 with its seven
    companions, every lane the scalar ixo_exp_lane's bits and within 5e-7 of
    the double exponential."""
    L = oracle_lib.lib()
    worst = 0.0
    k = 0
    while True:
        x = float(np.float32(np.float32(k) * np.float32(0.37) + np.float32(-87.0)))
        if x >= 10.0:
            break
        xs = np.float32(x)
        lanes = np.array([xs, xs + np.float32(0.111), xs + np.float32(0.222), xs + np.float32(0.333),
                          -xs * np.float32(0.1), xs * np.float32(0.5), xs + np.float32(0.777), 0.0], np.float32)
        got = exp8(L, lanes)
        want = np.array([L.ixo_exp_lane(float(v)) for v in lanes], np.float32)
        assert (bits(got) == bits(want)).all(), (x, got, want)
        for v, g in zip(lanes, got):
            ref = math.exp(float(v))
            if ref > 1e-30:
                worst = max(worst, abs(float(g) - ref) / ref)
        k += 1
    assert worst < 5e-7, worst


def test_exp8_nan_and_infinite_lanes():
    """exp_poly.rs:186: NaN of either sign and +inf clamp to the high end,
    -inf and -MAX to the low end, on every lane as in the scalar function."""
    L = oracle_lib.lib()
    big = float(np.finfo(np.float32).max)
    lanes = np.array([np.nan, np.inf, -np.inf, -np.nan, big, -big, 0.0, 1.0], np.float32)
    lanes[3] = np.float32(-1) * np.abs(lanes[3])  # a NaN with the sign bit set
    got = exp8(L, lanes)
    want = np.array([L.ixo_exp_lane(float(v)) for v in lanes], np.float32)
    assert (bits(got) == bits(want)).all()
    hi, lo = L.ixo_exp_lane(88.02875), L.ixo_exp_lane(-87.33654)
    assert bits(got[[0, 1, 3, 4]]).tolist() == [bits(hi)] * 4
    assert bits(got[[2, 5]]).tolist() == [bits(lo)] * 2
    assert got[6] == 1.0 and abs(got[7] - math.e) < 1e-6
    # the scalar-only build takes the scalar loop: the same bits
    assert (bits(exp8(oracle_lib.lib_scalar(), lanes)) == bits(want)).all()


def lcg(seed):
    """The reference test's generator (slab.rs:403-408): uniform in [-1, 1)."""
    state = seed

    def nxt():
        nonlocal state
        state = (state * 6364136223846793005 + 1) & 0xFFFFFFFFFFFFFFFF
        u = np.float32(np.float32(state >> 33) / np.float32(4294967295.0))
        return float(np.float32(float(u) * 2.0 - 1.0))
    return nxt


def scan_pair(planes, o, d):
    L = oracle_lib.lib()
    arr = (abi.FacetPlane * len(planes))(*planes)
    t1, t2 = C.c_float(-7.0), C.c_float(-7.0)
    f1 = L.ixo_intersect(arr, len(planes), f3(*o), f3(*d), C.byref(t1))
    f2 = L.ixo_slab_scan(arr, len(planes), f3(*o), f3(*d), C.byref(t2))
    return (f1, bits(t1.value)), (f2, bits(t2.value))


def test_slab_scan_matches_scalar_bitwise():
    """slab.rs:401: random unit planes, a plane exactly parallel to the ray
    every ninth (the early-reject branch), random directions and origins, at
    counts on both sides of every multiple of 8."""
    nxt = lcg(7)
    outcomes = {"hit": 0, "miss": 0}
    for count in (1, 5, 8, 9, 16, 17, 40, 100):
        for _ in range(20):
            d = np.array([nxt(), nxt(), nxt()], np.float32)
            d = d / np.float32(max(float(np.linalg.norm(d)), 1e-3))
            planes = []
            for i in range(count):
                if i % 9 == 3:
                    fb = np.array([1, 0, 0] if abs(d[0]) < 0.9 else [0, 1, 0], np.float32)
                    n = np.cross(d, fb)
                    n = (n / np.linalg.norm(n)).astype(np.float32)
                else:
                    n = np.array([nxt(), nxt(), nxt()], np.float32)
                ln = np.float32(max(float(np.sqrt(np.float32(n @ n))), 1e-3))
                planes.append(plane((n / ln).astype(np.float32), nxt()))
            o = [nxt() * 3.0, nxt() * 3.0, nxt() * 3.0]
            want, got = scan_pair(planes, o, d.tolist())
            assert want[0] == got[0]
            if want[0] >= 0:
                assert want[1] == got[1]
            # the same planes pushed out to enclose the origin: rays from near
            # the centre leave through one of them (the hit arm, both lanes' rules)
            inside = [plane(p.normal, -0.5 - abs(p.d)) for p in planes]
            o = [nxt() * 0.1, nxt() * 0.1, nxt() * 0.1]
            want, got = scan_pair(inside, o, d.tolist())
            assert want == got
            outcomes["hit" if got[0] >= 0 else "miss"] += 1
    assert outcomes["hit"] > 100, outcomes


def cube():
    return [plane(n, -1.0) for n in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                                                [0, 0, -1])]


def test_slab_scan_closed_forms_and_ties():
    """The unit cube seen from inside exits at t = 1 on the face the ray
    points at; from outside a parallel face rejects; duplicated planes (the
    same plane 8 and 3 slots later, so the copies fall in one lane and in
    different lanes) resolve to the lowest index, entering and exiting."""
    for axis in range(3):
        for sgn in (1.0, -1.0):
            d = [0.0, 0.0, 0.0]
            d[axis] = sgn
            want, got = scan_pair(cube(), [0.1, -0.2, 0.3], d)
            assert got == want
            face = 2 * axis + (0 if sgn > 0 else 1)
            assert got[0] == face
            expect = np.float32(1.0) - np.float32([0.1, -0.2, 0.3][axis]) * np.float32(sgn)
            assert got[1] == bits(expect)
    want, got = scan_pair(cube(), [0.0, 2.0, 0.0], [1.0, 0.0, 0.0])
    assert want[0] == got[0] == -1
    base = cube() + [plane([0.6, 0.8, 0.0], -2.0)] * 2
    dup = base + base[:3] + base[:8]   # 21 planes: every face again at +8 and the first three at +16
    for d in ([0.48, 0.6, 0.64], [-0.6, 0.0, 0.8], [0.0, -1.0, 0.0]):
        for o in ([0.0, 0.0, 0.0], [0.0, 0.0, -3.0]):
            want, got = scan_pair(dup, o, d)
            assert got == want
            if got[0] >= 0:
                assert got[0] < 8
    # the entering plane of a ray from outside: a hit on the -z face, lowest copy
    want, got = scan_pair(dup, [0.0, 0.0, -3.0], [0.0, 0.0, 1.0])
    assert got == want and got[0] == 5 and got[1] == bits(2.0)


def test_avx2_frames_equal_scalar_frames():
    """The film on C1's brilliant and C2's oval (both absorption arms, HG
    scattering) and the metric image with fire and scintillation: the AVX2
    build's sums and pixels are the scalar build's, bit for bit."""
    for sc, spp in ((workloads.film_c1_scene(24, 24), 8), (workloads.film_c2_scene(24, 24, 4), 4)):
        a = oracle_lib.accumulate(sc, 0, spp)
        b = np.zeros_like(a)
        oracle_lib.lib_scalar().ixo_accumulate(sc.ref(), 0, spp, b.ctypes.data_as(C.POINTER(C.c_double)),
                                               oracle_lib.threads())
        assert a.any() and np.array_equal(a, b)
    cut = tracer.standard_round_brilliant()
    planes = (abi.FacetPlane * len(cut))(*cut)
    cam = abi.Camera()
    oracle_lib.lib().ixo_camera_new(0.6, 0.45, 2.4, 42.0, 30, 19, 1, C.byref(cam))
    env = abi.Environment(abi.PRESETS["LightTent"], 1.0, 0.85, 0.95, 0.0, 1.0, 16.0)
    mat = tracer.GemMaterial.diamond().to_abi()
    req = metrics.make_request(cam, 30, 19, 2, planes, mat, env, want_fire=True, want_scintillation=True,
                               pose=(0.6, 0.45, 2.4))
    outs = []
    for L in (oracle_lib.lib(), oracle_lib.lib_scalar()):
        px = (abi.OpticalPixel * (30 * 19))()
        tot = abi.MetricTotals()
        assert L.ixo_metric_image(C.byref(req), px, C.byref(tot)) == 0
        outs.append((bytes(px), bytes(tot)))
    assert outs[0] == outs[1]


def test_flop_count_of_the_slab_scan():
    """The counting rule on the scan: 12 operations a plane (two 3-term dot
    products at 5, the offset add, the divide), 72 on the cube, and the scan's
    result unchanged."""
    L = oracle_lib.lib_flops()
    arr = (abi.FacetPlane * 6)(*cube())
    t, n = C.c_float(), C.c_uint64()
    f = L.ixo_flops_intersect(arr, 6, f3(0.1, -0.2, 0.3), f3(0.0, 0.0, 1.0), C.byref(t), C.byref(n))
    assert (f, n.value) == (4, 72) and t.value == pytest.approx(0.7)
    # a parallel face seen from outside returns at that plane: three planes scanned
    f = L.ixo_flops_intersect(arr, 6, f3(0.0, 2.0, 0.0), f3(1.0, 0.0, 0.0), C.byref(t), C.byref(n))
    assert (f, n.value) == (-1, 36)


def test_counting_build_is_the_oracle():
    """The counting build's film is the FP32 oracle's bit for bit (the
    counting type stores f32 and rounds every operation as f32 does), and its
    count is additive over sample ranges."""
    sc = workloads.film_c1_scene(16, 16)
    a_ref = oracle_lib.accumulate(sc, 0, 4, nthreads=1)
    a, n4 = oracle_lib.accumulate_flops(sc, 0, 4)
    assert np.array_equal(a, a_ref)
    a2a, n2a = oracle_lib.accumulate_flops(sc, 0, 2)
    a2b, n2b = oracle_lib.accumulate_flops(sc, 2, 2)
    _, n1a = oracle_lib.accumulate_flops(sc, 0, 1)
    _, n1b = oracle_lib.accumulate_flops(sc, 1, 1)
    # each call pays a fixed scene set-up once, the paths are additive
    setup = n2a + n2b - n4
    assert 0 < setup < 10_000, setup
    assert n1a + n1b + n2b - n4 == 2 * setup
    np.testing.assert_allclose(a2a + a2b, a, rtol=1e-12, atol=1e-12)
    # per path: at least one scan of the 57-facet cut (12 x 57), well under the bounce cap's worth
    per_path = n4 / (16 * 16 * 4)
    assert 12 * 57 < per_path < 12 * 57 * 64, per_path
