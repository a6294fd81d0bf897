"""CPU checks of the per-function probe tier: the ULP distance of the
reference's harness (its own unit cases, renderer/gpu/ulp.rs:70-169), the
record layouts the library and the oracle agree on, and the oracle's probe
dispatch pinned against the oracle's separately exported functions and a
closed form.  No GPU calls."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
import probe_cases
from indicatrix_amd import _lib, abi
from ulp import summarize, ulp_distance, within_tolerance

F32 = np.float32


def nxt(x, k=1):
    v = np.float32(x)
    for _ in range(k):
        v = np.nextafter(v, np.float32(np.inf), dtype=np.float32)
    return v


def test_ulp_distance_reference_cases():
    # ulp.rs:74-169
    assert ulp_distance(F32(1.0), F32(1.0)) == 0
    assert ulp_distance(F32(0.0), F32(0.0)) == 0 and ulp_distance(F32(0.0), F32(-0.0)) == 0
    a = F32(1.0)
    assert ulp_distance(a, nxt(a)) == 1 and ulp_distance(nxt(a), a) == 1
    tiny = np.float32(1.4e-45)
    assert ulp_distance(-tiny, tiny) == 2
    assert ulp_distance(F32(-0.0), F32(0.0)) == 0
    assert ulp_distance(F32(2.5), nxt(2.5, 5)) == 5 and ulp_distance(nxt(2.5, 5), F32(2.5)) == 5
    assert ulp_distance(F32(-2.5), -nxt(2.5, 5)) == 5
    one_bits = int(np.float32(1.0).view(np.uint32))
    assert ulp_distance(F32(-1.0), F32(1.0)) == 2 * one_bits
    d2 = np.uint32(2).view(np.float32)
    d3 = np.uint32(3).view(np.float32)
    assert ulp_distance(-d2, d3) == 5 and ulp_distance(d3, -d2) == 5
    assert ulp_distance(np.uint32(1).view(np.float32), np.uint32(8).view(np.float32)) == 7
    assert ulp_distance(-np.uint32(1).view(np.float32), -np.uint32(8).view(np.float32)) == 7
    mx = np.float32(np.finfo(np.float32).max)
    below = np.nextafter(mx, F32(0.0), dtype=np.float32)
    assert ulp_distance(mx, below) == 1 and ulp_distance(-mx, -below) == 1
    assert ulp_distance(-mx, mx) == 2 * int(mx.view(np.uint32))
    # the hybrid rule (ulp.rs:42-44): the floor admits a zero crossing
    assert within_tolerance(F32(1e-7), F32(-1e-7), 16, 1e-6)
    assert not within_tolerance(F32(1e-7), F32(-1e-7), 16, 1e-8)
    s = summarize(np.array([100.0, 1e-7], F32), np.array([nxt(100.0, 3), -1e-7], F32), 2, 1e-6)
    assert s["passed"] is False and s["max_genuine_ulp"] == 3 and s["exempted"] == 1


def test_probe_layouts_agree():
    L = oracle_lib.lib()
    G = _lib.lib()
    assert len(abi.PROBE_FNS) == 30
    for fid in range(len(abi.PROBE_FNS)):
        a, b, c, d = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        assert L.ixo_probe_shape(fid, C.byref(a), C.byref(b)) == 1
        assert G.ixg_probe_shape(fid, C.byref(c), C.byref(d)) == 1
        assert (a.value, b.value) == (c.value, d.value), abi.PROBE_FNS[fid]
    assert L.ixo_probe_shape(len(abi.PROBE_FNS), None, None) == 0
    assert G.ixg_probe_shape(-1, None, None) == 0
    for case in probe_cases.CASES:
        sc, rec = probe_cases.case_inputs(case, 0)
        w_in = C.c_uint32()
        L.ixo_probe_shape(abi.PROBE[case["fn"]], C.byref(w_in), None)
        assert rec.shape[1] == w_in.value, case["label"]


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_oracle_probes_are_the_exported_functions():
    L = oracle_lib.lib()
    sc = probe_cases.scene_c1()
    rng = np.random.default_rng(5)
    # sample draws and the hero comb
    rec = probe_cases.gen_sample_draws(rng, sc)[:512]
    out = oracle_lib.probe(sc, "sample_draws", rec)
    for r, o in zip(rec, out):
        seed, jx, jy, hero = C.c_uint32(), C.c_float(), C.c_float(), C.c_float()
        L.ixo_sample_draws(int(r[0:1].view(np.uint32)[0]), int(r[1:2].view(np.uint32)[0]), C.byref(seed),
                           C.byref(jx), C.byref(jy), C.byref(hero))
        lam = (C.c_float * 8)()
        L.ixo_wrapped_hero_wavelengths(hero.value, lam)
        want = np.array([jx.value, jy.value, hero.value, np.uint32(seed.value).view(np.float32)] + list(lam), F32)
        assert np.array_equal(_bits(o), _bits(want))
    # directions and normals
    rec = probe_cases.gen_cosine(rng, sc)[:256]
    out = oracle_lib.probe(sc, "cosine_hemisphere", rec)
    for r, o in zip(rec, out):
        want = (C.c_float * 3)()
        L.ixo_cosine_hemisphere(float(r[0]), float(r[1]), (C.c_float * 3)(*r[2:5]), want)
        assert np.array_equal(_bits(o), _bits(list(want)))
    rec = probe_cases.gen_hg(rng, sc)[::36]
    out = oracle_lib.probe(sc, "hg_direction", rec)
    for r, o in zip(rec, out):
        want = (C.c_float * 3)()
        L.ixo_hg_direction(float(r[0]), float(r[1]), float(r[2]), (C.c_float * 3)(*r[3:6]), want)
        assert np.array_equal(_bits(o), _bits(list(want)))
    rec = probe_cases.gen_shading(rng, sc)[:256]
    out = oracle_lib.probe(sc, "shading_normal", rec)
    for r, o in zip(rec, out):
        want = (C.c_float * 3)()
        L.ixo_shading_normal_near_edge(sc.planes, len(sc.planes), (C.c_float * 3)(*r[0:3]),
                                       int(r[3:4].view(np.int32)[0]), (C.c_float * 3)(*r[4:7]), float(r[7]), want)
        assert np.array_equal(_bits(o), _bits(list(want)))
    # spectra
    rec = probe_cases.gen_lambda(rng, sc)[::7]
    out = oracle_lib.probe(sc, "cie_cmf", rec)
    for r, o in zip(rec, out):
        want = (C.c_float * 3)()
        L.ixo_cie_cmf(float(r[0]), want)
        assert np.array_equal(_bits(o), _bits(list(want)))
    rec = probe_cases.gen_exp(rng, sc)[::13]
    out = oracle_lib.probe(sc, "exp_lane", rec)
    assert np.array_equal(_bits(out[:, 0]), _bits([L.ixo_exp_lane(float(x)) for x in rec[:, 0]]))
    out = oracle_lib.probe(sc, "exp", rec)
    with np.errstate(over="ignore"):
        assert np.array_equal(out[:, 0], np.exp(rec[:, 0].astype(np.float64)).astype(np.float32))
    # n(lambda) and the bands on the pleochroic material
    sa = probe_cases.scene_sapphire()
    rec = probe_cases.gen_channel_optics(rng, sa)[:128]
    out = oracle_lib.probe(sa, "channel_optics", rec)
    m = sa.material_abi
    for r, o in zip(rec, out):
        n_o = [L.ixo_dispersion_evaluate(C.byref(m.dispersion), float(x)) for x in r]
        ao = [sum_bands(L, m.bands_o, m.n_bands_o, x) for x in r]
        ae = [sum_bands(L, m.bands_e, m.n_bands_e, x) for x in r]
        assert np.array_equal(_bits(o), _bits(n_o + ao + ae))
    # the slab scan
    rec = probe_cases.gen_intersect(rng, sc)[::16]
    out = oracle_lib.probe(sc, "intersect", rec)
    for r, o in zip(rec, out):
        t = C.c_float()
        f = L.ixo_slab_scan(sc.planes, len(sc.planes), (C.c_float * 3)(*r[:3]), (C.c_float * 3)(*r[3:6]),
                            C.byref(t))
        assert int(o[0:1].view(np.int32)[0]) == f
        if f >= 0:
            assert np.float32(t.value) == o[1]


def sum_bands(L, bands, n, lam):
    s = np.float32(-0.0)
    for i in range(n):
        s = np.float32(s + np.float32(L.ixo_band_evaluate(C.byref(bands[i]), float(lam))))
    return float(s)


def test_oracle_finish_closed_forms():
    """PathState::finish (bounce_body.rs:724-764, color.rs:145-172): with every
    channel in one family of equal pdfs the balance weight is exactly 1 and
    the estimate is norm * sum_c L_c cmf(lambda_c) (white balance 1 under the
    D65 dome, whose f32 round trip through LMS moves a few 1e-6); a channel
    alone in its family carries 8 pdf_0 / pdf_c.  A dropped channel moves the
    sum by about 1/8, a wrong weight by far more than the tolerance."""
    L = oracle_lib.lib()
    sc = probe_cases.scene_c1()
    rng = np.random.default_rng(3)
    lam = probe_cases.gen_channel_optics(rng, sc)[:64]
    rad = rng.uniform(0.0, 2.0, (64, 8)).astype(np.float32)
    ones = np.uint32(0xFFFFFFFF).view(np.float32)
    rec = np.concatenate([np.full((64, 2), ones, np.float32), np.full((64, 8), 0.5, np.float32), lam, rad], axis=1)
    out = oracle_lib.probe(sc, "finish", rec).astype(np.float64)
    norm = (400.0 / 8.0) / 106.856
    for i in range(64):
        want = np.zeros(3)
        for c in range(8):
            cmf = (C.c_float * 3)()
            L.ixo_cie_cmf(float(lam[i, c]), cmf)
            want += norm * float(rad[i, c]) * np.array(list(cmf))
        assert np.allclose(out[i], want, rtol=2e-5, atol=1e-9)
    # singleton families: byte c = 1 << c
    lo = np.uint32(0x08040201).view(np.float32)
    hi = np.uint32(0x80402010).view(np.float32)
    pdf = rng.uniform(0.1, 1.0, (64, 8)).astype(np.float32)
    rec = np.concatenate([np.full((64, 1), lo), np.full((64, 1), hi), pdf, lam, rad], axis=1).astype(np.float32)
    out = oracle_lib.probe(sc, "finish", rec).astype(np.float64)
    for i in range(64):
        want = np.zeros(3)
        for c in range(8):
            cmf = (C.c_float * 3)()
            L.ixo_cie_cmf(float(lam[i, c]), cmf)
            w = 8.0 * pdf[i, 0] / pdf[i, c]
            want += norm * float(rad[i, c]) * w * np.array(list(cmf))
        assert np.allclose(out[i], want, rtol=2e-5, atol=1e-9)


def test_oracle_probes_finite_on_every_case():
    for case in probe_cases.CASES:
        for si in range(len(case["scenes"])):
            sc, rec = probe_cases.case_inputs(case, si)
            out = oracle_lib.probe(sc, case["fn"], rec)
            exact = case.get("exact", [])
            vals = np.delete(out, exact, axis=1) if exact else out
            assert np.all(np.isfinite(vals)), (case["label"], si)


def test_oracle_interface_closed_forms():
    """The probe inputs reach the physics the oracle is pinned by: a
    normal-incidence Fresnel reflection is ((n1 - n2) / (n1 + n2))^2 of an
    unpolarised beam, transmission 1 - R, and total internal reflection
    keeps the intensity."""
    sc = probe_cases.scene_c1()
    n = np.array([1.5, 2.417, 1.76], np.float32)
    rec = np.array([[1.0, x, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0] for x in n], np.float32)
    r = oracle_lib.probe(sc, "fresnel_reflection", rec)
    t = oracle_lib.probe(sc, "fresnel_transmission", rec)
    R = ((1.0 - n.astype(np.float64)) / (1.0 + n)) ** 2
    assert np.allclose(r[:, 0], R, rtol=1e-6) and np.allclose(t[:, 0], 1.0 - R, rtol=1e-6)
    rec = np.array([[2.417, 0.3, math.sqrt(1 - 0.09), 1.0, 0.4, -0.3, 0.5]], np.float32)
    out = oracle_lib.probe(sc, "tir_retardation", rec)
    assert out[0, 0] == np.float32(1.0) and out[0, 1] == np.float32(0.4)
    assert abs(float(np.hypot(out[0, 2], out[0, 3])) - math.hypot(-0.3, 0.5)) < 1e-6
