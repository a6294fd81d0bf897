"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-side helpers and routing predicate behave (no GPU calls)."""
import ctypes as C
import os

import numpy as np

import oracle_lib
from indicatrix_amd import _lib, abi, tracer, workloads


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in _lib.header_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert len(_lib.header_symbols()) >= 18


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libixg.so"))
    try:
        _lib.lib()
        raise AssertionError("expected ImportError")
    except ImportError as e:
        assert "missing" in str(e)


def test_camera_matches_oracle_bitwise():
    a, b = abi.Camera(), abi.Camera()
    for pose in [(0.35, 0.28, 5.0, 18.0), (0.6, 0.45, 2.4, 42.0), (2.0, -1.2, 3.0, 30.0),
                 (0.0, 1.5707964, 4.0, 20.0)]:
        _lib.lib().ixg_camera_new(*pose, 64, 48, 4, C.byref(a))
        oracle_lib.lib().ixo_camera_new(*pose, 64, 48, 4, C.byref(b))
        assert bytes(a) == bytes(b)


def test_cuts_match_oracle_bitwise():
    a = (abi.FacetPlane * 80)()
    b = (abi.FacetPlane * 80)()
    na = _lib.lib().ixg_cut_standard_round_brilliant(a)
    nb = oracle_lib.lib().ixo_cut_standard_round_brilliant(b)
    assert na == nb == 74
    assert bytes(a) == bytes(b)
    na = _lib.lib().ixg_cut_round_brilliant(34.5, 40.75, 0.57, a)
    nb = oracle_lib.lib().ixo_cut_round_brilliant(34.5, 40.75, 0.57, b)
    assert na == nb == 74 and bytes(a) == bytes(b)


def test_parametric_brilliant_reproduces_standard_geometry():
    """At the standard cut's own angles the parametric generator lands within a
    few 1e-4 of standard_round_brilliant's offsets (table 57.5 %)."""
    std = tracer.planes_array(tracer.standard_round_brilliant())
    par = tracer.planes_array(tracer.round_brilliant(34.5, 41.0, 0.5756))
    assert np.allclose(std[:, :3], par[:, :3], atol=1e-6)
    assert np.allclose(std[:, 3], par[:, 3], atol=2e-3)


def test_planes_are_unit_and_origin_interior():
    p = tracer.planes_array(workloads.film_c1_scene().planes)
    assert np.allclose(np.linalg.norm(p[:, :3], axis=1), 1.0, atol=1e-6)
    assert np.all(p[:, 3] < 0.0)


def test_route_predicate():
    sc = workloads.film_c1_scene(8, 8)
    assert tracer.route(sc) == abi.ROUTE_OK
    m = tracer.GemMaterial.diamond()
    m.cubic = False
    m.birefringence_delta = 0.009
    sc2 = tracer.FrameScene(sc.scene.camera, 8, 8, sc.planes, m, 16, sc.scene.environment)
    assert tracer.route(sc2) == abi.ROUTE_OK                  # uniaxial
    sc2.material_abi.biaxial = 1
    assert tracer.route(sc2) == abi.ROUTE_OK                  # biaxial
    sc2.material_abi.cubic = 1
    assert tracer.route(sc2) == abi.ROUTE_UNSUPPORTED_MATERIAL  # "biaxial" on a cubic stone
    for name in ("LightTent", "WhiteTray", "ShopLights", "WindowDaylight", "DaylightDome", "DaylightSun", "Aset"):
        env = tracer.LightingPreset.studio(name)
        sc3 = tracer.FrameScene(sc.scene.camera, 8, 8, sc.planes, tracer.GemMaterial.diamond(), 16, env)
        assert tracer.route(sc3) == abi.ROUTE_OK, name           # the lit models
    env.backdrop = 0.23                                              # a lit backdrop tray
    sc3 = tracer.FrameScene(sc.scene.camera, 8, 8, sc.planes, tracer.GemMaterial.diamond(), 16, env)
    assert tracer.route(sc3) == abi.ROUTE_OK
    for preset in (13, 14):             # the UV lamps (scene_routes_to_gpu declines them too)
        env = abi.Environment(preset, 1.0, 0.0, 0.5, 0.0, 1.0, 16.0)
        sc3 = tracer.FrameScene(sc.scene.camera, 8, 8, sc.planes, tracer.GemMaterial.diamond(), 16, env)
        assert tracer.route(sc3) == abi.ROUTE_UNSUPPORTED_ENVIRONMENT, preset
    env = abi.Environment(0, 1.0, 0.0, 0.5, float("nan"), 1.0, 16.0)
    sc3 = tracer.FrameScene(sc.scene.camera, 8, 8, sc.planes, tracer.GemMaterial.diamond(), 16, env)
    assert tracer.route(sc3) == abi.ROUTE_UNSUPPORTED_ENVIRONMENT
    big = np.tile(tracer.planes_array(sc.planes), (4, 1))
    sc4 = tracer.FrameScene(sc.scene.camera, 8, 8, big, tracer.GemMaterial.diamond(), 16,
                            sc.scene.environment)
    assert tracer.route(sc4) == abi.ROUTE_DEVICE_LIMITS
    assert tracer.route(workloads.film_absorbing_scene(8, 8)) == abi.ROUTE_OK
    sc5 = workloads.film_absorbing_scene(8, 8)
    sc5.material_abi.n_bands_o = abi.MAX_BANDS + 1      # more bands than a ray carries
    assert tracer.route(sc5) == abi.ROUTE_UNSUPPORTED_MATERIAL


def test_create_reports_status_without_crashing():
    ctx = C.c_void_p()
    st = _lib.lib().ixg_create(C.byref(abi.Opts(0, 0)), C.byref(ctx))
    assert st in (abi.IXG_OK, abi.IXG_DECLINED)
    if st == abi.IXG_OK:
        _lib.lib().ixg_destroy(ctx)


def test_sellmeier_from_nd_abbe_hits_targets():
    for nd, v in [(1.54, 20.0), (1.9, 40.0), (2.42, 60.0)]:
        b, c = tracer.sellmeier1_from_nd_abbe(nd, v)
        n = lambda l: (1 + b * l * l / (l * l - c)) ** 0.5
        assert abs(n(0.5893) - nd) < 1e-9
        assert abs((n(0.5893) - 1) / (n(0.4861) - n(0.6563)) - v) < 1e-6


def test_batch_entry_rejects_bad_arguments_without_reporting():
    """ixg_accumulate_batch: a null context or callback is IXG_INVALID and no
    item is reported (no device work is started)."""
    calls = []
    cb = abi.BatchDoneFn(lambda _u, i, s: calls.append((i, s)))
    items = (abi.BatchItem * 1)()
    L = _lib.lib()
    assert L.ixg_accumulate_batch(None, items, 1, None, cb, None) == abi.IXG_INVALID
    assert L.ixg_accumulate_batch(None, None, 0, None, cb, None) == abi.IXG_INVALID
    assert calls == []
