"""The oracle's biaxial indicatrix and biaxial transport, pinned.

The reference's own properties (crates/indicatrix/src/optics/birefringence/
tests/biaxial_reduction.rs): the uniaxial and isotropic reductions of
wave_indices, orthonormal transverse eigen-polarisations, the eigen-equation
residual over a dense sweep for the built-in biaxial materials, and the
walk-off facts of poynting_direction.  Closed forms the mathematics fixes:
the index pairs along the principal axes, the two roots meeting along the
optic axes at the textbook angle, a positive uniaxial degeneracy whose slow
mode walks off exactly as the uniaxial extraordinary formula
(birefringence/uniaxial.rs) says, Snell's law at the fixed point of
resolve_entry_mode, and the analytic white furnace through the biaxial
transport (a non-absorbing biaxial stone in a uniform panorama returns the
panorama's spectrum within the 16-bounce truncation, as the isotropic one
does: tests/test_hdr_env.py).
"""
import math

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, tracer
from oracle_lib import Biaxial


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def n_eff(n_o, n_e, theta):
    """BirefringenceParams::effective_extraordinary_index in float64."""
    s, c = math.sin(theta), math.cos(theta)
    return n_o * n_e / math.sqrt((n_e * c) ** 2 + (n_o * s) ** 2)


DIRS = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (0.2, 0.9, 0.1), (-0.5, 0.3, 0.8), (0.7, -0.6, 0.2),
        (0.1, 0.1, 0.99), (-0.9, -0.3, 0.2)]


def test_wave_indices_reduce_to_the_uniaxial_formula():
    # biaxial_reduction.rs:21-58 (n_alpha == n_beta: the closed-form branch)
    n_o, n_e = 1.65, 1.72
    c = unit([0.35, 0.82, -0.45])
    ind = Biaxial.from_gamma_axis(n_o, n_o, n_e, c)
    for d in DIRS + [tuple(c)]:
        k = unit(d)
        ns, nf = ind.wave_indices(k)
        th = math.acos(min(1.0, abs(float(np.dot(k, c)))))
        ne = n_eff(n_o, n_e, th)
        assert abs(nf - min(n_o, ne)) < 1e-3 and abs(ns - max(n_o, ne)) < 1e-3


def test_general_quadratic_tends_to_the_uniaxial_formula():
    # the general branch (n_beta - n_alpha above the degeneracy tolerance) a
    # little off the uniaxial case: roots within the perturbation of the
    # uniaxial ones (an index moves by at most the perturbation)
    n_o, n_e, eps = 1.65, 1.72, 2e-3
    c = unit([0.35, 0.82, -0.45])
    ind = Biaxial.from_gamma_axis(n_o - eps, n_o, n_e, c)
    for d in DIRS:
        k = unit(d)
        ns, nf = ind.wave_indices(k)
        th = math.acos(min(1.0, abs(float(np.dot(k, c)))))
        ne = n_eff(n_o, n_e, th)
        assert -eps - 1e-5 <= nf - min(n_o, ne) <= 1e-5
        assert -eps - 1e-5 <= ns - max(n_o, ne) <= 1e-5


def test_wave_indices_reduce_to_isotropic_when_all_equal():
    # biaxial_reduction.rs:63-84
    ind = Biaxial(1.9, 1.9, 1.9, np.eye(3))
    for d in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (0.3, 0.5, 0.8)]:
        ns, nf = ind.wave_indices(unit(d))
        assert abs(ns - 1.9) < 1e-4 and abs(nf - 1.9) < 1e-4


def test_principal_axis_index_pairs():
    # along a principal axis D lies in the other two: the two indices are
    # exactly those two principal indices
    na, nb, ng = 1.60, 1.65, 1.75
    frame = np.array([unit([0.6, 0.8, 0.0]), unit([-0.8, 0.6, 0.0]), [0.0, 0.0, 1.0]], np.float32)
    ind = Biaxial(na, nb, ng, frame)
    for axis, pair in [(frame[0], (ng, nb)), (frame[1], (ng, na)), (frame[2], (nb, na))]:
        ns, nf = ind.wave_indices(axis)
        assert abs(ns - pair[0]) < 2e-6 and abs(nf - pair[1]) < 2e-6


def test_roots_meet_along_the_optic_axes():
    # the wave-normal optic axes lie in the alpha-gamma plane at angle V from
    # gamma with sin^2 V = (1/nb^2 - 1/na^2) / (1/ng^2 - 1/na^2); there both
    # roots are n_beta, and the split opens away from it
    na, nb, ng = 1.60, 1.65, 1.75
    ind = Biaxial(na, nb, ng, np.eye(3))
    s2 = (nb ** -2 - na ** -2) / (ng ** -2 - na ** -2)
    V = math.asin(math.sqrt(s2))
    for sign in (1.0, -1.0):
        k = unit([sign * math.sin(V), 0.0, math.cos(V)])
        ns, nf = ind.wave_indices(k)
        assert abs(ns - nb) < 5e-5 and abs(nf - nb) < 5e-5
        k5 = unit([sign * math.sin(V + 0.09), 0.0, math.cos(V + 0.09)])
        a, b = ind.wave_indices(k5)
        assert a - b > 5e-3


def test_eigen_polarizations_orthonormal_and_transverse():
    # biaxial_reduction.rs:90-130
    ind = Biaxial(1.60, 1.65, 1.75, np.eye(3))
    for d in [(0.3, 0.5, 0.8), (0.9, 0.1, -0.2), (-0.4, 0.6, 0.5)]:
        k = unit(d)
        ds, df = ind.eigen_d(k, True), ind.eigen_d(k, False)
        assert np.all(np.isfinite(ds)) and np.all(np.isfinite(df))
        assert abs(np.linalg.norm(ds) - 1) < 1e-3 and abs(np.linalg.norm(df) - 1) < 1e-3
        assert abs(ds @ df) < 1e-2 and abs(ds @ k) < 1e-2 and abs(df @ k) < 1e-2


BUILT_IN = {  # (dispersion kind, coefficients, birefringence_delta, delta_beta_alpha)
    "Alexandrite": (abi.DISP_SELLMEIER3, [0.78522, 1.21202, 16.81, 0.0, 0.01262, 1000.0], 0.0076, 0.001951),
    "Topaz": (abi.DISP_CAUCHY, [1.614872, 0.004273, 0.0], 0.0080, 0.0024),
    "Tanzanite": (abi.DISP_CAUCHY, [1.674589, 0.009123, 0.0], 0.0130, 0.0025),
}


def built_in(name, lam=589.3):
    """GemMaterial::biaxial_indicatrix (materials/optics.rs:16-27) of a built-in."""
    kind, coeff, bd, dba = BUILT_IN[name]
    d = abi.Dispersion()
    d.kind = kind
    for i, v in enumerate(coeff):
        d.coeff[i] = v
    nb = np.float32(oracle_lib.lib().ixo_dispersion_evaluate(d, lam))
    na = np.float32(nb - np.float32(dba))
    ng = np.float32(na + np.float32(bd))
    return Biaxial.from_gamma_axis(na, nb, ng, (0.0, 1.0, 0.0))


@pytest.mark.parametrize("name", list(BUILT_IN))
def test_eigenvectors_satisfy_the_eigen_equation_across_a_dense_sweep(name):
    # biaxial_reduction.rs:150-226: D_i (x - b_i) = -k_i (k . E), E_j = b_j D_j,
    # in the principal frame; unit length; continuous up to sign
    ind = built_in(name)
    b = [1.0 / (n * n) for n in ind.n]
    u1, u2 = unit([0.4, 0.6, 0.69]), unit([0.8, -0.3, 0.2])
    prev = [None, None]
    for i in range(181):
        th = i / 180.0 * math.pi
        k = unit(math.cos(th) * u1.astype(np.float64) + math.sin(th) * u2.astype(np.float64))
        local = ind.axes @ k
        ns, nf = ind.wave_indices(k)
        for j, (n, slow) in enumerate([(ns, True), (nf, False)]):
            dh = ind.eigen_d(k, slow)
            assert abs(np.linalg.norm(dh) - 1.0) < 1e-4
            x = 1.0 / (float(n) * float(n))
            dl = ind.axes @ dh
            e_dot_k = sum(local[a] * b[a] * dl[a] for a in range(3))
            res = np.array([dl[a] * (x - b[a]) + local[a] * e_dot_k for a in range(3)])
            assert np.linalg.norm(res) < 2e-4, (name, th, np.linalg.norm(res))
            if prev[j] is not None:
                assert abs(float(prev[j] @ dh)) > 0.98
            prev[j] = dh


def test_poynting_direction_without_and_with_walk_off():
    # biaxial_reduction.rs:301-330
    k = unit([0.0, 0.0, 1.0])
    assert np.array_equal(oracle_lib.poynting_direction(k, unit([1.0, 0.0, 0.0])), k)
    e = unit([1.0, 0.0, 0.1])
    s = oracle_lib.poynting_direction(k, e)
    assert np.all(np.isfinite(s)) and abs(np.linalg.norm(s) - 1.0) < 1e-5
    assert abs(s @ e) < 1e-6 and s @ k > 0.99 and abs(s[0]) > 1e-3


def test_positive_uniaxial_degeneracy_walks_off_as_the_uniaxial_formula():
    # n_alpha == n_beta = n_o, n_gamma = n_e > n_o: the slow mode is the
    # extraordinary wave, whose ray direction is k turned toward / away from
    # the optic axis by tan(rho) = (n_o^2 - n_e^2) sin cos / (n_e^2 cos^2 +
    # n_o^2 sin^2) in the k-c plane; the fast (ordinary) mode does not walk off
    n_o, n_e = 1.65, 1.72
    c = unit([0.35, 0.82, -0.45]).astype(np.float64)
    ind = Biaxial.from_gamma_axis(n_o, n_o, n_e, c)
    for d in DIRS[3:]:
        k = unit(d).astype(np.float64)
        ct = float(k @ c)
        ac, st = abs(ct), math.sqrt(max(0.0, 1.0 - ct * ct))
        rho = math.atan((n_o ** 2 - n_e ** 2) * st * ac / (n_e ** 2 * ac * ac + n_o ** 2 * st * st))
        cp = c - k * ct
        cp /= np.linalg.norm(cp)
        expect = k * math.cos(rho) - cp * math.copysign(1.0, ct) * math.sin(rho)
        s_slow = ind.mode_poynting(k.astype(np.float32), True).astype(np.float64)
        assert np.linalg.norm(s_slow - expect) < 2e-5, (d, s_slow, expect)
        s_fast = ind.mode_poynting(k.astype(np.float32), False).astype(np.float64)
        assert np.linalg.norm(s_fast - k) < 2e-6


def test_mode_ray_directions_are_transverse_to_their_e_fields():
    ind = built_in("Alexandrite")
    rng = np.random.default_rng(7)
    for _ in range(200):
        k = unit(rng.normal(size=3))
        for slow in (True, False):
            e = ind.d_to_e(ind.eigen_d(k, slow))
            s = ind.mode_poynting(k, slow)
            assert abs(float(s @ e)) < 2e-6 and float(s @ k) > 0.999


def test_d_to_e_on_the_principal_axes():
    ind = Biaxial(1.60, 1.65, 1.75, np.eye(3))
    for a in range(3):
        d = np.zeros(3, np.float32)
        d[a] = 1.0
        assert np.array_equal(ind.d_to_e(d), d)
    d = unit([1.0, 1.0, 0.0])
    e = ind.d_to_e(d)
    ratio = (1.0 / 1.60 ** 2) / (1.0 / 1.65 ** 2)
    assert abs(e[0] / e[1] - ratio) < 1e-5 and e[2] == 0.0


@pytest.mark.parametrize("slow", [True, False])
def test_resolve_entry_mode_obeys_snell_at_its_fixed_point(slow):
    ind = built_in("Topaz")
    rng = np.random.default_rng(3 + slow)
    normal = unit([0.1, 1.0, -0.2])
    for _ in range(100):
        k = unit(-normal.astype(np.float64) + 0.9 * rng.normal(size=3))
        cos_i = float(np.clip(-(k @ normal), 0.0, 1.0))
        if cos_i < 0.05:
            continue
        n, w = ind.resolve_entry(k, normal, cos_i, ind.n[1], slow)
        ns, nf = ind.wave_indices(w)
        assert n == (ns if slow else nf)
        # w is Snell's refraction of k for the previous step's index: the
        # implied index is within a small fraction of the birefringence of n
        # (two fixed-point steps from n_beta); w stays in the plane of incidence
        sin_i = math.sqrt(1.0 - cos_i * cos_i)
        sin_t = np.linalg.norm(np.cross(w.astype(np.float64), normal.astype(np.float64)))
        spread = ind.n[2] - ind.n[0]
        resid = abs(sin_i / sin_t - float(n))
        assert resid < 0.05 * spread
        assert abs(float(np.cross(k, normal).astype(np.float64) @ w)) < 1e-6


def biaxial_scene(material, grid=12, spp=32, preset="RingLights", bounces=16):
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, grid, grid, spp)
    env = tracer.LightingPreset.studio(preset, 1.0, 0.85, 0.95)
    return tracer.FrameScene(cam, grid, grid, tracer.standard_round_brilliant(), material, bounces, env)


def test_biaxial_white_furnace():
    from test_hdr_env import furnace_target
    m = tracer.GemMaterial("furnace probe", abi.DISP_CAUCHY, [1.5, 0.0, 0.0], birefringence_delta=0.02,
                           cubic=False, c_axis=(0.3, 0.8, 0.52))
    m.biaxial_delta_beta_alpha = 0.008
    sc = biaxial_scene(m, spp=96)
    sc.set_hdr_map(np.full((1, 1, 3), 2.5, np.float32))
    acc = oracle_lib.accumulate(sc, 0, 96)
    mean = acc.sum(axis=0) / (12 * 12 * 96)
    target = furnace_target(2.5)
    assert (np.abs(mean - target) / target).max() < 0.01, (mean, target)


def test_biaxial_frames_are_finite_and_differ_from_the_beta_index_stone():
    alex = biaxial_scene(tracer.GemMaterial.alexandrite(), spp=16)
    acc = oracle_lib.accumulate(alex, 0, 16)
    assert np.all(np.isfinite(acc)) and acc.sum() > 0
    # the same stone with the trichroic beta bands dropped absorbs differently
    m = tracer.GemMaterial.alexandrite()
    m.bands_beta = None
    acc2 = oracle_lib.accumulate(biaxial_scene(m, spp=16), 0, 16)
    assert abs(acc2.sum() - acc.sum()) / acc.sum() > 1e-3


def test_biaxial_side_channel_records_both_modes():
    sc = biaxial_scene(tracer.GemMaterial.topaz(), spp=8)
    xyz, ev = oracle_lib.trace_paths(sc, 0, 8)
    assert np.all(np.isfinite(xyz))
    assert (ev["n_refract"] > 0).mean() > 0.3


def test_route_for_biaxial_stones():
    assert tracer.route(biaxial_scene(tracer.GemMaterial.alexandrite())) == abi.ROUTE_OK
    m = tracer.GemMaterial.topaz().to_abi()
    m.scattering_sigma_s = 0.5
    assert tracer.route(biaxial_scene(m)) == abi.ROUTE_OK        # scattering biaxial stones route
    m = tracer.GemMaterial.topaz()
    m.cubic = True
    assert tracer.route(biaxial_scene(m)) == abi.ROUTE_UNSUPPORTED_MATERIAL
    sc = biaxial_scene(tracer.GemMaterial.topaz())
    sc.set_finishes([1] + [0] * (len(sc.planes) - 1))
    assert tracer.route(sc) == abi.ROUTE_OK                      # frosted biaxial stones route
    sc.set_hdr_map(np.full((2, 4, 3), 1.0, np.float32))
    assert tracer.route(sc) == abi.ROUTE_UNSUPPORTED_ENVIRONMENT  # ... but not under a panorama


def test_biaxial_probe_dispatch_is_the_exported_functions():
    # the oracle's probe records go through the same functions the exported
    # entry points call, at the scene's indicatrix for each record's lambda
    import probe_cases
    sc = probe_cases.scene_strong_biaxial()
    rng = np.random.default_rng(5)
    rec = probe_cases.gen_biax_entry(rng, sc)[:64]
    m = sc.material_abi
    d = m.dispersion
    for r in rec:
        nb = np.float32(oracle_lib.lib().ixo_dispersion_evaluate(d, float(r[0])))
        na = np.float32(nb - np.float32(m.biaxial_delta_beta_alpha))
        ind = Biaxial.from_gamma_axis(na, nb, np.float32(na + np.float32(m.birefringence_delta)), m.c_axis[:])
        k, normal, slow = r[1:4], r[4:7], bool(r[7])
        wi = oracle_lib.probe(sc, "biax_wave_indices", np.concatenate([r[:4]])[None, :])[0]
        assert wi.tobytes() == np.array(ind.wave_indices(k), np.float32).tobytes()
        ed = oracle_lib.probe(sc, "biax_eigen", np.concatenate([r[:4], [r[7]]])[None, :].astype(np.float32))[0]
        assert ed.tobytes() == ind.eigen_d(k, slow).tobytes()
        sp = oracle_lib.probe(sc, "biax_mode_poynting", np.concatenate([r[:4], [r[7]]])[None, :].astype(np.float32))[0]
        assert sp.tobytes() == ind.mode_poynting(k, slow).tobytes()
        w = oracle_lib.probe(sc, "biax_resolve_entry", r[None, :].astype(np.float32))[0]
        dots = np.float32(np.float32(np.float32(-k[0]) * normal[0]) + np.float32(np.float32(-k[1]) * normal[1]))
        cos_i = np.float32(np.clip(np.float32(dots + np.float32(np.float32(-k[2]) * normal[2])), 0.0, 1.0))
        _, w2 = ind.resolve_entry(k, normal, cos_i, nb, slow)
        assert w.tobytes() == w2.tobytes()
