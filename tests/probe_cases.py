"""Inputs and budgets of the per-function probes (ixg_probe / ixo_probe).

Each case names the reference's tier-2 check it mirrors (the labels and
budgets of crates/indicatrix/docs/gpu_harness_last_run.json:24-73; the
budgets and absolute floors are the constants of the *_check.rs modules under
renderer/gpu/, e.g. transport_check/fresnel.rs:27-28, frame_and_tir.rs:44-45,
p2_uniaxial_fresnel.rs:99-100), the probe function, the scenes whose camera,
material, environment or facet set it reads, and a seeded input generator
shaped like what the film path feeds the function.  `checks` lists output
column groups with their (budget, abs_floor); `exact` lists integer columns
that must match bit for bit.
"""
import math

import numpy as np

from indicatrix_amd import abi, tracer, workloads

STOKES = np.array([[1, 0, 0, 0], [1, 1, 0, 0], [1, -1, 0, 0], [1, 0, 1, 0], [1, 0, 0, 1], [1, 0, 0, -1],
                   [0.7, 0.3, -0.2, 0.1], [1.0, 0.5, 0.5, 0.5], [0.25, -0.1, 0.15, -0.05]], np.float32)


def _u32(a):
    return np.asarray(a, np.uint32).view(np.float32)


def _unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _stokes(rng, n):
    """Physically valid Stokes vectors: I >= |(Q, U, V)|, with the fixed set first."""
    s = np.zeros((n, 4), np.float32)
    k = min(n, len(STOKES))
    s[:k] = STOKES[:k]
    m = n - k
    if m:
        i = rng.uniform(0.05, 2.0, m)
        d = _unit(rng, m) * (i * rng.uniform(0.0, 1.0, m))[:, None]
        s[k:, 0] = i
        s[k:, 1:] = d
    return s


def _facing(k, normal):
    """Flip normals so that they face the incident side: (-k) . n >= 0."""
    s = np.sign(np.einsum("ij,ij->i", -k.astype(np.float64), normal))
    s[s == 0] = 1.0
    return (normal * s[:, None]).astype(np.float32)


# --- scenes ---------------------------------------------------------------
def scene_c1():
    return workloads.film_c1_scene(128, 128)


def scene_c2():
    return workloads.film_c2_scene(64, 64, 1)


def scene_absorbing():
    return workloads.film_absorbing_scene(16, 16)


def scene_sapphire():
    return workloads.golden_scene("Daylight", 0.3, 0.6, material=tracer.GemMaterial.sapphire())


def scene_e_dispersion():
    """A uniaxial material with its own extraordinary Cauchy curve."""
    m = tracer.GemMaterial("cauchy-uniaxial", abi.DISP_CAUCHY, [1.70, 0.012, 0.0003], birefringence_delta=0.02,
                           cubic=False, c_axis=(0.3, 0.8, -0.52)).to_abi()
    m.has_e_dispersion = 1
    m.e_dispersion.kind = abi.DISP_SELLMEIER1
    b, c = tracer.sellmeier1_from_nd_abbe(1.74, 42.0)
    m.e_dispersion.coeff[0] = b
    m.e_dispersion.coeff[1] = c
    base = workloads.film_c1_scene(16, 16)
    return tracer.FrameScene(base.scene.camera, 16, 16, base.planes, m, 16, base.scene.environment)


def scene_preset(preset, yaw=0.4, pitch=0.7, backdrop=0.0):
    sc = workloads.golden_scene(preset, yaw, pitch)
    if backdrop:
        sc.scene.environment = tracer.LightingPreset.studio(preset, 1.0, yaw, pitch, backdrop=backdrop)
    return sc


def scene_hdr():
    sc = workloads.golden_scene("Daylight", 0.0, 0.0)
    rng = np.random.default_rng(7)
    pano = rng.uniform(0.0, 1.5, (32, 64, 3)).astype(np.float32)
    pano[8:11, 20:24] = (30.0, 28.0, 25.0)  # a bright lamp
    sc.set_hdr_map(pano)
    return sc


ENV_SCENES = [lambda p=p: scene_preset(p) for p in abi.PRESETS] + [
    lambda: scene_preset("Daylight", 2.1, 0.2, backdrop=0.35),
    lambda: scene_preset("WhiteTray", 1.0, 1.1, backdrop=0.2),
    lambda: scene_preset("DaylightSun", 0.9, 0.75)]


# --- input generators -------------------------------------------------------
def gen_sample_draws(rng, sc):
    n = 32768
    pix = np.concatenate([rng.integers(0, 2 ** 32, n // 2, dtype=np.uint64),
                          rng.integers(0, 1 << 20, n // 2, dtype=np.uint64)]).astype(np.uint32)
    smp = np.concatenate([rng.integers(0, 1 << 16, n - 3, dtype=np.uint64), [0, 1, 2 ** 32 - 1]]).astype(np.uint32)
    return np.stack([_u32(pix), _u32(smp)], axis=1)


def gen_camera(rng, sc):
    n = 32768
    pix = rng.integers(0, sc.width * sc.height, n).astype(np.uint32)
    smp = rng.integers(0, 4096, n).astype(np.uint32)
    return np.stack([_u32(pix), _u32(smp)], axis=1)


def gen_lambda(rng, sc):
    return np.concatenate([np.linspace(360.0, 830.0, 4701), rng.uniform(380.0, 780.0, 4000)]).astype(
        np.float32)[:, None]


def gen_environment(rng, sc):
    n = 16384
    d = _unit(rng, n)
    obs = _unit(rng, n)
    lam = rng.uniform(380.0, 780.0, n).astype(np.float32)
    return np.concatenate([d, obs, lam[:, None]], axis=1)


def gen_frame_rotation(rng, sc):
    psi = (np.arange(401) / 400.0 * 4.0 * math.pi - 2.0 * math.pi)
    rec = []
    for s in STOKES:
        scale = rng.uniform(0.5, 1.0, psi.size)  # unnormalised pairs, as the path's dot products
        rec.append(np.stack([np.cos(psi) * scale, np.sin(psi) * scale] + [np.full(psi.size, v) for v in s], axis=1))
    r = np.concatenate(rec).astype(np.float32)
    r[:, :2] = np.clip(r[:, :2], -1.0, 1.0)
    return r


def _interface(rng, n, tir=False):
    """n1, n2, cos_i, cos_t (f32, formed as the bounce forms them) for
    refraction or, with tir, n1 and cos_i / sin_i beyond the critical angle."""
    nn = rng.uniform(1.3, 2.9, n).astype(np.float32)
    if tir:
        crit = np.sqrt(1.0 - 1.0 / nn.astype(np.float64) ** 2)
        ci = (rng.uniform(0.0, 1.0, n) * crit * 0.999).astype(np.float32)
        si = np.sqrt(np.maximum(np.float32(1.0) - ci * ci, 0.0)).astype(np.float32)
        return nn, ci, si
    entering = rng.uniform(size=n) < 0.5
    n1 = np.where(entering, np.float32(1.0), nn).astype(np.float32)
    n2 = np.where(entering, nn, np.float32(1.0)).astype(np.float32)
    ci = rng.uniform(0.0, 1.0, n).astype(np.float32)
    ci[:8] = [1.0, 0.999, 0.5, 0.1, 0.02, 0.001, 1e-5, 0.0]
    eta = (n1 / n2).astype(np.float32)
    s2 = (eta * eta * (np.float32(1.0) - ci * ci)).astype(np.float32)
    ok = s2 <= 1.0
    ct = np.sqrt(np.maximum(np.float32(1.0) - s2, 0.0)).astype(np.float32)
    return n1[ok], n2[ok], ci[ok], ct[ok]


def gen_fresnel(rng, sc):
    n1, n2, ci, ct = _interface(rng, 16384)
    return np.concatenate([np.stack([n1, n2, ci, ct], axis=1), _stokes(rng, len(n1))], axis=1)


def gen_tir(rng, sc):
    n1, ci, si = _interface(rng, 8192, tir=True)
    return np.concatenate([np.stack([n1, ci, si], axis=1), _stokes(rng, len(n1))], axis=1)


def gen_channel_optics(rng, sc):
    n = 4096
    hero = rng.uniform(0.0, 1.0, n)
    lam = 380.0 + np.mod(hero[:, None] * 400.0 + np.arange(8)[None, :] * 50.0, 400.0)
    return lam.astype(np.float32)


def gen_pleochroic(rng, sc):
    n = 8192
    return np.concatenate([_unit(rng, n), rng.uniform(0.0, 2.0, (n, 16)).astype(np.float32)], axis=1)


def gen_two_vectors(rng, sc):
    n = 8192
    k = _unit(rng, n)
    c = _unit(rng, n)
    c[:64] = k[:64]  # a wave along the axis: the fallback basis
    c[64:128] = (k[64:128] + 1e-5 * _unit(rng, 64)).astype(np.float32)
    return np.concatenate([k, c], axis=1)


def gen_poynting(rng, sc):
    n = 8192
    no = rng.uniform(1.5, 2.5, n).astype(np.float32)
    d = rng.uniform(-0.3, 0.3, n).astype(np.float32)
    d[:32] = 0.0
    return np.concatenate([_unit(rng, n), _unit(rng, n), no[:, None], (no + d)[:, None]], axis=1)


def gen_theta_c(rng, sc):
    n = 8192
    lam = rng.uniform(380.0, 780.0, n).astype(np.float32)
    no = rng.uniform(1.55, 1.85, n).astype(np.float32)
    k = _unit(rng, n)
    normal = _facing(k, _unit(rng, n))
    inside = (rng.uniform(size=n) < 0.3).astype(np.float32)
    return np.concatenate([lam[:, None], no[:, None], k, normal, inside[:, None]], axis=1)


def gen_indices(rng, sc):
    n = 8192
    return np.stack([rng.uniform(380.0, 780.0, n), rng.uniform(1.5, 2.5, n),
                     rng.uniform(0.0, math.pi / 2, n)], axis=1).astype(np.float32)


def _uniaxial_geometry(rng, n):
    no = rng.uniform(1.5, 2.5, n).astype(np.float32)
    d = rng.choice([-1.0, 1.0], n) * rng.uniform(0.003, 0.3, n)
    ne = (no + d).astype(np.float32)
    c = _unit(rng, n)
    k = _unit(rng, n)
    normal = _facing(k, _unit(rng, n))
    cos_i = np.einsum("ij,ij->i", -k.astype(np.float64), normal)
    keep = cos_i > 0.02
    return no[keep], ne[keep], c[keep], k[keep], normal[keep]


def gen_entry(rng, sc):
    no, ne, c, k, normal = _uniaxial_geometry(rng, 8192)
    return np.concatenate([no[:, None], ne[:, None], c, k, normal], axis=1)


def _internal(rng, n):
    no, ne, c, k, normal = _uniaxial_geometry(rng, n)
    ordinary = (rng.uniform(size=len(no)) < 0.5)
    th = np.arccos(np.clip(np.abs(np.einsum("ij,ij->i", k.astype(np.float64), c)), 0.0, 1.0))
    neff = no * ne / np.sqrt((ne * np.cos(th)) ** 2 + (no * np.sin(th)) ** 2)
    n_inc = np.where(ordinary, no, neff).astype(np.float32)
    return np.concatenate([n_inc[:, None], no[:, None], ne[:, None], c, k, normal,
                           ordinary.astype(np.float32)[:, None]], axis=1)


def gen_internal(rng, sc):
    return _internal(rng, 8192)


def gen_exit(rng, sc):
    r = _internal(rng, 8192)
    m = len(r)
    return np.concatenate([r, rng.uniform(0.0, 1.0, m).astype(np.float32)[:, None],
                           rng.uniform(0.02, 0.98, m).astype(np.float32)[:, None]], axis=1)


def gen_hg(rng, sc):
    n = 9 * 1024
    g = np.repeat(np.array([-0.9, -0.5, -0.3, 0.0, 0.0005, 0.3, 0.5, 0.85, 0.95], np.float32), 1024)
    return np.concatenate([rng.uniform(0.0, 1.0, (n, 2)).astype(np.float32), g[:, None], _unit(rng, n)], axis=1)


def gen_cosine(rng, sc):
    n = 8192
    nrm = _unit(rng, n)
    nrm[:16] = [0.0, 1.0, 0.0]
    nrm[16:32] = [1.0, 0.0, 0.0]
    return np.concatenate([rng.uniform(0.0, 1.0, (n, 2)).astype(np.float32), nrm], axis=1)


def _outside_rays(rng, planes, n):
    """Rays from a sphere of radius 4 toward random points inside the stone."""
    pts = np.zeros((n, 3))
    i = 0
    while i < n:
        p = rng.uniform(-1.0, 1.0, (4 * n, 3))
        inside = np.all(p @ planes[:, :3].T.astype(np.float64) + planes[:, 3] < 0.0, axis=1)
        take = p[inside][: n - i]
        pts[i:i + len(take)] = take
        i += len(take)
    org = _unit(rng, n).astype(np.float64) * 4.0
    d = pts - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32), pts.astype(np.float32)


def gen_intersect(rng, sc):
    planes = tracer.planes_array(sc.planes)
    o1, d1, pts = _outside_rays(rng, planes, 8192)
    d2 = _unit(rng, 8192)  # from inside
    return np.concatenate([np.concatenate([o1, pts]), np.concatenate([d1, d2])], axis=1)


def gen_shading(rng, sc):
    """Hit points on the facets (first hits of outside rays, found by the
    oracle's own slab scan) with their facet and normal, and a radius."""
    import oracle_lib
    planes = tracer.planes_array(sc.planes)
    o, d, _ = _outside_rays(rng, planes, 8192)
    hit = oracle_lib.probe(sc, "intersect", np.concatenate([o, d], axis=1))
    f = hit[:, 0].view(np.int32)
    ok = f >= 0
    o, d, t, f = o[ok], d[ok], hit[ok, 1], f[ok]
    p = (o + d * t[:, None]).astype(np.float32)
    hn = planes[f, :3]
    radius = rng.choice([0.01, 0.03, 0.08], len(f)).astype(np.float32)
    return np.concatenate([p, _u32(f.astype(np.uint32))[:, None], hn, radius[:, None]], axis=1)


def gen_finish(rng, sc):
    n = 8192
    fam = rng.integers(0, 256, (n, 8)).astype(np.uint32) | (1 << np.arange(8, dtype=np.uint32))[None, :]
    fam[:64] = 0xFF
    lo = fam[:, 0] | (fam[:, 1] << 8) | (fam[:, 2] << 16) | (fam[:, 3] << 24)
    hi = fam[:, 4] | (fam[:, 5] << 8) | (fam[:, 6] << 16) | (fam[:, 7] << 24)
    pdf = rng.uniform(0.01, 1.0, (n, 8)).astype(np.float32)
    pdf[64:128, 1:] = 0.0
    lam = gen_channel_optics(rng, sc)[:n] if n <= 4096 else np.concatenate(
        [gen_channel_optics(rng, sc) for _ in range(n // 4096)])
    rad = rng.uniform(0.0, 3.0, (n, 8)).astype(np.float32)
    return np.concatenate([_u32(lo)[:, None], _u32(hi)[:, None], pdf, lam, rad], axis=1)


def gen_exp(rng, sc):
    return np.concatenate([np.linspace(-100.0, 100.0, 20001), rng.uniform(-90.0, 90.0, 20000),
                           [0.0, -0.0, 1e-30, -1e-30]]).astype(np.float32)[:, None]


def scene_alexandrite():
    return tracer.FrameScene(tracer.camera_new(0.35, 0.28, 5.0, 18.0, 16, 16, 1), 16, 16,
                             tracer.standard_round_brilliant(), tracer.GemMaterial.alexandrite(), 16,
                             tracer.LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0))


def scene_strong_biaxial():
    """A strongly biaxial stone (n_alpha ~ 1.60, n_beta ~ 1.65, n_gamma ~ 1.75)
    with a tilted gamma axis, so the general quadratic and large walk-offs
    are exercised away from the principal axes."""
    m = tracer.GemMaterial("strong biaxial", abi.DISP_CAUCHY, [1.64, 0.004, 0.0], birefringence_delta=0.15,
                           cubic=False, c_axis=(0.35, 0.82, -0.45))
    m.biaxial_delta_beta_alpha = 0.05
    m.bands_o = [(500.0, 40.0, 1.0, 0)]
    m.bands_e = [(600.0, 40.0, 2.0, 0)]
    m.bands_beta = [(450.0, 30.0, 1.5, 0)]
    return tracer.FrameScene(tracer.camera_new(0.35, 0.28, 5.0, 18.0, 16, 16, 1), 16, 16,
                             tracer.standard_round_brilliant(), m, 16,
                             tracer.LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0))


def _biax_k(rng, n):
    lam = rng.uniform(380.0, 780.0, n).astype(np.float32)[:, None]
    return lam, _unit(rng, n)


def gen_biax_indices(rng, sc):
    lam, k = _biax_k(rng, 4000)
    return np.concatenate([lam, k], axis=1)


def gen_biax_mode(rng, sc):
    lam, k = _biax_k(rng, 4000)
    slow = (rng.uniform(size=4000) < 0.5).astype(np.float32)[:, None]
    return np.concatenate([lam, k, slow], axis=1)


def gen_biax_entry(rng, sc):
    lam, k = _biax_k(rng, 4000)
    normal = _facing(k, _unit(rng, 4000))
    slow = (rng.uniform(size=4000) < 0.5).astype(np.float32)[:, None]
    return np.concatenate([lam, k, normal, slow], axis=1)


def gen_biax_alpha(rng, sc):
    rec = gen_biax_mode(rng, sc)
    alphas = rng.uniform(0.0, 3.0, (len(rec), 3)).astype(np.float32)
    return np.concatenate([rec, alphas], axis=1)


def gen_exp_rn(rng, sc):
    return np.concatenate([np.linspace(-103.0, 88.0, 20001), rng.uniform(-20.0, 20.0, 20000),
                           [0.0, -0.0, 1e-30, -1e-30]]).astype(np.float32)[:, None]


ALL = list(range(1000))

CASES = [
    dict(label="jx/jy/hero_rand/lambdas", fn="sample_draws", scenes=[scene_c1], gen=gen_sample_draws,
         checks=[([0, 1, 2] + list(range(4, 12)), 4, 1e-4)], exact=[3]),
    dict(label="camera ray generation", fn="camera_ray", scenes=[scene_c1], gen=gen_camera,
         checks=[(ALL, 64, 1e-4)]),
    dict(label="cie_1931_cmf", fn="cie_cmf", scenes=[scene_c1], gen=gen_lambda, checks=[(ALL, 32, 1e-6)]),
    dict(label="blackbody_spectrum / D65 illuminant", fn="illuminant",
         scenes=[scene_c1] + [lambda p=p: scene_preset(p) for p in ("Incandescent", "IlluminantA", "Daylight",
                                                                     "DarkSpotlight")],
         gen=gen_lambda, checks=[(ALL, 48, 1e-5)]),
    dict(label="sample_studio_environment", fn="environment", scenes=ENV_SCENES, gen=gen_environment,
         checks=[(ALL, 8192, 1e-4)]),
    dict(label="hdr environment", fn="environment", scenes=[scene_hdr], gen=gen_environment,
         checks=[(ALL, 4096, 1e-4)]),
    dict(label="frame_rotation + apply_matrix", fn="frame_rotation", scenes=[scene_c1], gen=gen_frame_rotation,
         checks=[(ALL, 16, 1e-6)]),
    dict(label="fresnel_reflection + apply_matrix", fn="fresnel_reflection", scenes=[scene_c1], gen=gen_fresnel,
         checks=[(ALL, 16, 1e-6)]),
    dict(label="fresnel_transmission + apply_matrix", fn="fresnel_transmission", scenes=[scene_c1],
         gen=gen_fresnel, checks=[([0, 1, 2, 3], 16, 1e-6), ([4, 5], 16, 1e-6)]),
    dict(label="tir_retardation + apply_matrix", fn="tir_retardation", scenes=[scene_c1], gen=gen_tir,
         checks=[(ALL, 16, 1e-6)]),
    dict(label="DispersionModel::evaluate", fn="channel_optics",
         scenes=[scene_c1, scene_c2, scene_absorbing, scene_sapphire, scene_e_dispersion], gen=gen_channel_optics,
         checks=[(list(range(8)), 24, 1e-6)]),
    dict(label="spectral_absorption", fn="channel_optics", scenes=[scene_c2, scene_absorbing, scene_sapphire],
         gen=gen_channel_optics, checks=[(list(range(8, 24)), 24, 1e-6)]),
    dict(label="pleochroic_channel_alpha", fn="pleochroic_alpha", scenes=[scene_c2, scene_absorbing, scene_sapphire],
         gen=gen_pleochroic, checks=[(ALL, 48, 1e-5)]),
    dict(label="ordinary_eigen_polarization", fn="ordinary_eigen", scenes=[scene_c1], gen=gen_two_vectors,
         checks=[(ALL, 32, 1e-5)]),
    dict(label="extraordinary_poynting_dir", fn="extraordinary_poynting", scenes=[scene_c1], gen=gen_poynting,
         checks=[(ALL, 32, 1e-5)]),
    dict(label="theta_c_for_bounce", fn="theta_c", scenes=[scene_c2, scene_sapphire, scene_e_dispersion],
         gen=gen_theta_c, checks=[(ALL, 96, 1e-5)]),
    dict(label="per_channel_uniaxial_indices", fn="uniaxial_indices", scenes=[scene_c2, scene_e_dispersion],
         gen=gen_indices, checks=[(ALL, 32, 1e-6)]),
    dict(label="entry_solve_pair (P2 full uniaxial Fresnel)", fn="entry_solve", scenes=[scene_c2], gen=gen_entry,
         checks=[(ALL, 4096, 1e-4)]),
    dict(label="internal_solve (P2 full uniaxial Fresnel)", fn="internal_solve", scenes=[scene_c2],
         gen=gen_internal, checks=[(ALL, 4096, 1e-4)]),
    dict(label="compute_uniaxial_exit_transmission", fn="exit_transmission", scenes=[scene_c2], gen=gen_exit,
         checks=[(ALL, 64, 1e-6)]),
    dict(label="sample_henyey_greenstein_direction", fn="hg_direction", scenes=[scene_c1], gen=gen_hg,
         checks=[(ALL, 128, 2e-4)]),
    dict(label="cosine_weighted_hemisphere", fn="cosine_hemisphere", scenes=[scene_c1], gen=gen_cosine,
         checks=[(ALL, 48, 1e-5)]),
    dict(label="shading_normal_near_edge", fn="shading_normal", scenes=[scene_c1], gen=gen_shading,
         checks=[(ALL, 1024, 1e-5)]),
    dict(label="PathState::finish (balance heuristic, CIE, von Kries)", fn="finish",
         scenes=[scene_c1, lambda: scene_preset("Incandescent")], gen=gen_finish, checks=[(ALL, 64, 1e-5)]),
    dict(label="slab intersection (facet exact, t)", fn="intersect", scenes=[scene_c1, scene_c2], gen=gen_intersect,
         checks=[([1], 512, 1e-6)], exact=[0]),
    dict(label="exp_f32x8 lane", fn="exp_lane", scenes=[scene_c1], gen=gen_exp, checks=[(ALL, 2, 0.0)]),
    dict(label="exp (correctly rounded)", fn="exp", scenes=[scene_c1], gen=gen_exp_rn, checks=[(ALL, 1, 0.0)]),
    dict(label="BiaxialIndicatrix::wave_indices", fn="biax_wave_indices",
         scenes=[scene_alexandrite, scene_strong_biaxial], gen=gen_biax_indices, checks=[(ALL, 64, 1e-5)]),
    dict(label="BiaxialIndicatrix::eigen_polarizations", fn="biax_eigen",
         scenes=[scene_alexandrite, scene_strong_biaxial], gen=gen_biax_mode, checks=[(ALL, 48, 5e-5)]),
    dict(label="BiaxialIndicatrix::mode_poynting_dir", fn="biax_mode_poynting",
         scenes=[scene_alexandrite, scene_strong_biaxial], gen=gen_biax_mode, checks=[(ALL, 48, 1e-5)]),
    dict(label="BiaxialIndicatrix::resolve_entry_mode", fn="biax_resolve_entry",
         scenes=[scene_alexandrite, scene_strong_biaxial], gen=gen_biax_entry, checks=[(ALL, 4096, 1e-5)]),
    dict(label="assigned_mode_alpha (biaxial, P1)", fn="biax_assigned_alpha",
         scenes=[scene_alexandrite, scene_strong_biaxial], gen=gen_biax_alpha, checks=[(ALL, 1200, 1e-5)]),
]


def case_inputs(case, scene_index, seed=1234):
    rng = np.random.default_rng(seed + 17 * scene_index + 1000 * abi.PROBE[case["fn"]])
    sc = case["scenes"][scene_index]()
    return sc, np.ascontiguousarray(case["gen"](rng, sc), np.float32)


def columns(cols, width):
    return [c for c in cols if c < width]
