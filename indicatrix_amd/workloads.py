"""The named workloads: scenes built from their stated parameters.

film_c1: 128x128 @ 64 spp, max depth 16; a 57-facet round brilliant (74
half-spaces with the 16-facet girdle band and culet) at crown 34.5 deg,
pavilion 40.75 deg, table 57 % of the girdle diameter; Diamond (Peter 1923
Sellmeier, n_D ~ 2.417); a uniform D65 hemispherical dome (IsoHemisphere with
the observer head shadow off); camera yaw 0.35, pitch 0.28, distance 5, fov 18
deg (the reference's estimator-check camera, renderer/gpu/estimator_check/mod.rs:142-146).
"""
import math

import numpy as np

from . import abi, designs, tracer

C1 = dict(width=128, height=128, spp=64, max_bounces=16, crown_deg=34.5, pavilion_deg=40.75,
          table_ratio=0.57, camera=(0.35, 0.28, 5.0, 18.0))


def film_c1_scene(width=None, height=None, max_bounces=None):
    w = width or C1["width"]
    h = height or C1["height"]
    planes = tracer.round_brilliant(C1["crown_deg"], C1["pavilion_deg"], C1["table_ratio"])
    cam = tracer.camera_new(*C1["camera"], w, h, C1["spp"])
    env = tracer.LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0, head_shadow_deg=0.0)
    return tracer.FrameScene(cam, w, h, planes, tracer.GemMaterial.diamond(),
                             max_bounces or C1["max_bounces"], env)


def golden_scene(preset, light_yaw, light_pitch, max_bounces=12, width=1, height=1, material=None):
    """The reference golden-regression scene: standard round brilliant, Diamond
    (or the given material, e.g. GemMaterial.sapphire())."""
    planes = tracer.standard_round_brilliant()
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, width, height, 1)
    env = tracer.LightingPreset.studio(preset, 1.0, light_yaw, light_pitch)
    return tracer.FrameScene(cam, width, height, planes, material or tracer.GemMaterial.diamond(),
                             max_bounces, env)


def film_absorbing_scene(width=32, height=24, max_bounces=16):
    """The C1 geometry and camera with an absorbing cubic material: two
    wavelength-domain bands and one wavenumber-domain band, different o / e
    coefficients and a tilted optic axis (the tensor's unpolarised mean over
    the two eigen-directions then depends on the ray direction)."""
    base = film_c1_scene(width, height, max_bounces)
    m = tracer.GemMaterial.diamond()
    m.c_axis = (0.48, 0.6, 0.64)
    m.bands_o = [(450.0, 35.0, 0.9, 0), (610.0, 1800.0, 0.6, 1)]
    m.bands_e = [(520.0, 55.0, 1.4, 0)]
    m.absorption_path_scale = 1.3
    return tracer.FrameScene(base.scene.camera, width, height, base.planes, m, max_bounces,
                             base.scene.environment)


C2 = dict(width=1024, height=1024, spp=256, max_bounces=24, facets=121, seed=2, aspect=1.3,
          camera=(0.35, 0.28, 5.0, 30.0))


def c2_material(seed=C2["seed"], birefringence=0.009, bands=True):
    """A seeded uniaxial, pleochroic material: a 3-term Sellmeier ordinary curve
    (n_D ~ U[1.60, 1.80], Abbe ~ U[30, 60], split over a visible, a UV and a
    weak IR pole), n_e = n_o + 0.009, the optic axis uniform on the sphere, and
    two Gaussian o-bands and two e-bands with centres U[420, 650] nm, widths
    U[20, 60] nm and peaks U[0, 2] /cm (= 0.1 x per-mm coefficients; the model
    unit is taken as 3.25 mm, half a 6.5 mm stone)."""
    rng = np.random.default_rng(20_000 + seed)
    nd = float(rng.uniform(1.60, 1.80))
    abbe = float(rng.uniform(30.0, 60.0))
    b, c = tracer.sellmeier1_from_nd_abbe(nd, abbe)
    coeff = [0.6 * b, 0.4 * b, 0.02, c, 0.8 * c, 100.0]
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    m = tracer.GemMaterial("uniaxial-pleochroic", abi.DISP_SELLMEIER3, coeff,
                           birefringence_delta=birefringence, cubic=False,
                           c_axis=tuple(float(v) for v in axis), absorption_path_scale=3.25)
    if bands:
        def band():
            return (float(rng.uniform(420.0, 650.0)), float(rng.uniform(20.0, 60.0)),
                    float(rng.uniform(0.0, 2.0)) * 0.1, 0)
        m.bands_o = [band(), band()]
        m.bands_e = [band(), band()]
    return m


def c2_planes(seed=C2["seed"]):
    """The 121-facet oval (tier angles U[30, 50] deg, 1.3:1) from the seed."""
    return designs.design_planes(designs.seeded_params(seed, n_facets=C2["facets"], angle_range=(30.0, 50.0),
                                                       aspect=C2["aspect"]))


def film_c2_scene(width=None, height=None, spp=None, max_bounces=None, material=None):
    w = width or C2["width"]
    h = height or C2["height"]
    cam = tracer.camera_new(*C2["camera"], w, h, spp or C2["spp"])
    env = tracer.LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0, head_shadow_deg=0.0)
    return tracer.FrameScene(cam, w, h, c2_planes(), material or c2_material(),
                             max_bounces or C2["max_bounces"], env)
