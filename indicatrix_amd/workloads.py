"""The named workloads: scenes built from their stated parameters.

film_c1: 128x128 @ 64 spp, max depth 16; a 57-facet round brilliant (74
half-spaces with the 16-facet girdle band and culet) at crown 34.5 deg,
pavilion 40.75 deg, table 57 % of the girdle diameter; Diamond (Peter 1923
Sellmeier, n_D ~ 2.417); a uniform D65 hemispherical dome (IsoHemisphere with
the observer head shadow off); camera yaw 0.35, pitch 0.28, distance 5, fov 18
deg (the reference's estimator-check camera, renderer/gpu/estimator_check/mod.rs:142-146).
"""
from . import tracer

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


def golden_scene(preset, light_yaw, light_pitch, max_bounces=12, width=1, height=1):
    """The reference golden-regression scene: standard round brilliant, Diamond."""
    planes = tracer.standard_round_brilliant()
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, width, height, 1)
    env = tracer.LightingPreset.studio(preset, 1.0, light_yaw, light_pitch)
    return tracer.FrameScene(cam, width, height, planes, tracer.GemMaterial.diamond(),
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
