"""ctypes binding of the CPU oracle (oracle/_build/libixo.so) for the tests.

The oracle is the checker only; nothing in indicatrix_amd imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from indicatrix_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "libixo.so")
LIB64 = os.path.join(ORACLE_DIR, "_build", "libixo64.so")
LIB_SCALAR = os.path.join(ORACLE_DIR, "_build", "libixo_scalar.so")
LIB_FLOPS = os.path.join(ORACLE_DIR, "_build", "libixo_flops.so")
_lib = None
_lib64 = None
_lib_scalar = None
_lib_flops = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        f3 = C.c_float * 3
        L.ixo_hash_u32.restype = C.c_uint32
        L.ixo_hash_u32.argtypes = [C.c_uint32]
        L.ixo_low_discrepancy_base2.restype = C.c_float
        L.ixo_low_discrepancy_base2.argtypes = [C.c_uint32]
        L.ixo_radical_inverse_base.restype = C.c_float
        L.ixo_radical_inverse_base.argtypes = [C.c_uint32, C.c_uint32]
        L.ixo_sample_draws.argtypes = [C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_float),
                                       P(C.c_float), P(C.c_float)]
        L.ixo_wrapped_hero_wavelengths.argtypes = [C.c_float, P(C.c_float)]
        L.ixo_dispersion_evaluate.restype = C.c_float
        L.ixo_dispersion_evaluate.argtypes = [P(abi.Dispersion), C.c_float]
        L.ixo_cie_cmf.argtypes = [C.c_float, P(C.c_float)]
        L.ixo_d65.restype = C.c_float
        L.ixo_d65.argtypes = [C.c_float]
        L.ixo_exp_lane.restype = C.c_float
        L.ixo_exp_lane.argtypes = [C.c_float]
        L.ixo_uniaxial_energy.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, f3, f3, f3, C.c_int,
                                          P(C.c_float)]
        L.ixo_uniaxial_entry_r.argtypes = [C.c_float, C.c_float, f3, f3, f3, C.c_int, P(C.c_float)]
        L.ixo_band_evaluate.restype = C.c_float
        L.ixo_band_evaluate.argtypes = [P(abi.Band), C.c_float]
        L.ixo_sample_studio_environment.restype = C.c_float
        L.ixo_sample_studio_environment.argtypes = [P(C.c_float), C.c_float, C.c_int32,
                                                    C.c_float, C.c_float, C.c_float]
        L.ixo_camera_new.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32,
                                     C.c_uint32, C.c_uint32, P(abi.Camera)]
        L.ixo_cut_standard_round_brilliant.restype = C.c_uint32
        L.ixo_cut_standard_round_brilliant.argtypes = [P(abi.FacetPlane)]
        L.ixo_cut_round_brilliant.restype = C.c_uint32
        L.ixo_cut_round_brilliant.argtypes = [C.c_float, C.c_float, C.c_float, P(abi.FacetPlane)]
        L.ixo_intersect.restype = C.c_int32
        L.ixo_intersect.argtypes = [P(abi.FacetPlane), C.c_uint32, P(C.c_float), P(C.c_float),
                                    P(C.c_float)]
        L.ixo_slab_scan.restype = C.c_int32
        L.ixo_slab_scan.argtypes = [P(abi.FacetPlane), C.c_uint32, P(C.c_float), P(C.c_float),
                                    P(C.c_float)]
        L.ixo_exp_lane8.argtypes = [P(C.c_float), P(C.c_float)]
        L.ixo_trace_ray.argtypes = [P(abi.Scene), f3, f3, C.c_uint32, C.c_float, P(C.c_float),
                                    P(abi.PathEvents)]
        L.ixo_generate_ray.argtypes = [P(abi.Camera), C.c_float, C.c_float, C.c_float, C.c_float,
                                       P(C.c_float), P(C.c_float)]
        L.ixo_trace_sample.argtypes = [P(abi.Scene), C.c_uint32, C.c_uint32, P(C.c_float),
                                       P(abi.PathEvents)]
        L.ixo_accumulate.argtypes = [P(abi.Scene), C.c_uint32, C.c_uint32, P(C.c_double), C.c_int]
        L.ixo_trace_paths.argtypes = [P(abi.Scene), C.c_uint32, C.c_uint32, C.c_uint64,
                                      C.c_uint64, P(C.c_float), P(abi.PathEvents), C.c_int]
        L.ixo_metric_image.restype = C.c_int
        L.ixo_metric_image.argtypes = [P(abi.MetricRequest), P(abi.OpticalPixel), P(abi.MetricTotals)]
        L.ixo_image_scintillation_pct.restype = C.c_float
        L.ixo_image_scintillation_pct.argtypes = [P(abi.OpticalPixel), C.c_uint32]
        PX = P(abi.OpticalPixel)
        L.ixo_scintillation_spatial.argtypes = [PX, C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_float)]
        L.ixo_scintillation_spatial_at.restype = C.c_float
        L.ixo_scintillation_spatial_at.argtypes = [PX, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.ixo_scintillation_index.restype = C.c_int
        L.ixo_scintillation_index.argtypes = [PX, C.c_uint32, P(C.c_float), C.c_uint32, C.c_uint32,
                                              P(C.c_float)]
        L.ixo_scintillation_index_at.restype = C.c_int
        L.ixo_scintillation_index_at.argtypes = [PX, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                                 P(C.c_float)]
        L.ixo_scintillation_indices.argtypes = [PX, C.c_uint32, P(C.c_float), P(C.c_float)]
        L.ixo_image_offset.argtypes = [C.c_uint32, P(C.c_float), P(C.c_float)]
        L.ixo_temporal_cameras.argtypes = [P(C.c_float), C.c_float, C.c_uint32, C.c_uint32, P(abi.Camera)]
        L.ixo_classify_ray.restype = C.c_int
        L.ixo_classify_ray.argtypes = [P(abi.FacetPlane), C.c_uint32, P(abi.Dispersion), P(abi.Environment),
                                       P(C.c_float), P(C.c_float), P(C.c_float)]
        L.ixo_model_width.restype = C.c_int
        L.ixo_model_width.argtypes = [P(abi.FacetPlane), C.c_uint32, P(C.c_double)]
        L.ixo_girdle_half_width.restype = C.c_float
        L.ixo_girdle_half_width.argtypes = [P(abi.FacetPlane), C.c_uint32]
        L.ixo_pose_metrics.restype = C.c_int
        L.ixo_pose_metrics.argtypes = [P(abi.FacetPlane), C.c_uint32, P(abi.Material),
                                       P(abi.Environment), C.c_float, C.c_float, C.c_int,
                                       P(C.c_float)]
        L.ixo_shading_normal_near_edge.argtypes = [P(abi.FacetPlane), C.c_uint32, P(C.c_float), C.c_int32,
                                                   P(C.c_float), C.c_float, P(C.c_float)]
        L.ixo_cosine_hemisphere.argtypes = [C.c_float, C.c_float, P(C.c_float), P(C.c_float)]
        L.ixo_hg_direction.argtypes = [C.c_float, C.c_float, C.c_float, P(C.c_float), P(C.c_float)]
        L.ixo_rgb_to_spectral.restype = C.c_float
        L.ixo_rgb_to_spectral.argtypes = [P(C.c_float), C.c_float]
        L.ixo_asc_planes.restype = C.c_uint32
        L.ixo_asc_planes.argtypes = [P(C.c_double), P(C.c_double), P(C.c_uint32), C.c_uint32, P(C.c_double),
                                     C.c_int32, C.c_double, P(abi.FacetPlane), C.c_uint32]
        L.ixo_solid_outline.restype = C.c_int
        L.ixo_solid_outline.argtypes = [P(abi.FacetPlane), C.c_uint32, P(C.c_double), P(C.c_uint32)]
        f9 = C.c_float * 9
        for name in ("ixo_biax_eigen_d", "ixo_biax_mode_poynting"):
            getattr(L, name).argtypes = [C.c_float, C.c_float, C.c_float, f9, f3, C.c_int, P(C.c_float)]
        L.ixo_biax_from_gamma_axis.argtypes = [f3, P(C.c_float)]
        L.ixo_biax_wave_indices.argtypes = [C.c_float, C.c_float, C.c_float, f9, f3, P(C.c_float)]
        L.ixo_biax_d_to_e.argtypes = [C.c_float, C.c_float, C.c_float, f9, f3, P(C.c_float)]
        L.ixo_biax_resolve_entry.restype = C.c_float
        L.ixo_biax_resolve_entry.argtypes = [C.c_float, C.c_float, C.c_float, f9, f3, f3, C.c_float, C.c_float,
                                             C.c_int, P(C.c_float)]
        L.ixo_poynting_direction.argtypes = [f3, f3, P(C.c_float)]
        L.ixo_sun_cone_direction.argtypes = [f3, C.c_float, C.c_float, P(C.c_float)]
        L.ixo_probe_shape.restype = C.c_int
        L.ixo_probe_shape.argtypes = [C.c_int32, P(C.c_uint32), P(C.c_uint32)]
        L.ixo_probe.restype = C.c_int
        L.ixo_probe.argtypes = [P(abi.Scene), C.c_int32, P(C.c_float), C.c_uint32, P(C.c_float)]
        _lib = L
    return _lib


def probe(scene, fn, records):
    """ixo_probe: the oracle's function for probe `fn` (a name of
    abi.PROBE_FNS) on float32 records (n, in_floats) -> (n, out_floats)."""
    L = lib()
    fid = abi.PROBE[fn]
    w_in, w_out = C.c_uint32(), C.c_uint32()
    assert L.ixo_probe_shape(fid, C.byref(w_in), C.byref(w_out))
    rec = np.ascontiguousarray(records, np.float32).reshape(-1, w_in.value)
    out = np.zeros((len(rec), w_out.value), np.float32)
    assert L.ixo_probe(scene.ref(), fid, rec.ctypes.data_as(C.POINTER(C.c_float)), len(rec),
                       out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def lib64():
    """The FP64 twin (oracle/ixo64.c): the same estimator in double precision."""
    global _lib64
    if _lib64 is None:
        if not os.path.exists(LIB64):
            build()
        L = C.CDLL(LIB64)
        P = C.POINTER
        L.ixo64_accumulate.argtypes = [P(abi.Scene), C.c_uint32, C.c_uint32, P(C.c_double), C.c_int]
        L.ixo64_trace_paths.argtypes = [P(abi.Scene), C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64,
                                        P(C.c_double), P(abi.PathEvents), C.c_int]
        _lib64 = L
    return _lib64


def lib_scalar():
    """The FP32 oracle built without the AVX2 scan and exponential."""
    global _lib_scalar
    if _lib_scalar is None:
        if not os.path.exists(LIB_SCALAR):
            build()
        L = C.CDLL(LIB_SCALAR)
        P = C.POINTER
        L.ixo_accumulate.argtypes = [P(abi.Scene), C.c_uint32, C.c_uint32, P(C.c_double), C.c_int]
        L.ixo_metric_image.restype = C.c_int
        L.ixo_metric_image.argtypes = [P(abi.MetricRequest), P(abi.OpticalPixel), P(abi.MetricTotals)]
        L.ixo_exp_lane8.argtypes = [P(C.c_float), P(C.c_float)]
        _lib_scalar = L
    return _lib_scalar


def lib_flops():
    """The FLOP-counting build (oracle/ixo_flops.cpp)."""
    global _lib_flops
    if _lib_flops is None:
        if not os.path.exists(LIB_FLOPS):
            build()
        L = C.CDLL(LIB_FLOPS)
        P = C.POINTER
        L.ixo_flops_accumulate.argtypes = [P(abi.Scene), C.c_uint32, C.c_uint32, P(C.c_double),
                                           P(C.c_uint64)]
        L.ixo_flops_metric_image.restype = C.c_int
        L.ixo_flops_metric_image.argtypes = [P(abi.MetricRequest), P(abi.OpticalPixel), P(abi.MetricTotals),
                                             P(C.c_uint64)]
        L.ixo_flops_intersect.restype = C.c_int32
        L.ixo_flops_intersect.argtypes = [P(abi.FacetPlane), C.c_uint32, P(C.c_float), P(C.c_float),
                                          P(C.c_float), P(C.c_uint64)]
        _lib_flops = L
    return _lib_flops


def accumulate_flops(scene, sample_offset, spp):
    """(per-pixel sums, FP operations) of the counting build, one thread."""
    acc = np.zeros((scene.width * scene.height, 3), np.float64)
    n = C.c_uint64(0)
    lib_flops().ixo_flops_accumulate(scene.ref(), sample_offset, spp,
                                     acc.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n))
    return acc, n.value


def accumulate64(scene, sample_offset, spp, nthreads=None):
    """Per-pixel FP64 sums of the FP64 twin, (pixels, 3)."""
    n = scene.width * scene.height
    acc = np.zeros((n, 3), np.float64)
    lib64().ixo64_accumulate(scene.ref(), sample_offset, spp, acc.ctypes.data_as(C.POINTER(C.c_double)),
                             nthreads or threads())
    return acc


def trace_paths64(scene, sample_offset, spp, first=0, count=None, nthreads=None):
    """Per-path XYZ (float64) and events of the FP64 twin."""
    from indicatrix_amd import tracer
    total = scene.width * scene.height * spp
    count = total - first if count is None else count
    xyz = np.zeros((count, 3), np.float64)
    ev = (abi.PathEvents * count)()
    lib64().ixo64_trace_paths(scene.ref(), sample_offset, spp, first, count,
                              xyz.ctypes.data_as(C.POINTER(C.c_double)), ev, nthreads or threads())
    return xyz, tracer.events_to_numpy(ev)


def metric_image(request):
    """The oracle's evaluate_optical_image -> (pixels ndarray, totals, scintillation %)."""
    from indicatrix_amd.metrics import PIXEL_DTYPE
    n = request.width * request.height
    px = (abi.OpticalPixel * n)()
    tot = abi.MetricTotals()
    st = lib().ixo_metric_image(C.byref(request), px, C.byref(tot))
    if st != 0:
        raise RuntimeError(f"oracle metric_image status {st}")
    scint = lib().ixo_image_scintillation_pct(px, n)
    return np.frombuffer(bytes(px), PIXEL_DTYPE).copy(), tot, scint


def pixels_ctypes(px):
    """A PIXEL_DTYPE ndarray as a ctypes OpticalPixel array (copied)."""
    n = len(px)
    arr = (abi.OpticalPixel * max(n, 1))()
    C.memmove(arr, np.ascontiguousarray(px).tobytes(), n * C.sizeof(abi.OpticalPixel))
    return arr


def scintillation_spatial(px, width, height):
    """scintillation_spatial (image_scint.rs:175-199) over a pixel array."""
    n = len(px)
    out = (C.c_float * max(n, 1))()
    lib().ixo_scintillation_spatial(pixels_ctypes(px), n, width, height, out)
    return np.frombuffer(bytes(out), np.float32)[:n].copy()


def scintillation_indices(px, spatial):
    """scintillation_index of every pixel, NaN for None."""
    n = len(px)
    sp = np.ascontiguousarray(spatial, np.float32)
    out = np.zeros(n, np.float32)
    lib().ixo_scintillation_indices(pixels_ctypes(px), n, sp.ctypes.data_as(C.POINTER(C.c_float)),
                                    out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def scintillation_spatial_at(px, width, height, i):
    return lib().ixo_scintillation_spatial_at(pixels_ctypes(px), len(px), width, height, i)


def scintillation_index(px, spatial, i):
    """scintillation_index (image_scint.rs:234-247): a float, or None."""
    out = C.c_float()
    sp = np.ascontiguousarray(spatial, np.float32)
    ok = lib().ixo_scintillation_index(pixels_ctypes(px), len(px), sp.ctypes.data_as(C.POINTER(C.c_float)),
                                       len(sp), i, C.byref(out))
    return out.value if ok else None


def scintillation_index_at(px, width, height, i):
    out = C.c_float()
    ok = lib().ixo_scintillation_index_at(pixels_ctypes(px), len(px), width, height, i, C.byref(out))
    return out.value if ok else None


def image_offset(s):
    jx, jy = C.c_float(), C.c_float()
    lib().ixo_image_offset(s, C.byref(jx), C.byref(jy))
    return jx.value, jy.value


def temporal_cameras(pose, fov_tan, width, height):
    out = (abi.Camera * 5)()
    lib().ixo_temporal_cameras((C.c_float * 3)(*pose), fov_tan, width, height, out)
    return list(out)


def classify_ray(planes, dispersion, environment, origin, direction, observer):
    arr = (abi.FacetPlane * len(planes))(*planes)
    f3 = C.c_float * 3
    return lib().ixo_classify_ray(arr, len(planes), C.byref(dispersion), C.byref(environment),
                                  f3(*origin), f3(*direction), f3(*observer))


def threads():
    return max(1, min(16, os.cpu_count() or 1))


def accumulate(scene, sample_offset, spp, nthreads=None):
    acc = np.zeros((scene.width * scene.height, 3), np.float64)
    lib().ixo_accumulate(scene.ref(), sample_offset, spp,
                         acc.ctypes.data_as(C.POINTER(C.c_double)), nthreads or threads())
    return acc


def trace_paths(scene, sample_offset, spp, first=0, count=None, nthreads=None):
    from indicatrix_amd.tracer import events_to_numpy
    total = scene.width * scene.height * spp
    count = total - first if count is None else count
    xyz = np.zeros((count, 3), np.float32)
    ev = (abi.PathEvents * count)()
    lib().ixo_trace_paths(scene.ref(), sample_offset, spp, first, count,
                          xyz.ctypes.data_as(C.POINTER(C.c_float)),
                          C.cast(ev, C.POINTER(abi.PathEvents)), nthreads or threads())
    return xyz, events_to_numpy(ev)


def sweep_pose(tilt_deg, distance):
    """(yaw, pitch) of a tilt-sweep camera: yaw 0 / 180 deg by the tilt's sign,
    pitch 90 - |tilt| (MET/pose_batch.rs:118-131), in the f32 arithmetic the
    library uses."""
    to_rad = np.float32(np.float32(np.pi) / np.float32(180.0))
    t = np.float32(tilt_deg)
    yaw = float(np.float32(180.0) * to_rad) if t < 0 else 0.0
    pitch = float(np.float32(np.float32(90.0) - np.float32(abs(t))) * to_rad)
    return yaw, pitch


def sweep_requests(planes, designs, tilts, width, height, samples, distance, environment):
    """The metric requests of a sweep, one per (design, tilt), as
    (design index, tilt index, request); each request keeps its planes and
    material alive."""
    from indicatrix_amd import metrics, tracer
    planes = np.ascontiguousarray(planes, np.float32).reshape(-1, 4)
    jobs = []
    for i, (off, cnt, disp) in enumerate(designs):
        pl = (abi.FacetPlane * cnt).from_buffer_copy(planes[off:off + cnt].tobytes())
        mat = abi.Material()
        mat.dispersion = disp
        mat.cubic = 1
        mat.absorption_path_scale = 1.0
        for j, t in enumerate(tilts):
            yaw, pitch = sweep_pose(t, distance)
            cam = tracer.camera_new(yaw, pitch, distance, 42.0, width, height, samples)
            req = metrics.make_request(cam, width, height, samples, pl, mat, environment, True, True,
                                       (yaw, pitch, distance))
            req._keep = (pl, mat)
            jobs.append((i, j, req))
    return jobs


def metric_image_flops(request):
    """FP operations of the oracle's evaluate_optical_image on one request
    (the counting build, one thread)."""
    n = request.width * request.height
    px = (abi.OpticalPixel * n)()
    tot = abi.MetricTotals()
    count = C.c_uint64(0)
    st = lib_flops().ixo_flops_metric_image(C.byref(request), px, C.byref(tot), C.byref(count))
    if st != 0:
        raise RuntimeError(f"ixo_flops_metric_image returned {st}")
    return count.value


def metric_sweep(planes, designs, tilts, width, height, samples, distance, environment,
                 nthreads=None):
    """The oracle's per-image evaluation of every (design, tilt) pair: (n, t, 5)
    {brilliance, fire, scintillation, windowing, extinction}, images spread over
    host threads."""
    from concurrent.futures import ThreadPoolExecutor
    jobs = sweep_requests(planes, designs, tilts, width, height, samples, distance, environment)
    out = np.zeros((len(designs), len(tilts), 5), np.float32)

    def run(job):
        i, j, req = job
        _, tot, scint = metric_image(req)
        out[i, j] = (tot.brilliance_pct, tot.fire_index, scint, tot.windowing_pct, tot.extinction_pct)

    with ThreadPoolExecutor(max_workers=nthreads or threads()) as ex:
        list(ex.map(run, jobs))
    return out


def pose_metrics(planes, dispersion, environment, poses, want_all=True, nthreads=None):
    """The oracle's evaluate_pose for every (yaw, pitch): (n_poses, 5), poses
    spread over host threads (each call is independent)."""
    from concurrent.futures import ThreadPoolExecutor
    planes = np.ascontiguousarray(planes, np.float32).reshape(-1, 4)
    pl = (abi.FacetPlane * max(1, len(planes))).from_buffer_copy(
        planes.tobytes() if len(planes) else bytes(C.sizeof(abi.FacetPlane)))
    mat = abi.Material()
    mat.dispersion = dispersion
    mat.cubic = 1
    mat.absorption_path_scale = 1.0
    out = np.zeros((len(poses), 5), np.float32)

    def run(k):
        o = (C.c_float * 5)()
        st = lib().ixo_pose_metrics(pl, len(planes), C.byref(mat), C.byref(environment),
                                    float(poses[k][0]), float(poses[k][1]), 1 if want_all else 0, o)
        if st != 0:
            raise RuntimeError(f"oracle pose_metrics status {st}")
        out[k] = o[:]

    with ThreadPoolExecutor(max_workers=nthreads or threads()) as ex:
        list(ex.map(run, range(len(poses))))
    return out


def axis_poses(azimuth_deg):
    """axis_poses (color/metrics/pose_batch.rs:118-131) in f32: the pole, then
    pitch 0..89 deg at the azimuth, then at the azimuth + 180 deg."""
    to_rad = np.float32(np.pi) / np.float32(180.0)
    pos = np.float32(azimuth_deg) * to_rad
    neg = np.float32(np.float32(azimuth_deg) + np.float32(180.0)) * to_rad
    poses = [(pos, np.float32(90.0) * to_rad)]
    poses += [(pos, np.float32(i) * to_rad) for i in range(90)]
    poses += [(neg, np.float32(i) * to_rad) for i in range(90)]
    return poses


def asc_planes(schedule):
    """The oracle's planes for a parsed schedule (tiers read from its fields)."""
    n = len(schedule.tiers)
    ang = (C.c_double * max(1, n))(*[t.angle_deg for t in schedule.tiers])
    mast = (C.c_double * max(1, n))(*[t.mast for t in schedule.tiers])
    counts = (C.c_uint32 * max(1, n))(*[len(t.indices) for t in schedule.tiers])
    flat = [i for t in schedule.tiers for i in t.indices]
    idx = (C.c_double * max(1, len(flat)))(*flat)
    cap = max(1, schedule.facet_plane_count())
    out = (abi.FacetPlane * cap)()
    k = lib().ixo_asc_planes(ang, mast, counts, n, idx, schedule.gear_teeth,
                             schedule.gear_reference_angle, out, cap)
    return (abi.FacetPlane * k)(*out[:k])


def solid_outline(planes):
    """(length/width ratio, planes without a face) of the solid, or None."""
    lw = C.c_double()
    bare = C.c_uint32()
    ok = lib().ixo_solid_outline(planes, len(planes), C.byref(lw), C.byref(bare))
    return (lw.value, bare.value) if ok else None


# ---- the biaxial indicatrix (optics/birefringence/biaxial.rs) ----------
class Biaxial:
    """One BiaxialIndicatrix: principal indices and the alpha / beta / gamma
    principal directions (the columns of `axes`, a 3x3 array)."""

    def __init__(self, na, nb, ng, axes):
        self.n = (float(na), float(nb), float(ng))
        self.axes = np.asarray(axes, np.float32).reshape(3, 3)  # rows = columns of the Mat3

    @staticmethod
    def from_gamma_axis(na, nb, ng, gamma_axis):
        out = (C.c_float * 9)()
        lib().ixo_biax_from_gamma_axis((C.c_float * 3)(*gamma_axis), out)
        return Biaxial(na, nb, ng, np.array(out[:], np.float32))

    def _ax(self):
        return (C.c_float * 9)(*[float(v) for v in self.axes.reshape(-1)])

    def wave_indices(self, k):
        out = (C.c_float * 2)()
        lib().ixo_biax_wave_indices(*self.n, self._ax(), (C.c_float * 3)(*k), out)
        return np.float32(out[0]), np.float32(out[1])

    def eigen_d(self, k, want_slow):
        out = (C.c_float * 3)()
        lib().ixo_biax_eigen_d(*self.n, self._ax(), (C.c_float * 3)(*k), int(want_slow), out)
        return np.array(out[:], np.float32)

    def d_to_e(self, d):
        out = (C.c_float * 3)()
        lib().ixo_biax_d_to_e(*self.n, self._ax(), (C.c_float * 3)(*d), out)
        return np.array(out[:], np.float32)

    def mode_poynting(self, k, want_slow):
        out = (C.c_float * 3)()
        lib().ixo_biax_mode_poynting(*self.n, self._ax(), (C.c_float * 3)(*k), int(want_slow), out)
        return np.array(out[:], np.float32)

    def resolve_entry(self, incident, normal, cos_i, n_seed, want_slow):
        out = (C.c_float * 3)()
        n = lib().ixo_biax_resolve_entry(*self.n, self._ax(), (C.c_float * 3)(*incident),
                                         (C.c_float * 3)(*normal), float(cos_i), float(n_seed),
                                         int(want_slow), out)
        return np.float32(n), np.array(out[:], np.float32)


def poynting_direction(k, e):
    out = (C.c_float * 3)()
    lib().ixo_poynting_direction((C.c_float * 3)(*k), (C.c_float * 3)(*e), out)
    return np.array(out[:], np.float32)
