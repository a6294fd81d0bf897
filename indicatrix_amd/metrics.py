"""Host mirror of the reference's metric API (color/metrics/) over libixg.

``evaluate_optical_image`` follows IX/color/metrics/image.rs:187 (request in,
per-pixel classes + totals out); ``metric_sweep`` batches many designs and
tilts into one device pass (the tilt sweeps of MET/sweep.rs, MET/pose_batch.rs).
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import IxgError, lib

PIXEL_DTYPE = np.dtype([("hits", "<u4"), ("classes", "<u4", (4,)), ("fire_sum_deg", "<f4"),
                        ("fire_qualified", "<u4"), ("fire_rejected", "<u4"),
                        ("centre_entry_facet", "<i4"), ("centre_exit_facet", "<i4"),
                        ("centre_bounces", "<u4"), ("centre_reason", "<u4"),
                        ("mixed_facets", "<u4"), ("scint_returned", "u1"), ("scint_poses", "u1"),
                        ("pad", "u1", (2,))])
METRIC_FIELDS = ["brilliance_pct", "fire_index", "scintillation_pct", "windowing_pct",
                 "extinction_pct"]


def make_request(camera, width, height, samples, planes, material_abi, environment,
                 want_fire=True, want_scintillation=False, pose=None):
    r = abi.MetricRequest()
    r.camera = camera
    r.width, r.height, r.samples = width, height, samples
    r.planes = C.cast(planes, C.POINTER(abi.FacetPlane))
    r.n_planes = len(planes)
    r.material = C.pointer(material_abi)
    r.environment = environment
    r.want_fire = int(want_fire)
    r.want_scintillation = int(want_scintillation)
    if pose is not None:
        r.has_pose = 1
        r.pose[0], r.pose[1], r.pose[2] = pose
    r._keep = (planes, material_abi)
    return r


def evaluate_optical_image(backend, request, want_pixels=True):
    n = request.width * request.height
    px = (abi.OpticalPixel * n)() if want_pixels else None
    tot = abi.MetricTotals()
    st = lib().ixg_metric_image(backend.handle, C.byref(request),
                                C.cast(px, C.POINTER(abi.OpticalPixel)) if want_pixels else None,
                                C.byref(tot))
    if st != abi.IXG_OK:
        raise IxgError(st, backend.last_error())
    pixels = np.frombuffer(bytes(px), PIXEL_DTYPE).copy() if want_pixels else None
    return pixels, tot


def _pixel_array(pixels):
    px = np.ascontiguousarray(pixels, PIXEL_DTYPE)
    arr = (abi.OpticalPixel * max(len(px), 1)).from_buffer_copy(px.tobytes() or bytes(C.sizeof(abi.OpticalPixel)))
    return arr, len(px)


def scintillation_maps(backend, pixels, width, height):
    """scintillation_spatial and scintillation_index over a finished map
    (image_scint.rs:175-247) on the device: (spatial, index) float32 arrays,
    index NaN where the temporal term was not computed."""
    arr, n = _pixel_array(pixels)
    spatial = np.zeros(n, np.float32)
    index = np.zeros(n, np.float32)
    st = lib().ixg_scintillation_maps(backend.handle, arr, n, width, height,
                                      spatial.ctypes.data_as(C.POINTER(C.c_float)),
                                      index.ctypes.data_as(C.POINTER(C.c_float)))
    if st != abi.IXG_OK:
        raise IxgError(st, backend.last_error())
    return spatial, index


def scintillation_maps_device(backend, d_pixels_ptr, n, width, height, d_spatial_ptr, d_index_ptr):
    """The same on device buffers (raw device pointers)."""
    st = lib().ixg_scintillation_maps_device(backend.handle, C.c_void_p(d_pixels_ptr), n, width, height,
                                             C.c_void_p(d_spatial_ptr), C.c_void_p(d_index_ptr))
    if st != abi.IXG_OK:
        raise IxgError(st, backend.last_error())


def scintillation_spatial_at(pixels, width, height, i):
    """scintillation_spatial_at (image_scint.rs:201-232), the hover read-out."""
    arr, n = _pixel_array(pixels)
    return lib().ixg_scintillation_spatial_at(arr, n, width, height, i)


def scintillation_index_at(pixels, width, height, i):
    """scintillation_index_at (image_scint.rs:249-259): a float, or None."""
    arr, n = _pixel_array(pixels)
    out = C.c_float()
    return out.value if lib().ixg_scintillation_index_at(arr, n, width, height, i, C.byref(out)) else None


def metric_sweep(backend, planes, designs, tilts_deg, width, height, samples, distance,
                 environment):
    """planes: (P,4) float32 table; designs: list of (offset, count, abi.Dispersion).
    Returns (n_designs, n_tilts, 5) float32 metrics in METRIC_FIELDS order."""
    planes = np.ascontiguousarray(planes, np.float32).reshape(-1, 4)
    pl = (abi.FacetPlane * len(planes)).from_buffer_copy(planes.tobytes())
    ds = (abi.Design * len(designs))()
    for i, (off, cnt, disp) in enumerate(designs):
        ds[i] = abi.Design(off, cnt, disp)
    tilts = np.ascontiguousarray(tilts_deg, np.float32)
    out = np.zeros((len(designs), len(tilts), 5), np.float32)
    st = lib().ixg_metric_sweep(backend.handle, pl, len(planes), ds, len(designs),
                                tilts.ctypes.data_as(C.POINTER(C.c_float)), len(tilts), width,
                                height, samples, distance, C.byref(environment),
                                out.ctypes.data_as(C.POINTER(C.c_float)))
    if st != abi.IXG_OK:
        raise IxgError(st, backend.last_error())
    return out


def _design_table(planes, designs):
    planes = np.ascontiguousarray(planes, np.float32).reshape(-1, 4)
    pl = (abi.FacetPlane * max(1, len(planes))).from_buffer_copy(
        planes.tobytes() if len(planes) else bytes(C.sizeof(abi.FacetPlane)))
    ds = (abi.Design * len(designs))()
    for i, (off, cnt, disp) in enumerate(designs):
        ds[i] = abi.Design(off, cnt, disp)
    return pl, len(planes), ds


def model_width(planes):
    """measure_model_width (render_setup/stone_width.rs:27): the width in model
    units, or None when the planes do not bound a finite solid."""
    planes = np.ascontiguousarray(planes, np.float32).reshape(-1, 4)
    pl = (abi.FacetPlane * max(1, len(planes))).from_buffer_copy(
        planes.tobytes() if len(planes) else bytes(C.sizeof(abi.FacetPlane)))
    w = C.c_double()
    st = lib().ixg_model_width(pl, len(planes), C.byref(w))
    if st == abi.IXG_DECLINED:
        return None
    if st != abi.IXG_OK:
        raise IxgError(st, "model width")
    return w.value


def pose_metrics(backend, planes, designs, poses, environment, want_all=True):
    """evaluate_gem_optical_metrics (color/metrics/evaluate.rs:158) for every
    (design, (yaw, pitch)) pair in one device pass.  Returns
    (n_designs, n_poses, 5) float32 in METRIC_FIELDS order; want_all=False is
    the Fire-less pass (fire 25, scintillation 75)."""
    pl, n, ds = _design_table(planes, designs)
    poses = np.ascontiguousarray(poses, np.float32).reshape(-1, 2)
    out = np.zeros((len(designs), len(poses), 5), np.float32)
    st = lib().ixg_pose_metrics(backend.handle, pl, n, ds, len(designs),
                                poses.ctypes.data_as(C.POINTER(C.c_float)), len(poses),
                                C.byref(environment), 1 if want_all else 0,
                                out.ctypes.data_as(C.POINTER(C.c_float)))
    if st != abi.IXG_OK:
        raise IxgError(st, backend.last_error())
    return out


def axis_profiles(backend, planes, designs, azimuths_deg, environment):
    """evaluate_full_axis_profile_at_azimuth (color/metrics/profile.rs:200) per
    (design, azimuth): (n_designs, n_azimuths, 3, 181) float32 curves of
    brilliance, extinction and windowing over tilt -90..90 degrees."""
    pl, n, ds = _design_table(planes, designs)
    az = np.ascontiguousarray(azimuths_deg, np.float32)
    out = np.zeros((len(designs), len(az), 3, 181), np.float32)
    st = lib().ixg_axis_profiles(backend.handle, pl, n, ds, len(designs),
                                 az.ctypes.data_as(C.POINTER(C.c_float)), len(az),
                                 C.byref(environment), out.ctypes.data_as(C.POINTER(C.c_float)))
    if st != abi.IXG_OK:
        raise IxgError(st, backend.last_error())
    return out


def set_metric_mode(backend, use_graph=True, block=0, tile=0):
    """How metric_sweep launches: graph capture on/off; block 0 = autotune,
    -1 = the 2-D grid kernel, a queue shape (block lanes, tile pixels), or
    -64 / -128 / -256 = the flow kernel with that many lanes per block and
    tile pixels per item (0 = the whole image)."""
    st = lib().ixg_set_metric_mode(backend.handle, 1 if use_graph else 0, block, tile)
    if st != abi.IXG_OK:
        raise IxgError(st, backend.last_error())


def metric_tuning(backend):
    """(block, tile, grid_blocks, graphed) of the last metric_sweep."""
    b, t, g, gr = C.c_int(), C.c_int(), C.c_uint32(), C.c_int()
    lib().ixg_metric_tuning(backend.handle, C.byref(b), C.byref(t), C.byref(g), C.byref(gr))
    return b.value, t.value, g.value, bool(gr.value)


def metric_last_ms(backend):
    """Device time (ms) of the last metric_sweep's launches (HIP events on the
    library's stream; autotuning excluded)."""
    return float(lib().ixg_metric_last_ms(backend.handle))
