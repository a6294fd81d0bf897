"""Loader for the in-tree HIP library (libixg.so).

The product path has no CPU fallback: if the library is missing or cannot be
loaded this raises, naming the build step.
"""
import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libixg.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "ixg.h")

_lib = None


class IxgError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        name = abi.STATUS_NAMES.get(status, str(status))
        super().__init__(f"ixg status {name}: {message}" if message else f"ixg status {name}")


def _declare(lib):
    P = C.POINTER
    st = C.c_int
    sig = {
        "ixg_create": (st, [P(abi.Opts), P(C.c_void_p)]),
        "ixg_destroy": (None, [C.c_void_p]),
        "ixg_is_lost": (C.c_int, [C.c_void_p]),
        "ixg_adapter_label": (st, [C.c_void_p, C.c_char_p, C.c_size_t]),
        "ixg_last_error": (st, [C.c_void_p, C.c_char_p, C.c_size_t]),
        "ixg_try_recover": (st, [C.c_void_p]),
        "ixg_inject_fault": (st, [C.c_void_p, C.c_int32]),
        "ixg_route": (C.c_int, [P(abi.Scene)]),
        "ixg_accumulate": (st, [C.c_void_p, P(abi.Scene), C.c_uint32, C.c_uint32,
                                P(C.c_double), P(C.c_uint8)]),
        "ixg_accumulate_device": (st, [C.c_void_p, P(abi.Scene), C.c_uint32, C.c_uint32,
                                       C.c_void_p]),
        "ixg_accumulate_device_noalloc": (st, [C.c_void_p, P(abi.Scene), C.c_uint32, C.c_uint32,
                                               C.c_void_p]),
        "ixg_accumulate_batch": (st, [C.c_void_p, P(abi.BatchItem), C.c_uint32, P(C.c_uint8),
                                      abi.BatchDoneFn, C.c_void_p]),
        "ixg_trace_paths": (st, [C.c_void_p, P(abi.Scene), C.c_uint32, C.c_uint32,
                                 P(C.c_float), P(abi.PathEvents)]),
        "ixg_film_samples": (st, [C.c_void_p, P(abi.Scene), C.c_uint32, C.c_uint32, P(C.c_float)]),
        "ixg_trace_rays": (st, [C.c_void_p, P(abi.Scene), P(C.c_float), C.c_uint32,
                                P(C.c_float), P(abi.PathEvents)]),
        "ixg_stream": (C.c_void_p, [C.c_void_p]),
        "ixg_render_chunk": (st, [C.c_void_p, P(abi.Scene), C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_float),
                                  P(C.c_uint32), P(C.c_double), P(C.c_uint32), P(C.c_uint8)]),
        "ixg_next_subbatch": (C.c_uint32, [C.c_uint32, C.c_double]),
        "ixg_probe_shape": (C.c_int, [C.c_int32, P(C.c_uint32), P(C.c_uint32)]),
        "ixg_probe": (st, [C.c_void_p, P(abi.Scene), C.c_int32, P(C.c_float), C.c_uint32, P(C.c_float)]),
        "ixg_count_events": (st, [C.c_void_p, P(abi.Scene), C.c_uint32, C.c_uint32,
                                  P(C.c_uint64)]),
        "ixg_metric_image": (st, [C.c_void_p, P(abi.MetricRequest), P(abi.OpticalPixel),
                                  P(abi.MetricTotals)]),
        "ixg_metric_sweep": (st, [C.c_void_p, P(abi.FacetPlane), C.c_uint32, P(abi.Design),
                                  C.c_uint32, P(C.c_float), C.c_uint32, C.c_uint32, C.c_uint32,
                                  C.c_uint32, C.c_float, P(abi.Environment), P(C.c_float)]),
        "ixg_scintillation_maps": (st, [C.c_void_p, P(abi.OpticalPixel), C.c_uint32, C.c_uint32, C.c_uint32,
                                        P(C.c_float), P(C.c_float)]),
        "ixg_scintillation_maps_device": (st, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                               C.c_void_p, C.c_void_p]),
        "ixg_scintillation_spatial_at": (C.c_float, [P(abi.OpticalPixel), C.c_uint32, C.c_uint32, C.c_uint32,
                                                     C.c_uint32]),
        "ixg_scintillation_index_at": (C.c_int, [P(abi.OpticalPixel), C.c_uint32, C.c_uint32, C.c_uint32,
                                                 C.c_uint32, P(C.c_float)]),
        "ixg_set_metric_mode": (st, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
        "ixg_metric_last_ms": (C.c_double, [C.c_void_p]),
        "ixg_metric_tuning": (st, [C.c_void_p, P(C.c_int), P(C.c_int), P(C.c_uint32), P(C.c_int)]),
        "ixg_model_width": (st, [P(abi.FacetPlane), C.c_uint32, P(C.c_double)]),
        "ixg_pose_metrics": (st, [C.c_void_p, P(abi.FacetPlane), C.c_uint32, P(abi.Design), C.c_uint32,
                                  P(C.c_float), C.c_uint32, P(abi.Environment), C.c_int, P(C.c_float)]),
        "ixg_axis_profiles": (st, [C.c_void_p, P(abi.FacetPlane), C.c_uint32, P(abi.Design), C.c_uint32,
                                   P(C.c_float), C.c_uint32, P(abi.Environment), P(C.c_float)]),
        "ixg_asc_parse_text": (st, [C.c_char_p, C.c_size_t, P(C.c_void_p), P(abi.AscError)]),
        "ixg_asc_parse_bytes": (st, [C.c_char_p, C.c_size_t, P(C.c_void_p), P(abi.AscError)]),
        "ixg_asc_free": (None, [C.c_void_p]),
        "ixg_radiance_shuffle_device": (st, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
        "ixg_radiance_unshuffle_device": (st, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
        "ixg_radiance_add_device": (st, [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p,
                                         C.c_void_p]),
        "ixg_radiance_check_payload": (st, [C.c_int32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                            C.c_int32, P(abi.RadianceError)]),
        "ixg_radiance_check_frame": (st, [C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32, C.c_uint32,
                                          C.c_uint32, P(abi.RadianceError)]),
        "ixg_asc_info_get": (st, [C.c_void_p, P(abi.AscInfo)]),
        "ixg_asc_tier_get": (st, [C.c_void_p, C.c_uint32, P(abi.AscTier), P(C.c_double), C.c_uint32]),
        "ixg_asc_text": (C.c_size_t, [C.c_void_p, C.c_int32, C.c_uint32, C.c_char_p, C.c_size_t]),
        "ixg_asc_index_name": (C.c_size_t, [C.c_void_p, C.c_uint32, C.c_uint32, P(C.c_uint32),
                                            C.c_char_p, C.c_size_t]),
        "ixg_asc_planes": (st, [C.c_void_p, P(abi.FacetPlane), C.c_uint32, P(C.c_uint32)]),
        "ixg_asc_schedule_planes": (st, [P(abi.AscTier), C.c_uint32, P(C.c_double), C.c_uint32,
                                         C.c_int32, C.c_double, P(abi.FacetPlane), C.c_uint32,
                                         P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "ixg_mlp_param_count": (C.c_uint32, []),
        "ixg_mlp_create": (st, [C.c_void_p, C.c_uint32, C.c_uint64, P(C.c_void_p)]),
        "ixg_mlp_destroy": (None, [C.c_void_p]),
        "ixg_mlp_set_params": (st, [C.c_void_p, P(C.c_float), C.c_size_t]),
        "ixg_mlp_get_params": (st, [C.c_void_p, P(C.c_float), C.c_size_t]),
        "ixg_mlp_get_state": (st, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_size_t, P(C.c_uint64)]),
        "ixg_mlp_set_state": (st, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_size_t, C.c_uint64]),
        "ixg_mlp_infer": (st, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
        "ixg_mlp_gradients": (st, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, P(C.c_float),
                                   C.c_size_t, P(C.c_float)]),
        "ixg_mlp_train_step": (st, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float,
                                    P(C.c_float)]),
        "ixg_mlp_sync": (st, [C.c_void_p]),
        "ixg_last_kernel_ms": (C.c_double, [C.c_void_p]),
        "ixg_last_kernel_launches": (C.c_uint32, [C.c_void_p]),
        "ixg_camera_new": (None, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32,
                                  C.c_uint32, C.c_uint32, P(abi.Camera)]),
        "ixg_cut_standard_round_brilliant": (C.c_uint32, [P(abi.FacetPlane)]),
        "ixg_cut_round_brilliant": (C.c_uint32, [C.c_float, C.c_float, C.c_float,
                                                 P(abi.FacetPlane)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """The loaded libixg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `make -C indicatrix_amd` "
                              "(or __graft_entry__.build()) first")
        # torch's wheel bundles its own HIP runtime under the same sonames as
        # /opt/rocm's.  Loading torch first lets libixg bind to that one copy,
        # so device pointers and streams are shared; loading libixg first would
        # put two HIP runtimes in the process and torch then finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


def header_symbols():
    """Function names declared in include/ixg.h."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^[A-Za-z_][A-Za-z0-9_ \*]*?\b(ixg_[a-z0-9_]+)\s*\(", text,
                                 re.M)))
