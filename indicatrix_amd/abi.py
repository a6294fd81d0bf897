"""ctypes mirror of include/ixg.h (the library's C ABI structs and enums).

Field order and types follow the header byte for byte; ``check_layout`` asserts
the sizes the header implies so a drift fails loudly at import.
"""
import ctypes as C

IXG_OK, IXG_DECLINED, IXG_CANCELLED, IXG_DEVICE_LOST, IXG_INVALID, IXG_RESOURCE_LIMIT = range(6)
STATUS_NAMES = {0: "OK", 1: "DECLINED", 2: "CANCELLED", 3: "DEVICE_LOST", 4: "INVALID",
                5: "RESOURCE_LIMIT"}

ROUTE_OK, ROUTE_UNSUPPORTED_MATERIAL, ROUTE_UNSUPPORTED_ENVIRONMENT, ROUTE_DEVICE_LIMITS, \
    ROUTE_NO_DEVICE = range(5)

MAX_PLANES = 256

# ixg_probe_fn (include/ixg.h): per-function probe ids, in order.
PROBE_FNS = ("sample_draws", "camera_ray", "cie_cmf", "illuminant", "environment", "frame_rotation",
             "fresnel_reflection", "fresnel_transmission", "tir_retardation", "channel_optics",
             "pleochroic_alpha", "ordinary_eigen", "extraordinary_poynting", "theta_c", "uniaxial_indices",
             "entry_solve", "internal_solve", "exit_transmission", "hg_direction", "cosine_hemisphere",
             "shading_normal", "finish", "intersect", "exp_lane", "exp", "biax_wave_indices", "biax_eigen",
             "biax_mode_poynting", "biax_resolve_entry", "biax_assigned_alpha")
PROBE = {name: i for i, name in enumerate(PROBE_FNS)}
NOALLOC_MAX_PATHS = 1 << 20
NOALLOC_MAX_HDR_TEXELS = 1024 * 512
MAX_BANDS = 8
NUM_CHANNELS = 8

DISP_SELLMEIER1, DISP_SELLMEIER3, DISP_CAUCHY = 0, 1, 2

# LightingPreset persisted ids (IX/optics/raytracer/environment/mod.rs:470-515).
PRESETS = {
    "LightTent": 0,
    "IsoHemisphere": 1,
    "WhiteTray": 2,
    "ShopLights": 3,
    "WindowDaylight": 4,
    "DaylightDome": 5,
    "DaylightSun": 6,
    "Daylight": 7,
    "Incandescent": 8,
    "IlluminantA": 9,
    "RingLights": 10,
    "DarkSpotlight": 11,
    "Aset": 12,
}


class FacetPlane(C.Structure):
    _fields_ = [("normal", C.c_float * 3), ("d", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("fov_tan", C.c_float),
                ("forward", C.c_float * 3), ("width", C.c_float),
                ("right", C.c_float * 3), ("height", C.c_float),
                ("up", C.c_float * 3), ("num_samples", C.c_uint32)]


class Dispersion(C.Structure):
    _fields_ = [("kind", C.c_int32), ("coeff", C.c_float * 6)]


class Band(C.Structure):
    _fields_ = [("center_nm", C.c_float), ("width", C.c_float), ("peak", C.c_float),
                ("shape", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("dispersion", Dispersion),
                ("birefringence_delta", C.c_float),
                ("c_axis", C.c_float * 3),
                ("cubic", C.c_int32),
                ("has_e_dispersion", C.c_int32),
                ("e_dispersion", Dispersion),
                ("n_bands_o", C.c_int32), ("n_bands_e", C.c_int32),
                ("bands_o", Band * MAX_BANDS), ("bands_e", Band * MAX_BANDS),
                ("absorption_path_scale", C.c_float),
                ("edge_rounding_radius", C.c_float),
                ("scattering_sigma_s", C.c_float),
                ("biaxial", C.c_int32), ("scattering_g", C.c_float),
                ("biaxial_delta_beta_alpha", C.c_float),
                ("has_beta_bands", C.c_int32), ("n_bands_beta", C.c_int32),
                ("bands_beta", Band * MAX_BANDS)]


class Environment(C.Structure):
    _fields_ = [("preset", C.c_int32), ("exposure", C.c_float), ("light_yaw", C.c_float),
                ("light_pitch", C.c_float), ("backdrop", C.c_float),
                ("surface_glare", C.c_float), ("head_shadow_deg", C.c_float)]


class Scene(C.Structure):
    _fields_ = [("camera", Camera),
                ("width", C.c_uint32), ("height", C.c_uint32),
                ("planes", C.POINTER(FacetPlane)), ("n_planes", C.c_uint32),
                ("material", C.POINTER(Material)),
                ("max_bounces", C.c_uint32),
                ("environment", Environment),
                ("facet_finishes", C.POINTER(C.c_uint8)),
                ("hdr_map", C.c_void_p)]


class HdrMap(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.POINTER(C.c_float))]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.POINTER(Scene)), ("first_sample", C.c_uint32),
                ("samples", C.c_uint32), ("out", C.POINTER(C.c_double))]


BatchDoneFn = C.CFUNCTYPE(None, C.c_void_p, C.c_uint32, C.c_int)


class Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("chunk_paths", C.c_uint32)]


class PathEvents(C.Structure):
    _fields_ = [("bounces", C.c_uint32), ("termination", C.c_uint32),
                ("facet_hash", C.c_uint32),
                ("n_tir", C.c_uint16), ("n_reflect", C.c_uint16),
                ("n_refract", C.c_uint16), ("n_probe", C.c_uint16)]


def check_layout():
    assert C.sizeof(FacetPlane) == 16
    assert C.sizeof(Camera) == 64
    assert Camera.fov_tan.offset == 12 and Camera.width.offset == 28
    assert Camera.height.offset == 44 and Camera.num_samples.offset == 60
    assert C.sizeof(PathEvents) == 20


check_layout()


class OpticalPixel(C.Structure):
    _fields_ = [("hits", C.c_uint32), ("classes", C.c_uint32 * 4), ("fire_sum_deg", C.c_float),
                ("fire_qualified", C.c_uint32), ("fire_rejected", C.c_uint32),
                ("centre_entry_facet", C.c_int32), ("centre_exit_facet", C.c_int32),
                ("centre_bounces", C.c_uint32), ("centre_reason", C.c_uint32),
                ("mixed_facets", C.c_uint32), ("scint_returned", C.c_uint8),
                ("scint_poses", C.c_uint8), ("pad", C.c_uint8 * 2)]


class MetricRequest(C.Structure):
    _fields_ = [("camera", Camera), ("width", C.c_uint32), ("height", C.c_uint32),
                ("samples", C.c_uint32), ("projection", C.c_uint32 * 2),
                ("has_tile_origin", C.c_int32), ("tile_origin", C.c_uint32 * 2),
                ("planes", C.POINTER(FacetPlane)), ("n_planes", C.c_uint32),
                ("material", C.POINTER(Material)), ("environment", Environment),
                ("want_fire", C.c_int32), ("want_scintillation", C.c_int32),
                ("has_pose", C.c_int32), ("pose", C.c_float * 3)]


class MetricTotals(C.Structure):
    _fields_ = [("total", OpticalPixel), ("brilliance_pct", C.c_float),
                ("windowing_pct", C.c_float), ("extinction_pct", C.c_float),
                ("fire_index", C.c_float), ("scintillation_pct", C.c_float)]


class Design(C.Structure):
    _fields_ = [("plane_offset", C.c_uint32), ("n_planes", C.c_uint32),
                ("dispersion", Dispersion)]


class AscError(C.Structure):
    _fields_ = [("kind", C.c_int32), ("line", C.c_uint32), ("value", C.c_double),
                ("teeth", C.c_int32), ("field_count", C.c_uint32), ("field", C.c_char * 32),
                ("token", C.c_char * 128), ("message", C.c_char * 512)]


class AscInfo(C.Structure):
    _fields_ = [("gear_teeth", C.c_int32), ("symmetry_order", C.c_uint32), ("mirror", C.c_int32),
                ("line_ending", C.c_int32), ("gear_reference_angle", C.c_double),
                ("refractive_index", C.c_double), ("n_tiers", C.c_uint32),
                ("n_headers", C.c_uint32), ("n_footnotes", C.c_uint32), ("n_warnings", C.c_uint32),
                ("gear_line", C.c_uint32), ("facet_plane_count", C.c_uint32)]


class AscTier(C.Structure):
    _fields_ = [("angle_deg", C.c_double), ("mast", C.c_double), ("n_indices", C.c_uint32),
                ("n_index_names", C.c_uint32)]


# ixg_asc_error_kind, named as the reference's AscParseError variants.
ASC_ERROR_KINDS = {
    0: "None", 1: "EmptyInput", 2: "GearLineMissingFields", 3: "GearTeethNotNumeric",
    4: "GearReferenceAngleNotNumeric", 5: "SymmetryLineMissingFields",
    6: "SymmetryOrderNotNumeric", 7: "RefractiveIndexMissing", 8: "RefractiveIndexNotNumeric",
    9: "NotWholeNumber", 10: "GearTeethZero", 11: "GearTeethTooLarge", 12: "MirrorFlagInvalid",
    13: "MissingGearLine", 14: "MissingSymmetryLine", 15: "MissingRefractiveIndexLine",
    16: "NoTierRecords", 17: "TierRecordTooShort", 18: "AngleNotNumeric", 19: "MastNotNumeric",
    20: "NonFiniteValue", 21: "SymmetryOrderZero", 22: "RefractiveIndexOutOfRange",
    23: "InvalidUtf8",
}
ASC_MAX_GEAR_TEETH = 720


assert C.sizeof(OpticalPixel) == 56
assert C.sizeof(AscError) == 696
assert C.sizeof(AscInfo) == 56
assert C.sizeof(AscTier) == 24


# ---- radiance payloads (indicatrix-net/src/radiance) ----
MAX_FRAME_LEN = 512 * 1024 * 1024
PAYLOAD_RAW, PAYLOAD_SHUFFLE_ZSTD, PAYLOAD_SHUFFLE_LZ4 = 0, 1, 2


def payload_zstd(level: int) -> int:
    """PayloadEncoding::ShuffleZstd { level } as the ABI's int32
    (IXG_PAYLOAD_ZSTD: the family in bits 0-7, the u8 level in bits 8-15)."""
    return PAYLOAD_SHUFFLE_ZSTD | ((level & 0xFF) << 8)

RADIANCE_ERROR_KINDS = {
    1: "LengthMismatch", 2: "Misaligned", 3: "TooLarge", 4: "RawLenMismatch",
    5: "UnsupportedEncoding", 6: "DecompressFailed", 7: "ShortOutput",
    8: "FrameOutsideRange", 9: "SampleCountOverrun", 10: "ZeroSampleFrame",
}


class RadianceError(C.Structure):
    _fields_ = [("kind", C.c_int32), ("encoding", C.c_int32), ("width", C.c_uint32),
                ("height", C.c_uint32), ("raw_len", C.c_uint32), ("first_sample", C.c_uint32),
                ("samples", C.c_uint32), ("start", C.c_uint32), ("end", C.c_uint32),
                ("cumulative", C.c_uint32), ("range_len", C.c_uint32), ("pad", C.c_uint32),
                ("expected_bytes", C.c_uint64), ("got_bytes", C.c_uint64),
                ("message", C.c_char * 256)]
