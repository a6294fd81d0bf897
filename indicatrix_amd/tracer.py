"""Host-side mirror of the reference's film API over libixg's C ABI.

Names follow the reference: ``GpuBackend.try_accumulate`` /
``try_accumulate_cancellable`` (IX/renderer/gpu_backend/backend.rs:182-300),
``GpuAccumulate`` {Done, Declined, Cancelled} (IX/renderer/gpu_backend/mod.rs:307-326),
``FrameScene`` (IX/renderer/frame_scene.rs:22-40), ``Camera.new``
(IX/optics/raytracer/camera.rs:108), ``GemMaterial`` (IX/optics/materials/mod.rs:116)
and ``LightingPreset.studio`` (IX/optics/raytracer/environment/mod.rs:620).

Arrays are numpy; device work goes through libixg only -- there is no CPU
fallback in this module (a declined scene is reported as Declined, exactly
as the reference's GPU backend reports it to its caller).
"""
import ctypes as C
import enum

import numpy as np

from . import abi
from ._lib import IxgError, lib


class GpuAccumulate(enum.Enum):
    DONE = "Done"
    DECLINED = "Declined"
    CANCELLED = "Cancelled"


# --------------------------------------------------------------------------
# materials
# --------------------------------------------------------------------------
class GemMaterial:
    """The GemMaterial fields the transport reads."""

    def __init__(self, name, dispersion_kind, coeff, birefringence_delta=0.0, cubic=True,
                 c_axis=(0.0, 1.0, 0.0), absorption_path_scale=1.0):
        self.name = name
        self.dispersion_kind = dispersion_kind
        self.coeff = list(coeff) + [0.0] * (6 - len(coeff))
        self.birefringence_delta = birefringence_delta
        self.cubic = cubic
        self.c_axis = tuple(c_axis)
        self.absorption_path_scale = absorption_path_scale
        self.bands_o = []
        self.bands_e = []
        # GemMaterial::biaxial_delta_beta_alpha (None = not biaxial) and
        # AbsorptionTensor::beta_ray (None = no trichroic beta band set)
        self.biaxial_delta_beta_alpha = None
        self.bands_beta = None

    def to_abi(self):
        m = abi.Material()
        m.dispersion.kind = self.dispersion_kind
        for i, v in enumerate(self.coeff):
            m.dispersion.coeff[i] = v
        m.birefringence_delta = self.birefringence_delta
        for i in range(3):
            m.c_axis[i] = self.c_axis[i]
        m.cubic = 1 if self.cubic else 0
        m.n_bands_o = len(self.bands_o)
        m.n_bands_e = len(self.bands_e)
        for i, (c, w, p, s) in enumerate(self.bands_o):
            m.bands_o[i] = abi.Band(c, w, p, s)
        for i, (c, w, p, s) in enumerate(self.bands_e):
            m.bands_e[i] = abi.Band(c, w, p, s)
        m.absorption_path_scale = self.absorption_path_scale
        if self.biaxial_delta_beta_alpha is not None:
            m.biaxial = 1
            m.biaxial_delta_beta_alpha = self.biaxial_delta_beta_alpha
        if self.bands_beta is not None:
            m.has_beta_bands = 1
            m.n_bands_beta = len(self.bands_beta)
            for i, (c, w, p, s) in enumerate(self.bands_beta):
                m.bands_beta[i] = abi.Band(c, w, p, s)
        return m

    @staticmethod
    def diamond():
        """Diamond (IX/optics/materials/diamond_through_emerald.rs:24-57): Peter (1923)
        Sellmeier in the reference's term order."""
        return GemMaterial("Diamond", abi.DISP_SELLMEIER3,
                           [4.3356, 0.3306, 0.0, 0.011236, 0.030625, 1.0])

    @staticmethod
    def sapphire():
        """Sapphire (IX/optics/materials/diamond_through_emerald.rs:58-134):
        Malitson-o Sellmeier, n_e - n_o = -0.0081, optic axis +Y, and the
        Fe-Ti pleochroic band set -- three energy-domain bands per ray, each
        given as (centre nm, FWHM cm^-1, peak) and stored with sigma =
        FWHM / 2.35482 (AbsorptionBand::energy, IX/optics/absorption.rs:95)."""
        m = GemMaterial("Sapphire", abi.DISP_SELLMEIER3,
                        [1.431349, 0.650547, 5.341402, 0.00528, 0.01424, 325.015],
                        birefringence_delta=-0.0081, cubic=False)

        def energy(centre, fwhm, peak):
            return (centre, float(np.float32(fwhm) / np.float32(2.35482)), peak, 1)

        m.bands_o = [energy(585.0, 4878.0, 2.01), energy(744.0, 3142.0, 0.887),
                     energy(457.0, 12000.0, 0.433)]
        m.bands_e = [energy(612.0, 7466.0, 0.371), energy(729.0, 4955.0, 1.03),
                     energy(467.0, 11169.0, 0.524)]
        return m

    @staticmethod
    def alexandrite():
        """Alexandrite (IX/optics/materials/zircon_through_topaz.rs:105-137):
        orthorhombic, biaxial positive; Sellmeier-3 beta curve, n_gamma - n_alpha
        = 0.0076, n_beta - n_alpha = 0.001951, gamma along +Y, and the trichroic
        band sets alpha / beta / gamma (AbsorptionTensor::biaxial)."""
        m = GemMaterial("Alexandrite", abi.DISP_SELLMEIER3, [0.78522, 1.21202, 16.81, 0.0, 0.01262, 1000.0],
                        birefringence_delta=0.0076, cubic=False)
        m.bands_o = [(565.0, 40.0, 1.9, 0), (420.0, 32.0, 1.8, 0)]
        m.bands_beta = [(560.0, 40.0, 0.65, 0), (422.0, 32.0, 3.5, 0)]
        m.bands_e = [(595.0, 40.0, 3.1, 0), (430.0, 32.0, 1.5, 0)]
        m.biaxial_delta_beta_alpha = 0.001951
        return m

    @staticmethod
    def topaz():
        """Topaz (IX/optics/materials/zircon_through_topaz.rs:211-233): Cauchy
        beta curve, n_gamma - n_alpha = 0.008, n_beta - n_alpha = 0.0024,
        gamma along +Y, one isotropic band (no beta_ray)."""
        m = GemMaterial("Topaz", abi.DISP_CAUCHY, [1.614872, 0.004273, 0.0],
                        birefringence_delta=0.0080, cubic=False)
        m.bands_o = [(620.0, 70.0, 1.6, 0)]
        m.bands_e = [(620.0, 70.0, 1.6, 0)]
        m.biaxial_delta_beta_alpha = 0.0024
        return m

    @staticmethod
    def isotropic_from_nd_abbe(n_d, abbe, name="custom"):
        """A one-term Sellmeier curve through n(589.3 nm) = n_d with Abbe number
        V_d = (n_d - 1) / (n_F - n_C), solved in double precision."""
        b1, c1 = sellmeier1_from_nd_abbe(n_d, abbe)
        return GemMaterial(name, abi.DISP_SELLMEIER1, [b1, c1])


def sellmeier1_from_nd_abbe(n_d, abbe):
    """Solves n^2 = 1 + b l^2 / (l^2 - c) (l in um) for (b, c) so that n(D) = n_d
    and (n_d - 1)/(n(F) - n(C)) = abbe."""
    lD, lF, lC = 0.5893, 0.4861, 0.6563

    def n_of(b, c, l):
        return (1.0 + b * l * l / (l * l - c)) ** 0.5

    def b_for(c):
        return (n_d * n_d - 1.0) * (lD * lD - c) / (lD * lD)

    target = (n_d - 1.0) / abbe
    lo, hi = 1e-6, 0.2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        b = b_for(mid)
        disp = n_of(b, mid, lF) - n_of(b, mid, lC)
        if disp < target:
            lo = mid
        else:
            hi = mid
    c = 0.5 * (lo + hi)
    return b_for(c), c


# --------------------------------------------------------------------------
# lighting
# --------------------------------------------------------------------------
class LightingPreset:
    NAMES = dict(abi.PRESETS)

    @staticmethod
    def studio(preset, exposure=1.0, light_yaw=0.0, light_pitch=0.0, head_shadow_deg=16.0,
               surface_glare=1.0, backdrop=0.0):
        """LightingPreset::studio (environment/mod.rs:619-634); ``backdrop`` is
        EnvironmentSource::with_backdrop's lit tray radiance (source.rs:53-80)."""
        pid = abi.PRESETS[preset] if isinstance(preset, str) else int(preset)
        return abi.Environment(pid, exposure, light_yaw, light_pitch, backdrop, surface_glare,
                               head_shadow_deg)


# --------------------------------------------------------------------------
# camera and cuts (host arithmetic in libixg, FP32 + glibc libm)
# --------------------------------------------------------------------------
def camera_new(yaw, pitch, distance, fov_deg, width, height, spp=1):
    cam = abi.Camera()
    lib().ixg_camera_new(yaw, pitch, distance, fov_deg, width, height, spp, C.byref(cam))
    return cam


def planes_array(planes):
    """numpy (N,4) float32 [nx, ny, nz, d] from an abi.FacetPlane array."""
    return np.array([[p.normal[0], p.normal[1], p.normal[2], p.d] for p in planes], np.float32)


def planes_from_array(arr):
    arr = np.asarray(arr, np.float32).reshape(-1, 4)
    out = (abi.FacetPlane * len(arr))()
    for i, r in enumerate(arr):
        out[i] = abi.FacetPlane((C.c_float * 3)(*r[:3]), float(r[3]))
    return out


def standard_round_brilliant():
    buf = (abi.FacetPlane * 80)()
    n = lib().ixg_cut_standard_round_brilliant(buf)
    return (abi.FacetPlane * n)(*buf[:n])


def round_brilliant(crown_deg, pavilion_deg, table_ratio):
    buf = (abi.FacetPlane * 80)()
    n = lib().ixg_cut_round_brilliant(crown_deg, pavilion_deg, table_ratio, buf)
    return (abi.FacetPlane * n)(*buf[:n])


# --------------------------------------------------------------------------
# dispatch lanes (indicatrix-dispatch/src/lane.rs)
# --------------------------------------------------------------------------
class ChunkResult:
    """ChunkResult (lane.rs:67-111): the summed XYZ of the traced prefix
    [first_sample, first_sample + done) as float32 (pixels, 3), the lane's
    steady-state samples/s (None: let the pool time it) and why the chunk
    ended short, if it did."""

    def __init__(self, sum, done, rate=None, error=None):
        self.sum = sum
        self.done = done
        self.rate = rate
        self.error = error

    @property
    def complete(self):
        return self.error is None


class GpuLane:
    """WorkerLane (lane.rs:113-142) backed by one libixg context: each chunk
    is ixg_render_chunk, the adaptive sub-batch loop of the coordinator's own
    lane (OwnLane, apps/indicatrix-worker/src/coordinator/lanes/mod.rs:536-570).
    The converged sub-batch size is remembered per scene and seeds the next
    chunk of that scene (stream_emit/sizing.rs remember_converged_subbatch)."""

    def __init__(self, backend, name="local GPU"):
        self.backend = backend
        self._name = name
        self._subbatch = {}

    def name(self):
        return self._name

    def lost(self):
        return self.backend.is_lost()

    def render_chunk(self, scene, first_sample, samples, cancel=None):
        n = scene.width * scene.height
        out = np.zeros((n, 3), np.float32)
        done, rate, nxt = C.c_uint32(), C.c_double(), C.c_uint32()
        cflag = None if cancel is None else C.cast(cancel, C.POINTER(C.c_uint8))
        st = lib().ixg_render_chunk(self.backend.handle, scene.ref(), first_sample, samples,
                                    self._subbatch.get(id(scene), 0), out.ctypes.data_as(C.POINTER(C.c_float)),
                                    C.byref(done), C.byref(rate), C.byref(nxt), cflag)
        if done.value:
            self._subbatch[id(scene)] = nxt.value
        r = rate.value if rate.value > 0.0 else None
        if st == abi.IXG_OK:
            return ChunkResult(out, done.value, r)
        if st == abi.IXG_CANCELLED:
            return ChunkResult(out, done.value, None, "cancelled")
        if st == abi.IXG_DECLINED and done.value == 0:
            return ChunkResult(np.zeros((0, 3), np.float32), 0, None,
                               "declined by the device route: " + self.backend.last_error())
        return ChunkResult(out, done.value, None, self.backend.last_error() or f"ixg status {st}")


def next_subbatch(prev, elapsed_s):
    """stream_emit/sizing.rs next_batch_size, as ixg_next_subbatch."""
    return lib().ixg_next_subbatch(prev, elapsed_s)


# --------------------------------------------------------------------------
# scenes and the backend
# --------------------------------------------------------------------------
class FrameScene:
    """Owns the ctypes storage behind one ixg_scene."""

    def __init__(self, camera, width, height, planes, material, max_bounces, environment,
                 facet_finishes=None):
        """facet_finishes: per plane 0 (polished) / 1 (frosted), or None."""
        self.planes = planes if isinstance(planes, C.Array) else planes_from_array(planes)
        self.material_abi = material.to_abi() if isinstance(material, GemMaterial) else material
        self.scene = abi.Scene()
        self.scene.camera = camera
        self.scene.width = width
        self.scene.height = height
        self.scene.planes = C.cast(self.planes, C.POINTER(abi.FacetPlane))
        self.scene.n_planes = len(self.planes)
        self.scene.material = C.pointer(self.material_abi)
        self.scene.max_bounces = max_bounces
        self.scene.environment = environment
        self.set_finishes(facet_finishes)

    def set_hdr_map(self, rgb):
        """Sets the environment to an equirectangular panorama ((h, w, 3)
        linear RGB), or back to the preset with None."""
        if rgb is None:
            self.hdr = None
            self.scene.hdr_map = None
            return
        import numpy as np
        arr = np.ascontiguousarray(rgb, np.float32)
        if arr.ndim != 3 or arr.shape[2] != 3:
            raise ValueError("rgb must be (height, width, 3)")
        self._hdr_pixels = arr
        self.hdr = abi.HdrMap(arr.shape[1], arr.shape[0], arr.ctypes.data_as(C.POINTER(C.c_float)))
        self.scene.hdr_map = C.cast(C.pointer(self.hdr), C.c_void_p)

    def set_finishes(self, facet_finishes):
        """Sets the FacetFinish of every plane (None: all polished)."""
        if facet_finishes is None:
            self.finishes = None
            self.scene.facet_finishes = None
        else:
            vals = [1 if f else 0 for f in facet_finishes]
            if len(vals) != len(self.planes):
                raise ValueError("one finish per plane")
            self.finishes = (C.c_uint8 * len(vals))(*vals)
            self.scene.facet_finishes = C.cast(self.finishes, C.POINTER(C.c_uint8))

    @property
    def width(self):
        return self.scene.width

    @property
    def height(self):
        return self.scene.height

    def ref(self):
        return C.byref(self.scene)


class GpuBackend:
    """A libixg context (GpuBackend in the reference)."""

    def __init__(self, device=0, chunk_paths=0):
        self._ctx = C.c_void_p()
        opts = abi.Opts(device, chunk_paths)
        st = lib().ixg_create(C.byref(opts), C.byref(self._ctx))
        if st != abi.IXG_OK:
            raise IxgError(st, "ixg_create failed (no HIP device?)")

    def close(self):
        if self._ctx:
            lib().ixg_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._ctx

    def last_error(self):
        buf = C.create_string_buffer(512)
        lib().ixg_last_error(self._ctx, buf, 512)
        return buf.value.decode()

    def adapter_label(self):
        buf = C.create_string_buffer(256)
        lib().ixg_adapter_label(self._ctx, buf, 256)
        return buf.value.decode()

    def is_lost(self):
        return bool(lib().ixg_is_lost(self._ctx))

    def try_recover(self):
        return lib().ixg_try_recover(self._ctx) == abi.IXG_OK

    def inject_fault(self, fail_on_chunk):
        lib().ixg_inject_fault(self._ctx, fail_on_chunk)

    def _check(self, st):
        if st not in (abi.IXG_OK, abi.IXG_DECLINED, abi.IXG_CANCELLED):
            raise IxgError(st, self.last_error())
        return st

    def try_accumulate_cancellable(self, scene, sample_offset, spp, accum, cancel=None):
        """ADDS per-pixel FP64 XYZ sums into ``accum`` (float64, (h*w, 3) or
        (h, w, 3), C-contiguous) only when the result is Done."""
        if accum.dtype != np.float64 or not accum.flags.c_contiguous:
            raise ValueError("accum must be a C-contiguous float64 array")
        if accum.size != scene.width * scene.height * 3:
            raise ValueError("accum length must be width*height*3")
        cflag = None if cancel is None else C.cast(cancel, C.POINTER(C.c_uint8))
        st = lib().ixg_accumulate(self._ctx, scene.ref(), sample_offset, spp,
                                  accum.ctypes.data_as(C.POINTER(C.c_double)), cflag)
        st = self._check(st)
        return {abi.IXG_OK: GpuAccumulate.DONE, abi.IXG_DECLINED: GpuAccumulate.DECLINED,
                abi.IXG_CANCELLED: GpuAccumulate.CANCELLED}[st]

    def try_accumulate_batch_cancellable(self, items, cancel=None, on_done=None):
        """try_accumulate_batch_cancellable (renderer/gpu_backend/backend.rs:451).

        ``items``: (scene, first_sample, samples, out) with ``out`` a C-contiguous
        float64 array of width*height*3 that receives exactly what
        try_accumulate_cancellable would ADD for that scene, only when its
        outcome is Done.  ``on_done(index, GpuAccumulate)`` runs once per item,
        in input order.  Returns the outcomes, in item order."""
        arr = (abi.BatchItem * max(1, len(items)))()
        for i, (scene, first, samples, out) in enumerate(items):
            if out.dtype != np.float64 or not out.flags.c_contiguous:
                raise ValueError("out must be a C-contiguous float64 array")
            if out.size != scene.width * scene.height * 3:
                raise ValueError("batch output buffer must have one entry per pixel")
            arr[i] = abi.BatchItem(C.pointer(scene.scene), first, samples,
                                   out.ctypes.data_as(C.POINTER(C.c_double)))
        outcomes = [None] * len(items)
        names = {abi.IXG_OK: GpuAccumulate.DONE, abi.IXG_DECLINED: GpuAccumulate.DECLINED,
                 abi.IXG_CANCELLED: GpuAccumulate.CANCELLED}

        def done(_user, index, status):
            if outcomes[index] is not None:
                raise RuntimeError(f"picture {index} reported more than once")
            outcomes[index] = names.get(status, GpuAccumulate.DECLINED)
            if on_done is not None:
                on_done(index, outcomes[index])

        cb = abi.BatchDoneFn(done)
        cflag = None if cancel is None else C.cast(cancel, C.POINTER(C.c_uint8))
        st = lib().ixg_accumulate_batch(self._ctx, arr, len(items), cflag, cb, None)
        if st != abi.IXG_OK:
            raise IxgError(st, self.last_error())
        return outcomes

    def try_accumulate(self, scene, sample_offset, spp, accum):
        return self.try_accumulate_cancellable(scene, sample_offset, spp, accum) \
            is GpuAccumulate.DONE

    def accumulate_device(self, scene, sample_offset, spp, d_accum_ptr):
        st = lib().ixg_accumulate_device(self._ctx, scene.ref(), sample_offset, spp,
                                         C.c_void_p(d_accum_ptr))
        if st != abi.IXG_OK:
            raise IxgError(st, self.last_error())

    def accumulate_device_noalloc(self, scene, sample_offset, spp, d_accum_ptr):
        """ixg_accumulate_device_noalloc: the same sums on the library's fixed
        device scratch, with no allocation (spp <= abi.NOALLOC_MAX_PATHS)."""
        st = lib().ixg_accumulate_device_noalloc(self._ctx, scene.ref(), sample_offset, spp,
                                                 C.c_void_p(d_accum_ptr))
        if st != abi.IXG_OK:
            raise IxgError(st, self.last_error())

    def last_kernel_ms(self):
        return lib().ixg_last_kernel_ms(self._ctx)

    def stream(self):
        """The hipStream_t (as an int) the context launches its work on."""
        return int(lib().ixg_stream(self._ctx) or 0)

    def trace_paths(self, scene, sample_offset, spp, want_xyz=True, want_events=True):
        n = scene.width * scene.height * spp
        xyz = np.zeros((n, 3), np.float32) if want_xyz else None
        ev = (abi.PathEvents * n)() if want_events else None
        st = lib().ixg_trace_paths(
            self._ctx, scene.ref(), sample_offset, spp,
            xyz.ctypes.data_as(C.POINTER(C.c_float)) if want_xyz else None,
            C.cast(ev, C.POINTER(abi.PathEvents)) if want_events else None)
        if st != abi.IXG_OK:
            raise IxgError(st, self.last_error())
        return xyz, (events_to_numpy(ev) if want_events else None)

    def film_samples(self, scene, sample_offset, spp):
        """The film kernel's own per-path XYZ (the launch ixg_accumulate
        times), shape (pixels * spp, 3) in (pixel, sample) order."""
        n = scene.width * scene.height * spp
        xyz = np.zeros((n, 3), np.float32)
        st = lib().ixg_film_samples(self._ctx, scene.ref(), sample_offset, spp,
                                    xyz.ctypes.data_as(C.POINTER(C.c_float)))
        if st != abi.IXG_OK:
            raise IxgError(st, self.last_error())
        return xyz

    def trace_rays(self, scene, origins, dirs, seeds, heros):
        """trace_spectral_ray for explicit rays -> ((n,3) float32 XYZ, events)."""
        origins = np.asarray(origins, np.float32).reshape(-1, 3)
        dirs = np.asarray(dirs, np.float32).reshape(-1, 3)
        n = len(origins)
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3] = origins
        rays[:, 3:6] = dirs
        rays[:, 6] = np.asarray(heros, np.float32)
        rays[:, 7] = np.asarray(seeds, np.uint32).view(np.float32)
        xyz = np.zeros((n, 3), np.float32)
        ev = (abi.PathEvents * n)()
        st = lib().ixg_trace_rays(self._ctx, scene.ref(), rays.ctypes.data_as(C.POINTER(C.c_float)),
                                  n, xyz.ctypes.data_as(C.POINTER(C.c_float)),
                                  C.cast(ev, C.POINTER(abi.PathEvents)))
        if st != abi.IXG_OK:
            raise IxgError(st, self.last_error())
        return xyz, events_to_numpy(ev)

    def probe(self, scene, fn, records):
        """ixg_probe: the device function `fn` (a name of abi.PROBE_FNS or its
        id) on float32 input records, shape (n, in_floats) -> (n, out_floats)."""
        fid = abi.PROBE[fn] if isinstance(fn, str) else int(fn)
        w_in, w_out = C.c_uint32(), C.c_uint32()
        if not lib().ixg_probe_shape(fid, C.byref(w_in), C.byref(w_out)):
            raise ValueError(f"unknown probe function {fn!r}")
        rec = np.ascontiguousarray(records, np.float32).reshape(-1, w_in.value)
        out = np.zeros((len(rec), w_out.value), np.float32)
        st = lib().ixg_probe(self._ctx, scene.ref(), fid, rec.ctypes.data_as(C.POINTER(C.c_float)), len(rec),
                             out.ctypes.data_as(C.POINTER(C.c_float)))
        if st != abi.IXG_OK:
            raise IxgError(st, self.last_error())
        return out

    def count_events(self, scene, sample_offset, spp):
        tot = (C.c_uint64 * len(EVENT_SLOTS))()
        st = lib().ixg_count_events(self._ctx, scene.ref(), sample_offset, spp, tot)
        if st != abi.IXG_OK:
            raise IxgError(st, self.last_error())
        return dict(zip(EVENT_SLOTS, list(tot)))


EVENT_SLOTS = ["paths", "bounces", "tir", "reflect", "refract", "probe", "escaped", "rotations",
               "scans", "probe_scans"]


def route(scene):
    return lib().ixg_route(scene.ref())


def events_to_numpy(ev):
    dt = np.dtype([("bounces", "<u4"), ("termination", "<u4"), ("facet_hash", "<u4"),
                   ("n_tir", "<u2"), ("n_reflect", "<u2"), ("n_refract", "<u2"),
                   ("n_probe", "<u2")])
    return np.frombuffer(bytes(ev), dtype=dt).copy()
