"""GemCAD ``.asc`` designs through libixg.

Mirrors the reference's reader (``indicatrix-formats/src/asc``: ``parse_asc``,
``parse_asc_with_gear_line``, ``parse_asc_bytes``, ``AscSchedule``,
``AscTier``, ``AscParseError``) and the plane builder
``StandardGemCuts::from_asc_schedule`` with its two bookkeeping views
(``indicatrix/src/geometry/cuts/asc_schedule.rs:20,29,60``).  Parsing and
plane construction run in the library's C++; this module only moves the
results into Python objects.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

from . import abi
from ._lib import lib

MAX_GEAR_TEETH = abi.ASC_MAX_GEAR_TEETH


class AscParseError(ValueError):
    """A rejected file: ``kind`` names the AscParseError variant (asc/error.rs),
    ``str(err)`` is its message; ``line``, ``field``, ``token``, ``value``,
    ``teeth`` and ``field_count`` are the variant's fields where it has them."""

    def __init__(self, e: abi.AscError):
        self.kind = abi.ASC_ERROR_KINDS.get(e.kind, str(e.kind))
        self.line = e.line
        self.field = e.field.decode("utf-8", "replace")
        self.token = e.token.decode("utf-8", "replace")
        self.value = e.value
        self.teeth = e.teeth
        self.field_count = e.field_count
        super().__init__(e.message.decode("utf-8", "replace"))


@dataclass
class AscTier:
    """One ``a`` record (asc/tier.rs).  A culet has ``angle_deg == -0.0``."""
    angle_deg: float = 0.0
    mast: float = 0.0
    name: str = ""
    indices: List[float] = field(default_factory=list)
    index_names: List[Tuple[int, str]] = field(default_factory=list)
    notes: str = ""

    def names(self) -> List[str]:
        return self.name.split("/") if self.name else []

    def is_culet(self) -> bool:
        import math
        return self.angle_deg == 0.0 and math.copysign(1.0, self.angle_deg) < 0


@dataclass
class AscSchedule:
    """AscSchedule (asc/schedule.rs); ``line_ending`` is ``"\\n"`` or ``"\\r\\n"``."""
    gemcad_version: str = ""
    gear_teeth: int = 0
    gear_reference_angle: float = 0.0
    symmetry_order: int = 0
    mirror: bool = False
    refractive_index: float = 0.0
    headers: List[str] = field(default_factory=list)
    footnotes: List[str] = field(default_factory=list)
    tiers: List[AscTier] = field(default_factory=list)
    warnings: List[str] = field(default_factory=list)
    line_ending: str = "\r\n"

    def gear_teeth_abs(self) -> int:
        return abs(self.gear_teeth)

    def facet_plane_count(self) -> int:
        return sum(max(len(t.indices), 1) for t in self.tiers)


def _text(handle, what, i=0):
    L = lib()
    n = L.ixg_asc_text(handle, what, i, None, 0)
    buf = C.create_string_buffer(n + 1)
    L.ixg_asc_text(handle, what, i, buf, n + 1)
    return buf.raw[:n].decode("utf-8")


def _read(handle) -> Tuple[AscSchedule, Optional[int]]:
    L = lib()
    info = abi.AscInfo()
    L.ixg_asc_info_get(handle, C.byref(info))
    s = AscSchedule(
        gemcad_version=_text(handle, 0),
        gear_teeth=info.gear_teeth,
        gear_reference_angle=info.gear_reference_angle,
        symmetry_order=info.symmetry_order,
        mirror=bool(info.mirror),
        refractive_index=info.refractive_index,
        headers=[_text(handle, 1, i) for i in range(info.n_headers)],
        footnotes=[_text(handle, 2, i) for i in range(info.n_footnotes)],
        warnings=[_text(handle, 3, i) for i in range(info.n_warnings)],
        line_ending="\r\n" if info.line_ending == 1 else "\n",
    )
    for t in range(info.n_tiers):
        rec = abi.AscTier()
        L.ixg_asc_tier_get(handle, t, C.byref(rec), None, 0)
        idx = (C.c_double * max(1, rec.n_indices))()
        L.ixg_asc_tier_get(handle, t, C.byref(rec), idx, rec.n_indices)
        names = []
        for k in range(rec.n_index_names):
            pos = C.c_uint32()
            n = L.ixg_asc_index_name(handle, t, k, C.byref(pos), None, 0)
            buf = C.create_string_buffer(n + 1)
            L.ixg_asc_index_name(handle, t, k, C.byref(pos), buf, n + 1)
            names.append((pos.value, buf.raw[:n].decode("utf-8")))
        s.tiers.append(AscTier(angle_deg=rec.angle_deg, mast=rec.mast, name=_text(handle, 4, t),
                               indices=list(idx[:rec.n_indices]), index_names=names,
                               notes=_text(handle, 5, t)))
    return s, (info.gear_line or None)


def _parse(entry, data: bytes):
    handle = C.c_void_p()
    err = abi.AscError()
    st = entry(data, len(data), C.byref(handle), C.byref(err))
    if st != 0:
        if err.kind:
            raise AscParseError(err)
        raise ValueError(f"ixg_asc_parse failed with status {st}")
    try:
        return _read(handle)
    finally:
        lib().ixg_asc_free(handle)


def parse_asc_with_gear_line(text: str) -> Tuple[AscSchedule, Optional[int]]:
    """parse_asc_with_gear_line (asc/parse.rs:68): the schedule and the
    1-based line of the gear record the reader kept."""
    return _parse(lib().ixg_asc_parse_text, text.encode("utf-8"))


def parse_asc(text: str) -> AscSchedule:
    """parse_asc (asc/parse.rs:54) of UTF-8 text."""
    return parse_asc_with_gear_line(text)[0]


def parse_asc_bytes(data: bytes) -> AscSchedule:
    """parse_asc_bytes (asc/decode.rs): raw file bytes, UTF-8 or Windows-1252."""
    return _parse(lib().ixg_asc_parse_bytes, bytes(data))[0]


def _schedule_planes(schedule: AscSchedule, gear_reference_angle=None):
    tiers = (abi.AscTier * max(1, len(schedule.tiers)))()
    flat = []
    for t, tier in enumerate(schedule.tiers):
        tiers[t].angle_deg = tier.angle_deg
        tiers[t].mast = tier.mast
        tiers[t].n_indices = len(tier.indices)
        flat.extend(tier.indices)
    idx = (C.c_double * max(1, len(flat)))(*flat)
    ref = schedule.gear_reference_angle if gear_reference_angle is None else gear_reference_angle
    n = C.c_uint32()
    raw = (C.c_uint32 * max(1, schedule.facet_plane_count()))()
    ends = (C.c_uint32 * max(1, len(schedule.tiers)))()
    L = lib()
    args = (tiers, len(schedule.tiers), idx, len(flat), schedule.gear_teeth, ref)
    st = L.ixg_asc_schedule_planes(*args, None, 0, C.byref(n), raw, ends)
    if st != 0:
        raise ValueError(f"ixg_asc_schedule_planes failed with status {st}")
    out = (abi.FacetPlane * max(1, n.value))()
    st = L.ixg_asc_schedule_planes(*args, out, n.value, C.byref(n), None, None)
    if st != 0:
        raise ValueError(f"ixg_asc_schedule_planes failed with status {st}")
    planes = (abi.FacetPlane * n.value)(*out[:n.value])
    return planes, list(raw[:schedule.facet_plane_count()]), list(ends[:len(schedule.tiers)])


def from_asc_schedule(schedule: AscSchedule):
    """StandardGemCuts::from_asc_schedule (asc_schedule.rs:20): the
    deduplicated facet planes, a ctypes FacetPlane array like the cut builders'."""
    return _schedule_planes(schedule)[0]


def asc_plane_provenance(schedule: AscSchedule) -> List[List[int]]:
    """asc_plane_provenance (asc_schedule.rs:29): for each tier, the output
    plane each of its raw planes became."""
    _, raw, ends = _schedule_planes(schedule)
    out, start, at = [], 0, 0
    for t in schedule.tiers:
        n = max(len(t.indices), 1)
        out.append(raw[at:at + n])
        at += n
    return out


def asc_tier_plane_boundaries(schedule: AscSchedule, gear_reference_angle: float) -> List[int]:
    """asc_tier_plane_boundaries (asc_schedule.rs:60): output planes up to the
    end of each tier, at the given gear reference angle."""
    return _schedule_planes(schedule, gear_reference_angle)[2]
