"""Radiance payloads: the body of a FRAME / PREVIEW message, on the device.

Mirrors ``indicatrix-net/src/radiance`` (``shuffle_into``, ``unshuffle_into``,
``add_unshuffled_valid``, ``PayloadDecoder`` and ``RadianceError``) and the
admission checks of ``Accumulator::apply_frame``
(``client/accumulate/mod.rs:414-465``).  A payload is ``width * height``
pixels of three little-endian f32; the compressed encodings start from its
byte-shuffled planes.  The byte work runs in libixg's HIP kernels on device
buffers (torch tensors here); the header checks run in its C++.
"""
from __future__ import annotations

import ctypes as C

from . import abi
from ._lib import IxgError, lib

BYTES_PER_PIXEL = 12
PLANES = 4


class RadianceError(ValueError):
    """A rejected payload or frame: ``kind`` names the RadianceError variant
    (radiance/mod.rs:66-119), ``str(err)`` is its Display text, ``fields``
    the variant's fields."""

    def __init__(self, e: abi.RadianceError):
        self.kind = abi.RADIANCE_ERROR_KINDS.get(e.kind, str(e.kind))
        self.fields = {f: getattr(e, f) for f, _ in abi.RadianceError._fields_
                       if f not in ("kind", "pad", "message")}
        super().__init__(e.message.decode("utf-8", "replace"))


def _ptr(t):
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("radiance kernels take contiguous device tensors")
    return C.c_void_p(t.data_ptr())


def _nbytes(t):
    return t.numel() * t.element_size()


def shuffle(backend, src, dst):
    """shuffle_into (shuffle.rs:35) of ``src``'s bytes into ``dst``."""
    if _nbytes(src) != _nbytes(dst):
        raise ValueError("shuffle output must match the input length")
    st = lib().ixg_radiance_shuffle_device(backend.handle, _ptr(src), _nbytes(src), _ptr(dst))
    if st != abi.IXG_OK:
        raise IxgError(st, "ixg_radiance_shuffle_device failed")


def unshuffle(backend, src, dst):
    """unshuffle_into (shuffle.rs:60)."""
    if _nbytes(src) != _nbytes(dst):
        raise ValueError("unshuffle output must match the input length")
    st = lib().ixg_radiance_unshuffle_device(backend.handle, _ptr(src), _nbytes(src), _ptr(dst))
    if st != abi.IXG_OK:
        raise IxgError(st, "ixg_radiance_unshuffle_device failed")


def add_valid(backend, payload, acc, shuffled: bool, dropped) -> None:
    """The summing arms of PayloadDecoder::decode_and_add (payload.rs:225-250)
    onto ``acc`` (f32, pixels * 3); ``dropped`` (one int32 device element)
    receives the count of pixels that failed is_valid_sample."""
    pixels = acc.numel() // 3
    if acc.numel() != pixels * 3 or _nbytes(payload) != pixels * BYTES_PER_PIXEL:
        raise ValueError("planes must hold exactly three f32 per accumulator pixel")
    st = lib().ixg_radiance_add_device(backend.handle, _ptr(payload), int(bool(shuffled)), pixels,
                                       _ptr(acc), _ptr(dropped))
    if st != abi.IXG_OK:
        raise IxgError(st, "ixg_radiance_add_device failed")


def check_payload(encoding: int, raw_len: int, wire_len: int, width: int, height: int,
                  compression: bool = False) -> None:
    """PayloadDecoder::prepare's header checks (payload.rs:282-320)."""
    err = abi.RadianceError()
    st = lib().ixg_radiance_check_payload(encoding, raw_len, wire_len, width, height,
                                          int(compression), C.byref(err))
    if st != abi.IXG_OK:
        if err.kind:
            raise RadianceError(err)
        raise ValueError(f"unknown payload encoding {encoding}")


def frame_within_range(first_sample: int, samples: int, start: int, end: int) -> bool:
    """frame_within_range (client/accumulate/mod.rs:198): inclusive start,
    exclusive end, an end past u32::MAX never within."""
    frame_end = first_sample + samples
    return frame_end <= 0xFFFFFFFF and first_sample >= start and frame_end <= end


def check_frame(first_sample: int, samples: int, expected_range=None, samples_done: int = 0) -> None:
    """Accumulator::apply_frame's admission checks (accumulate/mod.rs:414-465)."""
    err = abi.RadianceError()
    has = expected_range is not None
    start, end = expected_range if has else (0, 0)
    st = lib().ixg_radiance_check_frame(first_sample, samples, int(has), start, end, samples_done,
                                        C.byref(err))
    if st != abi.IXG_OK:
        raise RadianceError(err)
