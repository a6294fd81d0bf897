"""The ABI's struct layouts, held to the compiler's own.

A C program compiled against include/ixg.h (tests/native/abi_layout.c)
prints every struct's size and field offsets.  The ctypes mirror
(indicatrix_amd/abi.py) and the Rust binding a maintainer would paste from
INTEGRATION.md (its `#[repr(C)]` structs, laid out here by Rust's repr(C)
rules) must agree field by field, so a header change that forgets either
mirror fails here.  CPU only: nothing is loaded but the header.
"""
import ctypes as C
import json
import os
import re
import subprocess

import pytest

from indicatrix_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# C struct -> ctypes mirror
MIRRORS = {
    "ixg_facet_plane": abi.FacetPlane, "ixg_camera": abi.Camera, "ixg_dispersion": abi.Dispersion,
    "ixg_band": abi.Band, "ixg_material": abi.Material, "ixg_environment": abi.Environment,
    "ixg_hdr_map": abi.HdrMap, "ixg_scene": abi.Scene, "ixg_opts": abi.Opts,
    "ixg_path_events": abi.PathEvents, "ixg_batch_item": abi.BatchItem,
    "ixg_optical_pixel": abi.OpticalPixel, "ixg_metric_request": abi.MetricRequest,
    "ixg_metric_totals": abi.MetricTotals, "ixg_design": abi.Design, "ixg_asc_error": abi.AscError,
    "ixg_asc_info": abi.AscInfo, "ixg_asc_tier": abi.AscTier, "ixg_radiance_error": abi.RadianceError,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("abi") / "abi_layout")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "abi_layout.c")], check=True)
    return json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)


def test_the_layout_program_lists_every_field(c_layout):
    # the fields tile the struct: no gap wider than alignment padding, and
    # nothing after the last field but tail padding
    for name, s in c_layout.items():
        spans = sorted(s["fields"].values())
        end = 0
        for off, size in spans:
            assert 0 <= off - end < 8, (name, off, end)
            end = off + size
        assert 0 <= s["size"] - end < 8, (name, s["size"], end)


def test_ctypes_mirror_matches_the_header(c_layout):
    assert set(MIRRORS) == set(c_layout)
    for name, cls in MIRRORS.items():
        s = c_layout[name]
        assert C.sizeof(cls) == s["size"], name
        assert [f for f, _ in cls._fields_] == list(s["fields"]), name
        for f, (off, size) in s["fields"].items():
            d = getattr(cls, f)
            assert (d.offset, d.size) == (off, size), (name, f)


# ---- the Rust binding of INTEGRATION.md --------------------------------

_PRIM = {"u8": (1, 1), "i8": (1, 1), "u16": (2, 2), "i16": (2, 2), "u32": (4, 4), "i32": (4, 4),
         "f32": (4, 4), "u64": (8, 8), "i64": (8, 8), "f64": (8, 8), "usize": (8, 8)}


def rust_structs(text):
    """Every `#[repr(C)] pub struct` of the guide's rust blocks, as
    {name: [(field, type)]} in declaration order."""
    blocks = re.findall(r"```rust\n(.*?)```", text, re.S)
    out = {}
    for b in blocks:
        b = re.sub(r"//[^\n]*", "", b)
        for name, body in re.findall(r"#\[repr\(C\)\]\s*pub struct (\w+)\s*\{(.*?)\}", b, re.S):
            out[name] = [(f, t.strip()) for f, t in re.findall(r"pub (\w+):\s*([^,]+?)\s*(?:,|$)", body.strip())]
    return out


def snake(name):
    return re.sub(r"(?<!^)(?=[A-Z])", "_", name).lower()


def rust_layout(structs, name, memo):
    """(size, align, {field: (offset, size)}) of a repr(C) struct."""
    if name in memo:
        return memo[name]

    def ty(t):
        t = t.strip()
        if t in _PRIM:
            return _PRIM[t]
        if t.startswith("*const ") or t.startswith("*mut "):
            return 8, 8
        m = re.fullmatch(r"\[(.+);\s*(\d+)\]", t)
        if m:
            size, align = ty(m.group(1))
            return size * int(m.group(2)), align
        assert t in structs, f"unknown Rust type {t!r} in {name}"
        size, align, _ = rust_layout(structs, t, memo)
        return size, align

    off, align, fields = 0, 1, {}
    for f, t in structs[name]:
        size, a = ty(t)
        off = (off + a - 1) // a * a
        fields[f] = (off, size)
        off += size
        align = max(align, a)
    memo[name] = ((off + align - 1) // align * align, align, fields)
    return memo[name]


def test_integration_guide_rust_binding_matches_the_header(c_layout):
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        structs = rust_structs(f.read())
    # every struct a caller passes by pointer is declared
    for needed in ("IxgFacetPlane", "IxgCamera", "IxgDispersion", "IxgBand", "IxgMaterial",
                   "IxgEnvironment", "IxgHdrMap", "IxgScene", "IxgOpts", "IxgBatchItem",
                   "IxgDesign", "IxgAscError", "IxgAscTier", "IxgRadianceError"):
        assert needed in structs, f"INTEGRATION.md declares no {needed}"
    memo = {}
    for name in structs:
        c = c_layout[snake(name)]
        size, _, fields = rust_layout(structs, name, memo)
        assert list(fields) == list(c["fields"]), (name, list(fields))
        assert size == c["size"], (name, size, c["size"])
        for f, span in fields.items():
            assert list(span) == c["fields"][f], (name, f, span, c["fields"][f])


def test_the_rust_check_catches_a_dropped_field(c_layout):
    # a planted drift (a dropped trailing field) fails
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        structs = rust_structs(f.read())
    structs["IxgMaterial"] = [ft for ft in structs["IxgMaterial"] if ft[0] != "bands_beta"]
    size, _, fields = rust_layout(structs, "IxgMaterial", {})
    assert size != c_layout["ixg_material"]["size"] or list(fields) != list(c_layout["ixg_material"]["fields"])
