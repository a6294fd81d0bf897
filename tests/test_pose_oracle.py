"""Pins the oracle's fan-grid pose metrics and stone-width measurement, and
checks the library's host-side width measurement against it (no GPU needed).

evaluate_gem_optical_metrics is pinned by the reference's own golden bits
(golden_regression.rs:207-233); the width measurement by closed forms
(boxes, regular prisms, scale invariance of the power-of-two normalisation)
and by the cases measure_solid rejects (unbounded, flat, empty).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import abi, designs, metrics, workloads
from test_oracle import GOLDEN, bits, diamond_material


def _planes(rows):
    rows = np.asarray(rows, np.float32).reshape(-1, 4)
    return (abi.FacetPlane * max(1, len(rows))).from_buffer_copy(
        rows.tobytes() if len(rows) else bytes(C.sizeof(abi.FacetPlane))), len(rows)


def oracle_width(rows):
    pl, n = _planes(rows)
    w = C.c_double()
    ok = oracle_lib.lib().ixo_model_width(pl, n, C.byref(w))
    return w.value if ok else None


def srb_rows():
    planes = (abi.FacetPlane * 80)()
    n = oracle_lib.lib().ixo_cut_standard_round_brilliant(planes)
    return np.frombuffer(bytes(planes), np.float32).reshape(-1, 4)[:n].copy()


def box(hx, hy, hz):
    return [[1, 0, 0, -hx], [-1, 0, 0, -hx], [0, 1, 0, -hy], [0, -1, 0, -hy], [0, 0, 1, -hz], [0, 0, -1, -hz]]


def prism(k, r, h, phase=0.0):
    rows = [[math.cos(phase + 2 * math.pi * i / k), 0.0, math.sin(phase + 2 * math.pi * i / k), -r]
            for i in range(k)]
    return rows + [[0, 1, 0, -h], [0, -1, 0, -h]]


def studio_rings():
    return abi.Environment(abi.PRESETS["RingLights"], 1.0, 0.85, 0.95, 0.0, 1.0, 16.0)


def oracle_pose(rows, material, env, yaw, pitch, want_all=True):
    pl, n = _planes(rows)
    out = (C.c_float * 5)()
    st = oracle_lib.lib().ixo_pose_metrics(pl, n, C.byref(material), C.byref(env), yaw, pitch,
                                           1 if want_all else 0, out)
    assert st == 0
    return np.array(out[:], np.float32)


# ---------------------------------------------------------------- pose metrics
def test_pose_metrics_match_reference_golden_bits():
    """golden_regression.rs:207-233: Diamond SRB, yaw 0, pitch 0.45,
    RingLights.studio(1.0, 0.85, 0.95) -- all five metrics bit-exact."""
    g = GOLDEN["golden_regression"]["optical_metrics"]
    m = oracle_pose(srb_rows(), diamond_material(), studio_rings(), g["cam_yaw"], g["cam_pitch"])
    assert [bits(v) for v in m] == g["bits"]


def test_fireless_pass_keeps_the_three_percentages():
    """evaluate.rs:214-226: the core pass reports brilliance / extinction /
    windowing bit-identical to the full pass and the 25 / 75 placeholders."""
    rows, mat, env = srb_rows(), diamond_material(), studio_rings()
    for yaw, pitch in [(0.0, 0.45), (0.7, 1.2), (2.0, 0.1)]:
        full = oracle_pose(rows, mat, env, yaw, pitch, True)
        core = oracle_pose(rows, mat, env, yaw, pitch, False)
        assert [bits(v) for v in core[[0, 3, 4]]] == [bits(v) for v in full[[0, 3, 4]]]
        assert core[1] == 25.0 and core[2] == 75.0


def test_empty_design_reports_placeholders():
    """evaluate.rs:259-267: no planes -> the display defaults."""
    m = oracle_pose(np.zeros((0, 4), np.float32), diamond_material(), studio_rings(), 0.0, 0.45)
    assert m.tolist() == [85.0, 25.0, 75.0, 5.0, 5.0]


def test_pose_percentages_partition_the_hits():
    """Every hit is windowed, extinct or returned (no entry-blocked rays for a
    polished stone above n = 1.1), so the three shares sum to 100."""
    rows, mat, env = srb_rows(), diamond_material(), studio_rings()
    for yaw, pitch in [(0.0, 0.45), (0.3, 1.5), (1.1, 0.8)]:
        m = oracle_pose(rows, mat, env, yaw, pitch)
        assert abs(float(m[0] + m[3] + m[4]) - 100.0) < 1e-3
        assert 0.1 <= m[1] and 0.0 <= m[2] <= 100.0


def test_flat_index_has_no_fire():
    """A dispersion-free material: the F and C traces coincide, every Fire angle
    is acos(1) = 0 and the index floors at 0.1 (accumulate.rs:151)."""
    flat = diamond_material()
    flat.dispersion.kind = abi.DISP_CAUCHY
    flat.dispersion.coeff[0], flat.dispersion.coeff[1], flat.dispersion.coeff[2] = 2.4, 0.0, 0.0
    m = oracle_pose(srb_rows(), flat, studio_rings(), 0.0, 0.45)
    assert m[1] == np.float32(0.1)


# ---------------------------------------------------------------- stone width
def test_width_of_boxes_and_prisms():
    assert oracle_width(box(1, 1, 1)) == 2.0
    assert oracle_width(box(0.7, 0.3, 1.5)) == pytest.approx(1.4, abs=1e-7)
    assert oracle_width(box(1.5, 0.3, 0.7)) == pytest.approx(1.4, abs=1e-7)
    # a regular octagonal prism with a face normal along x: x and z extents are 2r
    assert oracle_width(prism(8, 0.8, 0.5)) == pytest.approx(1.6, abs=1e-6)
    # rotated by half a sector: vertices on the axes, extents 2r / cos(pi/8)
    assert oracle_width(prism(8, 0.8, 0.5, math.pi / 8)) == pytest.approx(1.6 / math.cos(math.pi / 8), abs=1e-6)


def test_width_is_scale_invariant():
    """pow2_scale_norm (stone_metrics/mod.rs:152): the epsilons act on the
    normalised arrangement, so a 1000x box measures 1000x exactly."""
    assert oracle_width(box(1000, 1000, 1000)) == 2000.0
    assert oracle_width(box(1e-3, 1e-3, 1e-3)) == pytest.approx(2e-3, rel=1e-6)


def test_width_rejects_what_measure_solid_rejects():
    assert oracle_width(np.zeros((0, 4))) is None                       # no planes
    assert oracle_width(box(1, 1, 1)[:5]) is None                         # open: reaches the blank
    assert oracle_width([[0, 1, 0, 0], [0, -1, 0, 0]] + box(1, 1, 1)[:2] + box(1, 1, 1)[4:]) is None  # flat
    dup = box(1, 1, 1) + box(1, 1, 1)[:2]                                 # duplicates are dropped
    assert oracle_width(dup) == 2.0


def test_srb_width_and_half_width():
    """The standard round brilliant's girdle radius is 1 (round_brilliant.rs:125)."""
    w = oracle_width(srb_rows())
    assert w == pytest.approx(2.0, abs=1e-6)
    pl, n = _planes(srb_rows())
    assert oracle_lib.lib().ixo_girdle_half_width(pl, n) == np.float32(w * 0.5)
    pl, n = _planes(box(1, 1, 1)[:5])
    assert oracle_lib.lib().ixo_girdle_half_width(pl, n) == 1.0   # fallback (fan.rs:31)


@pytest.mark.parametrize("seed", [0, 3, 11])
def test_library_width_matches_oracle(seed):
    """The library's pruned f64 walk (host code, no GPU) against the oracle's
    exhaustive one: bit-identical widths on seeded designs of 32-256 planes."""
    rows = designs.design_planes(designs.seeded_params(seed))
    assert metrics.model_width(rows) == oracle_width(rows)


def test_library_width_matches_oracle_on_fixed_cuts():
    for rows in [srb_rows(), workloads.c2_planes(), np.asarray(box(0.7, 0.3, 1.5), np.float32),
                 np.asarray(prism(8, 0.8, 0.5, math.pi / 8), np.float32)]:
        assert metrics.model_width(rows) == oracle_width(rows)
    assert metrics.model_width(np.asarray(box(1, 1, 1)[:5], np.float32)) is None
    assert metrics.model_width(np.zeros((0, 4), np.float32)) is None
