"""The FP64 twin of the oracle (oracle/ixo64.c) and the film's error against it.

The twin is oracle/ixo.c compiled with double arithmetic: the same estimator,
sample streams and operation order.  These tests check that it is the twin it
claims to be (double precision for real, f32 agreement wherever the two make
the same decisions, the analytic furnace anchor) and state the FP32 oracle's
error against it on the C1 frame and a C2 sub-frame.  The GPU's own error
against the twin is stated by tests/test_film_gpu.py.

What dominates the FP32 error is not arithmetic drift but decisions the
reference's algorithm takes against f32 thresholds -- a companion channel's
refracted direction matching the hero's (dot >= 1 - 1e-6, inside f32's
rounding of the dot product on a few percent of exit events), the branch
draws against Fresnel terms -- which flip a path's channel set or event
sequence on a few percent of paths.  The per-pixel error therefore falls with
the sample count while the frame total stays within 1e-3.
"""
import numpy as np
import pytest

import oracle_lib
from indicatrix_amd import workloads

EVENT_KEYS = ("bounces", "termination", "facet_hash", "n_tir", "n_reflect", "n_refract", "n_probe")


def pixel_error(a, ref):
    """Per-pixel relative error of the Y sums over pixels the FP64 sum lights,
    and the frame total's: (mean, p99, max, frame)."""
    y, r = a[:, 1], ref[:, 1]
    m = r > 0
    rel = np.abs(y[m] - r[m]) / r[m]
    return float(rel.mean()), float(np.quantile(rel, 0.99)), float(rel.max()), float(abs(y.sum() - r.sum()) / r.sum())


def agreement(scene, spp):
    x32, e32 = oracle_lib.trace_paths(scene, 0, spp)
    x64, e64 = oracle_lib.trace_paths64(scene, 0, spp)
    same = np.ones(len(x32), bool)
    for k in EVENT_KEYS:
        same &= e32[k] == e64[k]
    rel = np.abs(x32 - x64).max(axis=1) / np.maximum(np.abs(x64).max(axis=1), 1e-12)
    return x32, x64, same, rel


def test_the_twin_computes_in_double():
    sc = workloads.film_c1_scene(32, 32)
    x32, x64, same, rel = agreement(sc, 4)
    lit = x64.max(axis=1) > 0
    # not a second f32 build: lit paths whose decisions agree differ in the low bits
    assert np.mean((x32[lit & same] != x64[lit & same]).any(axis=1)) > 0.9
    # and only there: agreeing paths within f32 rounding accumulated over a path
    assert np.median(rel[same & lit]) < 2e-6, np.median(rel[same & lit])
    assert np.quantile(rel[same & lit], 0.9) < 1e-4, np.quantile(rel[same & lit], 0.9)


@pytest.mark.parametrize("case", ["c1", "c2"])
def test_fp32_error_against_fp64(case):
    """The error table: C1 = the full 128 x 128 @ 64 frame, C2 = the oval and
    its uniaxial pleochroic material on a 128 x 128 @ 16 sub-frame."""
    if case == "c1":
        sc, spp = workloads.film_c1_scene(), 64
    else:
        sc, spp = workloads.film_c2_scene(128, 128, 16), 16
    a32 = oracle_lib.accumulate(sc, 0, spp)
    a64 = oracle_lib.accumulate64(sc, 0, spp)
    mean, p99, mx, frame = pixel_error(a32, a64)
    _, _, same, rel = agreement(sc, spp)
    flips = 1.0 - same.mean()
    channel = float((rel[same] > 1e-3).mean())
    print(f"{case}: FP32 oracle vs FP64 twin, per-pixel Y: mean {mean:.3e} p99 {p99:.3e} max {mx:.3e}; "
          f"frame {frame:.3e}; event-sequence flips {flips:.3e}; agreeing paths beyond 1e-3 {channel:.3e}")
    assert frame < 1e-3
    assert flips < 0.05 and channel < 0.03
    assert mean < (0.03 if case == "c1" else 0.08)
    # more samples per pixel, less per-pixel error (the flips are noise, not a bias)
    if case == "c1":
        lo32 = oracle_lib.accumulate(sc, 0, 8)
        lo64 = oracle_lib.accumulate64(sc, 0, 8)
        assert pixel_error(lo32, lo64)[0] > 1.5 * mean


def test_furnace_anchor_in_fp64():
    """The analytic white furnace (tests/test_hdr_env.py) through the twin:
    within the 16-bounce truncation the reference measures (1 %), and the
    FP32 estimate within 0.1 % of the twin's."""
    from test_hdr_env import furnace_scene, furnace_target
    sc = furnace_scene()
    n = 12 * 12 * 96
    m64 = oracle_lib.accumulate64(sc, 0, 96).sum(axis=0) / n
    m32 = oracle_lib.accumulate(sc, 0, 96).sum(axis=0) / n
    target = furnace_target(2.5)
    rel64 = np.abs(m64 - target) / target
    print(f"furnace: FP64 {((m64 - target) / target).tolist()}, FP32 {((m32 - target) / target).tolist()}")
    assert rel64.max() < 0.01
    assert (np.abs(m32 - m64) / m64).max() < 1e-3
