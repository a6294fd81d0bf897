"""Fan-grid pose metrics and tilt-axis profiles on the GPU.

Against the reference's golden bits (golden_regression.rs:207-233) and the
oracle's evaluate_pose: the class shares and the scintillation are formed
from integer counts and must agree exactly; the Fire index agrees to the
device's acos (a few ULP per returned ray).
"""
import math

import numpy as np
import pytest

from indicatrix_amd import designs, metrics, tracer, workloads
from test_oracle import GOLDEN, bits, diamond_material
from test_pose_oracle import oracle_pose, srb_rows, studio_rings

pytestmark = pytest.mark.gpu

FIRE_RTOL = 1e-5


def _srb():
    rows = srb_rows()
    return rows, [(0, len(rows), diamond_material().dispersion)]


def _check_against_oracle(dev, rows, material, env, poses, want_all):
    for p, (yaw, pitch) in enumerate(poses):
        ref = oracle_pose(rows, material, env, yaw, pitch, want_all)
        got = dev[p]
        assert [bits(v) for v in got[[0, 2, 3, 4]]] == [bits(v) for v in ref[[0, 2, 3, 4]]], (yaw, pitch)
        assert abs(float(got[1]) - float(ref[1])) <= FIRE_RTOL * max(1.0, float(ref[1])), (yaw, pitch)


def test_pose_metrics_match_reference_golden_bits(backend):
    g = GOLDEN["golden_regression"]["optical_metrics"]
    rows, des = _srb()
    print("pose golden: SRB diamond, RingLights", flush=True)
    m = metrics.pose_metrics(backend, rows, des, [(g["cam_yaw"], g["cam_pitch"])], studio_rings())[0, 0]
    want = np.array(g["bits"], np.uint32).view(np.float32)
    assert [bits(v) for v in m[[0, 2, 3, 4]]] == [int(b) for b in np.array(g["bits"])[[0, 2, 3, 4]]]
    assert abs(float(m[1]) - float(want[1])) <= FIRE_RTOL * float(want[1])


@pytest.mark.parametrize("want_all", [True, False])
def test_pose_metrics_match_oracle(backend, want_all):
    poses = [(0.0, 0.45), (0.7, 1.2), (2.0, 0.1), (0.0, math.pi / 2), (3.5, 0.9)]
    env = studio_rings()
    rows, des = _srb()
    print(f"pose vs oracle: SRB, want_all={want_all}", flush=True)
    dev = metrics.pose_metrics(backend, rows, des, poses, env, want_all)[0]
    _check_against_oracle(dev, rows, diamond_material(), env, poses, want_all)
    # seeded designs of uneven size with their own materials, one batch
    table, ds, _ = designs.sweep_batch(3, first_seed=5)
    print("pose vs oracle: 3 seeded designs", flush=True)
    dev = metrics.pose_metrics(backend, table, ds, poses, env, want_all)
    for d, (off, cnt, disp) in enumerate(ds):
        mat = diamond_material()
        mat.dispersion = disp
        _check_against_oracle(dev[d], table[off:off + cnt], mat, env, poses, want_all)


def test_pose_metrics_of_catalogue_designs_match_oracle(backend):
    """The four committed GemCAD files, read and cut by the library, as one
    batch of designs; each must equal the oracle's evaluate_pose."""
    import os
    from indicatrix_amd import asc
    here = os.path.dirname(os.path.abspath(__file__))
    rows, ds = [], []
    mat = diamond_material()
    for name in ["trichecker12", "forfun", "texasstar", "shah"]:
        with open(os.path.join(here, "golden", "asc", name + ".asc"), "rb") as f:
            planes = np.frombuffer(bytes(asc.from_asc_schedule(asc.parse_asc_bytes(f.read()))),
                                   np.float32).reshape(-1, 4)
        ds.append((sum(len(r) for r in rows), len(planes), mat.dispersion))
        rows.append(planes)
    table = np.concatenate(rows)
    poses = [(0.0, 0.45), (0.9, 1.1), (2.6, 0.2)]
    env = studio_rings()
    print("pose vs oracle: 4 catalogue .asc designs", flush=True)
    dev = metrics.pose_metrics(backend, table, ds, poses, env, True)
    for d, (off, cnt, _) in enumerate(ds):
        _check_against_oracle(dev[d], table[off:off + cnt], mat, env, poses, True)


def test_pose_metrics_other_lighting(backend):
    env = tracer.LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0)
    rows = workloads.c2_planes()
    mat = diamond_material()
    poses = [(0.2, 0.5), (1.0, 1.3)]
    print("pose vs oracle: C2 oval, IsoHemisphere", flush=True)
    dev = metrics.pose_metrics(backend, rows, [(0, len(rows), mat.dispersion)], poses, env)[0]
    _check_against_oracle(dev, rows, mat, env, poses, True)


def test_batched_poses_equal_single_poses(backend):
    """pose_batch.rs:258-285: a batch returns what one-at-a-time calls do, bit for bit."""
    rows, des = _srb()
    poses = [(0.3 * i, 0.06 * i) for i in range(23)]
    print("pose batch vs singles", flush=True)
    batch = metrics.pose_metrics(backend, rows, des, poses, studio_rings(), False)[0]
    for p, pose in enumerate(poses[::5]):
        one = metrics.pose_metrics(backend, rows, des, [pose], studio_rings(), False)[0, 0]
        assert batch[p * 5].tobytes() == one.tobytes()


def test_axis_profiles_follow_the_tilt_slots(backend):
    """profile.rs:146-200 / pose_batch.rs:118-156: tilt t at azimuth A is the pose
    (A or A + 180 deg, 90 - |t| deg); the three curves are the Fire-less pass."""
    rows, des = _srb()
    env = studio_rings()
    print("axis profiles: SRB, 2 azimuths", flush=True)
    curves = metrics.axis_profiles(backend, rows, des, [0.0, 45.0], env)[0]
    assert curves.shape == (2, 3, 181)
    to_rad = np.float32(math.pi) / np.float32(180.0)
    for a, az in enumerate([0.0, 45.0]):
        for tilt in (-90, -31, 0, 1, 45, 90):
            yaw = np.float32(az if tilt >= 0 else az + 180.0) * to_rad
            pitch = np.float32(90 - abs(tilt)) * to_rad
            m = metrics.pose_metrics(backend, rows, des, [(yaw, pitch)], env, False)[0, 0]
            got = curves[a, :, tilt + 90]
            assert got.tobytes() == m[[0, 4, 3]].tobytes(), (az, tilt)
    assert np.isfinite(curves).all()
    assert abs(float(curves[0, 0, 90] + curves[0, 1, 90] + curves[0, 2, 90]) - 100.0) < 1e-3


def test_pose_empty_design_and_validation(backend):
    rows, des = _srb()
    table = np.concatenate([rows, rows])
    ds = [(0, len(rows), des[0][2]), (len(rows), 0, des[0][2])]
    out = metrics.pose_metrics(backend, table, ds, [(0.0, 0.45)], studio_rings())
    assert out[1, 0].tolist() == [85.0, 25.0, 75.0, 5.0, 5.0]
    bad = rows.copy()
    bad[3, :3] *= 2.0
    with pytest.raises(tracer.IxgError):
        metrics.pose_metrics(backend, bad, des, [(0.0, 0.45)], studio_rings())
    with pytest.raises(tracer.IxgError):
        metrics.pose_metrics(backend, rows, [(0, len(rows) + 1, des[0][2])], [(0.0, 0.45)], studio_rings())
