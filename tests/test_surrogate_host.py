"""Host side of the surrogate: the parametric cut family and the input features."""
import numpy as np
import pytest

from indicatrix_amd import surrogate as S


def _face_hits(planes, n_dirs=120_000, seed=0):
    """Rays from the centre: which plane each one leaves through (a face census)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n_dirs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = planes.astype(np.float64)
    nd = d @ p[:, :3].T
    t = np.where(nd > 0, -p[:, 3][None, :] / np.where(nd > 0, nd, 1.0), np.inf)
    return np.bincount(np.argmin(t, axis=1), minlength=len(p))


@pytest.mark.parametrize("case", ["low", "high", "mid"] + [f"rand{i}" for i in range(12)])
def test_every_family_plane_is_a_face(case):
    if case == "low":
        u = np.zeros(13)
    elif case == "high":
        u = np.ones(13)
    elif case == "mid":
        u = np.full(13, 0.5)
    else:
        u = np.random.default_rng(int(case[4:])).random(13)
    planes = S.family_planes(u)
    assert planes.shape == (50, 4)
    np.testing.assert_allclose(np.linalg.norm(planes[:, :3], axis=1), 1.0, atol=1e-6)
    assert (planes[:, 3] < 0).all()          # the centre is inside
    assert (_face_hits(planes) > 0).all()    # no redundant plane


def test_features_scale_to_unit_box():
    rng = np.random.default_rng(3)
    u, nd, abbe = S.sample_candidates(rng, 500)
    tilt = rng.uniform(-30, 30, 500)
    x = S.features(u, nd, abbe, tilt)
    assert x.shape == (500, 16) and x.dtype == np.float32
    assert x.min() >= -1.0 - 1e-6 and x.max() <= 1.0 + 1e-6
    np.testing.assert_allclose(x[:, 15], tilt / 30.0, rtol=1e-6)


def test_candidate_batch_layout():
    rng = np.random.default_rng(4)
    u, nd, abbe = S.sample_candidates(rng, 5)
    planes, ds = S.candidate_batch(u, nd, abbe)
    assert planes.shape == (250, 4)
    assert [d[0] for d in ds] == [0, 50, 100, 150, 200]
    assert all(d[1] == 50 for d in ds)
