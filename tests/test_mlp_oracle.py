"""The surrogate's CPU oracle pinned by mathematics: central finite
differences of the plain FP64 network, an independent BF16 rounding (torch's
float32 -> bfloat16 cast), and the padding invariants of the parameter layout."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402


def _offsets():
    offs, o = [], 0
    for r, c in mo.W_SHAPES:
        offs.append((o, r, c))
        o += r * c
    boffs = []
    for n in mo.B_SHAPES:
        boffs.append((o, n))
        o += n
    return offs, boffs


def _problem(seed=1, b=6):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed).astype(np.float64)
    offs, boffs = _offsets()
    for (o, n), rows in zip(boffs, [mo.H] * mo.L + [mo.OUT]):
        p[o:o + rows] = rng.normal(0, 0.1, rows)      # non-zero biases exercise their gradient
    x = rng.uniform(-1, 1, (b, mo.IN))
    t = rng.normal(0, 1, (b, mo.OUT))
    return p, x, t


def test_param_count_matches_layout():
    assert mo.N_PARAMS == 64 * 256 + 3 * 256 * 256 + 64 * 256 + 4 * 256 + 64
    logical = 16 * 256 + 3 * 256 * 256 + 256 * 3 + 4 * 256 + 3
    assert logical == 202_499


def test_exact_gradients_match_central_differences():
    p, x, t = _problem()
    _, g = mo.loss_and_grads(p, x, t, exact=True)
    offs, boffs = _offsets()
    rng = np.random.default_rng(5)
    idx = []
    for l, (o, r, c) in enumerate(offs):
        rows = mo.OUT if l == mo.L else r
        cols = mo.IN if l == 0 else c
        for _ in range(6):
            idx.append(o + int(rng.integers(rows)) * c + int(rng.integers(cols)))
    for l, (o, n) in enumerate(boffs):
        rows = mo.OUT if l == mo.L else n
        idx += [o + int(rng.integers(rows)) for _ in range(3)]
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_only(pp, x, t, exact=True) - mo.loss_only(pm, x, t, exact=True)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def test_gelu_derivative_matches_differences():
    z = np.linspace(-6, 6, 97)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(z), fd, rtol=1e-7, atol=1e-9)
    # closed forms: GELU(0) = 0, GELU(x) -> x for large x, GELU'(0) = 1/2
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    assert mo.gelu(np.array([8.0]))[0] == pytest.approx(8.0, abs=1e-12)
    assert mo.gelu_grad(np.array([0.0]))[0] == 0.5


def test_bf16_rounding_matches_torch_cast():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    v = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 2000), rng.normal(0, 300, 2000),
                        1.0 + np.arange(16) * 2.0 ** -8,  # every tie and midpoint around 1
                        [0.0, -0.0, 1e-40, 3.0e38]]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).float().numpy().astype(np.float64)
    np.testing.assert_array_equal(mo.bf16(v), ref)
    assert mo.bf16(np.float32(1 + 2 ** -8)) == 1.0          # tie to even (down)
    assert mo.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6  # tie to even (up)


def test_padding_receives_zero_gradient():
    p, x, t = _problem(seed=3)
    _, g = mo.loss_and_grads(p, x, t)
    offs, boffs = _offsets()
    o, r, c = offs[0]
    assert not g[o:o + r * c].reshape(r, c)[:, mo.IN:].any()
    o, r, c = offs[mo.L]
    assert not g[o:o + r * c].reshape(r, c)[mo.OUT:, :].any()
    o, n = boffs[mo.L]
    assert not g[o + mo.OUT:o + n].any()


def test_bf16_gradients_close_to_exact():
    p, x, t = _problem(seed=4, b=64)
    l1, g1 = mo.loss_and_grads(p, x, t)
    l0, g0 = mo.loss_and_grads(p, x, t, exact=True)
    assert l1 == pytest.approx(l0, rel=2e-2)
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 5e-2


def test_planted_transpose_is_caught():
    """A transposed hidden weight (a plausible indexing slip) fails the difference check."""
    p, x, t = _problem(seed=6, b=4)
    offs, _ = _offsets()
    o, r, c = offs[2]
    q = p.copy()
    q[o:o + r * c] = p[o:o + r * c].reshape(r, c).T.reshape(-1)
    _, g = mo.loss_and_grads(p, x, t, exact=True)   # gradient of the true network
    i = o + 5 * c + 9
    h = 1e-6
    qp, qm = q.copy(), q.copy()
    qp[i] += h
    qm[i] -= h
    fd_wrong = (mo.loss_only(qp, x, t, exact=True) - mo.loss_only(qm, x, t, exact=True)) / (2 * h)
    assert abs(g[i] - fd_wrong) > 1e-4 * max(abs(g[i]), 1e-6)
