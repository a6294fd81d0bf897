"""The BF16 MFMA surrogate kernels against the CPU oracle (oracle/mlp_oracle.py):
forward outputs, every parameter's gradient, the Adam update, and training."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(backend):
    from indicatrix_amd.surrogate import SurrogateMLP
    n = SurrogateMLP(backend, max_batch=4096, seed=11)
    yield n
    n.close()


def _problem(seed, b):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    o = mo.N_WEIGHTS
    for n in mo.B_SHAPES[:-1]:
        p[o:o + n] = rng.normal(0, 0.1, n)
        o += n
    p[o:o + mo.OUT] = rng.normal(0, 0.1, mo.OUT)
    x = rng.uniform(-1, 1, (b, mo.IN)).astype(np.float32)
    t = rng.normal(0, 1, (b, mo.OUT)).astype(np.float32)
    return p, x, t


def _layer_slices():
    out, o = [], 0
    for i, (r, c) in enumerate(mo.W_SHAPES):
        out.append((f"W{i}", slice(o, o + r * c)))
        o += r * c
    for i, n in enumerate(mo.B_SHAPES):
        out.append((f"b{i}", slice(o, o + n)))
        o += n
    return out


def test_param_count(net):
    assert net.n_params == mo.N_PARAMS


def test_forward_matches_oracle(net):
    print("mlp forward B=256", flush=True)
    p, x, _ = _problem(21, 256)
    net.set_params(p)
    y = net.infer_raw(x).cpu().numpy()
    ref = mo.forward(p, x)[0][:, :mo.OUT]
    scale = np.abs(ref).max()
    err = np.abs(y - ref).max()
    # FP32 accumulation order and device erf flip a few BF16 activation roundings
    assert err <= 1e-2 * scale, (err, scale)
    assert np.corrcoef(y.ravel(), ref.ravel())[0, 1] > 0.9999


def test_forward_ragged_batch_is_padded(net):
    p, x, _ = _problem(22, 200)
    net.set_params(p)
    y = net.infer_raw(x).cpu().numpy()
    y64 = net.infer_raw(x[:192]).cpu().numpy()
    np.testing.assert_array_equal(y[:192], y64)   # rows are independent and deterministic


def test_gradients_match_oracle(net):
    print("mlp gradients B=512", flush=True)
    p, x, t = _problem(23, 512)
    net.set_params(p)
    loss, g = net.gradients(x, t)
    rl, rg = mo.loss_and_grads(p, x, t)
    assert loss == pytest.approx(rl, rel=1e-3)
    for name, s in _layer_slices():
        a, b = g[s].astype(np.float64), rg[s]
        nb = np.linalg.norm(b)
        if nb == 0.0:
            assert not a.any(), name
            continue
        assert np.linalg.norm(a - b) / nb < 2e-2, name
    # padding stays exactly zero
    w0 = g[_layer_slices()[0][1]].reshape(256, 64)
    assert not w0[:, mo.IN:].any()
    w4 = g[_layer_slices()[4][1]].reshape(64, 256)
    assert not w4[mo.OUT:].any()


def test_gradients_are_deterministic(net):
    p, x, t = _problem(24, 256)
    net.set_params(p)
    l1, g1 = net.gradients(x, t)
    l2, g2 = net.gradients(x, t)
    np.testing.assert_array_equal(g1, g2)


def test_first_adam_step(net):
    """Step 1 of Adam is w - lr * g / (|g| + eps) (bias corrections cancel)."""
    import torch
    p, x, t = _problem(25, 256)
    net.set_params(p)
    _, g = net.gradients(x, t)
    xd = torch.as_tensor(x, device="cuda")
    td = torch.as_tensor(t, device="cuda")
    torch.cuda.synchronize()
    lr = 1e-3
    net.train_step(xd, td, lr)
    w = net.get_params()
    g64 = g.astype(np.float64)
    expect = p.astype(np.float64) - lr * g64 / (np.abs(g64) + 1e-8)
    np.testing.assert_allclose(w, expect, rtol=0, atol=2e-6)


def test_training_fits_a_smooth_target(net):
    print("mlp fit synthetic", flush=True)
    rng = np.random.default_rng(26)
    x = rng.uniform(-1, 1, (16384, mo.IN)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2], np.cos(x[:, 3] + x[:, 4]),
                  x[:, 5] ** 2 - 0.5 * x[:, 6]], axis=1).astype(np.float32)
    net.set_params(mo.init_params(26))
    hist = net.fit(x, t, epochs=30, batch=1024, lr=1e-3, seed=1)
    assert hist[-1] < 0.1 * hist[0], hist
    pred = net.predict(x[:1024]).cpu().numpy()
    assert np.abs(pred - t[:1024]).mean() < 0.1


def test_checkpoint_resumes_adam_exactly(net, tmp_path):
    """save() keeps the Adam moments and step: two steps, save, two more, and
    a fresh network loaded from the checkpoint taking the same two steps lands
    on the same parameters bit for bit."""
    import torch
    from indicatrix_amd.surrogate import SurrogateMLP
    p, x, t = _problem(27, 256)
    xd = torch.as_tensor(x, device="cuda")
    td = torch.as_tensor(t, device="cuda")
    net.set_params(p)
    torch.cuda.synchronize()
    for _ in range(2):
        net.train_step(xd, td, 1e-3, want_loss=False)
    path = str(tmp_path / "ckpt.npz")
    net.save(path)
    m, v, step = net.get_state()
    assert step == 2 and np.abs(m).sum() > 0 and np.abs(v).sum() > 0
    for _ in range(2):
        net.train_step(xd, td, 1e-3, want_loss=False)
    want = net.get_params()
    other = SurrogateMLP(net.backend, max_batch=256, seed=99)
    try:
        other.load(path)
        assert other.get_state()[2] == 2
        for _ in range(2):
            other.train_step(xd, td, 1e-3, want_loss=False)
        np.testing.assert_array_equal(other.get_params(), want)
    finally:
        other.close()
