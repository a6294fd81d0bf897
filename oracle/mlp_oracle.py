"""CPU oracle of the metric surrogate (test infrastructure only).

An FP64 numpy forward/backward of the network libixg trains on the GPU
(16 -> 4 x 256 GELU -> 3, padded parameter layout of ixg_mlp_param_count),
rounding to BF16 at exactly the points the device stores BF16: the inputs,
the weights' working copy, every hidden activation and every back-propagated
gradient.  Everything else (accumulation, GELU, loss) is FP64, so the GPU
differs from it only by FP32 accumulation order and the device erf/exp.

The reference has no surrogate, so nothing here restates reference code;
parity is pinned by (a) a central finite-difference check of this oracle's
own gradients in FP64 (tests/test_mlp_oracle.py) and (b) the device
gradients against these (tests/test_mlp_gpu.py).
"""
import math

import numpy as np

IN, IN_P, H, L, OUT, OUT_P = 16, 64, 256, 4, 3, 64
W_SHAPES = [(H, IN_P)] + [(H, H)] * 3 + [(OUT_P, H)]
B_SHAPES = [H] * L + [OUT_P]
N_WEIGHTS = sum(r * c for r, c in W_SHAPES)
N_PARAMS = N_WEIGHTS + sum(B_SHAPES)


def bf16(x):
    """Round float64/float32 values to the nearest BF16 (ties to even), as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(v):
    return np.asarray(v, np.float64)


def unpack(params, exact=False):
    p = np.asarray(params, np.float64)
    rnd = _ident if exact else bf16
    ws, bs, off = [], [], 0
    for r, c in W_SHAPES:
        ws.append(rnd(p[off:off + r * c].reshape(r, c)))
        off += r * c
    for n in B_SHAPES:
        bs.append(p[off:off + n].copy())
        off += n
    return ws, bs


def _erf(x):
    return np.vectorize(math.erf, otypes=[np.float64])(x)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(params, x, exact=False):
    """Outputs [B x 64] (first 3 columns are the metrics), activations, pre-activations.

    exact=True skips every BF16 rounding: the plain FP64 network."""
    rnd = _ident if exact else bf16
    ws, bs = unpack(params, exact)
    a = np.zeros((len(x), IN_P))
    a[:, :IN] = rnd(x)
    acts, pres = [a], []
    for l in range(L):
        z = acts[-1] @ ws[l].T + bs[l]
        pres.append(z)
        acts.append(rnd(gelu(z)))
    y = acts[-1] @ ws[L].T + bs[L]
    return y, acts, pres


def loss_and_grads(params, x, t, exact=False):
    """MSE over batch x 3 and d loss / d params (padded layout)."""
    rnd = _ident if exact else bf16
    ws, _ = unpack(params, exact)
    y, acts, pres = forward(params, x, exact)
    b = len(x)
    d = y[:, :OUT] - np.asarray(t, np.float64)
    loss = float(np.mean(d * d))
    dy = np.zeros((b, OUT_P))
    dy[:, :OUT] = 2.0 * d / (b * OUT)
    dy = rnd(dy)
    gw = [None] * (L + 1)
    gb = [None] * (L + 1)
    gw[L] = dy.T @ acts[L]
    gb[L] = dy.sum(axis=0)
    dz = rnd((dy @ ws[L]) * gelu_grad(pres[L - 1]))
    for l in range(L - 1, -1, -1):
        gw[l] = dz.T @ acts[l]
        gb[l] = dz.sum(axis=0)
        if l > 0:
            dz = rnd((dz @ ws[l]) * gelu_grad(pres[l - 1]))
    grads = np.concatenate([g.reshape(-1) for g in gw] + [g.reshape(-1) for g in gb])
    return loss, grads


def loss_only(params, x, t, exact=False):
    y, _, _ = forward(params, x, exact)
    d = y[:, :OUT] - np.asarray(t, np.float64)
    return float(np.mean(d * d))


def init_params(seed):
    """He-uniform weights (half-scale output layer), zero biases, zero padding."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    off = 0
    for li, (r, c) in enumerate(W_SHAPES):
        rows = OUT if li == L else r
        cols = IN if li == 0 else c
        lim = math.sqrt(6.0 / cols) * (0.5 if li == L else 1.0)
        w = np.zeros((r, c), np.float32)
        w[:rows, :cols] = rng.uniform(-lim, lim, (rows, cols))
        p[off:off + r * c] = w.reshape(-1)
        off += r * c
    return p
