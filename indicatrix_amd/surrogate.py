"""Metric surrogate: a BF16 MLP from cut parameters, dispersion and tilt to
brilliance / fire / scintillation, trained and evaluated with libixg's MFMA
kernels on the tracer's own GPU metric sweep.

The reference has no surrogate (its Optimize evaluates every candidate with
the metric tracer, crates/indicatrix-cut-core/src/optimize/objective.rs:303);
this module is new work, and the path tracer stays the judge of its error.

Inputs (16, each scaled to [-1, 1]):
  0..11  the parametric cut family below (tier angles, heights, phases,
         table, culet, aspect)
  12     girdle thickness
  13     n_D      (1.54 .. 2.42)
  14     Abbe V_d (20 .. 60)
  15     tilt     (-30 .. +30 deg)
Outputs (3): brilliance %, fire index, scintillation % (the tracer's
metric_sweep columns 0..2), standardised by the training set's mean / std.
"""
import ctypes as C
import math

import numpy as np

from . import abi
from ._lib import IxgError, lib
from .designs import TILTS_C3, design_planes
from .metrics import metric_sweep
from .tracer import LightingPreset, sellmeier1_from_nd_abbe

N_IN, N_OUT = 16, 3

# the parametric family: (name, lo, hi) of the 12 cut inputs.  Tier heights
# and the table / culet are placed relative to the tier angles so that every
# one of the 50 planes is a face of the stone for every input in the box.
CUT_RANGES = [
    ("table_ratio", 0.45, 0.70),      # table radius / girdle radius
    ("crown_lower_deg", 28.0, 42.0),  # crown tier touching the girdle
    ("crown_step_deg", 6.0, 16.0),    # upper crown tier = lower - step
    ("crown_break", 0.2, 0.8),        # where the crown tiers meet, between table and girdle
    ("crown_phase", 0.0, 1.0),        # upper crown tier azimuth phase (x pi/8)
    ("pav_lower_deg", 40.0, 47.0),    # pavilion tier touching the girdle
    ("pav_step_deg", 1.5, 6.0),       # pavilion tier towards the culet = lower - step
    ("pav_break", 0.2, 0.7),          # where the pavilion tiers meet, between culet and girdle
    ("pav_phase", 0.0, 1.0),          # (x pi/8)
    ("culet_frac", 0.90, 0.99),       # culet depth / pavilion apex depth
    ("aspect", 1.0, 1.5),             # oval stretch along x
    ("girdle_phase", 0.0, 1.0),       # (x pi/16)
]
GIRDLE_RANGE = (0.01, 0.06)           # total girdle band height
ND_RANGE = (1.54, 2.42)
ABBE_RANGE = (20.0, 60.0)
TILT_RANGE = (-30.0, 30.0)
METRIC_COLUMNS = ("brilliance_pct", "fire_index", "scintillation_pct")


def _lerp(u, lo, hi):
    return lo + (hi - lo) * float(u)


def family_params(u):
    """Design dict (designs.design_planes convention) from 13 unit-interval values
    (12 cut parameters + girdle thickness).  Every design has 50 planes: table,
    2 x 8 crown, 16 girdle, 2 x 8 pavilion, culet.

    A tier at angle a through the girdle radius at height h is the surface
    y = h + tan(a) (1 - r) along its facet centre line; two tiers with
    tan a1 > tan a2 and a height step b meet at r* = 1 - b / (tan a1 - tan a2),
    so b is set as a fraction of the span that keeps r* inside the crown
    (table edge .. girdle) or the pavilion (centre .. girdle)."""
    v = {name: _lerp(u[i], lo, hi) for i, (name, lo, hi) in enumerate(CUT_RANGES)}
    half = 0.5 * _lerp(u[12], *GIRDLE_RANGE)
    tau = v["table_ratio"]
    c1 = math.radians(v["crown_lower_deg"])
    c2 = math.radians(v["crown_lower_deg"] - v["crown_step_deg"])
    c_break = v["crown_break"] * (1.0 - tau) * (math.tan(c1) - math.tan(c2))
    table_h = half + c_break + math.tan(c2) * (1.0 - tau)
    p1 = math.radians(v["pav_lower_deg"])
    p2 = math.radians(v["pav_lower_deg"] - v["pav_step_deg"])
    p_break = v["pav_break"] * (math.tan(p1) - math.tan(p2))
    culet_h = v["culet_frac"] * (half + p_break + math.tan(p2))
    crown = [(8, v["crown_lower_deg"], 0.0, half),
             (8, math.degrees(c2), v["crown_phase"] * math.pi / 8.0, half + c_break)]
    pav = [(8, v["pav_lower_deg"], 0.0, half),
           (8, math.degrees(p2), v["pav_phase"] * math.pi / 8.0, half + p_break)]
    return dict(n=50, table_h=table_h, culet_h=culet_h, girdle=16,
                girdle_phase=v["girdle_phase"] * math.pi / 16.0, crown=crown, pavilion=pav,
                aspect=v["aspect"])


def family_planes(u):
    return design_planes(family_params(u))


def features(u13, nd, abbe, tilt_deg):
    """(n, 16) float32 network inputs in [-1, 1]."""
    u13 = np.atleast_2d(np.asarray(u13, np.float64))
    n = len(u13)
    x = np.empty((n, N_IN), np.float32)
    x[:, :13] = 2.0 * u13 - 1.0
    x[:, 13] = 2.0 * (np.broadcast_to(nd, n) - ND_RANGE[0]) / (ND_RANGE[1] - ND_RANGE[0]) - 1.0
    x[:, 14] = 2.0 * (np.broadcast_to(abbe, n) - ABBE_RANGE[0]) / (ABBE_RANGE[1] - ABBE_RANGE[0]) - 1.0
    x[:, 15] = np.broadcast_to(tilt_deg, n) / 30.0
    return x


def sample_candidates(rng, n):
    """n seeded candidates: (u13 (n,13), n_D (n,), Abbe (n,))."""
    u = rng.random((n, 13))
    nd = rng.uniform(*ND_RANGE, n)
    abbe = rng.uniform(*ABBE_RANGE, n)
    return u, nd, abbe


def candidate_batch(u, nd, abbe):
    """The metric_sweep inputs (plane table, design list) for a candidate set."""
    tables, ds, off = [], [], 0
    for i in range(len(u)):
        p = family_planes(u[i])
        b, c = sellmeier1_from_nd_abbe(float(nd[i]), float(abbe[i]))
        d = abi.Dispersion()
        d.kind = abi.DISP_SELLMEIER1
        d.coeff[0], d.coeff[1] = b, c
        tables.append(p)
        ds.append((off, len(p), d))
        off += len(p)
    return np.concatenate(tables), ds


def scoring_environment():
    return LightingPreset.studio("IsoHemisphere", 1.0, 0.0, 0.0)


def trace_metrics(backend, u, nd, abbe, tilts=TILTS_C3, width=64, height=64, samples=32,
                  distance=2.4, chunk=256):
    """Tracer metrics (n, n_tilts, 3) for a candidate set (metric_sweep, chunked by design)."""
    env = scoring_environment()
    out = np.empty((len(u), len(tilts), N_OUT), np.float32)
    for s in range(0, len(u), chunk):
        planes, ds = candidate_batch(u[s:s + chunk], nd[s:s + chunk], abbe[s:s + chunk])
        m = metric_sweep(backend, planes, ds, tilts, width, height, samples, distance, env)
        out[s:s + chunk] = m[..., :N_OUT]
    return out


def trace_dataset(backend, n_designs, seed, tilts=TILTS_C3, width=64, height=64, samples=32,
                  chunk=256):
    """Training data from the GPU metric sweep: X (n*t, 16), T (n*t, 3)."""
    rng = np.random.default_rng(seed)
    u, nd, abbe = sample_candidates(rng, n_designs)
    m = trace_metrics(backend, u, nd, abbe, tilts, width, height, samples, chunk=chunk)
    return design_features(u, nd, abbe, tilts), m.reshape(-1, N_OUT)


class SurrogateMLP:
    """The libixg BF16 MLP (ixg_mlp_*).  Device inputs are torch CUDA float32
    tensors; nothing here computes on the host."""

    def __init__(self, backend, max_batch=4096, seed=0):
        import torch
        self._torch = torch
        self.backend = backend
        self.max_batch = int(max_batch)
        if self.max_batch % 64:
            raise ValueError("max_batch must be a multiple of 64")
        self._h = C.c_void_p()
        st = lib().ixg_mlp_create(backend.handle, self.max_batch, seed, C.byref(self._h))
        if st != abi.IXG_OK:
            raise IxgError(st, "ixg_mlp_create failed")
        self.n_params = int(lib().ixg_mlp_param_count())
        self.t_mean = np.zeros(N_OUT, np.float32)
        self.t_std = np.ones(N_OUT, np.float32)

    def close(self):
        if self._h:
            lib().ixg_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != abi.IXG_OK:
            raise IxgError(st, what)

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self._check(lib().ixg_mlp_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                             self.n_params), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._check(lib().ixg_mlp_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)),
                                             p.size), "set_params")

    def get_state(self):
        """The Adam moments and step count: (m, v, step)."""
        m = np.empty(self.n_params, np.float32)
        v = np.empty(self.n_params, np.float32)
        step = C.c_uint64()
        self._check(lib().ixg_mlp_get_state(self._h, m.ctypes.data_as(C.POINTER(C.c_float)),
                                            v.ctypes.data_as(C.POINTER(C.c_float)), self.n_params,
                                            C.byref(step)), "get_state")
        return m, v, int(step.value)

    def set_state(self, m, v, step):
        m = np.ascontiguousarray(m, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        self._check(lib().ixg_mlp_set_state(self._h, m.ctypes.data_as(C.POINTER(C.c_float)),
                                            v.ctypes.data_as(C.POINTER(C.c_float)), m.size, int(step)),
                    "set_state")

    def _dev(self, a, cols):
        torch = self._torch
        t = torch.as_tensor(a, dtype=torch.float32, device="cuda").contiguous()
        if t.dim() != 2 or t.shape[1] != cols:
            raise ValueError(f"expected (batch, {cols})")
        return t

    def infer_raw(self, x):
        """Network outputs (standardised units) for x (B, 16); B is padded to 64."""
        torch = self._torch
        x = self._dev(x, N_IN)
        n = x.shape[0]
        y = torch.empty((n, N_OUT), dtype=torch.float32, device="cuda")
        for s in range(0, n, self.max_batch):
            xb = x[s:s + self.max_batch]
            b = xb.shape[0]
            bp = (b + 63) // 64 * 64
            if bp != b:
                xb = torch.cat([xb, xb.new_zeros((bp - b, N_IN))])
            yb = torch.empty((bp, N_OUT), dtype=torch.float32, device="cuda")
            torch.cuda.synchronize()
            self._check(lib().ixg_mlp_infer(self._h, C.c_void_p(xb.data_ptr()), bp,
                                            C.c_void_p(yb.data_ptr())), "infer")
            y[s:s + b] = yb[:b]
        return y

    def predict(self, x):
        """Metrics (B, 3) in tracer units."""
        torch = self._torch
        y = self.infer_raw(x)
        return y * torch.as_tensor(self.t_std, device="cuda") + torch.as_tensor(self.t_mean, device="cuda")

    def gradients(self, x, t):
        """(loss, grads) of the MSE for a batch (B a multiple of 64), no update."""
        x = self._dev(x, N_IN)
        t = self._dev(t, N_OUT)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        self._torch.cuda.synchronize()
        self._check(lib().ixg_mlp_gradients(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                            x.shape[0], g.ctypes.data_as(C.POINTER(C.c_float)),
                                            self.n_params, C.byref(loss)), "gradients")
        return float(loss.value), g

    def sync(self):
        """Wait for the network's queued work (ixg_mlp_sync)."""
        self._check(lib().ixg_mlp_sync(self._h), "sync")

    def train_step(self, x, t, lr, want_loss=True):
        loss = C.c_float()
        self._check(lib().ixg_mlp_train_step(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()),
                                             x.shape[0], lr, C.byref(loss) if want_loss else None),
                    "train_step")
        return float(loss.value) if want_loss else None

    def fit(self, x, t, epochs=10, batch=4096, lr=1e-3, lr_final=None, seed=0, log=None,
            weights=None, restandardise=True):
        """Adam on (x, t) in tracer units.  With restandardise (a fresh fit) the
        targets are standardised by their mean / std and the statistics kept;
        without it (a fine-tune of a loaded network) the stored statistics are
        used as they are.  weights (3,) weights each output's squared error: the
        standardising std of output j is divided by sqrt(weights[j]), which is
        the same loss up to a constant.  The learning rate follows a cosine from
        lr to lr_final (constant when None).  Returns the mean training loss of
        each epoch (standardised units)."""
        torch = self._torch
        if batch % 64 or batch > self.max_batch:
            raise ValueError("batch must be a multiple of 64 and at most max_batch")
        t = np.asarray(t, np.float32)
        if restandardise:
            self.t_mean = t.mean(axis=0).astype(np.float32)
            self.t_std = (t.std(axis=0) + 1e-6).astype(np.float32)
            if weights is not None:
                self.t_std = (self.t_std / np.sqrt(np.asarray(weights, np.float32))).astype(np.float32)
        xd = self._dev(x, N_IN)
        td = self._dev((t - self.t_mean) / self.t_std, N_OUT)
        n = (xd.shape[0] // batch) * batch
        if n == 0:
            raise ValueError("fewer samples than one batch")
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        self.rng_state = None
        steps_per_epoch = n // batch
        total = epochs * steps_per_epoch
        history = []
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(xd.shape[0], device="cuda", generator=gen)[:n]
            xs = xd[perm].contiguous()
            ts = td[perm].contiguous()
            torch.cuda.synchronize()
            tot = 0.0
            for s in range(0, n, batch):
                r = lr if lr_final is None else lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * k / total))
                tot += self.train_step(xs[s:s + batch], ts[s:s + batch], r, want_loss=True)
                k += 1
            history.append(tot / steps_per_epoch)
            self.rng_state = gen.get_state().cpu().numpy()   # the batch shuffler, for a resume
            if log:
                log(f"epoch {ep}: loss {history[-1]:.5f}")
        return history

    def fine_tune(self, x, t, steps, lr, batch=4096, seed=0):
        """Adam steps on (x, t) in tracer units, keeping the output statistics
        of the original fit (used to refine the network on verified designs)."""
        torch = self._torch
        t = (np.asarray(t, np.float32) - self.t_mean) / self.t_std
        n = len(x)
        b = min(batch, self.max_batch, n) // 64 * 64
        if b == 0 or steps <= 0:
            return None
        xd = self._dev(x, N_IN)
        td = self._dev(t, N_OUT)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        loss = None
        for k in range(steps):
            idx = torch.randperm(n, device="cuda", generator=gen)[:b]
            xs = xd[idx].contiguous()
            ts = td[idx].contiguous()
            torch.cuda.synchronize()
            loss = self.train_step(xs, ts, lr, want_loss=(k == steps - 1))
        return loss

    # output layer in the padded parameter layout (mlp_api.cpp): W4 [64 x 256]
    # after W0 [256 x 64] and W1..W3 [256 x 256]; b4 [64] after b0..b3 [256]
    _W4_OFF = 256 * 64 + 3 * 256 * 256
    _B4_OFF = _W4_OFF + 64 * 256 + 4 * 256

    def restandardise(self, t_mean, t_std):
        """Move the network to new output statistics without changing what it
        predicts: output j's weights and bias are scaled by old_std / new_std
        and the bias shifted by (old_mean - new_mean) / new_std.  The Adam
        moments restart (they belonged to the old scale)."""
        t_mean = np.asarray(t_mean, np.float32)
        t_std = np.asarray(t_std, np.float32)
        p = self.get_params().astype(np.float64)
        for j in range(N_OUT):
            a = float(self.t_std[j]) / float(t_std[j])
            c = (float(self.t_mean[j]) - float(t_mean[j])) / float(t_std[j])
            w = slice(self._W4_OFF + 256 * j, self._W4_OFF + 256 * (j + 1))
            p[w] *= a
            p[self._B4_OFF + j] = p[self._B4_OFF + j] * a + c
        self.set_params(p.astype(np.float32))
        self.t_mean = t_mean.copy()
        self.t_std = t_std.copy()
        z = np.zeros(self.n_params, np.float32)
        self.set_state(z, z, 0)

    STATE_KEYS = ("params", "t_mean", "t_std", "adam_m", "adam_v", "adam_step", "rng_state")

    def save(self, path, **meta):
        """A resumable checkpoint as .npz: parameters (padded FP32 layout),
        output statistics, the Adam moments and step, and the state of the
        batch shuffler of the last fit (when there was one)."""
        m, v, step = self.get_state()
        extra = {} if getattr(self, "rng_state", None) is None else {"rng_state": self.rng_state}
        np.savez(path, params=self.get_params(), t_mean=self.t_mean, t_std=self.t_std, adam_m=m, adam_v=v,
                 adam_step=np.uint64(step), **extra, **{k: np.asarray(v) for k, v in meta.items()})

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])
            self.t_mean = z["t_mean"].astype(np.float32)
            self.t_std = z["t_std"].astype(np.float32)
            if "adam_m" in z.files:
                self.set_state(z["adam_m"], z["adam_v"], int(z["adam_step"]))
            self.rng_state = z["rng_state"] if "rng_state" in z.files else None
            return {k: z[k][()] for k in z.files if k not in self.STATE_KEYS}


def design_features(u, nd, abbe, tilts=TILTS_C3):
    """Network inputs for every (design, tilt) pair, design-major: (n * t, 16)."""
    nt = len(tilts)
    return features(np.repeat(u, nt, axis=0), np.repeat(nd, nt), np.repeat(abbe, nt),
                    np.tile(np.asarray(tilts, np.float64), len(u)))


def predict_designs(net, u, nd, abbe, tilts=TILTS_C3):
    """Surrogate metrics (n, t, 3) for a candidate set, in one batched pass."""
    y = net.predict(design_features(u, nd, abbe, tilts))
    return y.cpu().numpy().reshape(len(u), len(tilts), N_OUT)


def relative_error(pred, ref):
    """Per-design relative error of the tilt-averaged metrics: (n, 3)."""
    p = np.asarray(pred, np.float64).mean(axis=-2)
    r = np.asarray(ref, np.float64).mean(axis=-2)
    return np.abs(p - r) / np.maximum(np.abs(r), 1e-6)
