"""Train the metric surrogate on the tracer's own GPU metric sweep.

    python -m indicatrix_amd.train_surrogate --designs 95239 --epochs 60 \\
        --out indicatrix_amd/data/surrogate.npz

Data: seeded candidates from the parametric family (surrogate.sample_candidates)
traced by ixg_metric_sweep at 64 x 64, 32 samples, 21 tilts (-30..30 deg), one
sample per (design, tilt): 95 239 designs give 2.0 M samples.  A held-out set
of designs (other seed) measures the tilt-averaged relative error against the
same tracer, and a subset of it against the 128 x 128, 64-sample verification
render the screened optimiser uses.
"""
import argparse
import json
import time

import numpy as np

from . import surrogate as S
from .tracer import GpuBackend


def build_dataset(backend, n_designs, seed, width=64, height=64, samples=32, chunk=2048, log=print):
    rng = np.random.default_rng(seed)
    u, nd, abbe = S.sample_candidates(rng, n_designs)
    m = np.empty((n_designs, len(S.TILTS_C3), S.N_OUT), np.float32)
    t0 = time.perf_counter()
    for s in range(0, n_designs, chunk):
        e = min(n_designs, s + chunk)
        m[s:e] = S.trace_metrics(backend, u[s:e], nd[s:e], abbe[s:e], S.TILTS_C3, width, height, samples)
        dt = time.perf_counter() - t0
        log(f"traced {e}/{n_designs} designs, {dt:.1f} s, {e / dt:.1f} design evals/s")
    return u, nd, abbe, m, time.perf_counter() - t0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--designs", type=int, default=95239)
    ap.add_argument("--val-designs", type=int, default=2048)
    ap.add_argument("--verify-designs", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-final", type=float, default=1e-5)
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    b = GpuBackend()
    u, nd, abbe, m, t_data = build_dataset(b, a.designs, a.seed, log=log)
    x = S.design_features(u, nd, abbe)
    t = m.reshape(-1, S.N_OUT)
    log(f"dataset {x.shape[0]} samples in {t_data:.1f} s")
    uv, ndv, abv, mv, _ = build_dataset(b, a.val_designs, a.seed + 1_000_003, log=log)

    net = S.SurrogateMLP(b, max_batch=max(a.batch, 64 * ((a.val_designs * 21 + 63) // 64)), seed=a.seed)
    t0 = time.perf_counter()
    hist = net.fit(x, t, epochs=a.epochs, batch=a.batch, lr=a.lr, lr_final=a.lr_final, seed=a.seed,
                   log=log)
    t_fit = time.perf_counter() - t0
    steps = a.epochs * (len(x) // a.batch)
    log(f"fit {a.epochs} epochs ({steps} Adam steps) in {t_fit:.1f} s, {1e3 * t_fit / steps:.3f} ms/step")

    pv = S.predict_designs(net, uv, ndv, abv)
    err_val = S.relative_error(pv, mv)
    nv = a.verify_designs
    m128 = S.trace_metrics(b, uv[:nv], ndv[:nv], abv[:nv], S.TILTS_C3, 128, 128, 64)
    err_128 = S.relative_error(pv[:nv], m128)
    err_res = S.relative_error(mv[:nv], m128)
    rep = dict(
        designs=a.designs, samples=int(x.shape[0]), epochs=a.epochs, batch=a.batch, lr=a.lr,
        lr_final=a.lr_final, seed=a.seed, data_seconds=t_data, fit_seconds=t_fit,
        ms_per_step=1e3 * t_fit / steps, final_loss=hist[-1],
        metrics=list(S.METRIC_COLUMNS),
        val_rel_err_median=np.median(err_val, axis=0).tolist(),
        val_rel_err_mean=err_val.mean(axis=0).tolist(),
        val_rel_err_p90=np.percentile(err_val, 90, axis=0).tolist(),
        verify128_rel_err_median=np.median(err_128, axis=0).tolist(),
        verify128_rel_err_mean=err_128.mean(axis=0).tolist(),
        tracer_64_vs_128_rel_err_median=np.median(err_res, axis=0).tolist(),
    )
    log(json.dumps(rep))
    net.save(a.out, report=json.dumps(rep))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return rep


if __name__ == "__main__":
    main()
