"""Algorithmic FLOP counts of the benchmark kernels (for roofline figures).

Film and metric counts are exact: the oracle's counting build (every FP32
add, subtract, multiply, divide, square root and transcendental counted 1,
fused multiply-add 2; negation, abs, min/max, compares, conversions and
integer work 0) was run on the benchmark's own inputs, and the totals are kept
in data/flop_counts.json together with the inputs they belong to.

The device film kernel does less than the algorithm in one place: a ray
leaving the convex stone outward through facet f provably misses the stone,
so the kernel resolves that scan from f's plane alone (12 FLOPs) where the
algorithm scans all N planes (12 N).  The device count is therefore the
algorithm's minus 12 (N - 1) for every such scan, taken from the device's own
exact per-launch event totals (ixg_count_events: bounces and probes that
needed a scan, and the full scans run).

The MLP counts are the GEMMs' 2 M N K over the logical layer shapes
(16 -> 256 -> 256 -> 256 -> 256 -> 3), not the zero padding the kernels carry.
"""
import json
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
COUNTS_PATH = os.path.join(_HERE, "data", "flop_counts.json")
PER_PLANE = 12   # two 3-term dot products (5 each), the offset add, the divide

_counts = None


def counts():
    global _counts
    if _counts is None:
        with open(COUNTS_PATH) as f:
            _counts = json.load(f)
    return _counts


def skipped_scans(totals):
    """Scans the device resolved from one plane: every bounce and exit-split
    probe needs a scan; `scans` and `probe_scans` are the ones run in full."""
    return (totals["bounces"] - totals["scans"]) + (totals["probe"] - totals["probe_scans"])


def film_device_flops(entry, totals):
    """(device FLOPs, algorithm FLOPs) of the film call a flop_counts.json
    entry describes, given the device's event totals for that same call."""
    algo = entry["flops"]
    return algo - PER_PLANE * (entry["planes"] - 1) * skipped_scans(totals), algo


def sweep_flops():
    """FLOPs of the whole benchmark metric sweep (512 designs x 21 tilts)."""
    return counts()["sweep_c3"]["sweep_flops"]


MLP_LAYERS = [(16, 256), (256, 256), (256, 256), (256, 256), (256, 3)]


def mlp_forward_flops(batch):
    return sum(2 * batch * k * n for k, n in MLP_LAYERS)


def mlp_train_step_flops(batch):
    """Forward, the backward data GEMMs (every layer but the first), and the
    weight-gradient GEMMs, at 2 M N K each."""
    fwd = mlp_forward_flops(batch)
    bwd_data = sum(2 * batch * k * n for k, n in MLP_LAYERS[1:])
    bwd_weight = fwd
    return fwd + bwd_data + bwd_weight
