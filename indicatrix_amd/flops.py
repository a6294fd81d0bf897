"""Algorithmic FLOP model of the film transport kernel.

FLOPs are counted from the arithmetic the reference's tracer performs per
event (FMA = 2, every other add/mul/div/sqrt/transcendental = 1; integer hash
work and compares are not counted), multiplied by the per-path event counts
the kernel itself reports (ixg_count_events), so the model is exact in the
counts and approximate only in the per-event constants below.
"""

# per path
PATH_SETUP = 220       # stratified draws, camera ray, 8-lambda comb, 8 x Sellmeier-3
FINALIZE = 285         # split commit, 8 MIS families, CMF lerp, XYZ, von Kries
# per slab scan
PER_PLANE = 12         # two 3-dots + offset add + the t division
CONVEX_EXIT_TEST = 12  # one plane's test standing in for a scan
# per facet event (common part: hit point, plane-of-incidence frame, rotation, geometry)
FACET_COMMON = 120
REFLECT = 359          # hero Fresnel + 8 channels' Fresnel/Mueller + reflect vector
TIR = 200              # 8 channels' TIR phase / Mueller + reflect vector
REFRACT = 760          # hero Snell + 8 channels' direction/transmission/MIS (+ compat on entry)
PROBE_ENV = 45         # exit-split environment lookup (plus a slab scan)
ESCAPE = 140           # escape environment lookup over 8 channels


def film_flops(totals, n_planes):
    """FLOPs for one film call given ixg_count_events totals."""
    t = totals
    # only the slab scans the kernel runs: a ray leaving a convex stone through
    # a facet provably misses it, so those scans are resolved from one plane
    scans = t["scans"] + t["probe_scans"]
    return (t["paths"] * (PATH_SETUP + FINALIZE)
            + scans * PER_PLANE * n_planes
            + (t["reflect"] + t["tir"] + t["refract"]) * FACET_COMMON
            + t["reflect"] * REFLECT + t["tir"] * TIR + t["refract"] * REFRACT
            + t["probe"] * PROBE_ENV
            + t["escaped"] * ESCAPE
            + (t["bounces"] - t["scans"] + t["probe"] - t["probe_scans"]) * CONVEX_EXIT_TEST)
