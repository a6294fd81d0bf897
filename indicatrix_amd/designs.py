"""Seeded convex gem designs for the batched metric sweep and the optimiser.

The reference builds facet planes from a GemCAD design through its meet
solver (IX/geometry/meet_solver/, crates/indicatrix-cut-core); this module
replaces that pipeline for synthetic workloads with a direct parametric
construction of the same convention (unit outward normals, n.x + d <= 0
inside, origin interior, girdle radius ~1):

  table (y = +h_t) | crown tiers | girdle band (g vertical facets, radius 1)
  | pavilion tiers | culet (y = -h_c)

Each tier is m facets equally spaced in azimuth (with a phase), all at one
angle from the girdle plane; a tier's offset is set so that it meets the
girdle radius at the tier's own height, which keeps every tier active.
"""
import math

import numpy as np

from . import abi

TILTS_C3 = np.arange(-30.0, 30.0 + 1e-6, 3.0, dtype=np.float32)  # 21 tilts


def _tier(m, angle_deg, phase, height, sign):
    a = math.radians(angle_deg)
    rows = []
    for i in range(m):
        phi = phase + 2.0 * math.pi * i / m
        n = np.array([math.sin(a) * math.cos(phi), sign * math.cos(a), math.sin(a) * math.sin(phi)])
        d = -(math.sin(a) + math.cos(a) * height)   # passes through (r=1, y=sign*height)
        rows.append([n[0], n[1], n[2], d])
    return rows


def design_planes(params):
    """Planes (N,4) float32 from a parameter dict (see seeded_design)."""
    rows = [[0.0, 1.0, 0.0, -params["table_h"]]]
    for (m, ang, ph, h) in params["crown"]:
        rows += _tier(m, ang, ph, h, +1.0)
    g = params["girdle"]
    for i in range(g):
        phi = params["girdle_phase"] + 2.0 * math.pi * i / g
        rows.append([math.cos(phi), 0.0, math.sin(phi), -1.0])
    for (m, ang, ph, h) in params["pavilion"]:
        rows += _tier(m, ang, ph, h, -1.0)
    rows.append([0.0, -1.0, 0.0, -params["culet_h"]])
    arr = np.array(rows, np.float64)
    arr[:, :3] /= np.linalg.norm(arr[:, :3], axis=1, keepdims=True)
    aspect = params.get("aspect", 1.0)
    if aspect != 1.0:
        # an oval: stretch x by `aspect` (planes transform by the inverse transpose)
        n = arr[:, :3] * np.array([1.0 / aspect, 1.0, 1.0])
        s = np.linalg.norm(n, axis=1, keepdims=True)
        arr = np.concatenate([n / s, arr[:, 3:] / s], axis=1)
    return arr.astype(np.float32)


def seeded_params(seed, n_facets=None, angle_range=(30.0, 50.0), aspect=1.0):
    """A design of exactly n_facets planes (uniform 32..256 when None)."""
    rng = np.random.default_rng(seed)
    n = int(n_facets if n_facets is not None else rng.integers(32, 257))
    girdle = int(min(max(8, 4 * (n // 16)), 32))
    girdle = girdle - girdle % 4
    rest = n - 2 - girdle
    crown_total = rest // 2
    pav_total = rest - crown_total

    def tiers(total, lo_h, hi_h, sign):
        out = []
        left = total
        heights = []
        while left > 0:
            m = int(min(left, rng.choice([8, 8, 16, 16, 12, 24]))) if left >= 8 else left
            if left - m < 3 and left - m > 0:
                m = left
            out.append(m)
            left -= m
        heights = np.sort(rng.uniform(lo_h, hi_h, len(out)))
        res = []
        for m, h in zip(out, heights):
            res.append((m, float(rng.uniform(*angle_range)), float(rng.uniform(0, 2 * math.pi / m)),
                        float(h)))
        return res

    return dict(n=n, table_h=float(rng.uniform(0.28, 0.38)), culet_h=float(rng.uniform(0.80, 0.95)),
                girdle=girdle, girdle_phase=float(rng.uniform(0, math.pi / girdle)),
                crown=tiers(crown_total, 0.01, 0.12, +1), pavilion=tiers(pav_total, 0.005, 0.08, -1),
                aspect=aspect)


def seeded_material(seed):
    """Isotropic Sellmeier-1 material with n_D ~ U[1.54, 2.42], Abbe ~ U[20, 60]."""
    from .tracer import sellmeier1_from_nd_abbe
    rng = np.random.default_rng(10_000 + seed)
    nd = float(rng.uniform(1.54, 2.42))
    abbe = float(rng.uniform(20.0, 60.0))
    b, c = sellmeier1_from_nd_abbe(nd, abbe)
    d = abi.Dispersion()
    d.kind = abi.DISP_SELLMEIER1
    d.coeff[0], d.coeff[1] = b, c
    return d, nd, abbe


def sweep_batch(n_designs, first_seed=0):
    """The batched-sweep inputs: a shared plane table and per-design ranges."""
    tables, designs, meta = [], [], []
    off = 0
    for s in range(first_seed, first_seed + n_designs):
        p = design_planes(seeded_params(s))
        d, nd, abbe = seeded_material(s)
        tables.append(p)
        designs.append((off, len(p), d))
        meta.append(dict(seed=s, n_planes=len(p), n_d=nd, abbe=abbe))
        off += len(p)
    return np.concatenate(tables), designs, meta
