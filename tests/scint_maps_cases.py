"""Shared inputs for the scintillation-map tests: hand-built maps of the
reference's tests and seeded random maps with ragged sizes."""
import numpy as np

from indicatrix_amd import metrics


def map_from(size, brilliant):
    px = np.zeros(size * size, metrics.PIXEL_DTYPE)
    for y in range(size):
        for x in range(size):
            px[y * size + x]["hits"] = 1
            px[y * size + x]["classes"][2 if brilliant(x, y) else 1] = 1
    return px


def random_map(w, h, seed, hit_rate=0.7, pose_rate=0.6):
    """A map whose pixels have 0..64 hits split over the four classes and,
    on some hit pixels, 0..5 of 5 poses returned."""
    rng = np.random.default_rng(seed)
    n = w * h
    px = np.zeros(n, metrics.PIXEL_DTYPE)
    hits = np.where(rng.random(n) < hit_rate, rng.integers(1, 65, n), 0)
    for i in range(n):
        if hits[i]:
            cut = np.sort(rng.integers(0, hits[i] + 1, 3))
            px[i]["classes"][:] = np.diff(np.concatenate([[0], cut, [hits[i]]]))
            px[i]["hits"] = hits[i]
            if rng.random() < pose_rate:
                px[i]["scint_poses"] = 5
                px[i]["scint_returned"] = rng.integers(0, 6)
    return px


def cases():
    """(name, pixels, width, height) with the sizes the maps must handle."""
    out = [("uniform9", map_from(9, lambda x, y: True), 9, 9),
           ("checker9", map_from(9, lambda x, y: (x + y) % 2 == 0), 9, 9)]
    hole = map_from(5, lambda x, y: (x + y) % 2 == 0)
    hole[12] = np.zeros(1, metrics.PIXEL_DTYPE)[0]
    out.append(("hole5", hole, 5, 5))
    out.append(("mismatch", hole.copy(), 6, 5))
    for k, (w, h) in enumerate([(1, 1), (1, 100), (100, 1), (17, 16), (16, 17), (37, 29), (64, 64), (513, 7)]):
        out.append((f"random{w}x{h}", random_map(w, h, 100 + k), w, h))
    return out
