"""Per-pixel z-score image equivalence, restating the reference's criteria
(crates/indicatrix/src/renderer/gpu/z_stats.rs): Welford moments per pixel,
z = (mean_b - mean_a) / sqrt(se_a^2 + se_b^2), at most max(5 x 0.27 %, 10)
pixels beyond 3 sigma and no 4-connected cluster of them above 6 pixels."""
import numpy as np

SIGMA_THRESHOLD = 3.0
NULL_OVER_3_SIGMA_RATE = 0.0027
MAX_CLUSTER_PIXELS = 6


def moments(samples):
    """(mean, standard error squared) per pixel of an (n_pixels, spp) array."""
    s = np.asarray(samples, np.float64)
    n = s.shape[1]
    mean = s.mean(axis=1)
    var = s.var(axis=1, ddof=1) if n > 1 else np.zeros_like(mean)
    return mean, var / n


def z_scores(a, b):
    ma, sa = moments(a)
    mb, sb = moments(b)
    se2 = sa + sb
    z = np.zeros_like(ma)
    ok = se2 > 0.0
    z[ok] = (mb[ok] - ma[ok]) / np.sqrt(se2[ok])
    return z


def cluster_sizes(z, width, height, threshold=SIGMA_THRESHOLD):
    """4-connected components of |z| > threshold, largest first."""
    over = np.abs(z.reshape(height, width)) > threshold
    seen = np.zeros_like(over)
    sizes = []
    for y0 in range(height):
        for x0 in range(width):
            if not over[y0, x0] or seen[y0, x0]:
                continue
            stack = [(y0, x0)]
            seen[y0, x0] = True
            size = 0
            while stack:
                y, x = stack.pop()
                size += 1
                for ny, nx in ((y, x - 1), (y, x + 1), (y - 1, x), (y + 1, x)):
                    if 0 <= ny < height and 0 <= nx < width and over[ny, nx] and not seen[ny, nx]:
                        seen[ny, nx] = True
                        stack.append((ny, nx))
            sizes.append(size)
    return sorted(sizes, reverse=True)


def passes(z, width, height):
    over = int((np.abs(z) > SIGMA_THRESHOLD).sum())
    allowed = max(NULL_OVER_3_SIGMA_RATE * z.size * 5.0, 10.0)
    sizes = cluster_sizes(z, width, height)
    return over <= allowed and (sizes[0] if sizes else 0) <= MAX_CLUSTER_PIXELS, over, sizes[:3]


def image_equivalent(xyz_a, xyz_b, width, height, spp):
    """Both (n_pixels*spp, 3) per-path XYZ in pixel-major order; every channel
    must pass."""
    results = []
    for ch in range(3):
        a = np.asarray(xyz_a)[:, ch].reshape(width * height, spp)
        b = np.asarray(xyz_b)[:, ch].reshape(width * height, spp)
        results.append(passes(z_scores(a, b), width, height))
    return all(r[0] for r in results), results
