"""ULP distance between f32 results, as the reference's GPU checks define it
(renderer/gpu/ulp.rs:15-67): the bit patterns mapped onto one monotonic
integer line (negative floats reflected through INT32_MIN, so +0 and -0
coincide and a pair straddling zero counts every step between them), and the
hybrid tolerance that also accepts an absolute difference under a floor
(within_tolerance, ulp.rs:42-44).  Vectorised over numpy arrays."""
import numpy as np


def to_ordered(x):
    """ulp.rs:59-67 to_ordered: i32 bits, negatives as INT32_MIN - bits."""
    b = np.asarray(x, np.float32).view(np.int32).astype(np.int64)
    return np.where(b < 0, np.int64(-(2 ** 31)) - b, b)


def ulp_distance(a, b):
    """ulp.rs:15-26: |ordered(a) - ordered(b)|, saturated at u32::MAX."""
    d = np.abs(to_ordered(a) - to_ordered(b))
    return np.minimum(d, np.int64(2 ** 32 - 1))


def within_tolerance(cpu, gpu, budget, abs_floor):
    """ulp.rs:42-44."""
    cpu = np.asarray(cpu, np.float32)
    gpu = np.asarray(gpu, np.float32)
    return (ulp_distance(cpu, gpu) <= budget) | (np.abs(cpu.astype(np.float64) - gpu) < abs_floor)


def summarize(cpu, gpu, budget, abs_floor):
    """The reference harness's per-check record (gpu_harness_last_run.json):
    comparisons; max raw ULP; the count beyond the budget that the absolute
    floor exempts; max_genuine_ulp, the largest distance among comparisons
    that neither the budget nor the floor accepts (0 when all pass); and
    max_ulp_above_floor, the largest distance among comparisons whose
    absolute difference reaches the floor."""
    cpu = np.asarray(cpu, np.float32).ravel()
    gpu = np.asarray(gpu, np.float32).ravel()
    d = ulp_distance(cpu, gpu)
    floor_ok = np.abs(cpu.astype(np.float64) - gpu) < abs_floor
    fail = (d > budget) & ~floor_ok
    above = d[~floor_ok]
    return dict(comparisons=int(d.size), max_raw_ulp=int(d.max()) if d.size else 0,
                exempted=int(((d > budget) & floor_ok).sum()),
                max_genuine_ulp=int(d[fail].max()) if fail.any() else 0,
                max_ulp_above_floor=int(above.max()) if above.size else 0,
                budget=budget, passed=not bool(fail.any()))
