"""Error metrics used by the parity tests and the benchmark report.

max_rel   : max |a - b| / max(|b|, 1e-6)         (the per-voxel relative error)
max_rel_fl: max |a - b| / max(|b|, 1e-3 * (max b - min b))  (floored by the range)
"""
import numpy as np


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)))) if np.size(a) else 0.0


def max_rel(a, b, floor=1e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def max_rel_fl(a, b):
    b64 = np.asarray(b, np.float64)
    if b64.size == 0:
        return 0.0
    rng = float(b64.max() - b64.min())
    return max_rel(a, b64, floor=max(1e-3 * rng, 1e-12))


def report(a, b):
    return {"max_abs": max_abs(a, b), "max_rel": max_rel(a, b), "max_rel_floored": max_rel_fl(a, b)}
