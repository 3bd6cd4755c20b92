"""Error metrics used by the parity tests and the benchmark report.

max_rel   : max |a - b| / max(|b|, 1e-6)         (the per-voxel relative error)
max_rel_fl: max |a - b| / max(|b|, 1e-3 * (max b - min b))  (floored by the range)

phantom_parity: the bar for phantom workloads, whose denoised background has a
few outputs near 0 where a plain relative error is ill-conditioned:
  * on voxels with |ref64| >= TAU_FRAC * (range of ref64), the unfloored
    relative error against the CPU FP32 path and against FP64 is <= tol
    (1e-5);
  * at EVERY voxel, |got - ref64| <= K_COND * 2^-24 * (A(x) + B(x)) / Z(x)
    (plus, where the caller passes it, the underflow allowance of
    underflow_bound), where Z = sum_d w_d, A = sum_d w_d |u~(x+d)| and
    B = sum_d w_d a_d (|u~(x+d)| + |y|) come from the FP64 oracle (a_d the
    exponent of w_d = exp(-a_d)): A / Z is the size of the terms whose
    weighted mean y is, so 2^-24 A / Z is the rounding unit of the sums; B / Z
    carries the error of the weights themselves, whose FP32 exponents hold a
    relative error of a few 2^-24, i.e. an absolute one of a few a_d 2^-24.
    Callers that pass no B get the A-only bar (stricter: it holds on the
    floating-point phantoms, where the dominant weights have small a_d; on
    8-bit images with h = 10 they reach a_d ~ 80 and the CPU FP32 path itself
    needs 69 units of A/Z).  K_COND = 32; the CPU FP32 path itself needs 8-18
    units of A/Z on the configuration phantoms (256^3 p5 s11, 96x96x48 p3 s7,
    2048^2 p5 s11);
  * the count of voxels whose unfloored relative error exceeds tol is
    reported (they all sit below TAU_FRAC of the range).
"""
import numpy as np

TAU_FRAC = 5e-3
K_COND = 32.0


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


def underflow_bound(pairs, umax, y64, Z64):
    """The absolute error FP32 weights below the normal range may bring: a
    weight w < 2^-126 may flush to 0 (or lose its low bits as a denormal), so
    A and Z each move by at most pairs * 2^-126 * (umax, resp. 1), and y = A/Z
    by at most pairs * 2^-126 * (umax + |y|) / Z.  It matters only where every
    candidate but the centre is far out (8-bit images with flat black regions:
    there y is ~1e-58 in FP64 and exactly 0 in FP32)."""
    return pairs * 2.0 ** -126 * (umax + np.abs(np.asarray(y64, np.float64))) / np.asarray(Z64, np.float64)


def fp32_units(err, A64, Z64, underflow=0.0, B64=None):
    """|err| in units of 2^-24 (A + B)/Z (B = 0 when not given), after the
    underflow allowance (see underflow_bound) is taken off.  Where A = 0
    (every candidate of the voxel is 0) y is exactly 0: an error of 0 there is
    0 units, any other error infinitely many."""
    err = np.maximum(np.abs(np.asarray(err, np.float64)) - underflow, 0.0)
    mag = np.asarray(A64, np.float64) + (0.0 if B64 is None else np.asarray(B64, np.float64))
    unit = 2.0 ** -24 * mag / np.asarray(Z64, np.float64)
    out = np.zeros_like(err)
    pos = unit > 0.0
    out[pos] = err[pos] / unit[pos]
    out[~pos & (err > 0.0)] = np.inf
    return out


def phantom_parity(got, ref32, ref64, Z64, A64, tol=1e-5, tol32=None, underflow=0.0, B64=None):
    """The phantom-workload bar (module notes).  tol32 (default tol) is the
    bar against the CPU FP32 path above tau; underflow the allowance of
    underflow_bound (0: none); B64 the weight-error term (None: the A-only
    bar).  Returns the report; the caller asserts report["ok"]."""
    tol32 = tol if tol32 is None else tol32
    g = np.asarray(got, np.float64)
    r32 = np.asarray(ref32, np.float64)
    r64 = np.asarray(ref64, np.float64)
    tau = TAU_FRAC * float(r64.max() - r64.min())
    big = np.abs(r64) >= tau
    rel32 = np.abs(g - r32) / np.maximum(np.abs(r32), 1e-30)
    rel64 = np.abs(g - r64) / np.maximum(np.abs(r64), 1e-30)
    cond = fp32_units(g - r64, A64, Z64, underflow, B64)
    rep = {
        "tau": tau,
        "voxels_below_tau": int((~big).sum()),
        "max_rel_vs_fp32_above_tau": float(rel32[big].max()) if big.any() else 0.0,
        "max_rel_vs_fp64_above_tau": float(rel64[big].max()) if big.any() else 0.0,
        "max_err_in_fp32_cond_units": float(cond.max()),
        "cond_terms": "A" if B64 is None else "A + B",
        "k_cond": K_COND,
        "voxels_unfloored_rel_over_tol": int((rel64 > tol).sum()),
        "max_abs_vs_fp64": float(np.abs(g - r64).max()),
    }
    rep["ok"] = bool(rep["max_rel_vs_fp32_above_tau"] <= tol32 and rep["max_rel_vs_fp64_above_tau"] <= tol and
                     rep["max_err_in_fp32_cond_units"] <= K_COND)
    return rep
