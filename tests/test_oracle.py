"""Pins the CPU oracle (oracle/nlm_oracle.c) without the reference build.

The reference (CUDA + OpenCV) cannot be built here and ships no golden
outputs, so each oracle function is checked against what the mathematics
fixes, chosen so that a plausible slip (wrong normalisation P, a dropped
patch term, a wrong offset sign, a transposed axis, a wrong border rule)
fails at least one of them:
  * reflect-101 index rule        vs numpy.pad(mode="reflect") (whole-sample
                                     symmetric = OpenCV BORDER_REFLECT_101)
  * h -> infinity                 vs the box mean over the search cube on the
                                     mirror extension (scipy uniform_filter,
                                     mode="mirror")
  * h -> 0                        returns the input (only the self weight)
  * constant input                returns the constant
  * x-step volume, r_p = r_s = 1  closed form y = (a(1+q) + q b)/(1+2q),
                                     q = exp(-(b-a)^2/(3 h^2)) at the step
  * single impulse, r_p = r_s = 1 closed forms at the impulse and at its
                                     x neighbours (worked by hand below)
  * z-invariant volume            3D output equals the 2D output per plane
  * flips / axis transposes       output flips / transposes with the input
  * convex combination            min u <= y <= max u
  * FP32 paths                    separable == brute force bit for bit, and
                                     within 1e-5 of FP64
  * adjoint                       FP64 central finite differences
"""
import numpy as np
import pytest
from scipy.ndimage import uniform_filter

import oracle


def rnd(shape, seed, dtype=np.float64):
    return np.random.default_rng(seed).random(shape).astype(dtype)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_reflect101_matches_numpy_pad(n):
    base = np.arange(n)
    pad = 3 * n + 2
    ref = np.pad(base, pad, mode="reflect") if n > 1 else np.zeros(n + 2 * pad, int)
    got = [oracle.reflect101(i, n) for i in range(-pad, n + pad)]
    assert list(ref) == got


def test_constant_volume():
    u = np.full((6, 7, 8), 0.37)
    for f in (lambda a: oracle.nlm3d_brute(a, 1, 2, 0.1),
              lambda a: oracle.nlm3d_brute(a.astype(np.float32), 1, 2, 0.1),
              lambda a: oracle.nlm3d_separable_f32(a, 1, 2, 0.1),
              lambda a: oracle.nlm3d_separable_f64(a, 1, 2, 0.1)):
        y = f(u)
        assert np.allclose(y, 0.37, rtol=1e-6, atol=0)
    img = np.full((9, 11), 0.6)
    assert np.allclose(oracle.nlm2d_brute(img, 2, 2, 0.1), 0.6)
    assert np.allclose(oracle.nlm2d_separable_f32(img, 2, 2, 0.1), 0.6, rtol=1e-6)


@pytest.mark.parametrize("rp,rs", [(1, 1), (1, 2), (2, 1), (0, 3)])
def test_large_h_is_box_mean(rp, rs):
    u = rnd((7, 9, 10), 1)
    y = oracle.nlm3d_brute(u, rp, rs, 1e6)
    box = uniform_filter(u, size=2 * rs + 1, mode="mirror")
    assert np.max(np.abs(y - box)) < 1e-9
    img = rnd((12, 13), 2)
    y2 = oracle.nlm2d_brute(img, rp, rs, 1e6)
    assert np.max(np.abs(y2 - uniform_filter(img, size=2 * rs + 1, mode="mirror"))) < 1e-9


def test_small_h_returns_input():
    u = rnd((6, 7, 8), 3)
    assert np.array_equal(oracle.nlm3d_brute(u, 1, 2, 1e-9), u)
    u32 = u.astype(np.float32)
    assert np.array_equal(oracle.nlm3d_brute(u32, 1, 2, 1e-9), u32)
    assert np.array_equal(oracle.nlm3d_separable_f32(u32, 1, 2, 1e-9), u32)
    img = rnd((8, 9), 4)
    assert np.array_equal(oracle.nlm2d_brute(img, 1, 2, 1e-9), img)


def test_step_closed_form():
    a, b, h, m = 0.2, 0.7, 0.3, 6
    f = np.where(np.arange(12) < m, a, b)
    u = np.broadcast_to(f, (5, 5, 12)).copy()
    y = oracle.nlm3d_brute(u, 1, 1, h)
    q = np.exp(-(b - a) ** 2 / (3 * h * h))
    want_left = (a * (1 + q) + q * b) / (1 + 2 * q)
    want_right = (b * (1 + q) + q * a) / (1 + 2 * q)
    assert np.allclose(y[:, :, m - 1], want_left, rtol=1e-13)
    assert np.allclose(y[:, :, m], want_right, rtol=1e-13)
    assert np.allclose(y[:, :, 2:m - 1], a, rtol=1e-13)
    assert np.allclose(y[:, :, m + 1:10], b, rtol=1e-13)


def test_impulse_closed_form():
    """A single impulse c at x0 in a zero 9^3 volume, patch and search radius
    1 (P = 27), worked by hand:
      * at x0 every d != 0 has S_d = 2 c^2 (the impulse enters the patch
        difference twice, at p = 0 and p = -d), so y = c / (1 + 26 w2);
      * at x1 = x0 + e (e = +x or -x) only d = -e reaches the impulse, with
        S = 2 c^2; of the other offsets, the 17 with d.x != e.x have S =
        2 c^2 and the 9 with d.x = e.x have S = c^2 (the second copy of the
        impulse falls outside the patch), so y = c w2 / (1 + 17 w2 + 9 w1);
      * two or more voxels from the impulse (outside every search window that
        holds it), y = 0 exactly;
    with w_k = exp(-k c^2 / (27 h^2))."""
    c, h = 0.8, 0.35
    u = np.zeros((9, 9, 9))
    u[4, 4, 4] = c
    y = oracle.nlm3d_brute(u, 1, 1, h)
    w1 = np.exp(-c * c / (27 * h * h))
    w2 = np.exp(-2 * c * c / (27 * h * h))
    assert abs(y[4, 4, 4] - c / (1 + 26 * w2)) <= 1e-14
    for x in (3, 5):
        assert abs(y[4, 4, x] - c * w2 / (1 + 17 * w2 + 9 * w1)) <= 1e-14, x
    assert y[4, 4, 6] == 0.0 and y[4, 4, 2] == 0.0 and y[4, 6, 4] == 0.0 and y[6, 4, 4] == 0.0
    y32 = oracle.nlm3d_brute(u.astype(np.float32), 1, 1, h)
    assert abs(y32[4, 4, 4] - y[4, 4, 4]) <= 1e-6 * y[4, 4, 4]


def test_z_invariant_3d_equals_2d():
    img = rnd((11, 13), 5)
    vol = np.repeat(img[None], 6, axis=0)
    for rp, rs in [(1, 1), (1, 2), (2, 2)]:
        y3 = oracle.nlm3d_brute(vol, rp, rs, 0.25)
        y2 = oracle.nlm2d_brute(img, rp, rs, 0.25)
        for z in range(6):
            assert np.max(np.abs(y3[z] - y2)) < 1e-13


def test_flip_and_transpose_equivariance():
    u = rnd((6, 7, 9), 6)
    y = oracle.nlm3d_brute(u, 1, 2, 0.3)
    for ax in range(3):
        yf = oracle.nlm3d_brute(np.ascontiguousarray(np.flip(u, ax)), 1, 2, 0.3)
        assert np.max(np.abs(yf - np.flip(y, ax))) < 1e-13
    for perm in [(0, 2, 1), (2, 1, 0), (1, 0, 2)]:
        yt = oracle.nlm3d_brute(np.ascontiguousarray(u.transpose(perm)), 1, 2, 0.3)
        assert np.max(np.abs(yt - y.transpose(perm))) < 1e-13
    img = rnd((9, 12), 7)
    y2 = oracle.nlm2d_brute(img, 1, 2, 0.3)
    yt2 = oracle.nlm2d_brute(np.ascontiguousarray(img.T), 1, 2, 0.3)
    assert np.max(np.abs(yt2 - y2.T)) < 1e-13


def test_convex_combination():
    u = rnd((7, 8, 9), 8) * 3 - 1
    y = oracle.nlm3d_brute(u, 1, 2, 0.5)
    assert y.min() >= u.min() and y.max() <= u.max()


@pytest.mark.parametrize("shape,rp,rs", [((7, 9, 10), 1, 2), ((5, 6, 7), 2, 3), ((4, 11, 13), 0, 2)])
def test_fp32_paths(shape, rp, rs):
    u = rnd(shape, 9, np.float32)
    b32 = oracle.nlm3d_brute(u, rp, rs, 0.2)
    s32 = oracle.nlm3d_separable_f32(u, rp, rs, 0.2)
    assert np.array_equal(b32, s32)
    b64 = oracle.nlm3d_brute(u.astype(np.float64), rp, rs, 0.2)
    s64 = oracle.nlm3d_separable_f64(u, rp, rs, 0.2)
    assert np.max(np.abs(b64 - s64)) < 1e-13
    assert np.max(np.abs(b32 - b64) / np.abs(b64)) < 1e-5
    # partial z-range of the separable path equals the full run's planes
    part = oracle.nlm3d_separable_f32(u, rp, rs, 0.2, z0=1, z1=shape[0] - 1)
    assert np.array_equal(part, s32[1:shape[0] - 1])
    img = rnd((17, 19), 10, np.float32)
    i32 = oracle.nlm2d_separable_f32(img, rp, rs, 0.2)
    assert np.array_equal(oracle.nlm2d_brute(img, rp, rs, 0.2), i32)
    # the 2D FP32 paths against 2D FP64 (brute force and separable)
    i64 = oracle.nlm2d_brute(img.astype(np.float64), rp, rs, 0.2)
    assert np.max(np.abs(i32 - i64) / np.abs(i64)) < 1e-5
    s64, _, _ = oracle.nlm2d_separable_f64(img, rp, rs, 0.2)
    assert np.max(np.abs(s64 - i64)) < 1e-13


@pytest.mark.parametrize("dims", [(6, 7, 8), None])
def test_normaliser_and_magnitude_sums(dims):
    """Z and A = sum_d w_d |u~(x+d)| of the FP64 separable paths: Z equals the
    brute force's normaliser; for non-negative input A = y Z exactly (every
    term is its own magnitude); A is even in u (w depends on differences
    only) while y is odd; and A >= |y| Z for mixed signs."""
    if dims is None:
        u = rnd((21, 23), 12, np.float32) * 2 - 0.5
        y, Z, A = oracle.nlm2d_separable_f64(u, 1, 3, 0.3)
        yb, Zb = oracle.nlm2d_brute(u.astype(np.float64), 1, 3, 0.3, want_Z=True)
        f = lambda v: oracle.nlm2d_separable_f64(v, 1, 3, 0.3)  # noqa: E731
    else:
        u = rnd(dims, 12, np.float32) * 2 - 0.5
        y, Z, A = oracle.nlm3d_separable_f64_za(u, 1, 2, 0.3)
        yb, Zb = oracle.nlm3d_brute(u.astype(np.float64), 1, 2, 0.3, want_Z=True)
        f = lambda v: oracle.nlm3d_separable_f64_za(v, 1, 2, 0.3)  # noqa: E731
    assert np.max(np.abs(y - yb)) < 1e-13 and np.max(np.abs(Z - Zb) / Zb) < 1e-13
    assert (A >= np.abs(y) * Z * (1 - 1e-12)).all()
    yn, Zn, An = f(-u)
    assert np.max(np.abs(An - A) / A) < 1e-13 and np.max(np.abs(yn + y)) < 1e-13
    up = np.abs(u)
    yp, Zp, Ap = f(up)
    assert np.max(np.abs(Ap - yp * Zp) / Ap) < 1e-13


def _fd_loss(u, clean, rp, rs, h):
    return oracle.nlm3d_mse_f64(u, clean, rp, rs, h)


def test_adjoint_vs_finite_differences():
    rng = np.random.default_rng(11)
    shape = (5, 6, 7)
    clean = rng.random(shape)
    u = clean + 0.1 * rng.standard_normal(shape)
    rp, rs, h = 1, 2, 0.4
    y = oracle.nlm3d_brute(u, rp, rs, h)
    N = u.size
    g = 2.0 * (y - clean) / N
    gu, gh = oracle.nlm3d_grad_f64(u, g, rp, rs, h)
    eps = 1e-6
    idx = [(0, 0, 0), (2, 3, 4), (4, 5, 6), (1, 0, 6), (3, 2, 1), (0, 5, 3)]
    for i in idx:
        up = u.copy(); up[i] += eps
        um = u.copy(); um[i] -= eps
        fd = (_fd_loss(up, clean, rp, rs, h) - _fd_loss(um, clean, rp, rs, h)) / (2 * eps)
        assert abs(fd - gu[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, gu[i])
    fdh = (_fd_loss(u, clean, rp, rs, h + eps) - _fd_loss(u, clean, rp, rs, h - eps)) / (2 * eps)
    assert abs(fdh - gh) <= 1e-6 * max(1.0, abs(fdh)) + 1e-9, (fdh, gh)
