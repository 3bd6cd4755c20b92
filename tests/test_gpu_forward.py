"""GPU parity of the forward kernels against the CPU oracle.

Tolerances (FP32), stated per comparison:
  * uniform random inputs (outputs bounded away from 0): max relative error
    |a-b|/max(|b|,1e-6) <= 1e-5 against both the oracle's FP32 path and FP64.
  * phantom workloads (background 0.05 + noise, so a few denoised voxels sit
    near 0 and the plain relative error is ill-conditioned there):
    nlm_metrics.phantom_parity -- unfloored relative error <= 1e-5 against
    the CPU FP32 path and FP64 on every voxel with |y| >= 0.5% of the output
    range, and at every voxel |gpu - fp64| <= 32 * 2^-24 * A/Z, the FP32
    rounding unit of a weighted mean of terms of size A/Z (A = sum_d w_d
    |u~(x+d)|, Z = sum_d w_d, both from the FP64 oracle).
  * 3D direct-sum tiles vs the generic path, padded vs reflect input, r_p = 4
    rows vs big tiles: bit-identical; the 2D tile vs the generic path:
    bit-identical on integer-valued input, within 2e-6 otherwise (tree vs
    left-to-right patch sums).
"""
import numpy as np
import pytest
import torch

import gnlm
import oracle
from nlm_metrics import max_rel, max_rel_fl, phantom_parity, report

pytestmark = pytest.mark.gpu

TOL = 1e-5       # vs the CPU FP32 path (the parity bar)
TOL_FP64 = 1e-4  # vs FP64 where FP32 rounding of 125-term patch sums dominates


def _run3(dev, u, rp, rs, h, path=gnlm.PATH_AUTO):
    t = torch.from_numpy(np.ascontiguousarray(u, np.float32)).to(dev)
    y = gnlm.nlm3d(t, rp, rs, h, path=path)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _run2(dev, u, rp, rs, h, path=gnlm.PATH_AUTO):
    t = torch.from_numpy(np.ascontiguousarray(u, np.float32)).to(dev)
    y = gnlm.nlm2d(t, rp, rs, h, path=path)
    torch.cuda.synchronize()
    return y.cpu().numpy()


CASES3 = [
    # (nz, ny, nx), rp, rs, h
    ((16, 16, 16), 1, 2, 0.3),
    ((17, 19, 23), 2, 3, 0.25),
    ((9, 33, 40), 1, 5, 0.2),
    ((5, 6, 7), 1, 3, 0.3),       # dims smaller than the halo: repeated reflection
    ((12, 12, 12), 0, 2, 0.3),    # 1-voxel patch
    ((12, 13, 14), 2, 0, 0.3),    # search window = self only
    ((10, 18, 21), 3, 2, 0.4),
    ((1, 20, 20), 1, 2, 0.3),     # single plane
]


SWEEP_DIMS = (1, 2, 3, 5, 7, 13, 17, 29, 31)


def _sweep3(n=32, seed=20261204):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(n):
        shape = tuple(int(v) for v in rng.choice(SWEEP_DIMS, 3))
        cases.append((shape, int(rng.integers(0, 5)), int(rng.integers(0, 8)), float(rng.uniform(0.1, 0.4))))
    return cases


@pytest.mark.parametrize("shape,rp,rs,h", _sweep3())
def test_3d_sweep(dev, shape, rp, rs, h):
    """Seeded shapes from 1 to 31 per axis (thin volumes, partial tiles on
    every axis, repeated reflection) and radii r_p 0-4, r_s 0-7 on the default
    path: within 1e-5 of the CPU FP32 path and of FP64."""
    rng = np.random.default_rng(sum(shape) * 13 + 5 * rp + rs)
    u = rng.random(shape, dtype=np.float32)
    print(f"case 3d sweep shape={shape} rp={rp} rs={rs} h={h:.3f}", flush=True)
    y = _run3(dev, u, rp, rs, h)
    ref32 = oracle.nlm3d_separable_f32(u, rp, rs, h)
    ref64 = oracle.nlm3d_separable_f64(u.astype(np.float64), rp, rs, h)
    assert max_rel(y, ref32) <= TOL, report(y, ref32)
    assert max_rel(y, ref64) <= TOL, report(y, ref64)


@pytest.mark.parametrize("shape,rp,rs,h", CASES3)
def test_3d_matches_oracle(dev, shape, rp, rs, h):
    rng = np.random.default_rng(hash((shape, rp, rs)) & 0xFFFF)
    u = rng.random(shape, dtype=np.float32)
    print(f"case 3d shape={shape} rp={rp} rs={rs} h={h}", flush=True)
    y = _run3(dev, u, rp, rs, h)
    ref32 = oracle.nlm3d_brute(u, rp, rs, h)
    ref64 = oracle.nlm3d_brute(u.astype(np.float64), rp, rs, h)
    assert max_rel(y, ref32) <= TOL, report(y, ref32)
    assert max_rel(y, ref64) <= TOL, report(y, ref64)


@pytest.mark.parametrize("shape,rp,rs,h", CASES3[:4] + CASES3[6:7])
def test_3d_direct_variants_equal_generic(dev, shape, rp, rs, h):
    """The direct-sum tile variants keep the canonical summation order: they
    match the generic path (and each other) bit for bit."""
    rng = np.random.default_rng(7)
    u = rng.random(shape, dtype=np.float32)
    b = _run3(dev, u, rp, rs, h, gnlm.PATH_GENERIC)
    try:
        for v in (gnlm.VARIANT_16x16x16, gnlm.VARIANT_16x16x8):
            gnlm.set_variant(rp, rs, v)
            a = _run3(dev, u, rp, rs, h, gnlm.PATH_TILED)
            assert np.array_equal(a, b), (v, report(a, b))
    finally:
        gnlm.reset_tuning()


SLAB_CASES = [
    ((28, 24, 32), 2, 5, 0.1),    # exactly 2x2x1 slab tiles for r_p = 2
    ((37, 29, 33), 2, 5, 0.2),    # partial tiles on every axis
    ((30, 14, 16), 1, 3, 0.3),
    ((26, 20, 17), 3, 2, 0.4),
    ((9, 8, 7), 1, 4, 0.3),       # smaller than one tile, reflection folds
    ((32, 16, 16), 0, 3, 0.3),
]


@pytest.mark.parametrize("variant", ["slab", "slab2"])
@pytest.mark.parametrize("shape,rp,rs,h", SLAB_CASES)
def test_3d_slab_variant_matches_oracle(dev, shape, rp, rs, h, variant):
    rng = np.random.default_rng(sum(shape) + 31 * rp + rs)
    u = rng.random(shape, dtype=np.float32)
    print(f"case {variant} shape={shape} rp={rp} rs={rs}", flush=True)
    v = gnlm.VARIANT_SLAB if variant == "slab" else gnlm.VARIANT_SLAB2
    try:
        gnlm.set_variant(rp, rs, v)
        y = _run3(dev, u, rp, rs, h, gnlm.PATH_TILED)
    finally:
        gnlm.reset_tuning()
    ref32 = oracle.nlm3d_brute(u, rp, rs, h)
    ref64 = oracle.nlm3d_brute(u.astype(np.float64), rp, rs, h)
    print("slab vs fp32", report(y, ref32), "vs fp64", report(y, ref64))
    assert max_rel(y, ref32) <= TOL, report(y, ref32)
    assert max_rel(y, ref64) <= TOL_FP64, report(y, ref64)


@pytest.mark.parametrize("shape,rp,rs,h", SLAB_CASES)
def test_3d_slab2_equals_slab(dev, shape, rp, rs, h):
    """The wide-access slab kernel keeps the slab kernel's arithmetic order."""
    rng = np.random.default_rng(5)
    u = rng.random(shape, dtype=np.float32)
    outs = []
    try:
        for v in (gnlm.VARIANT_SLAB, gnlm.VARIANT_SLAB2):
            gnlm.set_variant(rp, rs, v)
            outs.append(_run3(dev, u, rp, rs, h, gnlm.PATH_TILED))
    finally:
        gnlm.reset_tuning()
    assert np.array_equal(outs[0], outs[1]), report(outs[1], outs[0])


ROWS_CASES = [
    ((28, 24, 32), 2, 5, 0.1),    # whole 16 x 12 x 28 tiles
    ((37, 29, 33), 2, 5, 0.2),    # partial tiles on every axis
    ((17, 19, 23), 2, 3, 0.25),
    ((12, 12, 12), 2, 6, 0.3),
    ((13, 14, 15), 2, 7, 0.3),    # largest instantiated r_s for r_p = 2 (40-sample rows)
    ((30, 14, 16), 1, 3, 0.3),
    ((20, 20, 20), 1, 7, 0.3),
    ((18, 17, 21), 1, 8, 0.3),    # largest instantiated r_s for r_p = 1
    ((16, 16, 16), 1, 1, 0.3),
    ((9, 8, 7), 1, 4, 0.3),       # smaller than one tile, reflection folds
    ((5, 6, 7), 1, 3, 0.3),       # dims smaller than the halo
    ((1, 20, 20), 1, 2, 0.3),     # single plane
    ((70000, 1, 1), 1, 1, 0.3),   # more padded planes than one grid dimension holds
    ((37, 29, 33), 3, 4, 0.2),    # r_p = 3: 16 x 10 x 10 tiles, partial on every axis
    ((20, 20, 20), 3, 5, 0.3),
    ((17, 15, 13), 3, 6, 0.3),
    ((14, 15, 13), 3, 7, 0.3),    # largest instantiated r_s for r_p = 3
    ((9, 8, 7), 3, 1, 0.3),       # smaller than one tile
    ((20, 21, 22), 4, 3, 0.3),    # r_p = 4: 16 x 8 x 8 tiles
    ((9, 8, 7), 4, 2, 0.3),
]
# short 16 x 12 x 12 tiles (r_p = 2 only)
ROWS_SHORT_CASES = [
    ((28, 24, 24), 2, 5, 0.1),    # whole tiles
    ((37, 29, 33), 2, 5, 0.2),    # partial tiles on every axis
    ((17, 19, 23), 2, 1, 0.25),
    ((12, 12, 12), 2, 6, 0.3),
    ((19, 14, 17), 2, 7, 0.3),    # largest instantiated r_s (40-sample rows)
    ((9, 8, 7), 2, 3, 0.3),       # smaller than one tile
]


@pytest.mark.parametrize("shape,rp,rs,h,variant",
                         [c + ("rows",) for c in ROWS_CASES] + [c + ("short",) for c in ROWS_SHORT_CASES])
def test_3d_rows_variant_matches_oracle(dev, shape, rp, rs, h, variant):
    """Padded-copy "rows" variants: pairwise-tree box sums, so they agree with
    the oracle's left-to-right order to rounding (<= 1e-5 relative, uniform
    random input, against the FP32 path and FP64)."""
    rng = np.random.default_rng(sum(shape) + 17 * rp + rs)
    u = rng.random(shape, dtype=np.float32)
    print(f"case {variant} shape={shape} rp={rp} rs={rs}", flush=True)
    try:
        gnlm.set_variant(rp, rs, gnlm.VARIANT_ROWS if variant == "rows" else gnlm.VARIANT_ROWS_SHORT)
        y = _run3(dev, u, rp, rs, h, gnlm.PATH_TILED)
    finally:
        gnlm.reset_tuning()
    ref32 = oracle.nlm3d_brute(u, rp, rs, h)
    ref64 = oracle.nlm3d_brute(u.astype(np.float64), rp, rs, h)
    print("rows vs fp32", report(y, ref32), "vs fp64", report(y, ref64))
    assert max_rel(y, ref32) <= TOL, report(y, ref32)
    assert max_rel(y, ref64) <= TOL, report(y, ref64)


# large patches (variant 6): 16 x (32 - 2 r_p)^2 tiles on 1024 threads
ROWS_BIG_CASES = [
    ((24, 22, 21), 8, 6, 0.3),    # the reference's default NS = 17, SS = 13
    ((40, 36, 33), 4, 3, 0.3),    # several tiles per axis, partial tiles
    ((33, 35, 21), 6, 5, 0.3),    # NS = 13, SS = 11 (the reference's sweep)
    ((17, 19, 23), 5, 7, 0.25),   # largest instantiated r_s
    ((18, 16, 19), 7, 2, 0.3),
    ((10, 9, 8), 8, 1, 0.3),      # smaller than one tile: reflection folds more than once
]


@pytest.mark.parametrize("shape,rp,rs,h", ROWS_BIG_CASES)
def test_3d_rows_big_variant_matches_oracle(dev, shape, rp, rs, h):
    """Big-tile rows variant (r_p 4-8): shared-pair trees for patch widths up
    to 17; <= 1e-5 relative against the separable FP64 oracle and the CPU
    FP32 path (brute force is too slow at 17^3 patches)."""
    rng = np.random.default_rng(sum(shape) + 31 * rp + rs)
    u = rng.random(shape, dtype=np.float32)
    print(f"case rows-big shape={shape} rp={rp} rs={rs}", flush=True)
    try:
        gnlm.set_variant(rp, rs, gnlm.VARIANT_ROWS_BIG)
        y = _run3(dev, u, rp, rs, h, gnlm.PATH_TILED)
    finally:
        gnlm.reset_tuning()
    ref64 = oracle.nlm3d_separable_f64(u.astype(np.float64), rp, rs, h)
    ref32 = oracle.nlm3d_separable_f32(u, rp, rs, h)
    print("rows-big vs fp32", report(y, ref32), "vs fp64", report(y, ref64))
    assert max_rel(y, ref64) <= TOL, report(y, ref64)
    assert max_rel(y, ref32) <= TOL, report(y, ref32)


def test_3d_rows_big_padded_equals_reflect(dev):
    """A caller-padded input (the drop-in layout) gives the same bits as the
    reflect-on-read input through the big-tile variant."""
    rng = np.random.default_rng(12)
    u = rng.random((21, 26, 35), dtype=np.float32)
    rp, rs, h = 6, 5, 0.3
    R = rp + rs
    up = np.pad(u, R, mode="reflect")
    try:
        gnlm.set_variant(rp, rs, gnlm.VARIANT_ROWS_BIG)
        a = _run3(dev, u, rp, rs, h, gnlm.PATH_TILED)
        t = torch.from_numpy(np.ascontiguousarray(up)).to(dev)
        nz, ny, nx = u.shape
        b = gnlm.nlm3d_padded(t, R, nx, ny, nz, rp, rs, h, gnlm.PATH_TILED).cpu().numpy()
    finally:
        gnlm.reset_tuning()
    assert np.array_equal(a, b), report(b, a)


@pytest.mark.parametrize("rs", [2, 5, 7])
def test_3d_rp4_rows_equal_big_tiles(dev, rs):
    """r_p = 4 runs on rows tiles (16 x 8 x 8, the default) or big tiles
    (16 x 24 x 24).  Their patch-sum trees pair outputs the same way (every
    run starts at an even index), so the two agree bit for bit."""
    rng = np.random.default_rng(40 + rs)
    u = rng.random((27, 30, 35), dtype=np.float32)
    outs = []
    try:
        for v in (gnlm.VARIANT_ROWS, gnlm.VARIANT_ROWS_BIG):
            gnlm.set_variant(4, rs, v)
            outs.append(_run3(dev, u, 4, rs, 0.3, gnlm.PATH_TILED))
    finally:
        gnlm.reset_tuning()
    assert gnlm.tuned_variant(4, rs) == gnlm.VARIANT_ROWS
    assert np.array_equal(outs[0], outs[1]), report(outs[0], outs[1])


def test_3d_large_patch_default_path_is_tiled(dev):
    """The reference's default geometry (SS = 13, NS = 17) takes the tiled
    path by default and agrees with the any-radius generic path."""
    rng = np.random.default_rng(4)
    u = rng.random((20, 21, 22), dtype=np.float32)
    a = _run3(dev, u, 8, 6, 0.3)
    b = _run3(dev, u, 8, 6, 0.3, gnlm.PATH_GENERIC)
    assert gnlm.tuned_variant(8, 6) == gnlm.VARIANT_ROWS_BIG
    assert max_rel(a, b) <= TOL, report(a, b)


def test_3d_rows_padded_equals_reflect(dev):
    """A caller-padded input (the drop-in layout) gives the same bits as the
    reflect-on-read input through the rows variant."""
    rng = np.random.default_rng(11)
    u = rng.random((19, 26, 35), dtype=np.float32)
    rp, rs, h = 2, 4, 0.2
    R = rp + rs
    up = np.pad(u, R, mode="reflect")
    try:
        gnlm.set_variant(rp, rs, gnlm.VARIANT_ROWS)
        a = _run3(dev, u, rp, rs, h, gnlm.PATH_TILED)
        t = torch.from_numpy(np.ascontiguousarray(up)).to(dev)
        nz, ny, nx = u.shape
        b = gnlm.nlm3d_padded(t, R, nx, ny, nz, rp, rs, h, gnlm.PATH_TILED).cpu().numpy()
    finally:
        gnlm.reset_tuning()
    assert np.array_equal(a, b), report(b, a)


def test_3d_large_patch_generic(dev):
    # the reference's default geometry class (NS = 17 -> r_p = 8), small volume
    u, _ = gnlm.phantom3d(20, 18, 12, n_ellipsoids=3, axis_lo=3, axis_hi=7, sigma=0.05, seed=3)
    y = _run3(dev, u, 8, 2, 0.2)
    ref32 = oracle.nlm3d_separable_f32(u, 8, 2, 0.2)
    ref64, Z64, A64 = oracle.nlm3d_separable_f64_za(u, 8, 2, 0.2)
    rep = phantom_parity(y, ref32, ref64, Z64, A64, TOL)
    assert rep["ok"], rep


def test_3d_config1(dev):
    """96x96x48 ellipsoid phantom, patch 3^3, search 7^3, h = 0.05."""
    u, _ = gnlm.phantom3d(96, 96, 48, n_ellipsoids=12, axis_lo=4, axis_hi=20, background=0.05,
                          sigma=0.05, seed=20261204)
    y = _run3(dev, u, 1, 3, 0.05)
    ref32 = oracle.nlm3d_separable_f32(u, 1, 3, 0.05)
    ref64, Z64, A64 = oracle.nlm3d_separable_f64_za(u, 1, 3, 0.05)
    rep = phantom_parity(y, ref32, ref64, Z64, A64, TOL)
    print("cfg1", rep)
    assert rep["ok"], rep


def test_3d_rows_rp1_at_128_cubed(dev):
    """r_p = 1 at 128^3: 800 tiles of 16 x 14 x 14, more than 3 per CU fit at
    once on a 256-CU device, so the launch takes the 128-VGPR build (4 tiles
    per CU, one round).  Slabs at both z ends (the last tile layer is partial)
    and in the middle against the CPU FP32 path and FP64 (<= 1e-5 relative)."""
    rng = np.random.default_rng(128)
    u = rng.random((128, 128, 128), dtype=np.float32)
    rp, rs, h = 1, 4, 0.25
    try:
        gnlm.set_variant(rp, rs, gnlm.VARIANT_ROWS)
        y = _run3(dev, u, rp, rs, h, gnlm.PATH_TILED)
    finally:
        gnlm.reset_tuning()
    for z0 in (0, 61, 125):
        ref32 = oracle.nlm3d_separable_f32(u, rp, rs, h, z0=z0, z1=z0 + 3)
        ref64 = oracle.nlm3d_separable_f64(u, rp, rs, h, z0=z0, z1=z0 + 3)
        print("rp1 128^3 slab", z0, "vs fp32", report(y[z0:z0 + 3], ref32), "vs fp64", report(y[z0:z0 + 3], ref64))
        assert max_rel(y[z0:z0 + 3], ref32) <= TOL
        assert max_rel(y[z0:z0 + 3], ref64) <= TOL


def test_3d_padded_equals_reflect(dev):
    rng = np.random.default_rng(11)
    u = rng.random((13, 17, 19), dtype=np.float32)
    rp, rs = 2, 3
    R = rp + rs
    up = np.pad(u, R, mode="reflect")  # numpy 'reflect' is reflect-101
    a = _run3(dev, u, rp, rs, 0.3)
    t = torch.from_numpy(up).to(dev)
    b = gnlm.nlm3d_padded(t, R, 19, 17, 13, rp, rs, 0.3).cpu().numpy()
    assert np.array_equal(a, b)


def test_3d_deterministic(dev):
    u, _ = gnlm.phantom3d(40, 36, 34, n_ellipsoids=5, axis_lo=3, axis_hi=9, sigma=0.05, seed=5)
    a = _run3(dev, u, 2, 5, 0.05)
    b = _run3(dev, u, 2, 5, 0.05)
    assert np.array_equal(a, b)


@pytest.mark.slow
def test_3d_config3_slabs(dev):
    """256^3, patch 5^3, search 11^3: full GPU run, slabs checked against the
    CPU FP32 path and FP64; whole-volume properties checked everywhere."""
    u, _ = gnlm.phantom3d(256, 256, 256, n_ellipsoids=64, axis_lo=5, axis_hi=40, background=0.05,
                          sigma=0.05, seed=20261206)
    y = _run3(dev, u, 2, 5, 0.05)
    assert np.isfinite(y).all()
    assert y.min() >= u.min() - 1e-6 and y.max() <= u.max() + 1e-6  # convex combination
    for z0 in (0, 128, 253):
        ref32 = oracle.nlm3d_separable_f32(u, 2, 5, 0.05, z0=z0, z1=z0 + 3)
        ref64, Z64, A64 = oracle.nlm3d_separable_f64_za(u, 2, 5, 0.05, z0=z0, z1=z0 + 3)
        rep = phantom_parity(y[z0:z0 + 3], ref32, ref64, Z64, A64, TOL)
        print("cfg3 slab", z0, rep)
        assert rep["ok"], rep


CASES2 = [
    ((64, 64), 1, 2, 0.2),
    ((37, 101), 2, 5, 0.1),
    ((70, 130), 5, 10, 0.3),   # the reference's own 2D geometry (nb = 11, sr = 21)
    ((5, 9), 2, 3, 0.3),
    ((40, 50), 0, 3, 0.2),
]


def _sweep2(n=24, seed=20261205):
    rng = np.random.default_rng(seed)
    dims = (1, 2, 3, 5, 17, 31, 64, 97)
    cases = []
    for _ in range(n):
        shape = tuple(int(v) for v in rng.choice(dims, 2))
        cases.append((shape, int(rng.integers(0, 7)), int(rng.integers(0, 12)), float(rng.uniform(0.1, 0.4))))
    return cases


@pytest.mark.parametrize("shape,rp,rs,h", _sweep2())
def test_2d_sweep(dev, shape, rp, rs, h):
    """Seeded 2D shapes from 1 to 97 per axis and radii r_p 0-6, r_s 0-11
    (the tiled kernels, their register-window builds and the generic path):
    within 1e-5 of the CPU FP32 path and of FP64."""
    rng = np.random.default_rng(sum(shape) * 17 + 3 * rp + rs)
    u = rng.random(shape, dtype=np.float32)
    print(f"case 2d sweep shape={shape} rp={rp} rs={rs} h={h:.3f}", flush=True)
    y = _run2(dev, u, rp, rs, h)
    ref32 = oracle.nlm2d_separable_f32(u, rp, rs, h)
    ref64 = oracle.nlm2d_separable_f64(u.astype(np.float64), rp, rs, h)[0]
    assert max_rel(y, ref32) <= TOL, report(y, ref32)
    assert max_rel(y, ref64) <= TOL, report(y, ref64)


@pytest.mark.parametrize("shape,rp,rs,h", CASES2)
def test_2d_matches_oracle(dev, shape, rp, rs, h):
    rng = np.random.default_rng(sum(shape) + rp + rs)
    u = rng.random(shape, dtype=np.float32)
    print(f"case 2d shape={shape} rp={rp} rs={rs} h={h}", flush=True)
    y = _run2(dev, u, rp, rs, h)
    ref32 = oracle.nlm2d_brute(u, rp, rs, h)
    ref64 = oracle.nlm2d_brute(u.astype(np.float64), rp, rs, h)
    assert max_rel(y, ref32) <= TOL, report(y, ref32)
    assert max_rel(y, ref64) <= TOL, report(y, ref64)


@pytest.mark.parametrize("shape,rp,rs,h", CASES2[:3])
def test_2d_tiled_equals_generic(dev, shape, rp, rs, h):
    """The tiled kernel sums patches in trees, the generic path left to right:
    on integer-valued input (the reference's uint8 images, nb <= 11) every
    patch distance is an exact FP32 integer in both, so the results are
    identical bit for bit; on general input they agree to rounding."""
    rng = np.random.default_rng(3)
    u8 = rng.integers(0, 256, shape).astype(np.float32)
    a = _run2(dev, u8, rp, rs, 255.0 * h, gnlm.PATH_TILED)
    b = _run2(dev, u8, rp, rs, 255.0 * h, gnlm.PATH_GENERIC)
    assert np.array_equal(a, b), report(a, b)
    u = rng.random(shape, dtype=np.float32)
    a = _run2(dev, u, rp, rs, h, gnlm.PATH_TILED)
    b = _run2(dev, u, rp, rs, h, gnlm.PATH_GENERIC)
    assert max_rel(a, b) <= 2e-6, report(a, b)


def test_2d_config2(dev):
    """2048^2 phantom, patch 5x5, search 11x11, h = 0.08."""
    u, _ = gnlm.phantom2d(2048, 2048, sigma=0.08, seed=20261205)
    y = _run2(dev, u, 2, 5, 0.08)
    ref32 = oracle.nlm2d_separable_f32(u, 2, 5, 0.08)
    ref64, Z64, A64 = oracle.nlm2d_separable_f64(u, 2, 5, 0.08)
    rep = phantom_parity(y, ref32, ref64, Z64, A64, TOL)
    print("cfg2", rep)
    assert rep["ok"], rep


def test_z_invariant_3d_equals_2d(dev):
    """A volume constant along z: 3D NLM (NS, SS, h) equals 2D NLM (NS, SS, h),
    since S_3D = NS * S_2D and P_3D = NS * P_2D (the reference's own driver
    builds exactly such volumes, 3DNLM/nlm.cpp:55-61)."""
    rng = np.random.default_rng(9)
    img = rng.random((30, 34), dtype=np.float32)
    vol = np.repeat(img[None], 9, axis=0)
    y3 = _run3(dev, vol, 1, 2, 0.3)
    y2 = _run2(dev, img, 1, 2, 0.3)
    for z in range(9):
        assert max_rel(y3[z], y2) <= 2e-6


def test_dropin_3d_u8(dev):
    """The reference's drop-in input: padded uint8, h = PARAMETER = 10."""
    rng = np.random.default_rng(4)
    X, Y, D, SS, NS = 20, 18, 10, 5, 3
    pad = (SS + NS - 2) // 2
    img = rng.integers(0, 256, (D, Y, X), dtype=np.uint8)
    padded = np.pad(img, pad, mode="reflect")
    out = gnlm.nl_means_3d_u8(padded, X, Y, D, SS, NS, 10.0)
    ref = oracle.nlm3d_brute(img.astype(np.float64), (NS - 1) // 2, (SS - 1) // 2, 10.0)
    assert max_rel(out, ref) <= TOL, report(out, ref)


def test_dropin_2d_u8(dev):
    rng = np.random.default_rng(5)
    n, sr, nb = 48, 7, 5
    pad = (sr + nb - 2) // 2
    img = rng.integers(0, 256, (n, n), dtype=np.uint8)
    padded = np.pad(img, pad, mode="reflect")
    W = H = n + 2 * pad
    out = gnlm.nl_means_2d_u8(padded, W, H, sr, nb, n, n, 10.0)
    ref = oracle.nlm2d_brute(img.astype(np.float64), (nb - 1) // 2, (sr - 1) // 2, 10.0)
    assert max_rel(out, ref) <= TOL, report(out, ref)
