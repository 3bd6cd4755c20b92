"""The CPU-baseline sample of bench.py (bench.cpu_sample), on the CPU.

For large patches the sample is one x-y window of a plane, computed by the
CPU path on the window grown by R = r_p + r_s.  The window's outputs must be
exactly those of the CPU path run on the whole volume.  The separable path
reads only samples within R of an output, and a clipped face of the crop is
the volume's own face.
"""
import numpy as np

import bench
import oracle


def test_headline_sample_is_a_slab_of_whole_planes():
    """About 4 s of CPU work per timed run (the baseline is the median of
    three): a centred slab of full 256^2 planes; the CPU path runs on the
    whole volume's planes, so reflect-101 applies at the true faces."""
    (z0, n, y0, wy, x0, wx), crop = bench.cpu_sample(256, 256, 256, 2, 5)
    assert 64 <= n <= 256 and (y0, wy, x0, wx) == (0, 256, 0, 256)
    assert z0 == 128 - n // 2
    assert crop == (0, 256, 0, 256, 0, 256)


def test_large_patch_sample_is_a_window():
    nx, ny, nz, rp, rs = 512, 512, 128, 8, 6
    (z0, n, y0, wy, x0, wx), (cz0, cz1, cy0, cy1, cx0, cx1) = bench.cpu_sample(nx, ny, nz, rp, rs)
    R = rp + rs
    assert n == 1 and wx == wy and wx % 16 == 0 and wx < nx
    assert (cz0, cz1) == (z0 - R, z0 + 1 + R)
    assert (cy0, cy1, cx0, cx1) == (y0 - R, y0 + wy + R, x0 - R, x0 + wx + R)
    assert 0 <= cy0 and cy1 <= ny and 0 <= cx0 and cx1 <= nx


def test_window_equals_whole_volume_result():
    rng = np.random.default_rng(21)
    nz, ny, nx, rp, rs, h = 30, 40, 44, 2, 3, 0.2
    u = rng.random((nz, ny, nx), dtype=np.float32)
    # a tiny budget forces the window: 32 wide in x; y (40 < 32 + 2R) stays whole
    (z0, n, y0, wy, x0, wx), (cz0, cz1, cy0, cy1, cx0, cx1) = bench.cpu_sample(nx, ny, nz, rp, rs, budget_s=1e-6)
    assert n == 1 and wx == 32 and wy == ny
    crop = np.ascontiguousarray(u[cz0:cz1, cy0:cy1, cx0:cx1])
    win = (slice(None), slice(y0 - cy0, y0 - cy0 + wy), slice(x0 - cx0, x0 - cx0 + wx))
    got = oracle.nlm3d_separable_f32(crop, rp, rs, h, z0=z0 - cz0, z1=z0 - cz0 + n)[win]
    ref = oracle.nlm3d_separable_f32(u, rp, rs, h, z0=z0, z1=z0 + n)[:, y0:y0 + wy, x0:x0 + wx]
    assert np.array_equal(got, ref)
    got64 = oracle.nlm3d_separable_f64(crop, rp, rs, h, z0=z0 - cz0, z1=z0 - cz0 + n)[win]
    ref64 = oracle.nlm3d_separable_f64(u, rp, rs, h, z0=z0, z1=z0 + n)[:, y0:y0 + wy, x0:x0 + wx]
    assert np.array_equal(got64, ref64)


def test_traffic_figures_per_launch_shape():
    """The headline's byte counters are recorded per launch shape (the shapes
    give the same bits but not the same register spills); the figure the bench
    line carries is the chosen shape's, and an unknown shape or workload gives
    None rather than another shape's figure."""
    figs = [bench._traffic_from_profiles("nlm3d_256_p5_s11", c) for c in range(2)]
    assert all(isinstance(f, int) and f > 1.342e8 for f in figs), figs
    assert bench._traffic_from_profiles("nlm3d_256_p5_s11", 7) is None
    assert bench._traffic_from_profiles("no such workload") is None
    assert isinstance(bench._traffic_from_profiles("nlm3d_512x512x128_p17_s13_u8"), int)


def test_parity_units_where_every_candidate_is_zero():
    """The A/Z bar on 8-bit images with black regions: where A = 0 the output
    is exactly 0, so an exact 0 is 0 units (not 0/0) and anything else fails."""
    from nlm_metrics import fp32_units, phantom_parity
    A = np.array([0.0, 2.0, 0.0])
    Z = np.array([1.0, 1.0, 3.0])
    u = fp32_units(np.array([0.0, 2.0 ** -24, 1e-9]), A, Z)
    assert u[0] == 0.0 and u[1] == 0.5 and np.isinf(u[2])
    ref = np.array([0.0, 1.0, 2.0])
    rep = phantom_parity(ref.copy(), ref, ref, np.ones(3), np.array([0.0, 1.0, 2.0]))
    assert rep["ok"] and rep["max_err_in_fp32_units_of_A_over_Z"] == 0.0
