"""The CPU-baseline sample of bench.py (bench.cpu_sample), on the CPU.

For large patches the sample is one x-y window of a plane, computed by the
CPU path on the window grown by R = r_p + r_s.  The window's outputs must be
exactly those of the CPU path run on the whole volume.  The separable path
reads only samples within R of an output, and a clipped face of the crop is
the volume's own face.
"""
import numpy as np
import pytest

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
    assert rep["ok"] and rep["max_err_in_fp32_cond_units"] == 0.0


def test_parity_underflow_allowance_on_black_regions():
    """An 8-bit image with a flat black region next to a bright block: in FP64
    the black pixels near the block come out ~1e-58 (weights exp(-a) with
    a > 87, below FP32's range), in FP32 exactly 0, which the plain A/Z bar
    calls 2^24 units.  With the underflow allowance what is left is the
    weights' own error (a up to ~650 here: 33 units of A/Z); with the B term
    the CPU FP32 path sits well inside the bar."""
    import oracle
    from nlm_metrics import K_COND, fp32_units, phantom_parity, underflow_bound
    img = np.zeros((48, 48), np.float32)
    img[20:30, 26:40] = 255.0
    rp, rs, h = 5, 10, 10.0
    y32 = oracle.nlm2d_separable_f32(img, rp, rs, h, nthreads=2)
    y64, Z64, A64, B64 = oracle.nlm2d_separable_f64(img, rp, rs, h, nthreads=2, with_b=True)
    assert ((y32 == 0.0) & (y64 > 0.0) & (y64 < 1e-40)).any()
    e = y32 - y64
    assert fp32_units(e, A64, Z64).max() > 1e6
    uf = underflow_bound((2 * rs + 1) ** 2, 255.0, y64, Z64)
    assert fp32_units(e, A64, Z64, uf).max() > K_COND
    assert fp32_units(e, A64, Z64, uf, B64).max() <= 4.0
    rep = phantom_parity(y32, y32, y64, Z64, A64, underflow=uf, B64=B64)
    assert rep["ok"] and rep["cond_terms"] == "A + B"


def _reflect101(n, i):
    return -i if i < 0 else (2 * (n - 1) - i if i >= n else i)


@pytest.mark.parametrize("shape,rp,rs,h", [((1, 6, 7), 1, 2, 0.3), ((4, 5, 6), 1, 1, 0.25), ((1, 9, 7), 2, 3, 40.0)])
def test_oracle_b_term_brute_force(shape, rp, rs, h):
    """B = sum_d w_d a_d (|u~(x+d)| + |y|) of the FP64 separable paths against
    a brute-force loop over voxels and offsets (reflect-101 borders), 2D
    (nz = 1) and 3D; Z, A and y along with it."""
    import oracle
    nz, ny, nx = shape
    u = np.random.default_rng(nz * 100 + ny + rp).random(shape).astype(np.float32)
    if h > 1.0:
        u = np.rint(u * 255.0).astype(np.float32)
    P = (2 * rp + 1) ** (2 if nz == 1 else 3)
    zr = range(-rs, rs + 1) if nz > 1 else [0]
    pr = range(-rp, rp + 1)
    pz = pr if nz > 1 else [0]
    U = lambda z, y, x: float(u[_reflect101(nz, z), _reflect101(ny, y), _reflect101(nx, x)])  # noqa: E731
    want = np.zeros((4,) + shape)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                terms = []
                for dz in zr:
                    for dy in range(-rs, rs + 1):
                        for dx in range(-rs, rs + 1):
                            S = sum((U(z + qz, y + qy, x + qx) - U(z + dz + qz, y + dy + qy, x + dx + qx)) ** 2
                                    for qz in pz for qy in pr for qx in pr)
                            a = S / P / (h * h)
                            terms.append((np.exp(-a), a, U(z + dz, y + dy, x + dx)))
                Z = sum(w for w, _, _ in terms)
                yy = sum(w * c for w, _, c in terms) / Z
                want[:, z, y, x] = (yy, Z, sum(w * abs(c) for w, _, c in terms),
                                    sum(w * a * (abs(c) + abs(yy)) for w, a, c in terms))
    if nz == 1:
        got = oracle.nlm2d_separable_f64(u[0], rp, rs, h, with_b=True)
        got = [g[None] for g in got]
    else:
        got = oracle.nlm3d_separable_f64_za(u, rp, rs, h, with_b=True)
    for g, w in zip(got, want):
        assert np.allclose(g, w, rtol=1e-11, atol=0.0)
