"""GPU adjoint: dL/d(input) and dL/dh.

Both forms of the rows-organised adjoint (gather, symmetric; r_p 0-2) are
run; r_p = 3 runs on the LDS-block tile in either setting.
Tolerances:
  * vs the oracle's FP64 chain-rule adjoint on the same input and dL/dy:
    max |gpu - ref| <= 1e-4 * max |ref| for dL/du, and relative 1e-4 for dL/dh.
  * vs FP64 central finite differences (eps = 1e-4) of the MSE loss against
    the clean phantom on a 24^3 crop (the differentiable workload's check):
    |gpu - fd| <= 1e-4 * max(|fd|, max|grad|) at sampled voxels and for h.
"""
import numpy as np
import pytest
import torch

import gnlm
import oracle
from nlm_autograd import nlm3d

pytestmark = pytest.mark.gpu


def _gpu_grad(dev, u, g, rp, rs, h):
    ut = torch.from_numpy(np.ascontiguousarray(u, np.float32)).to(dev)
    y, Z = gnlm.nlm3d(ut, rp, rs, h, want_Z=True)
    gt = torch.from_numpy(np.ascontiguousarray(g, np.float32)).to(dev)
    gu, gh = gnlm.nlm3d_backward(ut, y, Z, gt, rp, rs, h)
    torch.cuda.synchronize()
    return gu.cpu().numpy().astype(np.float64), float(gh.item())


# the two forms of the rows adjoint, r_p 0-2 (gather over every offset; half of the
# offsets with the mirror terms scattered), forced for the whole test
@pytest.fixture(params=[gnlm.BWD_GATHER, gnlm.BWD_SYMMETRIC], ids=["gather", "symmetric"])
def form(request):
    gnlm.set_backward_form(request.param)
    yield request.param
    gnlm.set_backward_form(gnlm.BWD_AUTO)


@pytest.mark.parametrize("shape,rp,rs,h", [
    ((7, 8, 9), 1, 2, 0.3),
    ((10, 12, 11), 1, 3, 0.25),
    ((6, 9, 8), 0, 2, 0.3),
    ((12, 12, 12), 1, 4, 0.2),
    ((3, 17, 18), 1, 2, 0.3),   # thin volume: reflect-101 folds many positions
    ((26, 31, 29), 1, 4, 0.2),  # several tiles per axis: scatter regions overlap
    ((14, 15, 13), 1, 5, 0.25),
    ((9, 10, 11), 1, 7, 0.3),   # largest instantiated r_s
    ((5, 4, 23), 1, 1, 0.3),
    ((12, 12, 12), 2, 3, 0.2),  # patch radius 2: 20 x 16 phase-1 rows, quarter rows
    ((26, 23, 21), 2, 4, 0.25), # several tiles per axis
    ((9, 10, 11), 2, 7, 0.3),   # largest instantiated r_s
    ((14, 13, 6), 2, 1, 0.3),
    ((20, 21, 19), 0, 5, 0.3),  # single-voxel patches, several tiles
    ((12, 13, 14), 3, 2, 0.3),  # patch radius 3: the LDS-block tile (10 x 10 x 8)
    ((17, 16, 15), 3, 3, 0.3),  # largest r_s it fits
])
def test_backward_matches_oracle(dev, form, shape, rp, rs, h):
    rng = np.random.default_rng(sum(shape) * 7 + rs)
    u = rng.random(shape).astype(np.float32)
    g = rng.standard_normal(shape).astype(np.float32)
    print(f"case bwd form={form} shape={shape} rp={rp} rs={rs}", flush=True)
    gu, gh = _gpu_grad(dev, u, g, rp, rs, h)
    ref_gu, ref_gh = oracle.nlm3d_grad_f64(u.astype(np.float64), g.astype(np.float64), rp, rs, h)
    scale = np.abs(ref_gu).max()
    assert np.abs(gu - ref_gu).max() <= 1e-4 * scale, (np.abs(gu - ref_gu).max(), scale)
    assert abs(gh - ref_gh) <= 1e-4 * max(abs(ref_gh), 1e-8), (gh, ref_gh)


@pytest.mark.parametrize("shape,rp,rs,h", [
    ((9, 10, 11), 4, 2, 0.3),   # NS 9: past the on-chip adjoints
    ((8, 9, 10), 3, 4, 0.3),    # NS 7 with SS 9: past the LDS-block tile's image
    ((10, 9, 8), 8, 1, 0.4),    # NS 17
    ((7, 8, 6), 2, 8, 0.3),     # SS 17
    ((6, 7, 8), 8, 6, 10.0),    # the reference's default NS 17 / SS 13 (3DNLM/config.h:4-5), h on 0-255
])
def test_backward_any_radius(dev, shape, rp, rs, h):
    """The any-radius adjoint (per-offset passes through HBM) against the
    FP64 chain-rule adjoint: same bars as the on-chip forms."""
    nz, ny, nx = shape
    assert gnlm.backward_route(nx, ny, nz, rp, rs) == (gnlm.OK, gnlm.KERNEL_BWD_GENERIC)
    rng = np.random.default_rng(sum(shape) * 3 + rp)
    u = rng.random(shape).astype(np.float32) * (255.0 if h > 1 else 1.0)
    g = rng.standard_normal(shape).astype(np.float32)
    print(f"case bwd generic shape={shape} rp={rp} rs={rs}", flush=True)
    gu, gh = _gpu_grad(dev, u, g, rp, rs, h)
    ref_gu, ref_gh = oracle.nlm3d_grad_f64(u.astype(np.float64), g.astype(np.float64), rp, rs, h)
    scale = np.abs(ref_gu).max()
    assert np.abs(gu - ref_gu).max() <= 1e-4 * scale, (np.abs(gu - ref_gu).max(), scale)
    assert abs(gh - ref_gh) <= 1e-4 * max(abs(ref_gh), 1e-8), (gh, ref_gh)


def test_backward_any_radius_vs_finite_differences(dev):
    """NS 9, SS 5 on a 10^3 phantom crop: the any-radius adjoint of the MSE
    loss against FP64 central finite differences (eps = 1e-4)."""
    noisy, clean = gnlm.phantom3d(10, 10, 10, n_ellipsoids=3, axis_lo=2, axis_hi=5, sigma=0.05, seed=31)
    rp, rs, h = 4, 2, 0.15
    ut = torch.from_numpy(noisy).to(dev).requires_grad_(True)
    ht = torch.tensor(h, dtype=torch.float32, requires_grad=True)
    y = nlm3d(ut, ht, rp, rs)
    loss = ((y - torch.from_numpy(clean).to(dev)) ** 2).mean()
    loss.backward()
    gu = ut.grad.cpu().numpy().astype(np.float64)
    u64, c64 = noisy.astype(np.float64), clean.astype(np.float64)
    eps = 1e-4
    scale = np.abs(gu).max()
    for i in [(0, 0, 0), (5, 4, 3), (9, 9, 9), (2, 7, 1)]:
        up = u64.copy(); up[i] += eps
        um = u64.copy(); um[i] -= eps
        fd = (oracle.nlm3d_mse_f64(up, c64, rp, rs, h) - oracle.nlm3d_mse_f64(um, c64, rp, rs, h)) / (2 * eps)
        assert abs(fd - gu[i]) <= 1e-4 * max(abs(fd), scale), (i, fd, gu[i])
    fdh = (oracle.nlm3d_mse_f64(u64, c64, rp, rs, h + eps) - oracle.nlm3d_mse_f64(u64, c64, rp, rs, h - eps)) / (2 * eps)
    assert abs(float(ht.grad) - fdh) <= 1e-4 * abs(fdh), (float(ht.grad), fdh)


def test_backward_deterministic(dev, form):
    rng = np.random.default_rng(1)
    u = rng.random((20, 22, 24)).astype(np.float32)
    g = rng.standard_normal(u.shape).astype(np.float32)
    a = _gpu_grad(dev, u, g, 1, 3, 0.2)
    b = _gpu_grad(dev, u, g, 1, 3, 0.2)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]


def test_backward_forms_agree(dev):
    """Gather and symmetric forms on a 40^3 volume: the same gradient to FP32
    rounding (different summation orders)."""
    _forms_agree(dev, 1, 4)


def test_backward_forms_agree_rp2(dev):
    _forms_agree(dev, 2, 3)


def test_backward_forms_agree_whole_tiles(dev):
    """78^3: 735 tiles of the symmetric form at r_p 1, r_s 4, within one round
    of 768 workgroup slots (three per CU on a 256-CU device), so each
    workgroup takes a whole tile.  The 40^3 volumes above split each tile's
    offset rows over several workgroups; both layouts agree with the gather
    form."""
    _forms_agree(dev, 1, 4, n=78)


def test_backward_forms_agree_three_parts(dev):
    """70^3: 504 tiles at r_p 1, r_s 4, so on a 256-CU device (768 slots) each
    tile's 40 offset rows go to 3 workgroups of 13, 13 and 14 rows."""
    _forms_agree(dev, 1, 4, n=70)


@pytest.mark.parametrize("rp,rs", [(1, 3), (2, 2), (0, 4)])
def test_backward_forced_part_counts(dev, rp, rs):
    """Each tile's offset rows split over np = 1, 2, 3 and 4 workgroups
    (forced, whatever the device's compute-unit count): every split agrees
    with the gather form and with the FP64 chain-rule adjoint, and each is
    deterministic; the part count changes only the summation order."""
    rng = np.random.default_rng(40 + rp)
    u = rng.random((30, 28, 26)).astype(np.float32)
    g = rng.standard_normal(u.shape).astype(np.float32)
    ref_gu, ref_gh = oracle.nlm3d_grad_f64(u.astype(np.float64), g.astype(np.float64), rp, rs, 0.2)
    scale = np.abs(ref_gu).max()
    try:
        gnlm.set_backward_form(gnlm.BWD_GATHER)
        ga = _gpu_grad(dev, u, g, rp, rs, 0.2)
        gnlm.set_backward_form(gnlm.BWD_SYMMETRIC)
        for np_ in (1, 2, 3, 4):
            gnlm.set_backward_parts(np_)
            print(f"case bwd parts={np_} rp={rp} rs={rs}", flush=True)
            a = _gpu_grad(dev, u, g, rp, rs, 0.2)
            b = _gpu_grad(dev, u, g, rp, rs, 0.2)
            assert np.array_equal(a[0], b[0]) and a[1] == b[1], np_
            assert np.abs(a[0] - ga[0]).max() <= 2e-5 * scale, np_
            assert abs(a[1] - ga[1]) <= 2e-5 * abs(ga[1]), np_
            assert np.abs(a[0] - ref_gu).max() <= 1e-4 * scale, np_
            assert abs(a[1] - ref_gh) <= 1e-4 * abs(ref_gh), np_
    finally:
        gnlm.set_backward_parts(0)
        gnlm.set_backward_form(gnlm.BWD_AUTO)


def _forms_agree(dev, rp, rs, n=40):
    rng = np.random.default_rng(5 + rp)
    u = rng.random((n, n, n)).astype(np.float32)
    g = rng.standard_normal(u.shape).astype(np.float32)
    try:
        gnlm.set_backward_form(gnlm.BWD_GATHER)
        a = _gpu_grad(dev, u, g, rp, rs, 0.15)
        gnlm.set_backward_form(gnlm.BWD_SYMMETRIC)
        b = _gpu_grad(dev, u, g, rp, rs, 0.15)
    finally:
        gnlm.set_backward_form(gnlm.BWD_AUTO)
    scale = np.abs(a[0]).max()
    print("forms: max |diff| / max |g|", np.abs(a[0] - b[0]).max() / scale, "dh", a[1], b[1])
    assert np.abs(a[0] - b[0]).max() <= 2e-5 * scale
    assert abs(a[1] - b[1]) <= 2e-5 * abs(a[1])


def _crop24():
    noisy, clean = gnlm.phantom3d(128, 128, 128, n_ellipsoids=24, axis_lo=5, axis_hi=30, background=0.05,
                                  sigma=0.05, seed=20261207)
    s = slice(52, 76)
    return np.ascontiguousarray(noisy[s, s, s]), np.ascontiguousarray(clean[s, s, s])


def test_backward_vs_finite_differences_crop24(dev):
    """24^3 crop, patch 3^3, search 9^3, h = 0.1: MSE loss vs the clean phantom."""
    u, clean = _crop24()
    rp, rs, h = 1, 4, 0.1
    ut = torch.from_numpy(u).to(dev).requires_grad_(True)
    ht = torch.tensor(h, dtype=torch.float32, device=dev, requires_grad=True)
    ct = torch.from_numpy(clean).to(dev)
    y = nlm3d(ut, ht, rp, rs)
    loss = ((y - ct) ** 2).mean()
    loss.backward()
    gu = ut.grad.cpu().numpy().astype(np.float64)
    gh = float(ht.grad.item())
    u64, c64 = u.astype(np.float64), clean.astype(np.float64)
    eps = 1e-4
    rng = np.random.default_rng(3)
    idx = [tuple(rng.integers(0, 24, 3)) for _ in range(6)] + [(0, 0, 0), (23, 23, 23), (0, 12, 23)]
    scale = np.abs(gu).max()
    worst = 0.0
    for i in idx:
        up = u64.copy(); up[i] += eps
        um = u64.copy(); um[i] -= eps
        fd = (oracle.nlm3d_mse_f64(up, c64, rp, rs, h) - oracle.nlm3d_mse_f64(um, c64, rp, rs, h)) / (2 * eps)
        err = abs(fd - gu[i]) / max(abs(fd), scale)
        worst = max(worst, err)
        assert err <= 1e-4, (i, fd, gu[i])
    fdh = (oracle.nlm3d_mse_f64(u64, c64, rp, rs, h + eps) - oracle.nlm3d_mse_f64(u64, c64, rp, rs, h - eps)) / (2 * eps)
    print(f"crop24: worst voxel rel err {worst:.2e}, dL/dh gpu {gh:.6e} fd {fdh:.6e}")
    assert abs(gh - fdh) <= 1e-4 * abs(fdh)


def test_fit_h_decreases_loss(dev):
    """A few gradient steps on h (log-parameterised, a CPU leaf tensor) lower
    the MSE vs clean."""
    u, clean = _crop24()
    ut = torch.from_numpy(u).to(dev)
    ct = torch.from_numpy(clean).to(dev)
    # h lives on the host: its value reaches the C-ABI without a device wait
    logh = torch.tensor(np.log(0.1), dtype=torch.float32, requires_grad=True)
    opt = torch.optim.Adam([logh], lr=0.05)
    losses = []
    for _ in range(15):
        opt.zero_grad()
        y = nlm3d(ut, torch.exp(logh), 1, 4)
        loss = ((y - ct) ** 2).mean()
        loss.backward()
        opt.step()
        losses.append(loss.item())
    print("fit-h losses", losses[0], losses[-1], "h", float(torch.exp(logh.detach())))
    assert losses[-1] < losses[0]


SWEEP_DIMS = (1, 2, 3, 5, 7, 13, 17, 29, 31)


def _sweep_cases(n=24, seed=20261204):
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(n):
        shape = tuple(int(v) for v in rng.choice(SWEEP_DIMS, 3))
        cases.append((shape, int(rng.integers(0, 3)), int(rng.integers(1, 8))))
    return cases


@pytest.mark.parametrize("shape,rp,rs", _sweep_cases())
def test_backward_sweep(dev, shape, rp, rs):
    """Seeded shapes from 1 to 31 per axis (thin volumes, partial tiles on
    every axis, several reflections of one sample) and radii r_p 0-2, r_s
    1-7: the symmetric and gather forms agree to FP32 rounding, and on the
    smaller cases both match the FP64 chain-rule adjoint."""
    rng = np.random.default_rng(sum(shape) * 11 + 7 * rp + rs)
    u = rng.random(shape).astype(np.float32)
    g = rng.standard_normal(shape).astype(np.float32)
    print(f"case bwd sweep shape={shape} rp={rp} rs={rs}", flush=True)
    try:
        gnlm.set_backward_form(gnlm.BWD_GATHER)
        a = _gpu_grad(dev, u, g, rp, rs, 0.2)
        gnlm.set_backward_form(gnlm.BWD_SYMMETRIC)
        b = _gpu_grad(dev, u, g, rp, rs, 0.2)
    finally:
        gnlm.set_backward_form(gnlm.BWD_AUTO)
    scale = max(np.abs(a[0]).max(), 1e-30)
    assert np.abs(a[0] - b[0]).max() <= 2e-5 * scale, np.abs(a[0] - b[0]).max() / scale
    # dL/dh is a sum of terms of both signs over every position and offset;
    # on volumes of a few hundred voxels it cancels to a fraction of its
    # terms' size (1/200 on some of these), so the two forms are compared at
    # the oracle's bar
    assert abs(a[1] - b[1]) <= 1e-4 * max(abs(a[1]), 1e-8), (a[1], b[1])
    if u.size * (2 * rs + 1) ** 3 <= 10_000_000:
        ref_gu, ref_gh = oracle.nlm3d_grad_f64(u.astype(np.float64), g.astype(np.float64), rp, rs, 0.2)
        rscale = max(np.abs(ref_gu).max(), 1e-30)
        for gu, gh in (a, b):
            assert np.abs(gu - ref_gu).max() <= 1e-4 * rscale, np.abs(gu - ref_gu).max() / rscale
            assert abs(gh - ref_gh) <= 1e-4 * max(abs(ref_gh), 1e-8), (gh, ref_gh)
