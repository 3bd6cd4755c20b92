"""Batched persistent forward and the autotuner.

Bars: bit-identical outputs between the graph-replayed persistent launch, its
eager launch, and per-volume gnlm3d_forward calls (whatever tile variant the
table selects).  Against the oracle, nlm_metrics.phantom_parity: FP64
evaluated at the float32 h the FP32 paths receive, unfloored relative error
<= 1e-5 on voxels above 0.5% of the output range and |gpu - fp64| <= 32 *
2^-24 * A/Z everywhere; against the CPU FP32 path (above the same threshold)
<= 1e-5 plus that path's own distance from the same FP64 result.  These
volumes use h = sigma down to 0.02, where the weights exp(-S/(P h^2)) make the
output sensitive to the last bit of the exponent scale: the CPU FP32 path
rounds h^2 in float while the kernels fold log2(e)/(P h^2) in double, and the
two FP32 results can sit on opposite sides of the FP64 value, ~1e-5 apart,
while each is within ~6e-6 of it."""
import numpy as np
import pytest
import torch

import gnlm
import oracle
from nlm_metrics import max_rel, phantom_parity

pytestmark = pytest.mark.gpu


def _volumes(dev, n, seed, dims=(16, 24, 32, 40, 48), rps=(1, 2)):
    rng = np.random.default_rng(seed)
    vols = []
    for i in range(n):
        nx, ny, nz = (int(rng.choice(dims)) for _ in range(3))
        sigma = float(rng.uniform(0.02, 0.10))
        u, _ = gnlm.phantom3d(nx, ny, nz, n_ellipsoids=4, axis_lo=3, axis_hi=10, sigma=sigma, seed=seed * 100 + i)
        rp = int(rng.choice(rps))
        rs = int(rng.choice([3, 5]))
        vols.append((torch.from_numpy(u).to(dev), rp, rs, sigma))
    return vols


def test_batch_graph_equals_per_volume(dev):
    vols = _volumes(dev, 10, 1)
    outs = [torch.empty_like(u) for u, *_ in vols]
    plan = gnlm.BatchPlan([(u, o, rp, rs, h) for (u, rp, rs, h), o in zip(vols, outs)])
    ntiles, nblocks = plan.info()
    print(f"batch: {len(vols)} volumes, {ntiles} tiles, {nblocks} workgroups", flush=True)
    refs = [gnlm.nlm3d(u, rp, rs, h) for u, rp, rs, h in vols]
    torch.cuda.synchronize()
    for rep in range(3):
        for o in outs:
            o.fill_(-1.0)           # eager work between replays
        torch.cuda.synchronize()
        plan.run()
        torch.cuda.synchronize()
        for o, r in zip(outs, refs):
            assert torch.equal(o, r), f"replay {rep} differs"
    for o in outs:
        o.zero_()
    plan.run(eager=True)
    torch.cuda.synchronize()
    for o, r in zip(outs, refs):
        assert torch.equal(o, r)
    plan.close()


def test_batch_matches_oracle(dev):
    vols = _volumes(dev, 4, 2, dims=(16, 24))
    outs = [torch.empty_like(u) for u, *_ in vols]
    plan = gnlm.BatchPlan([(u, o, rp, rs, h) for (u, rp, rs, h), o in zip(vols, outs)])
    plan.run()
    torch.cuda.synchronize()
    for (u, rp, rs, h), o in zip(vols, outs):
        _check_volume(u.cpu().numpy(), o.cpu().numpy(), rp, rs, h)
    plan.close()


def _check_volume(un, got, rp, rs, h, z0=0, z1=None):
    """Planes [z0, z1) of one batched output against the oracle (module bars)."""
    z1 = un.shape[0] if z1 is None else z1
    h32 = float(np.float32(h))  # the value both FP32 paths receive
    ref32 = oracle.nlm3d_separable_f32(un, rp, rs, h32, z0=z0, z1=z1)
    ref64, Z64, A64 = oracle.nlm3d_separable_f64_za(un, rp, rs, h32, z0=z0, z1=z1)
    cpu = phantom_parity(ref32, ref32, ref64, Z64, A64)
    rep = phantom_parity(got[z0:z1], ref32, ref64, Z64, A64, 1e-5, 1e-5 + cpu["max_rel_vs_fp64_above_tau"])
    print("batch vol", un.shape, rp, rs, f"{h:.4f}", z0, z1, rep, "cpu fp32 vs fp64", cpu["max_rel_vs_fp64_above_tau"])
    assert rep["ok"], rep


def _config4_volumes(dev):
    """The 64-volume set of the batched workload (bench.py nlm3d_batch64: the
    same seed, draws and phantoms): dims from {32, 48, ..., 192} per axis,
    sigma ~ U[0.02, 0.10], h = sigma, r_p in {1, 2}, r_s in {3, 5}."""
    rng = np.random.default_rng(20261208)
    dims_set = np.arange(32, 193, 16)
    vols = []
    for i in range(64):
        nx, ny, nz = (int(v) for v in rng.choice(dims_set, 3))
        sigma = float(rng.uniform(0.02, 0.10))
        rp, rs = int(rng.choice([1, 2])), int(rng.choice([3, 5]))
        u, _ = gnlm.phantom3d(nx, ny, nz, n_ellipsoids=12, axis_lo=4, axis_hi=20, background=0.05, sigma=sigma,
                              seed=20261208 * 100 + i)
        vols.append((torch.from_numpy(u).to(dev), rp, rs, sigma))
    return vols


def test_batch_config4_bit_identical(dev):
    """The batched workload itself, 64 uneven volumes up to 192 per axis: the
    graph replays and the eager launch equal the 64 per-volume launches bit
    for bit (per-volume launches of large r_p = 1 volumes take the 4-per-CU
    build, which the persistent kernel never uses)."""
    vols = _config4_volumes(dev)
    outs = [torch.empty_like(u) for u, *_ in vols]
    plan = gnlm.BatchPlan([(u, o, rp, rs, h) for (u, rp, rs, h), o in zip(vols, outs)])
    ntiles, nblocks = plan.info()
    print(f"config 4: 64 volumes, {ntiles} tiles, {nblocks} workgroups", flush=True)
    refs = [gnlm.nlm3d(u, rp, rs, h) for u, rp, rs, h in vols]
    torch.cuda.synchronize()
    for rep in range(2):
        for o in outs:
            o.fill_(-1.0)
        torch.cuda.synchronize()
        plan.run()
        torch.cuda.synchronize()
        for i, (o, r) in enumerate(zip(outs, refs)):
            assert torch.equal(o, r), f"replay {rep}, volume {i} {tuple(o.shape)}"
    for o in outs:
        o.zero_()
    plan.run(eager=True)
    torch.cuda.synchronize()
    for i, (o, r) in enumerate(zip(outs, refs)):
        assert torch.equal(o, r), f"eager, volume {i}"
    plan.close()


@pytest.mark.slow
def test_batch_config4_matches_oracle(dev):
    """Every volume of the batched workload against the oracle: whole volumes
    up to 64^3, 3-plane slabs at both z ends and in the middle above."""
    vols = _config4_volumes(dev)
    outs = [torch.empty_like(u) for u, *_ in vols]
    plan = gnlm.BatchPlan([(u, o, rp, rs, h) for (u, rp, rs, h), o in zip(vols, outs)])
    plan.run()
    torch.cuda.synchronize()
    for (u, rp, rs, h), o in zip(vols, outs):
        un, got = u.cpu().numpy(), o.cpu().numpy()
        nz = un.shape[0]
        if un.size <= 64 ** 3:
            _check_volume(un, got, rp, rs, h)
        else:
            for z0 in (0, nz // 2 - 1, nz - 3):
                _check_volume(un, got, rp, rs, h, z0, z0 + 3)
    plan.close()


@pytest.mark.parametrize("rp,rs", [(1, 3), (2, 5), (1, 5)])
def test_autotune_variants(dev, rp, rs):
    """The tuner times every variant that fits and keeps the fastest; the
    direct-sum variants agree bit for bit, the slab variant to rounding."""
    try:
        v, ms = gnlm.autotune(rp, rs, probe_n=48)
        print(f"autotune rp={rp} rs={rs}: variant {v}, ms {ms}")
        assert v >= 0 and ms[v] > 0 and ms[v] == min(m for m in ms if m > 0)
        rng = np.random.default_rng(rp * 10 + rs)
        u = torch.from_numpy(rng.random((21, 35, 37), dtype=np.float32)).to(dev)
        outs = {}
        for var in range(len(ms)):
            if ms[var] < 0:
                continue
            gnlm.set_variant(rp, rs, var)
            outs[var] = gnlm.nlm3d(u, rp, rs, 0.2).cpu().numpy()
        direct = [outs[k] for k in (gnlm.VARIANT_16x16x16, gnlm.VARIANT_16x16x8) if k in outs]
        for o in direct[1:]:
            assert np.array_equal(o, direct[0])
        for sv in (gnlm.VARIANT_SLAB, gnlm.VARIANT_SLAB2, gnlm.VARIANT_ROWS, gnlm.VARIANT_ROWS_SHORT):
            if sv in outs:
                assert max_rel(outs[sv], direct[0]) <= 1e-5, sv
    finally:
        gnlm.reset_tuning()


def test_batch_slab_and_direct_variants(dev):
    """Per-volume and batched launches agree bit for bit whichever variant
    the table selects."""
    vols = _volumes(dev, 6, 3)
    try:
        for var in (gnlm.VARIANT_16x16x16, gnlm.VARIANT_SLAB, gnlm.VARIANT_SLAB2, gnlm.VARIANT_ROWS,
                    gnlm.VARIANT_ROWS_SHORT):
            for rp in (1, 2):
                for rs in (3, 5):
                    # short tiles exist for r_p = 2 only
                    gnlm.set_variant(rp, rs, gnlm.VARIANT_ROWS if (var, rp) == (gnlm.VARIANT_ROWS_SHORT, 1) else var)
            outs = [torch.empty_like(u) for u, *_ in vols]
            plan = gnlm.BatchPlan([(u, o, rp, rs, h) for (u, rp, rs, h), o in zip(vols, outs)])
            plan.run()
            refs = [gnlm.nlm3d(u, rp, rs, h) for u, rp, rs, h in vols]
            torch.cuda.synchronize()
            for o, r in zip(outs, refs):
                assert torch.equal(o, r), var
            plan.close()
    finally:
        gnlm.reset_tuning()


def test_batch_with_patch_radius_3(dev):
    """r_p = 3 volumes (16 x 10 x 10 rows tiles, the default for NS = 7) in a
    plan with r_p 1 and 2 volumes: bit-identical to per-volume launches."""
    vols = _volumes(dev, 9, 1, rps=(1, 2, 3))
    assert {rp for _, rp, _, _ in vols} == {1, 2, 3}
    for _, rp, rs, _ in vols:
        if rp == 3:
            assert gnlm.tuned_variant(rp, rs) == gnlm.VARIANT_ROWS
    outs = [torch.empty_like(u) for u, *_ in vols]
    plan = gnlm.BatchPlan([(u, o, rp, rs, h) for (u, rp, rs, h), o in zip(vols, outs)])
    refs = [gnlm.nlm3d(u, rp, rs, h) for u, rp, rs, h in vols]
    for rep in range(2):
        plan.run()
        torch.cuda.synchronize()
        for o, r in zip(outs, refs):
            assert torch.equal(o, r), rep
    plan.close()


def test_batch_with_patch_radius_4(dev):
    """r_p = 4 volumes run on the 16 x 8 x 8 rows tiles inside batch plans:
    bit-identical to per-volume launches."""
    vols = _volumes(dev, 5, 3, dims=(16, 24, 40), rps=(2, 4))
    assert {rp for _, rp, _, _ in vols} == {2, 4}
    outs = [torch.empty_like(u) for u, *_ in vols]
    plan = gnlm.BatchPlan([(u, o, rp, rs, h) for (u, rp, rs, h), o in zip(vols, outs)])
    refs = [gnlm.nlm3d(u, rp, rs, h) for u, rp, rs, h in vols]
    plan.run()
    torch.cuda.synchronize()
    for o, r in zip(outs, refs):
        assert torch.equal(o, r)
    plan.close()
