"""Launch-shape tuning of the 3D forward and of batch plans.

The shapes of one tile family (register budgets = waves per SIMD, and tile
depths in z = the share of halo rows) compute exactly the same arithmetic per
voxel, so every shape must give the same bits -- that is what lets the tuner
choose by timing at first use without the output ever depending on the
outcome.  Bars: bit-identical outputs across shapes, across tuned and untuned
calls, and across batch plans with and without tuning; one shape per family
also against the oracle (<= 1e-5 relative, random input)."""
import numpy as np
import pytest
import torch

import gnlm
import oracle
from nlm_metrics import max_rel, report

pytestmark = pytest.mark.gpu

FAMILIES = [
    # (nz, ny, nx), r_p, r_s: partial last layers for the deep tiles (24 / 28 planes)
    ((61, 44, 40), 1, 3),     # 16x14x14 at 3 or 4 per CU, 16x14x28
    ((30, 36, 40), 2, 5),     # 16x12x12, 16x12x24
    ((52, 27, 33), 2, 3),
    ((26, 20, 17), 3, 2),     # 128 or 168 VGPRs
    ((20, 24, 28), 4, 3),     # 16x8x8 rows tiles, 16x24x24 big tiles
]


@pytest.mark.parametrize("shape3,rp,rs", FAMILIES)
def test_shapes_are_bit_identical(dev, shape3, rp, rs):
    rng = np.random.default_rng(sum(shape3) + rp)
    u = rng.random(shape3, dtype=np.float32)
    ut = torch.from_numpy(u).to(dev)
    try:
        _, n = gnlm.forward_shape(shape3, rp, rs)
        assert n >= 2, n
        outs = []
        for c in range(n):
            gnlm.set_forward_shape(shape3, rp, rs, c)
            print(f"case shape={c} of {n} dims={shape3} rp={rp} rs={rs}", flush=True)
            outs.append(gnlm.nlm3d(ut, rp, rs, 0.25).cpu().numpy())
        for c in range(1, n):
            assert np.array_equal(outs[c], outs[0]), (c, report(outs[c], outs[0]))
        ref64 = oracle.nlm3d_brute(u.astype(np.float64), rp, rs, 0.25)
        assert max_rel(outs[0], ref64) <= 1e-5, report(outs[0], ref64)
    finally:
        gnlm.reset_tuning()


def test_lazy_tuning_at_first_use(dev):
    """The first forward of a size class times the shapes and records one;
    later calls keep it; the output equals an untuned (default-rule) call."""
    shape3, rp, rs = (48, 56, 64), 2, 5
    u, _ = gnlm.phantom3d(64, 56, 48, n_ellipsoids=6, axis_lo=3, axis_hi=12, sigma=0.05, seed=3)
    ut = torch.from_numpy(u).to(dev)
    try:
        gnlm.reset_tuning()
        gnlm.set_tuning(gnlm.TUNE_OFF)
        y_off = gnlm.nlm3d(ut, rp, rs, 0.05).cpu().numpy()
        assert gnlm.forward_shape(shape3, rp, rs)[0] == -1
        gnlm.set_tuning(gnlm.TUNE_LAZY)
        y1 = gnlm.nlm3d(ut, rp, rs, 0.05).cpu().numpy()
        s1, n = gnlm.forward_shape(shape3, rp, rs)
        print(f"tuned shape {s1} of {n}", flush=True)
        assert 0 <= s1 < n
        y2 = gnlm.nlm3d(ut, rp, rs, 0.05).cpu().numpy()
        assert gnlm.forward_shape(shape3, rp, rs)[0] == s1
        assert np.array_equal(y1, y_off) and np.array_equal(y2, y_off)
        # the same size class (voxel count within half an octave) shares the choice
        assert gnlm.forward_shape((48, 56, 66), rp, rs)[0] == s1
    finally:
        gnlm.set_tuning(gnlm.TUNE_LAZY)
        gnlm.reset_tuning()


def test_batch_plan_tuning_keeps_bits(dev):
    """A batch plan created with tuning on (workgroups per CU timed) and one
    created with it off give the same bits as per-volume launches."""
    rng = np.random.default_rng(7)
    vols = []
    for i in range(6):
        nx, ny, nz = (int(rng.choice([24, 40, 56])) for _ in range(3))
        u, _ = gnlm.phantom3d(nx, ny, nz, n_ellipsoids=4, axis_lo=3, axis_hi=9, sigma=0.05, seed=70 + i)
        vols.append((torch.from_numpy(u).to(dev), int(rng.choice([1, 2])), int(rng.choice([3, 5])), 0.05))
    refs = [gnlm.nlm3d(u, rp, rs, h) for u, rp, rs, h in vols]
    try:
        for mode in (gnlm.TUNE_LAZY, gnlm.TUNE_OFF):
            gnlm.set_tuning(mode)
            outs = [torch.empty_like(u) for u, *_ in vols]
            plan = gnlm.BatchPlan([(u, o, rp, rs, h) for (u, rp, rs, h), o in zip(vols, outs)])
            ntiles, nblocks = plan.info()
            print(f"tuning mode {mode}: {ntiles} tiles, {nblocks} workgroups", flush=True)
            for o in outs:
                o.fill_(-1.0)
            plan.run()
            torch.cuda.synchronize()
            for o, r in zip(outs, refs):
                assert torch.equal(o, r), mode
            plan.close()
    finally:
        gnlm.set_tuning(gnlm.TUNE_LAZY)


@pytest.mark.parametrize("shape3,rp,rs", [((48, 96, 96), 1, 3), ((64, 64, 64), 2, 5), ((17, 30, 23), 3, 4),
                                          ((40, 40, 40), 2, 3), ((9, 70, 21), 1, 2)])
def test_dz_split_same_bits(dev, shape3, rp, rs):
    """The dz split of small volumes (several workgroups per tile, each
    writing per-dz partial sums; a combine pass adds them in dz order) gives
    the unsplit launch's bits, y and Z, under the default rule and tuned."""
    u = np.random.default_rng(sum(shape3) * 3 + rp).random(shape3, dtype=np.float32)
    ut = torch.from_numpy(u).to(dev)
    outs = []
    try:
        for split in (True, False):
            gnlm.set_dz_split(split)
            for mode in (gnlm.TUNE_OFF, gnlm.TUNE_LAZY):
                gnlm.reset_tuning()
                gnlm.set_tuning(mode)
                print(f"case dz split={split} tuning={mode} dims={shape3} rp={rp} rs={rs}", flush=True)
                y, Z = gnlm.nlm3d(ut, rp, rs, 0.15, want_Z=True)
                y2 = gnlm.nlm3d(ut, rp, rs, 0.15)
                outs.append((y.cpu(), Z.cpu(), y2.cpu()))
    finally:
        gnlm.set_dz_split(True)
        gnlm.set_tuning(gnlm.TUNE_LAZY)
        gnlm.reset_tuning()
    for y, Z, y2 in outs:
        assert torch.equal(y, outs[0][0]) and torch.equal(Z, outs[0][1]) and torch.equal(y2, outs[0][0])
    ref64 = oracle.nlm3d_brute(u.astype(np.float64), rp, rs, 0.15)
    assert max_rel(outs[0][0].numpy(), ref64) <= 1e-5, report(outs[0][0].numpy(), ref64)
