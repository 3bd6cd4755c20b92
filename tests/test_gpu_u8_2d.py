"""The 8-bit 2D forward (gnlm2d_forward_u8, gnlm.nlm2d_u8).

The tiled kernel's 8-bit build stages the bytes itself and runs both patch box
sums as sliding windows.  On 8-bit samples every partial sum of squared
differences is an integer below 2^24, exact in FP32 in any order, so the build
must give the FP32 build's bits on the same samples as floats:
  * at radius pairs of every tiled build -- the register-window builds (patch
    11 / search 21, the reference's own 2D geometry, and patch 5 / search 11)
    and the run-time-radius build at r_p 0-5 -- and through the float-copy
    fallback past the tiled kernel (r_p 6);
  * on images that drive the sums to their largest values (0/255
    checkerboards, where every difference is 0 or 255);
  * on ragged sizes (one row, one column, sizes that are not multiples of
    the 32 x 64 tile), and with a padded input whose rows are longer than the
    padded width.
Against the oracle: the FP64 brute force within 1e-5 on random images, and
the reference's own geometry on the 2048^2 8-bit phantom to the bar of
nlm_metrics (A + B terms, underflow allowance).
"""
import ctypes

import numpy as np
import pytest
import torch

import gnlm
import oracle
from nlm_metrics import max_rel, phantom_parity, report, underflow_bound

pytestmark = pytest.mark.gpu

TOL = 1e-5

CASES = [
    ((77, 200), 5, 10, 10.0),   # patch 11 / search 21: register-window build
    ((130, 33), 5, 10, 25.0),
    ((64, 96), 2, 5, 12.0),     # patch 5 / search 11: register-window build
    ((33, 41), 0, 3, 8.0),      # run-time-radius builds
    ((50, 37), 1, 4, 15.0),
    ((70, 90), 3, 6, 20.0),
    ((45, 66), 4, 2, 30.0),
    ((40, 29), 5, 3, 10.0),
    ((1, 97), 2, 5, 10.0),
    ((97, 1), 1, 3, 10.0),
    ((5, 3), 1, 2, 10.0),
    ((31, 57), 6, 4, 20.0),     # past the tiled kernel: float copy + generic path
]


def _img(shape, seed):
    return np.random.default_rng(seed).integers(0, 256, shape, dtype=np.uint8)


@pytest.mark.parametrize("shape,rp,rs,h", CASES)
def test_u8_2d_same_bits_as_fp32(dev, shape, rp, rs, h):
    img = _img(shape, sum(shape) * 7 + rp + rs)
    print(f"case u8 2d shape={shape} rp={rp} rs={rs} h={h}", flush=True)
    t8 = torch.from_numpy(img).to(dev)
    y8, Z8 = gnlm.nlm2d_u8(t8, rp, rs, h, want_Z=True)
    yf, Zf = gnlm.nlm2d(t8.float(), rp, rs, h, want_Z=True)
    torch.cuda.synchronize()
    assert torch.equal(y8, yf) and torch.equal(Z8, Zf)


@pytest.mark.parametrize("shape,rp,rs,h", [c for c in CASES if min(c[0]) > c[1] + c[2]][:6])
def test_u8_2d_matches_fp64(dev, shape, rp, rs, h):
    img = _img(shape, sum(shape) + 11 * rp + rs)
    print(f"case u8 2d vs fp64 shape={shape} rp={rp} rs={rs} h={h}", flush=True)
    y = gnlm.nlm2d_u8(torch.from_numpy(img).to(dev), rp, rs, h).cpu().numpy()
    ref64 = oracle.nlm2d_brute(img.astype(np.float64), rp, rs, h)
    assert max_rel(y, ref64) <= TOL, report(y, ref64)


@pytest.mark.parametrize("rp,rs", [(5, 10), (2, 5), (5, 2), (3, 4)])
def test_u8_2d_extreme_sums(dev, rp, rs):
    """0/255 checkerboards (cell sizes 1 and 3) and a flat 255 block on 0: the
    largest squared differences and patch sums the 8-bit build meets."""
    yy, xx = np.mgrid[0:70, 0:83]
    imgs = [((yy + xx) % 2 * 255).astype(np.uint8), (((yy // 3) + (xx // 3)) % 2 * 255).astype(np.uint8)]
    block = np.zeros((70, 83), np.uint8)
    block[20:45, 30:60] = 255
    imgs.append(block)
    for k, img in enumerate(imgs):
        print(f"case u8 2d extreme image={k} rp={rp} rs={rs}", flush=True)
        t8 = torch.from_numpy(img).to(dev)
        for h in (10.0, 200.0):
            y8, Z8 = gnlm.nlm2d_u8(t8, rp, rs, h, want_Z=True)
            yf, Zf = gnlm.nlm2d(t8.float(), rp, rs, h, want_Z=True)
            torch.cuda.synchronize()
            assert torch.equal(y8, yf) and torch.equal(Z8, Zf), (k, h)


@pytest.mark.parametrize("shape,rp,rs", [((70, 90), 5, 10), ((33, 41), 2, 5), ((20, 64), 4, 3)])
def test_u8_2d_strided_padded_input(dev, shape, rp, rs):
    """A padded 8-bit image inside a larger array (rows 9 bytes longer than the
    padded width, two extra rows, the slack 255): the same bits as the
    unpadded image with reflect-101 on the device."""
    img = _img(shape, sum(shape) + rp)
    ny, nx = shape
    pad = rp + rs + 1
    padded = np.pad(img, pad, mode="reflect")
    py, px = padded.shape
    sy = px + 9
    big = np.full((py + 2, sy), 255, np.uint8)
    big[:py, :px] = padded
    print(f"case u8 2d strided shape={shape} rp={rp} rs={rs} sy={sy}", flush=True)
    t = torch.from_numpy(big.ravel()).to(dev)
    y = torch.empty(shape, dtype=torch.float32, device=dev)
    st = gnlm.lib().gnlm2d_forward_u8(ctypes.c_void_p(t.data_ptr()), pad, sy, ctypes.c_void_p(y.data_ptr()), None,
                                      nx, ny, rp, rs, 15.0,
                                      ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    ref = gnlm.nlm2d_u8(torch.from_numpy(img).to(dev), rp, rs, 15.0)
    assert torch.equal(y, ref)


def test_u8_2d_reference_geometry_full_size(dev):
    """The reference's own 2D geometry (patch 11, search 21, h = PARAMETER =
    10) on the 2048^2 8-bit phantom of the benchmark: the FP32 build's bits,
    and the oracle's FP64 separable path to the A + B bar."""
    u, _ = gnlm.phantom2d(2048, 2048, sigma=0.08, seed=20261205)
    img = np.clip(np.rint(u * 255.0), 0, 255).astype(np.uint8)
    print("case u8 2d 2048^2 p11 s21", flush=True)
    t8 = torch.from_numpy(img).to(dev)
    y8 = gnlm.nlm2d_u8(t8, 5, 10, 10.0)
    yf = gnlm.nlm2d(t8.float(), 5, 10, 10.0)
    torch.cuda.synchronize()
    assert torch.equal(y8, yf)
    y = y8.cpu().numpy()
    uf = img.astype(np.float32)
    ref32 = oracle.nlm2d_separable_f32(uf, 5, 10, 10.0)
    ref64, Z64, A64, B64 = oracle.nlm2d_separable_f64(uf, 5, 10, 10.0, with_b=True)
    rep = phantom_parity(y, ref32, ref64, Z64, A64, TOL,
                         underflow=underflow_bound(21 * 21, 255.0, ref64, Z64), B64=B64)
    print("u8 2d 2048", rep, flush=True)
    assert rep["ok"], rep


@pytest.mark.parametrize("shape,rp,rs,h", [((512, 512), 5, 10, 10.0), ((128, 128), 5, 10, 10.0), ((300, 170), 2, 11, 12.0),
                                           ((64, 200), 3, 12, 9.0), ((97, 33), 1, 10, 15.0), ((300, 170), 2, 5, 12.0)])
def test_2d_dy_split_same_bits(dev, shape, rp, rs, h):
    """The dy split of small images at search windows of 21 x 21 and more
    (several workgroups per tile, each writing per-dy partial sums; a combine
    pass adds them in dy order) gives the unsplit launch's bits, y and Z, on
    the 8-bit and the FP32 builds (the register-window build of patch 11 /
    search 21 and the run-time-radius builds; the last case takes no split)."""
    img = _img(shape, sum(shape) + 5 * rp + rs)
    t8 = torch.from_numpy(img).to(dev)
    tf = t8.float()
    outs = []
    try:
        for split in (True, False):
            gnlm.set_dz_split(split)
            print(f"case 2d dy split={split} shape={shape} rp={rp} rs={rs}", flush=True)
            y8, Z8 = gnlm.nlm2d_u8(t8, rp, rs, h, want_Z=True)
            yf, Zf = gnlm.nlm2d(tf, rp, rs, h, want_Z=True)
            y1 = gnlm.nlm2d_u8(t8, rp, rs, h)
            torch.cuda.synchronize()
            outs.append((y8.cpu(), Z8.cpu(), yf.cpu(), Zf.cpu(), y1.cpu()))
    finally:
        gnlm.set_dz_split(True)
    for o in outs:
        for a, b in zip(o, outs[0]):
            assert torch.equal(a, b)
    assert torch.equal(outs[0][0], outs[0][2]) and torch.equal(outs[0][4], outs[0][0])
