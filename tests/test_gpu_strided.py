"""Caller-padded inputs with row and plane strides larger than the dense ones.

The C-ABI takes a padded input with explicit strides (gnlm3d_forward_ex /
gnlm3d_forward_u8: in_pad, in_sy, in_sz; gnlm2d_forward_ex: in_pad, in_sy),
the layout a caller's own buffers may have (the reference's launcher receives
a dense padded block, 3DNLM/main.cu:7613-7620; its 2D kernels take a row
stride, "2D NLM/11x11/main.cu:472-539").  Here the padded volume sits inside a
larger array -- rows and planes longer than the padded extent, the slack
filled with NaN -- and every route must give the same bits as the unpadded
input with reflect-101 on the device (a read of the slack would bring NaN
in)."""
import ctypes

import numpy as np
import pytest
import torch

import gnlm

pytestmark = pytest.mark.gpu


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _embed(padded, extra_x, extra_y, fill):
    """padded [pz, py, px] inside an array with rows of px + extra_x and planes
    of (py + extra_y) rows; returns the flat array and the strides."""
    pz, py, px = padded.shape
    sy, rows = px + extra_x, py + extra_y
    big = np.full((pz, rows, sy), fill, dtype=padded.dtype)
    big[:, :py, :px] = padded
    return big.ravel(), sy, sy * rows


@pytest.mark.parametrize("shape,rp,rs", [((13, 17, 21), 2, 5), ((9, 11, 30), 1, 3), ((12, 20, 18), 6, 2),
                                         ((7, 9, 8), 9, 1), ((10, 14, 12), 4, 3)])
def test_3d_strided_padded_input(dev, shape, rp, rs):
    u = np.random.default_rng(sum(shape) + rp).random(shape, dtype=np.float32)
    nz, ny, nx = shape
    pad = rp + rs + 2
    flat, sy, sz = _embed(np.pad(u, pad, mode="reflect"), 7, 3, np.float32(np.nan))
    print(f"case strided 3d shape={shape} rp={rp} rs={rs} sy={sy} sz={sz}", flush=True)
    ut = torch.from_numpy(flat).to(dev)
    y = torch.empty(shape, dtype=torch.float32, device=dev)
    st = gnlm.lib().gnlm3d_forward_ex(ctypes.c_void_p(ut.data_ptr()), pad, sy, sz, ctypes.c_void_p(y.data_ptr()),
                                      None, nx, ny, nz, rp, rs, 0.2, gnlm.PATH_AUTO, _stream(dev))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    ref = gnlm.nlm3d(torch.from_numpy(u).to(dev), rp, rs, 0.2)
    assert torch.isfinite(y).all() and torch.equal(y, ref)


@pytest.mark.parametrize("shape,rp,rs", [((18, 21, 19), 8, 3), ((16, 23, 17), 6, 4), ((11, 15, 20), 2, 3)])
def test_u8_strided_padded_input(dev, shape, rp, rs):
    img = np.random.default_rng(sum(shape) + rp).integers(0, 256, shape, dtype=np.uint8)
    nz, ny, nx = shape
    pad = rp + rs + 1
    flat, sy, sz = _embed(np.pad(img, pad, mode="reflect"), 9, 2, np.uint8(255))
    print(f"case strided u8 shape={shape} rp={rp} rs={rs} sy={sy} sz={sz}", flush=True)
    t = torch.from_numpy(flat).to(dev)
    y = torch.empty(shape, dtype=torch.float32, device=dev)
    st = gnlm.lib().gnlm3d_forward_u8(ctypes.c_void_p(t.data_ptr()), pad, sy, sz, ctypes.c_void_p(y.data_ptr()),
                                      None, nx, ny, nz, rp, rs, 20.0, _stream(dev))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    ref = gnlm.nlm3d_u8(torch.from_numpy(img).to(dev), rp, rs, 20.0)
    assert torch.equal(y, ref)


@pytest.mark.parametrize("shape,rp,rs", [((70, 90), 5, 10), ((33, 41), 2, 5), ((20, 64), 6, 3)])
def test_2d_strided_padded_input(dev, shape, rp, rs):
    u = np.random.default_rng(sum(shape) + rp).random(shape, dtype=np.float32)
    ny, nx = shape
    pad = rp + rs
    flat, sy, _ = _embed(np.pad(u, pad, mode="reflect")[None], 5, 0, np.float32(np.nan))
    print(f"case strided 2d shape={shape} rp={rp} rs={rs} sy={sy}", flush=True)
    ut = torch.from_numpy(flat).to(dev)
    y = torch.empty(shape, dtype=torch.float32, device=dev)
    st = gnlm.lib().gnlm2d_forward_ex(ctypes.c_void_p(ut.data_ptr()), pad, sy, ctypes.c_void_p(y.data_ptr()), None,
                                      nx, ny, rp, rs, 0.2, gnlm.PATH_AUTO, _stream(dev))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    ref = gnlm.nlm2d(torch.from_numpy(u).to(dev), rp, rs, 0.2)
    assert torch.isfinite(y).all() and torch.equal(y, ref)
