"""gnlm3d_forward_noalloc: the 3D forward that takes its workspace from the
library's fixed per-device scratch array instead of allocating.  Called through
the C-ABI and compared bit for bit with gnlm3d_forward (the same kernels), on
two calls in a row -- the second on a smaller volume, which reuses the array
the first one left dirty -- plus the size and route refusals."""
import ctypes

import numpy as np
import pytest
import torch

import gnlm

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _both(dev, shape, rp, rs, h, want_Z, seed):
    nz, ny, nx = shape
    u, _ = gnlm.phantom3d(nx, ny, nz, n_ellipsoids=6, axis_lo=3, axis_hi=12, sigma=0.05, seed=seed)
    ut = torch.from_numpy(u).to(dev)
    L = gnlm.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    y_ref = torch.full_like(ut, -7.0)
    Z_ref = torch.full_like(ut, -7.0) if want_Z else None
    assert L.gnlm3d_forward(_ptr(ut), _ptr(y_ref), _ptr(Z_ref), nx, ny, nz, rp, rs, h, stream) == gnlm.OK
    y = torch.full_like(ut, -3.0)
    Z = torch.full_like(ut, -3.0) if want_Z else None
    print(f"case noalloc shape={shape} rp={rp} rs={rs} want_Z={want_Z}", flush=True)
    st = L.gnlm3d_forward_noalloc(_ptr(ut), _ptr(y), _ptr(Z), nx, ny, nz, rp, rs, h, stream)
    assert st == gnlm.OK, st
    torch.cuda.synchronize()
    assert torch.equal(y, y_ref)
    if want_Z:
        assert torch.equal(Z, Z_ref)
    return y


def test_noalloc_matches_forward_two_calls(dev):
    # first call: the 96x96x48 configuration, no Z buffer (the rows tiles then
    # write y alone)
    y1 = _both(dev, (48, 96, 96), 1, 3, 0.05, False, 20261204)
    assert torch.isfinite(y1).all()
    # second call, smaller and with other radii, on the array the first left dirty
    _both(dev, (30, 36, 40), 2, 5, 0.05, True, 11)
    # and a third on yet another shape class (big-patch tiles, no Z)
    _both(dev, (20, 24, 28), 8, 2, 0.2, False, 12)


def test_noalloc_capacity_without_Z(dev):
    """The rows tiles keep their totals in registers, so without a Z buffer
    the scratch array holds the padded copy alone: 232^3 at patch 5^3 /
    search 11^3 (a 66 MB copy) fits the 64 MiB array; with a Z buffer the
    caller's own Z takes the totals, so the same volume fits as well."""
    _both(dev, (232, 232, 232), 2, 5, 0.05, False, 13)
    _both(dev, (232, 232, 232), 2, 5, 0.05, True, 14)


def test_noalloc_refusals(dev):
    L = gnlm.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    # 256^3 at patch 5^3 / search 11^3 needs ~77 MB of padded copy: refused
    # before anything runs, the output untouched
    u = torch.zeros((256, 256, 256), device=dev)
    y = torch.full_like(u, -3.0)
    st = L.gnlm3d_forward_noalloc(_ptr(u), _ptr(y), None, 256, 256, 256, 2, 5, 0.05, stream)
    torch.cuda.synchronize()
    assert st == gnlm.ENOMEM
    assert bool((y == -3.0).all())
    # radii only the generic per-offset path handles (it allocates): unsupported
    small = torch.zeros((8, 8, 8), device=dev)
    ys = torch.empty_like(small)
    assert L.gnlm3d_forward_noalloc(_ptr(small), _ptr(ys), None, 8, 8, 8, 9, 2, 0.1, stream) == gnlm.EUNSUPPORTED
    assert L.gnlm3d_forward_noalloc(None, _ptr(ys), None, 8, 8, 8, 1, 2, 0.1, stream) == gnlm.EINVAL
