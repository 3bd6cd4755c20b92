"""Guard bands around every buffer a kernel family touches.

Each entry point runs on ragged shapes with its input inside a larger
allocation whose margins hold NaN (or 255 for 8-bit input), and each output
(y, Z, the gradient, dL/dh, batched outputs) inside a larger allocation whose
margins hold a sentinel.  Bars:
  * the output margins come back holding the sentinel, bit for bit (no write
    past either end of an output);
  * the outputs are finite and bit-identical to the same call on tightly
    allocated buffers (a read past either end of the input would bring the
    NaN / 255 margins into them).
Margins are 64 Ki elements on each side, far beyond any tile or halo.
"""
import ctypes

import numpy as np
import pytest
import torch

import gnlm

pytestmark = pytest.mark.gpu

G = 1 << 16
SENT = -1234.5


class Guarded:
    """n elements at offset G of a buffer of G + n + G elements."""

    def __init__(self, dev, n, dtype=torch.float32, fill=SENT):
        self.n, self.fill = n, fill
        self.buf = torch.full((G + n + G,), fill, dtype=dtype, device=dev)

    def view(self, shape):
        return self.buf[G:G + self.n].view(shape)

    def ptr(self):
        return ctypes.c_void_p(self.buf[G:].data_ptr())

    def margins_intact(self):
        lo, hi = self.buf[:G], self.buf[G + self.n:]
        if isinstance(self.fill, float) and np.isnan(self.fill):
            return bool(torch.isnan(lo).all()) and bool(torch.isnan(hi).all())
        return bool((lo == self.fill).all()) and bool((hi == self.fill).all())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _guarded_input(dev, u, fill=float("nan"), dtype=torch.float32):
    g = Guarded(dev, u.size, dtype=dtype, fill=fill)
    g.view(u.shape).copy_(torch.from_numpy(np.ascontiguousarray(u)).to(dev))
    return g


FWD = [
    # (nz, ny, nx), r_p, r_s: rows tiles (r_p 1-4), big tiles (r_p 5-8), generic (r_p 9)
    ((13, 17, 21), 1, 3),
    ((5, 40, 9), 2, 5),
    ((11, 9, 30), 3, 3),
    ((20, 13, 7), 4, 3),
    ((9, 31, 18), 6, 2),
    ((18, 18, 19), 8, 2),
    ((6, 7, 8), 9, 1),
    ((7, 6, 5), 0, 2),
]


@pytest.mark.parametrize("shape,rp,rs", FWD)
def test_forward_guard_bands(dev, shape, rp, rs):
    u = np.random.default_rng(sum(shape) + rp).random(shape, dtype=np.float32)
    nz, ny, nx = shape
    print(f"case guard forward shape={shape} rp={rp} rs={rs}", flush=True)
    gi = _guarded_input(dev, u)
    gy, gz = Guarded(dev, u.size), Guarded(dev, u.size)
    L = gnlm.lib()
    st = L.gnlm3d_forward(gi.ptr(), gy.ptr(), gz.ptr(), nx, ny, nz, rp, rs, 0.2, _stream(dev))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    assert gi.margins_intact() and gy.margins_intact() and gz.margins_intact()
    y_ref, Z_ref = gnlm.nlm3d(torch.from_numpy(u).to(dev), rp, rs, 0.2, want_Z=True)
    assert torch.isfinite(gy.view(shape)).all() and torch.isfinite(gz.view(shape)).all()
    assert torch.equal(gy.view(shape), y_ref) and torch.equal(gz.view(shape), Z_ref)


def test_forward_variants_guard_bands(dev):
    """Every tile variant (direct, slab, slab2, rows, short rows, big rows)
    for a radius pair they all take."""
    shape, rp, rs = (19, 23, 18), 1, 2
    u = np.random.default_rng(5).random(shape, dtype=np.float32)
    nz, ny, nx = shape
    L = gnlm.lib()
    try:
        for v in range(L.gnlm_num_variants()):
            if L.gnlm_set_variant(rp, rs, v) != gnlm.OK:
                continue
            print(f"case guard variant={v}", flush=True)
            gi = _guarded_input(dev, u)
            gy, gz = Guarded(dev, u.size), Guarded(dev, u.size)
            st = L.gnlm3d_forward(gi.ptr(), gy.ptr(), gz.ptr(), nx, ny, nz, rp, rs, 0.2, _stream(dev))
            torch.cuda.synchronize()
            assert st == gnlm.OK, (v, st)
            assert gi.margins_intact() and gy.margins_intact() and gz.margins_intact(), v
            y_ref = gnlm.nlm3d(torch.from_numpy(u).to(dev), rp, rs, 0.2)
            assert torch.equal(gy.view(shape), y_ref), v
    finally:
        gnlm.reset_tuning()


@pytest.mark.parametrize("shape,rp,rs", [((21, 19, 23), 5, 3), ((17, 33, 20), 6, 2), ((9, 20, 35), 7, 2),
                                         ((18, 17, 16), 8, 2)])
def test_u8_guard_bands(dev, shape, rp, rs):
    img = np.random.default_rng(sum(shape) + rp).integers(0, 256, shape, dtype=np.uint8)
    nz, ny, nx = shape
    print(f"case guard u8 shape={shape} rp={rp} rs={rs}", flush=True)
    assert gnlm.forward_u8_route(nx, ny, nz, rp, rs)[1] == gnlm.KERNEL_U8_INTEGER
    gi = _guarded_input(dev, img, fill=255, dtype=torch.uint8)
    gy, gz = Guarded(dev, img.size), Guarded(dev, img.size)
    st = gnlm.lib().gnlm3d_forward_u8(gi.ptr(), 0, 0, 0, gy.ptr(), gz.ptr(), nx, ny, nz, rp, rs, 30.0,
                                      _stream(dev))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    assert gi.margins_intact() and gy.margins_intact() and gz.margins_intact()
    y_ref, Z_ref = gnlm.nlm3d_u8(torch.from_numpy(img).to(dev), rp, rs, 30.0, want_Z=True)
    assert torch.equal(gy.view(shape), y_ref) and torch.equal(gz.view(shape), Z_ref)


@pytest.mark.parametrize("shape,rp,rs", [((77, 200), 5, 10), ((130, 33), 2, 5), ((9, 70), 1, 3), ((40, 41), 6, 4)])
def test_2d_guard_bands(dev, shape, rp, rs):
    u = np.random.default_rng(sum(shape) + rp).random(shape, dtype=np.float32)
    ny, nx = shape
    print(f"case guard 2d shape={shape} rp={rp} rs={rs}", flush=True)
    gi = _guarded_input(dev, u)
    gy, gz = Guarded(dev, u.size), Guarded(dev, u.size)
    st = gnlm.lib().gnlm2d_forward_ex(gi.ptr(), 0, 0, gy.ptr(), gz.ptr(), nx, ny, rp, rs, 0.2, gnlm.PATH_AUTO,
                                      _stream(dev))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    assert gi.margins_intact() and gy.margins_intact() and gz.margins_intact()
    y_ref = gnlm.nlm2d(torch.from_numpy(u).to(dev), rp, rs, 0.2)
    assert torch.equal(gy.view(shape), y_ref)


@pytest.mark.parametrize("shape,rp,rs", [((77, 200), 5, 10), ((130, 33), 2, 5), ((40, 41), 4, 3), ((31, 20), 6, 2)])
def test_2d_u8_guard_bands(dev, shape, rp, rs):
    img = np.random.default_rng(sum(shape) + rp).integers(0, 256, shape, dtype=np.uint8)
    ny, nx = shape
    print(f"case guard 2d u8 shape={shape} rp={rp} rs={rs}", flush=True)
    gi = _guarded_input(dev, img, fill=255, dtype=torch.uint8)
    gy, gz = Guarded(dev, img.size), Guarded(dev, img.size)
    st = gnlm.lib().gnlm2d_forward_u8(gi.ptr(), 0, 0, gy.ptr(), gz.ptr(), nx, ny, rp, rs, 20.0, _stream(dev))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    assert gi.margins_intact() and gy.margins_intact() and gz.margins_intact()
    y_ref, Z_ref = gnlm.nlm2d_u8(torch.from_numpy(img).to(dev), rp, rs, 20.0, want_Z=True)
    assert torch.equal(gy.view(shape), y_ref) and torch.equal(gz.view(shape), Z_ref)


@pytest.mark.parametrize("shape,rp,rs", [((15, 13, 17), 1, 3), ((11, 19, 9), 2, 3), ((9, 8, 10), 0, 2),
                                         ((12, 10, 11), 3, 2), ((7, 8, 6), 4, 1)])
def test_backward_guard_bands(dev, shape, rp, rs):
    rng = np.random.default_rng(sum(shape) + 3 * rp)
    u = rng.random(shape, dtype=np.float32)
    go = rng.standard_normal(shape).astype(np.float32)
    nz, ny, nx = shape
    h = 0.2
    print(f"case guard backward shape={shape} rp={rp} rs={rs}", flush=True)
    ut = torch.from_numpy(u).to(dev)
    y, Z = gnlm.nlm3d(ut, rp, rs, h, want_Z=True)
    gi, gyb, gzb, ggo = (_guarded_input(dev, a) for a in (u, y.cpu().numpy(), Z.cpu().numpy(), go))
    ggin, ggh = Guarded(dev, u.size), Guarded(dev, 1)
    st = gnlm.lib().gnlm3d_backward(gi.ptr(), gyb.ptr(), gzb.ptr(), ggo.ptr(), nx, ny, nz, rp, rs, h, ggin.ptr(),
                                    ggh.ptr(), _stream(dev))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    assert all(g.margins_intact() for g in (gi, gyb, gzb, ggo, ggin, ggh))
    gin_ref, gh_ref = gnlm.nlm3d_backward(ut, y, Z, torch.from_numpy(go).to(dev), rp, rs, h)
    assert torch.isfinite(ggin.view(shape)).all()
    assert torch.equal(ggin.view(shape), gin_ref)
    assert float(ggh.view((1,))[0]) == float(gh_ref)


def test_batch_guard_bands(dev):
    """Batched outputs, each inside its own guarded buffer, through the graph
    replay and the eager launch."""
    shapes = [((21, 30, 17), 1, 3, 0.1), ((33, 16, 25), 2, 5, 0.06), ((10, 12, 40), 2, 3, 0.08)]
    rng = np.random.default_rng(11)
    us = [rng.random(s, dtype=np.float32) for s, *_ in shapes]
    ins = [torch.from_numpy(u).to(dev) for u in us]
    outs = [Guarded(dev, u.size) for u in us]
    plan = gnlm.BatchPlan([(ins[i], outs[i].view(us[i].shape), rp, rs, h)
                           for i, (s, rp, rs, h) in enumerate(shapes)])
    try:
        for eager in (False, True):
            print(f"case guard batch eager={eager}", flush=True)
            for g in outs:
                g.view((g.n,)).fill_(0.0)
            plan.run(eager=eager)
            torch.cuda.synchronize()
            for i, (s, rp, rs, h) in enumerate(shapes):
                assert outs[i].margins_intact(), (eager, i)
                assert torch.equal(outs[i].view(s), gnlm.nlm3d(ins[i], rp, rs, h)), (eager, i)
    finally:
        plan.close()


def test_noalloc_guard_bands(dev):
    shape, rp, rs = (23, 29, 31), 2, 5
    u = np.random.default_rng(17).random(shape, dtype=np.float32)
    nz, ny, nx = shape
    gi = _guarded_input(dev, u)
    gy, gz = Guarded(dev, u.size), Guarded(dev, u.size)
    st = gnlm.lib().gnlm3d_forward_noalloc(gi.ptr(), gy.ptr(), gz.ptr(), nx, ny, nz, rp, rs, 0.2, _stream(dev))
    torch.cuda.synchronize()
    assert st == gnlm.OK, st
    assert gi.margins_intact() and gy.margins_intact() and gz.margins_intact()
    y_ref, Z_ref = gnlm.nlm3d(torch.from_numpy(u).to(dev), rp, rs, 0.2, want_Z=True)
    assert torch.equal(gy.view(shape), y_ref) and torch.equal(gz.view(shape), Z_ref)
