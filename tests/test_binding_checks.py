"""The Python front end refuses buffers the kernels would overrun, before any
launch: shapes are checked first, so these run on the CPU with CPU tensors
(a shape mismatch raises ValueError; a correct shape then fails the device
check with TypeError, which shows that no kernel could have been reached)."""
import pytest
import torch

import gnlm


def _t(*shape):
    return torch.zeros(shape, dtype=torch.float32)


def test_nlm3d_checks():
    with pytest.raises(ValueError):
        gnlm.nlm3d(_t(8, 8), 1, 2, 0.1)                      # not 3-D
    with pytest.raises(ValueError):
        gnlm.nlm3d(_t(0, 8, 8), 1, 2, 0.1)                   # empty
    with pytest.raises(ValueError):
        gnlm.nlm3d(_t(8, 8, 8), 1, 2, 0.1, out=_t(8, 8, 7))  # out too small
    with pytest.raises(TypeError):
        gnlm.nlm3d(_t(8, 8, 8), 1, 2, 0.1, out=_t(8, 8, 8))  # right shape, not on the GPU


def test_nlm2d_checks():
    with pytest.raises(ValueError):
        gnlm.nlm2d(_t(4, 8, 8), 1, 2, 0.1)
    with pytest.raises(TypeError):
        gnlm.nlm2d(_t(8, 8), 1, 2, 0.1)


@pytest.mark.parametrize("which", ["y", "Z", "gy"])
def test_backward_checks(which):
    bufs = {"y": _t(6, 7, 8), "Z": _t(6, 7, 8), "gy": _t(6, 7, 8)}
    bufs[which] = _t(6, 7, 7)
    with pytest.raises(ValueError):
        gnlm.nlm3d_backward(_t(6, 7, 8), bufs["y"], bufs["Z"], bufs["gy"], 1, 2, 0.1)
    with pytest.raises(TypeError):
        gnlm.nlm3d_backward(_t(6, 7, 8), _t(6, 7, 8), _t(6, 7, 8), _t(6, 7, 8), 1, 2, 0.1)


def test_padded_checks():
    pad = 3
    need = (4 + 2 * pad) * (5 + 2 * pad) * (6 + 2 * pad)
    with pytest.raises(ValueError):
        gnlm.nlm3d_padded(_t(need - 1), pad, 6, 5, 4, 1, 2, 0.1)
    with pytest.raises(ValueError):
        gnlm.nlm3d_padded(_t(need), -1, 6, 5, 4, 1, 2, 0.1)
    with pytest.raises(TypeError):
        gnlm.nlm3d_padded(_t(need), pad, 6, 5, 4, 1, 2, 0.1)


def test_batch_plan_checks():
    with pytest.raises(ValueError):
        gnlm.BatchPlan([])
    with pytest.raises(ValueError):
        gnlm.BatchPlan([(_t(8, 8, 8), _t(8, 8, 8), 1, 3, 0.1), (_t(8, 8, 9), _t(8, 9, 8), 1, 3, 0.1)])
    with pytest.raises(TypeError):
        gnlm.BatchPlan([(_t(8, 8, 8), _t(8, 8, 8), 1, 3, 0.1)])


def test_autograd_checks():
    with pytest.raises(ValueError):
        gnlm.nlm3d_autograd(_t(8, 8), 0.1, 1, 2)
    with pytest.raises(TypeError):
        gnlm.nlm3d_autograd(_t(8, 8, 8), 0.1, 1, 2)


@pytest.mark.parametrize("fn,ndim", [(gnlm.nlm3d_u8, 3), (gnlm.nlm2d_u8, 2)])
def test_u8_checks(fn, ndim):
    u8 = lambda *s: torch.zeros(s, dtype=torch.uint8)  # noqa: E731
    with pytest.raises(ValueError):
        fn(u8(*([5] * (ndim + 1))), 1, 2, 10.0)             # wrong rank
    with pytest.raises(ValueError):
        fn(u8(*([0] + [5] * (ndim - 1))), 1, 2, 10.0)       # empty
    with pytest.raises(TypeError):
        fn(_t(*([5] * ndim)), 1, 2, 10.0)                    # float32, not uint8
    with pytest.raises(TypeError):
        fn(u8(*([5] * ndim)), 1, 2, 10.0)                    # right shape, not on the GPU
