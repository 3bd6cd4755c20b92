"""CPU checks of the C-ABI library: it loads, exports every symbol the public
headers declare, and rejects bad arguments before touching the GPU."""
import ctypes
import os
import re

import pytest

import gnlm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # C-ABI functions are the lower-case gnlm_* / gnlm2d_* / gnlm3d_* names
        for m in re.finditer(r"\b(gnlm[0-9a-z_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_something():
    assert "gnlm3d_forward" in _declared_functions()


@pytest.mark.parametrize("name", _declared_functions())
def test_symbol_exported(name):
    L = ctypes.CDLL(gnlm.LIB_PATH)
    assert hasattr(L, name), f"{name} declared in include/ but not exported by libgnlm.so"


def test_status_strings():
    L = gnlm.lib()
    for s in range(6):
        assert L.gnlm_status_string(s)
    assert L.gnlm_version() >= 100


def test_invalid_arguments_rejected_without_gpu():
    L = gnlm.lib()
    fake = ctypes.c_void_p(0x1000)
    # NULL input
    assert L.gnlm3d_forward(None, fake, None, 8, 8, 8, 1, 2, 0.1, None) == gnlm.EINVAL
    # bad dims / radii / h
    assert L.gnlm3d_forward(fake, fake, None, 0, 8, 8, 1, 2, 0.1, None) == gnlm.EINVAL
    assert L.gnlm3d_forward(fake, fake, None, 8, 8, 8, -1, 2, 0.1, None) == gnlm.EINVAL
    assert L.gnlm3d_forward(fake, fake, None, 8, 8, 8, 1, 2, 0.0, None) == gnlm.EINVAL
    assert L.gnlm3d_forward(fake, fake, None, 8, 8, 8, 1, 2, float("nan"), None) == gnlm.EINVAL
    assert L.gnlm2d_forward(fake, fake, 8, -3, 1, 2, 0.1, None) == gnlm.EINVAL
    # padding smaller than the halo
    assert L.gnlm3d_forward_ex(fake, 2, 0, 0, fake, None, 8, 8, 8, 1, 2, 0.1, 0, None) == gnlm.EINVAL
    # unknown path
    assert L.gnlm3d_forward_ex(fake, 0, 0, 0, fake, None, 8, 8, 8, 1, 2, 0.1, 7, None) == gnlm.EINVAL
    # tiled path for radii it does not implement (patch radius 9, search radius 8 at patch radius 8)
    assert L.gnlm3d_forward_ex(fake, 0, 0, 0, fake, None, 8, 8, 8, 9, 6, 0.1, 1, None) == gnlm.EUNSUPPORTED
    assert L.gnlm3d_forward_ex(fake, 0, 0, 0, fake, None, 8, 8, 8, 8, 8, 0.1, 1, None) == gnlm.EUNSUPPORTED
    # even window sizes in the drop-in
    buf = (ctypes.c_ubyte * 16)()
    out = (ctypes.c_float * 16)()
    assert L.gnlm_nl_means_3d_u8(buf, out, 4, 4, 4, 4, 3, 10.0) == gnlm.EINVAL
    assert L.gnlm_nl_means_2d_u8(buf, out, 4, 4, 3, 4, 2, 2, 10.0) == gnlm.EINVAL
    # the dz-split switch takes 0 or 1
    assert L.gnlm_set_dz_split(2) == gnlm.EINVAL and L.gnlm_set_dz_split(-1) == gnlm.EINVAL
    assert L.gnlm_set_dz_split(1) == gnlm.OK
    # the 8-bit 2D entry: NULL input or output, bad h, padding smaller than the halo
    assert L.gnlm2d_forward_u8(None, 0, 0, fake, None, 8, 8, 1, 2, 10.0, None) == gnlm.EINVAL
    assert L.gnlm2d_forward_u8(fake, 0, 0, None, None, 8, 8, 1, 2, 10.0, None) == gnlm.EINVAL
    assert L.gnlm2d_forward_u8(fake, 0, 0, fake, None, 8, 8, 1, 2, 0.0, None) == gnlm.EINVAL
    assert L.gnlm2d_forward_u8(fake, 2, 12, fake, None, 8, 8, 1, 2, 10.0, None) == gnlm.EINVAL


def test_phantom_deterministic():
    a, ca = gnlm.phantom3d(20, 18, 16, n_ellipsoids=4, axis_lo=2, axis_hi=6, seed=42)
    b, cb = gnlm.phantom3d(20, 18, 16, n_ellipsoids=4, axis_lo=2, axis_hi=6, seed=42)
    assert gnlm.hash_f32(a) == gnlm.hash_f32(b)
    assert (ca == cb).all()
    c, _ = gnlm.phantom3d(20, 18, 16, n_ellipsoids=4, axis_lo=2, axis_hi=6, seed=43)
    assert gnlm.hash_f32(a) != gnlm.hash_f32(c)
    vals = set(ca.ravel().tolist())
    assert 0.05 in {round(v, 6) for v in vals}  # background present
    assert all(0.05 - 1e-6 <= v <= 0.9 + 1e-6 for v in vals)
