"""ctypes front end of the CPU oracle (oracle/nlm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the thing
measured or shipped.  The C file's header lists the reference lines each
function restates (3DNLM/nlm.cpp:36-61, 3DNLM/main.cu:52-54,58-68,107-111,125,
"2D NLM/11x11/main.cu:36-63").
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i = ctypes.c_int


def build(force=False):
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.oracle_reflect101.restype = ctypes.c_long
        L.oracle_reflect101.argtypes = [ctypes.c_long, ctypes.c_long]
        L.oracle_nlm3d_brute_f64.argtypes = [_f64p, _i, _i, _i, _i, _i, ctypes.c_double, _f64p, ctypes.c_void_p]
        L.oracle_nlm3d_brute_f32.argtypes = [_f32p, _i, _i, _i, _i, _i, ctypes.c_float, _f32p, ctypes.c_void_p]
        L.oracle_nlm2d_brute_f64.argtypes = [_f64p, _i, _i, _i, _i, ctypes.c_double, _f64p, ctypes.c_void_p]
        L.oracle_nlm2d_brute_f32.argtypes = [_f32p, _i, _i, _i, _i, ctypes.c_float, _f32p, ctypes.c_void_p]
        L.oracle_nlm3d_separable_f32.argtypes = [_f32p, _i, _i, _i, _i, _i, ctypes.c_float, _f32p, _i, _i, _i]
        L.oracle_nlm3d_separable_f32.restype = _i
        L.oracle_nlm3d_separable_f64.argtypes = [_f32p, _i, _i, _i, _i, _i, ctypes.c_double, _f64p, _i, _i, _i]
        L.oracle_nlm3d_separable_f64.restype = _i
        _vp = ctypes.c_void_p
        L.oracle_nlm3d_separable_f64_za.argtypes = [_f32p, _i, _i, _i, _i, _i, ctypes.c_double, _f64p, _vp, _vp,
                                                    _vp, _i, _i, _i]
        L.oracle_nlm3d_separable_f64_za.restype = _i
        L.oracle_nlm2d_separable_f64.argtypes = [_f32p, _i, _i, _i, _i, ctypes.c_double, _f64p, _vp, _vp, _vp, _i]
        L.oracle_nlm2d_separable_f64.restype = _i
        L.oracle_nlm2d_separable_f32.argtypes = [_f32p, _i, _i, _i, _i, ctypes.c_float, _f32p, _i]
        L.oracle_nlm2d_separable_f32.restype = _i
        L.oracle_nlm3d_grad_f64.argtypes = [_f64p, _f64p, _i, _i, _i, _i, _i, ctypes.c_double, _f64p,
                                            ctypes.POINTER(ctypes.c_double)]
        L.oracle_nlm3d_mse_f64.argtypes = [_f64p, _f64p, _i, _i, _i, _i, _i, ctypes.c_double]
        L.oracle_nlm3d_mse_f64.restype = ctypes.c_double
        _lib = L
    return _lib


def reflect101(i, n):
    return lib().oracle_reflect101(int(i), int(n))


def _zout(Z, dtype):
    if Z is None:
        return None
    assert Z.dtype == dtype and Z.flags.c_contiguous
    return Z.ctypes.data_as(ctypes.c_void_p)


def nlm3d_brute(u, rp, rs, h, want_Z=False):
    """Brute-force 3D NLM of u[z, y, x]; FP64 if u is float64 else FP32."""
    nz, ny, nx = u.shape
    if u.dtype == np.float64:
        y = np.empty_like(u)
        Z = np.empty_like(u) if want_Z else None
        lib().oracle_nlm3d_brute_f64(np.ascontiguousarray(u), nx, ny, nz, rp, rs, float(h), y, _zout(Z, np.float64))
    else:
        u = np.ascontiguousarray(u, dtype=np.float32)
        y = np.empty_like(u)
        Z = np.empty_like(u) if want_Z else None
        lib().oracle_nlm3d_brute_f32(u, nx, ny, nz, rp, rs, float(h), y, _zout(Z, np.float32))
    return (y, Z) if want_Z else y


def nlm2d_brute(u, rp, rs, h, want_Z=False):
    ny, nx = u.shape
    if u.dtype == np.float64:
        y = np.empty_like(u)
        Z = np.empty_like(u) if want_Z else None
        lib().oracle_nlm2d_brute_f64(np.ascontiguousarray(u), nx, ny, rp, rs, float(h), y, _zout(Z, np.float64))
    else:
        u = np.ascontiguousarray(u, dtype=np.float32)
        y = np.empty_like(u)
        Z = np.empty_like(u) if want_Z else None
        lib().oracle_nlm2d_brute_f32(u, nx, ny, rp, rs, float(h), y, _zout(Z, np.float32))
    return (y, Z) if want_Z else y


def nlm3d_separable_f32(u, rp, rs, h, z0=0, z1=None, nthreads=0):
    """The timed CPU path: separable FP32, OpenMP over z-slabs.

    Returns planes [z0, z1) of the output.  Same canonical FP32 order as
    nlm3d_brute on float32 input."""
    u = np.ascontiguousarray(u, dtype=np.float32)
    nz, ny, nx = u.shape
    z1 = nz if z1 is None else z1
    y = np.empty((z1 - z0, ny, nx), np.float32)
    rc = lib().oracle_nlm3d_separable_f32(u, nx, ny, nz, rp, rs, float(h), y, z0, z1, int(nthreads))
    if rc != 0:
        raise MemoryError("oracle_nlm3d_separable_f32 failed")
    return y


def nlm3d_separable_f64(u, rp, rs, h, z0=0, z1=None, nthreads=0):
    """FP64 re-run of planes [z0, z1) (input given as float32)."""
    u = np.ascontiguousarray(u, dtype=np.float32)
    nz, ny, nx = u.shape
    z1 = nz if z1 is None else z1
    y = np.empty((z1 - z0, ny, nx), np.float64)
    rc = lib().oracle_nlm3d_separable_f64(u, nx, ny, nz, rp, rs, float(h), y, z0, z1, int(nthreads))
    if rc != 0:
        raise MemoryError("oracle_nlm3d_separable_f64 failed")
    return y


def nlm3d_separable_f64_za(u, rp, rs, h, z0=0, z1=None, nthreads=0, with_b=False):
    """FP64 re-run of planes [z0, z1) with the normaliser Z = sum_d w_d and
    A = sum_d w_d |u~(x+d)|: returns (y, Z, A), float64; with_b also B =
    sum_d w_d a_d (|u~(x+d)| + |y|), a_d the exponent of w_d: (y, Z, A, B)."""
    u = np.ascontiguousarray(u, dtype=np.float32)
    nz, ny, nx = u.shape
    z1 = nz if z1 is None else z1
    y = np.empty((z1 - z0, ny, nx), np.float64)
    Z = np.empty_like(y)
    A = np.empty_like(y)
    B = np.empty_like(y) if with_b else None
    rc = lib().oracle_nlm3d_separable_f64_za(u, nx, ny, nz, rp, rs, float(h), y, Z.ctypes.data, A.ctypes.data,
                                             B.ctypes.data if with_b else None, z0, z1, int(nthreads))
    if rc != 0:
        raise MemoryError("oracle_nlm3d_separable_f64_za failed")
    return (y, Z, A, B) if with_b else (y, Z, A)


def nlm2d_separable_f64(u, rp, rs, h, nthreads=0, with_b=False):
    """2D FP64 separable path: returns (y, Z, A), or (y, Z, A, B) with with_b,
    as nlm3d_separable_f64_za."""
    u = np.ascontiguousarray(u, dtype=np.float32)
    ny, nx = u.shape
    y = np.empty((ny, nx), np.float64)
    Z = np.empty_like(y)
    A = np.empty_like(y)
    B = np.empty_like(y) if with_b else None
    rc = lib().oracle_nlm2d_separable_f64(u, nx, ny, rp, rs, float(h), y, Z.ctypes.data, A.ctypes.data,
                                          B.ctypes.data if with_b else None, int(nthreads))
    if rc != 0:
        raise MemoryError("oracle_nlm2d_separable_f64 failed")
    return (y, Z, A, B) if with_b else (y, Z, A)


def nlm2d_separable_f32(u, rp, rs, h, nthreads=0):
    u = np.ascontiguousarray(u, dtype=np.float32)
    ny, nx = u.shape
    y = np.empty_like(u)
    rc = lib().oracle_nlm2d_separable_f32(u, nx, ny, rp, rs, float(h), y, int(nthreads))
    if rc != 0:
        raise MemoryError("oracle_nlm2d_separable_f32 failed")
    return y


def nlm3d_grad_f64(u, g, rp, rs, h):
    """Analytic adjoint (FP64 chain rule): returns (dL/du, dL/dh)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    nz, ny, nx = u.shape
    gu = np.empty_like(u)
    gh = ctypes.c_double(0.0)
    lib().oracle_nlm3d_grad_f64(u, g, nx, ny, nz, rp, rs, float(h), gu, ctypes.byref(gh))
    return gu, gh.value


def nlm3d_mse_f64(u, clean, rp, rs, h):
    u = np.ascontiguousarray(u, dtype=np.float64)
    clean = np.ascontiguousarray(clean, dtype=np.float64)
    nz, ny, nx = u.shape
    return lib().oracle_nlm3d_mse_f64(u, clean, nx, ny, nz, rp, rs, float(h))
