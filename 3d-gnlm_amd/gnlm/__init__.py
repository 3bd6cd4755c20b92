"""Python front end of libgnlm.so (the MI355X NLM library), over ctypes.

The library's C-ABI (include/gnlm.h) takes plain device pointers and a
hipStream_t; this module passes torch tensors' data pointers and torch's
current stream.  torch is only the device-memory and stream plumbing here:
every computation runs in the HIP kernels of libgnlm.so, and nothing falls
back to a CPU or torch implementation -- if the library is missing the
calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgnlm.so")

OK, EINVAL, ENOMEM, EUNSUPPORTED, EHIP, ESTATE = range(6)
PATH_AUTO, PATH_TILED, PATH_GENERIC = 0, 1, 2

_lib = None


class GnlmError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = lib().gnlm_status_string(status).decode() if _lib is not None else str(status)
        super().__init__(f"{where}: {msg} (status {status})")


def lib():
    """Load libgnlm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libgnlm.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, f, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_longlong
    L.gnlm_status_string.restype = ctypes.c_char_p
    L.gnlm_status_string.argtypes = [i]
    L.gnlm_version.restype = i
    L.gnlm3d_forward.argtypes = [vp, vp, vp, i, i, i, i, i, f, vp]
    L.gnlm3d_forward_ex.argtypes = [vp, i, ll, ll, vp, vp, i, i, i, i, i, f, i, vp]
    L.gnlm2d_forward.argtypes = [vp, vp, i, i, i, i, f, vp]
    L.gnlm2d_forward_ex.argtypes = [vp, i, ll, vp, vp, i, i, i, i, f, i, vp]
    L.gnlm_u8_to_f32.argtypes = [vp, vp, ll, vp]
    L.gnlm3d_forward_u8.argtypes = [vp, i, ll, ll, vp, vp, i, i, i, i, i, f, vp]
    L.gnlm2d_forward_u8.argtypes = [vp, i, ll, vp, vp, i, i, i, i, f, vp]
    L.gnlm_set_dz_split.argtypes = [i]
    L.gnlm_set_dz_split.restype = i
    L.gnlm_nl_means_3d_u8.argtypes = [vp, vp, i, i, i, i, i, f]
    L.gnlm_nl_means_2d_u8.argtypes = [vp, vp, i, i, i, i, i, i, f]
    L.gnlm_phantom3d.argtypes = [vp, vp, i, i, i, i, f, f, f, f, ctypes.c_ulonglong]
    L.gnlm_phantom2d.argtypes = [vp, vp, i, i, i, f, f, f, f, f, f, ctypes.c_ulonglong]
    L.gnlm_hash_f32.restype = ctypes.c_ulonglong
    L.gnlm_hash_f32.argtypes = [vp, ll]
    for name in ("gnlm3d_forward", "gnlm3d_forward_ex", "gnlm2d_forward", "gnlm2d_forward_ex",
                 "gnlm_u8_to_f32", "gnlm3d_forward_u8", "gnlm2d_forward_u8", "gnlm_nl_means_3d_u8", "gnlm_nl_means_2d_u8", "gnlm_phantom3d",
                 "gnlm_phantom2d"):
        getattr(L, name).restype = i
    _extend(L)
    _lib = L
    return L


def _extend(L):
    """Argument types of the backward / batch / tuning entry points."""
    vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    if hasattr(L, "gnlm3d_backward"):
        L.gnlm3d_backward.argtypes = [vp, vp, vp, vp, i, i, i, i, i, f, vp, vp, vp]
        L.gnlm3d_backward.restype = i
    L.gnlm_batch_plan_create.argtypes = [ctypes.POINTER(BatchVolume), i, ctypes.POINTER(vp)]
    L.gnlm_batch_run.argtypes = [vp, vp]
    L.gnlm_batch_run_eager.argtypes = [vp, vp]
    L.gnlm_batch_plan_info.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i)]
    L.gnlm_batch_plan_destroy.argtypes = [vp]
    L.gnlm_autotune.argtypes = [i, i, i, vp, ctypes.POINTER(i), ctypes.POINTER(f)]
    L.gnlm_tuned_variant.argtypes = [i, i]
    L.gnlm_set_variant.argtypes = [i, i, i]
    L.gnlm_reset_tuning.argtypes = []
    L.gnlm_reset_tuning.restype = None
    for name in ("gnlm_batch_plan_create", "gnlm_batch_run", "gnlm_batch_run_eager", "gnlm_batch_plan_info",
                 "gnlm_batch_plan_destroy", "gnlm_autotune", "gnlm_tuned_variant", "gnlm_set_variant",
                 "gnlm_num_variants", "gnlm_set_backward_form"):
        getattr(L, name).restype = i
    L.gnlm_set_backward_form.argtypes = [i]
    L.gnlm_set_backward_parts.argtypes = [i]
    L.gnlm_set_backward_parts.restype = i
    L.gnlm_set_tuning.argtypes = [i]
    L.gnlm_set_tuning.restype = i
    L.gnlm_forward_shape.argtypes = [i, i, i, i, i, ctypes.POINTER(i), ctypes.POINTER(i)]
    L.gnlm_forward_shape.restype = i
    L.gnlm_set_forward_shape.argtypes = [i, i, i, i, i, i]
    L.gnlm_set_forward_shape.restype = i
    L.gnlm3d_forward_noalloc.argtypes = [vp, vp, vp, i, i, i, i, i, f, vp]
    L.gnlm3d_forward_noalloc.restype = i
    L.gnlm3d_forward_route.argtypes = [i, i, i, i, i, i, ctypes.POINTER(i)]
    L.gnlm3d_forward_route.restype = i
    L.gnlm3d_backward_route.argtypes = [i, i, i, i, i, ctypes.POINTER(i)]
    L.gnlm3d_backward_route.restype = i
    L.gnlm3d_forward_u8_route.argtypes = [i, i, i, i, i, ctypes.POINTER(i)]
    L.gnlm3d_forward_u8_route.restype = i


class BatchVolume(ctypes.Structure):
    """Mirror of gnlm_batch_volume (include/gnlm.h)."""
    _fields_ = [("d_in", ctypes.c_void_p), ("d_out", ctypes.c_void_p), ("d_Z", ctypes.c_void_p),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("r_p", ctypes.c_int), ("r_s", ctypes.c_int), ("h", ctypes.c_float)]


class BatchPlan:
    """Persistent, graph-captured forward over many volumes.

    vols: list of (u, out, r_p, r_s, h) with u/out float32 CUDA tensors
    [nz, ny, nx].  The tensors must stay alive while the plan is used."""

    def __init__(self, vols):
        arr = (BatchVolume * len(vols))()
        self._keep = []
        if len(vols) == 0:
            raise ValueError("BatchPlan needs at least one volume")
        for i, (u, out, rp, rs, h) in enumerate(vols):
            _check_vol(u, 3, f"vols[{i}].u")
            _check_like(out, u, f"vols[{i}].out", f"vols[{i}].u")
        _check_dev([t for u, out, *_ in vols for t in (u, out)],
                   [f"vols[{i}].{n}" for i in range(len(vols)) for n in ("u", "out")])
        for i, (u, out, rp, rs, h) in enumerate(vols):
            nz, ny, nx = u.shape
            arr[i] = BatchVolume(_dev_f32(u, "u").value, _dev_f32(out, "out").value, None, nx, ny, nz,
                                 int(rp), int(rs), float(h))
            self._keep += [u, out]
        self._plan = ctypes.c_void_p()
        check(lib().gnlm_batch_plan_create(arr, len(vols), ctypes.byref(self._plan)), "gnlm_batch_plan_create")
        self._dev = vols[0][0]

    def run(self, eager=False):
        fn = lib().gnlm_batch_run_eager if eager else lib().gnlm_batch_run
        check(fn(self._plan, _stream(self._dev)), "gnlm_batch_run")

    def info(self):
        nt, nb = ctypes.c_int(), ctypes.c_int()
        check(lib().gnlm_batch_plan_info(self._plan, ctypes.byref(nt), ctypes.byref(nb)), "info")
        return nt.value, nb.value

    def close(self):
        if self._plan:
            lib().gnlm_batch_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(r_p, r_s, probe_n=64, device_tensor=None):
    """Time the tile variants for (r_p, r_s); returns (chosen variant, [ms per variant])."""
    import torch
    L = lib()
    nv = L.gnlm_num_variants()
    ms = (ctypes.c_float * nv)()
    v = ctypes.c_int(-1)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(L.gnlm_autotune(int(r_p), int(r_s), int(probe_n), stream, ctypes.byref(v), ms), "gnlm_autotune")
    return v.value, list(ms)


def set_variant(r_p, r_s, variant):
    check(lib().gnlm_set_variant(int(r_p), int(r_s), int(variant)), "gnlm_set_variant")


def tuned_variant(r_p, r_s):
    """The variant the tiled 3D forward uses for (r_p, r_s); -1 if none fits."""
    return int(lib().gnlm_tuned_variant(int(r_p), int(r_s)))


def reset_tuning():
    lib().gnlm_reset_tuning()


(VARIANT_16x16x16, VARIANT_16x16x8, VARIANT_SLAB, VARIANT_SLAB2, VARIANT_ROWS, VARIANT_ROWS_SHORT,
 VARIANT_ROWS_BIG) = 0, 1, 2, 3, 4, 5, 6

BWD_AUTO, BWD_GATHER, BWD_SYMMETRIC = 0, 1, 2
KERNEL_GENERIC, KERNEL_BWD_ROWS_GATHER, KERNEL_BWD_ROWS_SYMMETRIC, KERNEL_BWD_BLOCK, KERNEL_BWD_GENERIC = (
    100, 201, 202, 203, 204)


def forward_route(nx, ny, nz, r_p, r_s, path=PATH_AUTO):
    """(status, kernel) a 3D forward of this geometry takes (gnlm3d_forward_route;
    host logic only): kernel is the tile variant 0-6 or KERNEL_GENERIC."""
    k = ctypes.c_int(-1)
    st = lib().gnlm3d_forward_route(int(nx), int(ny), int(nz), int(r_p), int(r_s), int(path), ctypes.byref(k))
    return st, k.value


def backward_route(nx, ny, nz, r_p, r_s):
    """(status, kernel) of the 3D adjoint (gnlm3d_backward_route; host logic only)."""
    k = ctypes.c_int(-1)
    st = lib().gnlm3d_backward_route(int(nx), int(ny), int(nz), int(r_p), int(r_s), ctypes.byref(k))
    return st, k.value


def set_backward_form(form):
    """Process-wide form of the rows-organised adjoint, r_p 0-2 (gnlm_set_backward_form)."""
    check(lib().gnlm_set_backward_form(int(form)), "gnlm_set_backward_form")


TUNE_LAZY, TUNE_OFF = 0, 1


def set_tuning(mode):
    """TUNE_LAZY (default): time the bit-identical launch shapes at first use per
    volume size class; TUNE_OFF: the default rule (gnlm_set_tuning)."""
    check(lib().gnlm_set_tuning(int(mode)), "gnlm_set_tuning")


def forward_shape(shape3, r_p, r_s):
    """(shape recorded for this volume's class or -1, number of shapes);
    shape3 = (nz, ny, nx)."""
    nz, ny, nx = (int(v) for v in shape3)
    sh, n = ctypes.c_int(), ctypes.c_int()
    check(lib().gnlm_forward_shape(nx, ny, nz, int(r_p), int(r_s), ctypes.byref(sh), ctypes.byref(n)),
          "gnlm_forward_shape")
    return sh.value, n.value


def set_forward_shape(shape3, r_p, r_s, shape):
    nz, ny, nx = (int(v) for v in shape3)
    check(lib().gnlm_set_forward_shape(nx, ny, nz, int(r_p), int(r_s), int(shape)), "gnlm_set_forward_shape")


def set_backward_parts(nparts):
    """Workgroups per tile of the symmetric adjoint: 0 = automatic, 1-4 forced
    (gnlm_set_backward_parts)."""
    check(lib().gnlm_set_backward_parts(int(nparts)), "gnlm_set_backward_parts")


def set_dz_split(enable):
    """dz-split forward of small volumes on (True, the default) or off
    (gnlm_set_dz_split; same bits either way)."""
    check(lib().gnlm_set_dz_split(1 if enable else 0), "gnlm_set_dz_split")


def check(status, where=""):
    if status != OK:
        raise GnlmError(status, where)


def _stream(t):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _dev_f32(t, name):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise TypeError(f"{name} must be a contiguous float32 CUDA tensor")
    return ctypes.c_void_p(t.data_ptr())


def _check_vol(u, ndim, name="u"):
    """u must be a non-empty ndim-D tensor (checked first, so that a bad shape
    is reported whatever the device), then a contiguous float32 CUDA tensor."""
    shape = tuple(getattr(u, "shape", ()))
    if len(shape) != ndim or 0 in shape:
        raise ValueError(f"{name} must be a non-empty {ndim}-D tensor, got shape {shape}")


def _check_like(t, u, name, uname="u"):
    """t must have u's shape: the kernels read or write exactly u.numel()
    elements through it."""
    if tuple(getattr(t, "shape", ())) != tuple(u.shape):
        raise ValueError(f"{name} must have {uname}'s shape {tuple(u.shape)}, got {tuple(getattr(t, 'shape', ()))}")


def _check_dev(ts, names):
    """Every tensor is a contiguous float32 CUDA tensor on the first one's device."""
    for t, name in zip(ts, names):
        _dev_f32(t, name)
        if t.device != ts[0].device:
            raise ValueError(f"{name} is on {t.device}, {names[0]} on {ts[0].device}")


def nlm3d(u, r_p, r_s, h, want_Z=False, path=PATH_AUTO, out=None):
    """3D NLM of u[z, y, x] (float32 CUDA tensor), reflect-101 borders."""
    import torch
    _check_vol(u, 3)
    if out is not None:
        _check_like(out, u, "out")
    _check_dev([u] + ([out] if out is not None else []), ["u", "out"])
    nz, ny, nx = u.shape
    y = torch.empty_like(u) if out is None else out
    Z = torch.empty_like(u) if want_Z else None
    st = lib().gnlm3d_forward_ex(_dev_f32(u, "u"), 0, 0, 0, _dev_f32(y, "out"),
                                 _dev_f32(Z, "Z") if Z is not None else None,
                                 nx, ny, nz, int(r_p), int(r_s), float(h), int(path), _stream(u))
    check(st, "gnlm3d_forward_ex")
    return (y, Z) if want_Z else y


def nlm3d_u8(u8, r_p, r_s, h, want_Z=False, out=None):
    """3D NLM of an 8-bit volume u8[z, y, x] (uint8 CUDA tensor), reflect-101
    borders, h on the 0-255 scale; returns float32 (gnlm3d_forward_u8)."""
    import torch
    _check_vol(u8, 3, "u8")
    if not (isinstance(u8, torch.Tensor) and u8.is_cuda and u8.dtype == torch.uint8 and u8.is_contiguous()):
        raise TypeError("u8 must be a contiguous uint8 CUDA tensor")
    nz, ny, nx = u8.shape
    if out is not None:
        _check_like(out, u8, "out", "u8")
        _check_dev([out], ["out"])
        if out.device != u8.device:
            raise ValueError(f"out is on {out.device}, u8 on {u8.device}")
    y = torch.empty(u8.shape, dtype=torch.float32, device=u8.device) if out is None else out
    Z = torch.empty_like(y) if want_Z else None
    st = lib().gnlm3d_forward_u8(ctypes.c_void_p(u8.data_ptr()), 0, 0, 0, _dev_f32(y, "out"),
                                 _dev_f32(Z, "Z") if Z is not None else None,
                                 nx, ny, nz, int(r_p), int(r_s), float(h), _stream(u8))
    check(st, "gnlm3d_forward_u8")
    return (y, Z) if want_Z else y


KERNEL_U8_INTEGER = 110


def forward_u8_route(nx, ny, nz, r_p, r_s):
    """(status, kernel) of gnlm3d_forward_u8 (gnlm3d_forward_u8_route; host
    logic only): KERNEL_U8_INTEGER for the integer-distance kernel, else the
    FP32 route's kernel on the float copy."""
    k = ctypes.c_int(-1)
    st = lib().gnlm3d_forward_u8_route(int(nx), int(ny), int(nz), int(r_p), int(r_s), ctypes.byref(k))
    return st, k.value


def nlm2d_u8(u8, r_p, r_s, h, want_Z=False, out=None):
    """2D NLM of an 8-bit image u8[y, x] (uint8 CUDA tensor), reflect-101
    borders, h on the 0-255 scale; returns float32 (gnlm2d_forward_u8; the
    same bits as nlm2d on the image as floats)."""
    import torch
    _check_vol(u8, 2, "u8")
    if not (isinstance(u8, torch.Tensor) and u8.is_cuda and u8.dtype == torch.uint8 and u8.is_contiguous()):
        raise TypeError("u8 must be a contiguous uint8 CUDA tensor")
    ny, nx = u8.shape
    if out is not None:
        _check_like(out, u8, "out", "u8")
        _check_dev([out], ["out"])
        if out.device != u8.device:
            raise ValueError(f"out is on {out.device}, u8 on {u8.device}")
    y = torch.empty(u8.shape, dtype=torch.float32, device=u8.device) if out is None else out
    Z = torch.empty_like(y) if want_Z else None
    st = lib().gnlm2d_forward_u8(ctypes.c_void_p(u8.data_ptr()), 0, 0, _dev_f32(y, "out"),
                                 _dev_f32(Z, "Z") if Z is not None else None,
                                 nx, ny, int(r_p), int(r_s), float(h), _stream(u8))
    check(st, "gnlm2d_forward_u8")
    return (y, Z) if want_Z else y


def nlm2d(u, r_p, r_s, h, want_Z=False, path=PATH_AUTO):
    """2D NLM of u[y, x] (float32 CUDA tensor), reflect-101 borders."""
    import torch
    _check_vol(u, 2)
    _check_dev([u], ["u"])
    ny, nx = u.shape
    y = torch.empty_like(u)
    Z = torch.empty_like(u) if want_Z else None
    st = lib().gnlm2d_forward_ex(_dev_f32(u, "u"), 0, 0, _dev_f32(y, "out"),
                                 _dev_f32(Z, "Z") if Z is not None else None,
                                 nx, ny, int(r_p), int(r_s), float(h), int(path), _stream(u))
    check(st, "gnlm2d_forward_ex")
    return (y, Z) if want_Z else y


def nlm3d_backward(u, y, Z, gy, r_p, r_s, h):
    """Adjoint of nlm3d: returns (dL/du, dL/dh) as (tensor like u, 0-dim tensor)."""
    import torch
    _check_vol(u, 3)
    for t, name in ((y, "y"), (Z, "Z"), (gy, "gy")):
        _check_like(t, u, name)
    _check_dev([u, y, Z, gy], ["u", "y", "Z", "gy"])
    nz, ny, nx = u.shape
    gu = torch.empty_like(u)
    gh = torch.empty((), dtype=torch.float32, device=u.device)
    st = lib().gnlm3d_backward(_dev_f32(u, "u"), _dev_f32(y, "y"), _dev_f32(Z, "Z"), _dev_f32(gy, "gy"),
                               nx, ny, nz, int(r_p), int(r_s), float(h), _dev_f32(gu, "gu"),
                               ctypes.c_void_p(gh.data_ptr()), _stream(u))
    check(st, "gnlm3d_backward")
    return gu, gh


def nlm3d_padded(up, pad, nx, ny, nz, r_p, r_s, h, path=PATH_AUTO):
    """3D NLM of a pre-padded volume up[(nz+2pad), (ny+2pad), (nx+2pad)]."""
    import torch
    nx, ny, nz, pad = int(nx), int(ny), int(nz), int(pad)
    if min(nx, ny, nz) < 1 or pad < 0:
        raise ValueError(f"bad dims ({nx}, {ny}, {nz}) or pad {pad}")
    need = (nz + 2 * pad) * (ny + 2 * pad) * (nx + 2 * pad)
    if up.numel() < need:
        raise ValueError(f"up holds {up.numel()} elements; the padded extent needs {need}")
    _check_dev([up], ["up"])
    y = torch.empty((nz, ny, nx), dtype=torch.float32, device=up.device)
    st = lib().gnlm3d_forward_ex(_dev_f32(up, "up"), int(pad), 0, 0, _dev_f32(y, "out"), None,
                                 nx, ny, nz, int(r_p), int(r_s), float(h), int(path), _stream(up))
    check(st, "gnlm3d_forward_ex")
    return y


def nl_means_3d_u8(padded_u8, X, Y, depth, SS, NS, h=10.0):
    """Host-pointer drop-in backing: padded uint8 volume in, X*Y*depth floats out."""
    a = np.ascontiguousarray(padded_u8, dtype=np.uint8)
    out = np.empty((depth, Y, X), np.float32)
    st = lib().gnlm_nl_means_3d_u8(a.ctypes.data, out.ctypes.data, X, Y, depth, SS, NS, float(h))
    check(st, "gnlm_nl_means_3d_u8")
    return out


def nl_means_2d_u8(padded_u8, W, H, sr, nb, X, Y, h=10.0):
    a = np.ascontiguousarray(padded_u8, dtype=np.uint8)
    out = np.zeros((Y, Y), np.float32)
    st = lib().gnlm_nl_means_2d_u8(a.ctypes.data, out.ctypes.data, W, H, sr, nb, X, Y, float(h))
    check(st, "gnlm_nl_means_2d_u8")
    return out


def phantom3d(nx, ny, nz, n_ellipsoids=12, axis_lo=4.0, axis_hi=20.0, background=0.05, sigma=0.05, seed=0):
    """Seeded ellipsoid phantom -> (noisy, clean) float32 arrays [nz, ny, nx]."""
    noisy = np.empty((nz, ny, nx), np.float32)
    clean = np.empty((nz, ny, nx), np.float32)
    st = lib().gnlm_phantom3d(noisy.ctypes.data, clean.ctypes.data, nx, ny, nz, n_ellipsoids,
                              axis_lo, axis_hi, background, sigma, seed)
    if st != 0:
        raise ValueError("gnlm_phantom3d: bad arguments")
    return noisy, clean


def phantom2d(nx, ny, n_shapes=40, size_lo=0.008, size_hi=0.125, sine_amp=0.2, period_lo=32.0,
              period_hi=256.0, sigma=0.08, seed=0):
    """Seeded 2D rectangles/discs + sinusoid phantom -> (noisy, clean) [ny, nx]."""
    noisy = np.empty((ny, nx), np.float32)
    clean = np.empty((ny, nx), np.float32)
    st = lib().gnlm_phantom2d(noisy.ctypes.data, clean.ctypes.data, nx, ny, n_shapes, size_lo, size_hi,
                              sine_amp, period_lo, period_hi, sigma, seed)
    if st != 0:
        raise ValueError("gnlm_phantom2d: bad arguments")
    return noisy, clean


def hash_f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return int(lib().gnlm_hash_f32(a.ctypes.data, a.size))


def nlm3d_autograd(u, h, r_p, r_s):
    """Differentiable 3D NLM (torch.autograd), gradients to u and h; see
    gnlm.autograd for where h should live."""
    from .autograd import nlm3d_autograd as f
    return f(u, h, r_p, r_s)
