"""Which kernel a forward or adjoint call runs, and the status it returns for
volumes past the rows kernels' 32-bit offsets -- host logic only, checked
without a GPU (the calls return before any allocation or launch)."""
import ctypes

import pytest

import gnlm


def _rows_fits(n, rp, rs, ty, tz):
    """Size of the padded copy the rows forward builds (16-wide tiles in x,
    halo R = r_p + r_s, x halo rounded up to a multiple of 4) vs 2^30 floats."""
    R = rp + rs
    lp = 4 * ((R + 3) // 4)
    px = -(-n // 16) * 16 + 2 * lp
    py = -(-n // ty) * ty + 2 * R
    pz = -(-n // tz) * tz + 2 * R
    return px * py * pz < 2 ** 30


def test_headline_takes_the_rows_tiles():
    assert gnlm.forward_route(256, 256, 256, 2, 5) == (gnlm.OK, gnlm.VARIANT_ROWS_SHORT)
    assert gnlm.forward_route(96, 96, 48, 1, 3) == (gnlm.OK, gnlm.VARIANT_ROWS)
    assert gnlm.forward_route(512, 512, 128, 8, 6) == (gnlm.OK, gnlm.VARIANT_ROWS_BIG)


def test_rows_limit_boundary():
    """The 16 x 12 x 12 tiles (r_p 2, r_s 5) of a 1000^3 volume still fit the
    32-bit padded copy, those of 1010^3 do not: the router falls back to the
    slab tiles, which read the input in place with 64-bit offsets."""
    assert _rows_fits(1000, 2, 5, 12, 12) and not _rows_fits(1010, 2, 5, 12, 12)
    assert gnlm.forward_route(1000, 1000, 1000, 2, 5) == (gnlm.OK, gnlm.VARIANT_ROWS_SHORT)
    st, k = gnlm.forward_route(1010, 1010, 1010, 2, 5)
    assert st == gnlm.OK and k in (gnlm.VARIANT_SLAB2, gnlm.VARIANT_SLAB, gnlm.VARIANT_16x16x8,
                                   gnlm.VARIANT_16x16x16), k


def test_forward_1100_cubed():
    for rp, rs in ((1, 3), (2, 5), (3, 3)):
        st, k = gnlm.forward_route(1100, 1100, 1100, rp, rs)
        assert st == gnlm.OK and k in (gnlm.VARIANT_SLAB2, gnlm.VARIANT_SLAB, gnlm.VARIANT_16x16x8,
                                       gnlm.VARIANT_16x16x16), (rp, rs, k)
    # large patches have no 64-bit tiles: the per-offset passes take them,
    # and a request for the tiled path alone is unsupported (not "invalid")
    assert gnlm.forward_route(1100, 1100, 1100, 8, 6) == (gnlm.OK, gnlm.KERNEL_GENERIC)
    st, _ = gnlm.forward_route(1100, 1100, 1100, 8, 6, gnlm.PATH_TILED)
    assert st == gnlm.EUNSUPPORTED
    # past 2^40 voxels nothing can hold the buffers
    st, _ = gnlm.forward_route(1 << 14, 1 << 14, 1 << 13, 1, 3)
    assert st == gnlm.EUNSUPPORTED


def test_forward_call_returns_route_status_before_the_device():
    L = gnlm.lib()
    fake = ctypes.c_void_p(0x1000)
    assert L.gnlm3d_forward_ex(fake, 0, 0, 0, fake, None, 1 << 14, 1 << 14, 1 << 13, 1, 3, 0.1, 0,
                               None) == gnlm.EUNSUPPORTED
    assert L.gnlm3d_forward_ex(fake, 0, 0, 0, fake, None, 1100, 1100, 1100, 8, 6, 0.1, gnlm.PATH_TILED,
                               None) == gnlm.EUNSUPPORTED


def test_backward_700_cubed():
    """The rows adjoint's (G, GY, u~, 0) copy (16 bytes per position) passes
    4 GiB near 620^3 at r_p 1 / r_s 4; the LDS-block adjoint takes over."""
    assert gnlm.backward_route(128, 128, 128, 1, 4) == (gnlm.OK, gnlm.KERNEL_BWD_ROWS_SYMMETRIC)
    assert gnlm.backward_route(700, 700, 700, 1, 4) == (gnlm.OK, gnlm.KERNEL_BWD_BLOCK)
    # radii the LDS-block tile cannot hold: the any-radius passes
    assert gnlm.backward_route(700, 700, 700, 2, 7) == (gnlm.OK, gnlm.KERNEL_BWD_GENERIC)
    # 2^31 voxels and up: unsupported (not invalid), returned before the device
    st, _ = gnlm.backward_route(2048, 1024, 1024, 2, 7)
    assert st == gnlm.EUNSUPPORTED
    L = gnlm.lib()
    fake = ctypes.c_void_p(0x1000)
    assert L.gnlm3d_backward(fake, fake, fake, fake, 2048, 1024, 1024, 2, 7, 0.1, fake, None,
                             None) == gnlm.EUNSUPPORTED


def test_backward_form_switch_routes():
    try:
        gnlm.set_backward_form(gnlm.BWD_GATHER)
        assert gnlm.backward_route(64, 64, 64, 2, 3) == (gnlm.OK, gnlm.KERNEL_BWD_ROWS_GATHER)
    finally:
        gnlm.set_backward_form(gnlm.BWD_AUTO)
    assert gnlm.backward_route(64, 64, 64, 2, 3) == (gnlm.OK, gnlm.KERNEL_BWD_ROWS_SYMMETRIC)
    assert gnlm.backward_route(40, 40, 40, 3, 3) == (gnlm.OK, gnlm.KERNEL_BWD_BLOCK)
    # patches past the on-chip adjoints, up to the reference's NS 17 / SS 13
    for rp, rs in ((4, 3), (3, 5), (8, 6)):
        assert gnlm.backward_route(40, 40, 40, rp, rs) == (gnlm.OK, gnlm.KERNEL_BWD_GENERIC)


@pytest.mark.parametrize("args", [(0, 8, 8, 1, 2), (8, 8, 8, -1, 2), (8, 8, 8, 1, -2)])
def test_route_rejects_bad_geometry(args):
    assert gnlm.forward_route(*args)[0] == gnlm.EINVAL
    assert gnlm.backward_route(*args)[0] == gnlm.EINVAL


def _u8_record_bytes(n_xy, n_z, rp, rs):
    """Size of the 8-byte (A, u) records the 8-bit forward builds: the volume
    rounded up to 16 x (32 - 2 r_p)^2 tiles (r_p 7-8; 16^3 tiles at r_p 5-6)
    and grown by r_s per side."""
    t = 32 - 2 * rp if rp >= 7 else 16
    rx = -(-n_xy // 16) * 16 + 2 * rs
    ry = -(-n_xy // t) * t + 2 * rs
    rz = -(-n_z // t) * t + 2 * rs
    return 8 * rx * ry * rz


def test_u8_route():
    """gnlm3d_forward_u8: integer patch distances for r_p 6-8 (the reference's
    NS 13-17) and r_p 5 with r_s >= 3, while the records fit 32-bit byte
    offsets; otherwise the FP32 route on a float copy, with its statuses."""
    assert gnlm.forward_u8_route(512, 512, 128, 8, 6) == (gnlm.OK, gnlm.KERNEL_U8_INTEGER)
    assert gnlm.forward_u8_route(512, 512, 16, 7, 5) == (gnlm.OK, gnlm.KERNEL_U8_INTEGER)
    assert gnlm.forward_u8_route(512, 512, 64, 6, 6) == (gnlm.OK, gnlm.KERNEL_U8_INTEGER)
    assert gnlm.forward_u8_route(512, 512, 64, 5, 3) == (gnlm.OK, gnlm.KERNEL_U8_INTEGER)
    for rp, rs in ((1, 3), (2, 5), (4, 3), (4, 5), (5, 2), (6, 8), (8, 8), (9, 2)):
        assert gnlm.forward_u8_route(64, 64, 64, rp, rs) == gnlm.forward_route(64, 64, 64, rp, rs), (rp, rs)
    # 780^3 at NS 17 / SS 13: records below 4 GiB; 820^3: past them
    assert _u8_record_bytes(780, 780, 8, 6) < 2 ** 32 <= _u8_record_bytes(820, 820, 8, 6)
    assert gnlm.forward_u8_route(780, 780, 780, 8, 6) == (gnlm.OK, gnlm.KERNEL_U8_INTEGER)
    assert gnlm.forward_u8_route(820, 820, 820, 8, 6) == gnlm.forward_route(820, 820, 820, 8, 6)
    st, _ = gnlm.forward_u8_route(1 << 14, 1 << 14, 1 << 13, 8, 6)
    assert st == gnlm.EUNSUPPORTED
    assert gnlm.forward_u8_route(0, 8, 8, 8, 6)[0] == gnlm.EINVAL


def test_u8_call_rejects_before_the_device():
    L = gnlm.lib()
    fake = ctypes.c_void_p(0x1000)
    assert L.gnlm3d_forward_u8(None, 0, 0, 0, fake, None, 8, 8, 8, 8, 6, 10.0, None) == gnlm.EINVAL
    assert L.gnlm3d_forward_u8(fake, 0, 0, 0, fake, None, 8, 8, 8, 8, 6, 0.0, None) == gnlm.EINVAL
    # padding smaller than the halo r_p + r_s
    assert L.gnlm3d_forward_u8(fake, 13, 0, 0, fake, None, 8, 8, 8, 8, 6, 10.0, None) == gnlm.EINVAL
    assert L.gnlm3d_forward_u8(fake, 0, 0, 0, fake, None, 1 << 14, 1 << 14, 1 << 13, 8, 6, 10.0,
                               None) == gnlm.EUNSUPPORTED
