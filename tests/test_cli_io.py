"""Image files of the nl-means command line (3d-gnlm_amd/cli/image_io.h), on
the CPU through the image_io_check program.

The reference's drivers read their input with OpenCV imread(path, 0)
(3DNLM/nlm.cpp:31, "2D NLM/11x11/nlm.cpp:8"); the fixture is the reference's
own 8-bit data file 3DNLM/data/128.bmp (identical to "2D NLM/11x11/data/
128.bmp"), kept as tests/golden/ref_data_128.bmp.  Checks:
  * the fixture decodes to the pixels an independent numpy reading of its
    header, gray palette and bottom-up rows gives;
  * colour pixels follow BT.601 luma in 14-bit fixed point: pure red, green
    and blue give 76, 150 and 29 (OpenCV's BGR2GRAY values), any gray (v, v, v)
    gives v, for 24- and 32-bit BMPs and for non-gray palettes;
  * top-down BMPs, binary PGM (with a comment), odd widths (row padding);
  * written BMPs decode back to the same pixels.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cli_io", "image_io_check")
NLM = os.path.join(ROOT, "3d-gnlm_amd", "cli", "nl-means")
FIXTURE = os.path.join(ROOT, "tests", "golden", "ref_data_128.bmp")


def _decode(path, tmp_path, rebmp=None):
    assert os.path.exists(CHECK), "tests/cli_io/image_io_check not built (run __graft_entry__.build())"
    raw = tmp_path / "img.raw"
    args = [CHECK, str(path), str(raw)] + ([str(rebmp)] if rebmp else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    data = raw.read_bytes()
    head, body = data.split(b"\n", 1)
    rows, cols = (int(v) for v in head.split())
    return np.frombuffer(body, np.uint8).reshape(rows, cols)


def _bmp(pixels, bpp, palette=None, top_down=False):
    """Writes a BMP independently of image_io.h: pixels is (rows, cols) for 8
    bpp (palette indices) or (rows, cols, 3) RGB."""
    rows, cols = pixels.shape[:2]
    bypp = bpp // 8
    stride = (cols * bypp + 3) // 4 * 4
    pal = b""
    if bpp == 8:
        pal = b"".join(bytes((int(b), int(g), int(r), 0)) for r, g, b in palette)
    off = 14 + 40 + len(pal)
    body = bytearray()
    order = range(rows) if top_down else range(rows - 1, -1, -1)
    for r in order:
        row = bytearray()
        for c in range(cols):
            if bpp == 8:
                row.append(int(pixels[r, c]))
            else:
                R, G, B = (int(v) for v in pixels[r, c])
                row += bytes((B, G, R)) + (b"\xff" if bpp == 32 else b"")
        row += b"\0" * (stride - len(row))
        body += row
    h = -rows if top_down else rows
    info = struct.pack("<IiiHHIIiiII", 40, cols, h, 1, bpp, 0, len(body), 2835, 2835,
                       len(palette) if palette is not None else 0, 0)
    return b"BM" + struct.pack("<IHHI", off + len(body), 0, 0, off) + info + pal + bytes(body)


def test_fixture_decodes_like_numpy(tmp_path):
    b = open(FIXTURE, "rb").read()
    off, = struct.unpack("<I", b[10:14])
    w, h = struct.unpack("<ii", b[18:26])
    bpp, = struct.unpack("<H", b[28:30])
    assert (w, h, bpp) == (128, 128, 8)
    pal = np.frombuffer(b[54:54 + 1024], np.uint8).reshape(256, 4)
    assert np.array_equal(pal[:, 0], np.arange(256)) and np.array_equal(pal[:, 1], pal[:, 0])
    idx = np.frombuffer(b[off:off + 128 * 128], np.uint8).reshape(128, 128)[::-1]
    got = _decode(FIXTURE, tmp_path)
    assert np.array_equal(got, pal[idx, 0])
    assert got.std() > 10  # a real image, not a constant


@pytest.mark.parametrize("bpp", [24, 32])
@pytest.mark.parametrize("top_down", [False, True])
def test_colour_bmp_luma(tmp_path, bpp, top_down):
    rng = np.random.default_rng(bpp)
    px = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)   # odd width: row padding
    px[0, 0], px[0, 1], px[0, 2] = (255, 0, 0), (0, 255, 0), (0, 0, 255)
    for v in range(4):
        px[1, v] = (60 * v + 7,) * 3
    p = tmp_path / "c.bmp"
    p.write_bytes(_bmp(px, bpp, top_down=top_down))
    got = _decode(p, tmp_path)
    assert got.shape == (5, 7)
    assert list(got[0, :3]) == [76, 150, 29]
    assert list(got[1, :4]) == [60 * v + 7 for v in range(4)]


def test_gray_values_exact_and_nongray_palette(tmp_path):
    # every gray level maps to itself; a colour palette goes through the luma
    gray = [(v, v, v) for v in range(256)]
    px = np.arange(256, dtype=np.uint8).reshape(16, 16)
    p = tmp_path / "g.bmp"
    p.write_bytes(_bmp(px, 8, palette=gray))
    assert np.array_equal(_decode(p, tmp_path), px)
    pal = [(255, 0, 0), (0, 255, 0), (0, 0, 255), (10, 10, 10)]
    px2 = np.array([[0, 1, 2, 3, 3]], np.uint8)
    p2 = tmp_path / "p.bmp"
    p2.write_bytes(_bmp(px2, 8, palette=pal))
    assert list(_decode(p2, tmp_path)[0]) == [76, 150, 29, 10, 10]


def test_pgm(tmp_path):
    px = np.random.default_rng(3).integers(0, 256, (9, 11), dtype=np.uint8)
    p = tmp_path / "a.pgm"
    p.write_bytes(b"P5\n# a comment\n11 9\n255\n" + px.tobytes())
    assert np.array_equal(_decode(p, tmp_path), px)


def test_bmp_write_roundtrip(tmp_path):
    out = tmp_path / "re.bmp"
    a = _decode(FIXTURE, tmp_path, rebmp=out)
    assert np.array_equal(_decode(out, tmp_path), a)
    # the written file is the gray-palette, bottom-up layout of the fixture
    assert out.read_bytes()[:2] == b"BM" and struct.unpack("<I", out.read_bytes()[10:14])[0] == 1078
    # odd width
    px = np.random.default_rng(9).integers(0, 256, (6, 13), dtype=np.uint8)
    p = tmp_path / "o.pgm"
    p.write_bytes(b"P5 13 6 255\n" + px.tobytes())
    out2 = tmp_path / "o.bmp"
    _decode(p, tmp_path, rebmp=out2)
    assert np.array_equal(_decode(out2, tmp_path), px)


def test_rejects_bad_input(tmp_path):
    p = tmp_path / "x.bmp"
    p.write_bytes(b"BM" + b"\0" * 10)
    r = subprocess.run([CHECK, str(p), str(tmp_path / "r")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "BMP" in r.stderr


def test_cli_usage_without_gpu():
    """The command line parses its arguments before touching the GPU."""
    assert os.path.exists(NLM), "3d-gnlm_amd/cli/nl-means not built (run __graft_entry__.build())"
    r = subprocess.run([NLM, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage" in r.stderr
    r = subprocess.run([NLM, FIXTURE, "--ss", "4"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "odd" in r.stderr
    r = subprocess.run([NLM], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
