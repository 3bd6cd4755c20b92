"""The nl-means command line (3d-gnlm_amd/cli/nl_means.cpp) end to end on the
reference's own 8-bit data file 3DNLM/data/128.bmp (tests/golden/
ref_data_128.bmp), as 3DNLM/nlm.cpp and "2D NLM/11x11/nlm.cpp" run it.

3D: output.bin = int depth, int rows, int cols, then the truncated 8-bit
result (3DNLM/nlm.cpp:17-24,66-73).  Expected: floor of the FP64 oracle on the
z-stacked image (h = 10 on 0-255, reflect-101).  uint8 input with NS = 3 keeps
every patch distance an exact integer in FP32, so the GPU's weights match the
oracle's to exp rounding; a byte may differ by one only where the FP64 value
lies within 1e-3 of an integer.  The slices must be identical (z-invariant
input) and the BMP slice files must hold the same bytes.
2D: the output BMP holds the result rounded to 8 bits; the padded-input BMP
is the reflect-101 padding of the image.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NLM = os.path.join(ROOT, "3d-gnlm_amd", "cli", "nl-means")
CHECK = os.path.join(ROOT, "tests", "cli_io", "image_io_check")
FIXTURE = os.path.join(ROOT, "tests", "golden", "ref_data_128.bmp")


def _gray(path, tmp_path):
    raw = tmp_path / "g.raw"
    r = subprocess.run([CHECK, str(path), str(raw)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    head, body = raw.read_bytes().split(b"\n", 1)
    rows, cols = (int(v) for v in head.split())
    return np.frombuffer(body, np.uint8).reshape(rows, cols)


def _close_bytes(got, ref, rounding):
    """got: uint8; ref: FP64 values.  Exact away from rounding boundaries."""
    q = np.floor(ref) if rounding == "trunc" else np.floor(ref + 0.5)
    frac = ref - np.floor(ref) if rounding == "trunc" else ref + 0.5 - np.floor(ref + 0.5)
    near = (frac < 1e-3) | (frac > 1 - 1e-3)
    d = np.abs(got.astype(np.int64) - q.astype(np.int64))
    assert d.max() <= 1, d.max()
    assert (d[~near] == 0).all(), int((d[~near] != 0).sum())


def test_cli_3d_output_bin_and_slices(tmp_path):
    assert os.path.exists(NLM), "3d-gnlm_amd/cli/nl-means not built (run __graft_entry__.build())"
    img = _gray(FIXTURE, tmp_path)
    depth, ss, ns = 4, 5, 3
    out = tmp_path / "output.bin"
    sl = tmp_path / "slices"
    sl.mkdir()
    print(f"case cli 3d ss={ss} ns={ns} depth={depth}", flush=True)
    r = subprocess.run([NLM, FIXTURE, "--ss", str(ss), "--ns", str(ns), "--depth", str(depth), "--out", str(out),
                        "--slices", str(sl)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    b = out.read_bytes()
    d, rows, cols = np.frombuffer(b[:12], np.int32)
    assert (d, rows, cols) == (depth, 128, 128)
    got = np.frombuffer(b[12:], np.uint8).reshape(depth, rows, cols)
    for k in range(1, depth):
        assert np.array_equal(got[k], got[0])
    for k in range(depth):
        assert np.array_equal(_gray(sl / f"i_{k}_result.bmp", tmp_path), got[k])
    vol = np.repeat(img[None].astype(np.float64), depth, axis=0)
    ref = oracle.nlm3d_brute(vol, ns // 2, ss // 2, 10.0)
    _close_bytes(got, ref, "trunc")


def test_cli_3d_reference_default_geometry(tmp_path):
    """The reference's default SS = 13 / NS = 17 (3DNLM/config.h:4-5) on its
    own 128.bmp, 6 slices deep: the command line reaches the 8-bit kernel
    with integer patch distances (gnlm_nl_means_3d_u8 -> gnlm3d_forward_u8),
    and output.bin holds the floor of the FP64 oracle (one byte of slack only
    within 1e-3 of an integer)."""
    import gnlm
    img = _gray(FIXTURE, tmp_path)
    depth, ss, ns = 6, 13, 17
    assert gnlm.forward_u8_route(128, 128, depth, ns // 2, ss // 2)[1] == gnlm.KERNEL_U8_INTEGER
    out = tmp_path / "output.bin"
    print(f"case cli 3d ss={ss} ns={ns} depth={depth}", flush=True)
    r = subprocess.run([NLM, FIXTURE, "--ss", str(ss), "--ns", str(ns), "--depth", str(depth), "--out", str(out)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    b = out.read_bytes()
    assert tuple(np.frombuffer(b[:12], np.int32)) == (depth, 128, 128)
    got = np.frombuffer(b[12:], np.uint8).reshape(depth, 128, 128)
    vol = np.repeat(img[None].astype(np.float32), depth, axis=0)
    ref = oracle.nlm3d_separable_f64(vol, ns // 2, ss // 2, 10.0)
    _close_bytes(got, ref, "trunc")


def test_cli_2d_output_and_padded_bmp(tmp_path):
    img = _gray(FIXTURE, tmp_path)
    sr, nb = 7, 3
    out, padded = tmp_path / "output.bmp", tmp_path / "input.bmp"
    print(f"case cli 2d sr={sr} nb={nb}", flush=True)
    r = subprocess.run([NLM, "--2d", FIXTURE, str(sr), str(nb), "--out", str(out), "--padded", str(padded)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    pad = (sr + nb - 2) // 2
    assert np.array_equal(_gray(padded, tmp_path), np.pad(img, pad, mode="reflect"))
    got = _gray(out, tmp_path)
    ref = oracle.nlm2d_brute(img.astype(np.float64), nb // 2, sr // 2, 10.0)
    _close_bytes(got, ref, "round")


def test_cli_2d_wide_image_runs_transposed(tmp_path):
    """A wider-than-tall image (the drop-in's 2D entry point takes X <= Y) runs
    transposed; the result is the oracle's on the image as given (the filter
    commutes with transposition; only the order of the offset sums changes)."""
    img = np.ascontiguousarray(_gray(FIXTURE, tmp_path)[:40, :96])
    pgm = tmp_path / "wide.pgm"
    pgm.write_bytes(b"P5 96 40 255\n" + img.tobytes())
    out = tmp_path / "w.bmp"
    r = subprocess.run([NLM, "--2d", str(pgm), "5", "3", "--out", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    got = _gray(out, tmp_path)
    assert got.shape == (40, 96)
    _close_bytes(got, oracle.nlm2d_brute(img.astype(np.float64), 1, 2, 10.0), "round")
