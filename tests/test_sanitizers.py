"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only;
gcc / g++ builds into a temporary directory, nothing of the GPU is built).

  * The CPU oracle (oracle/nlm_oracle.c) through tests/sanitize/
    oracle_sanitize.c: every entry point on ragged volumes (single-sample
    axes, axes shorter than the halo, partial z ranges) with exact-size heap
    buffers.  The driver also requires the FP32 brute force and separable
    paths to agree bit for bit and the FP64 ones to 1e-12.
  * The command line's image reader and writer (3d-gnlm_amd/cli/image_io.h)
    through tests/cli_io/image_io_check.cpp: the reference's 128.bmp fixture,
    colour and top-down BMPs, PGM, and malformed files (truncated, empty,
    headers claiming huge or negative sizes, bad magic), which must be
    rejected with an error rather than read past their end.
A sanitizer report aborts the program (-fno-sanitize-recover=all), so the
bar is: exit status as expected and no sanitizer output.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from test_cli_io import _bmp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="gcc / g++ not available")


def _clean(r):
    return "Sanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_sanitize"
    src = [os.path.join(ROOT, "tests", "sanitize", "oracle_sanitize.c"), os.path.join(ROOT, "oracle", "nlm_oracle.c")]
    b = subprocess.run(["gcc"] + SAN + ["-fopenmp", "-ffp-contract=off", "-std=c11", "-o", str(exe)] + src + ["-lm"],
                       capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0 and _clean(r), r.stderr[-4000:]
    assert "all cases ok" in r.stdout


@pytest.fixture(scope="module")
def io_check(tmp_path_factory):
    exe = tmp_path_factory.mktemp("io") / "image_io_check"
    src = os.path.join(ROOT, "tests", "cli_io", "image_io_check.cpp")
    b = subprocess.run(["g++"] + SAN + ["-std=c++17", "-I", os.path.join(ROOT, "3d-gnlm_amd", "cli"), "-o", str(exe),
                                        src], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr
    return str(exe)


def _run(exe, path, tmp_path, rebmp=False):
    args = [exe, str(path), str(tmp_path / "out.raw")] + ([str(tmp_path / "re.bmp")] if rebmp else [])
    return subprocess.run(args, capture_output=True, text=True, timeout=60, env=ENV)


def test_image_io_valid_inputs(io_check, tmp_path):
    rng = np.random.default_rng(4)
    files = {"fixture.bmp": open(os.path.join(ROOT, "tests", "golden", "ref_data_128.bmp"), "rb").read(),
             "gray8.bmp": _bmp(rng.integers(0, 256, (7, 13)), 8, palette=[(v, v, v) for v in range(256)]),
             "rgb24.bmp": _bmp(rng.integers(0, 256, (5, 9, 3)), 24),
             "rgb32_topdown.bmp": _bmp(rng.integers(0, 256, (6, 3, 3)), 32, top_down=True),
             "odd.pgm": b"P5\n# comment\n5 3\n255\n" + bytes(range(15))}
    for name, data in files.items():
        p = tmp_path / name
        p.write_bytes(data)
        r = _run(io_check, p, tmp_path, rebmp=True)
        assert r.returncode == 0 and _clean(r), (name, r.stderr[-4000:])


def test_image_io_malformed_inputs(io_check, tmp_path):
    good = _bmp(np.random.default_rng(5).integers(0, 256, (8, 11)), 8, palette=[(v, v, v) for v in range(256)])
    huge = bytearray(good)
    huge[18:26] = struct.pack("<ii", 1 << 30, 1 << 30)          # width, height far past the file
    neg = bytearray(good)
    neg[18:22] = struct.pack("<i", -11)                           # negative width
    badoff = bytearray(good)
    badoff[10:14] = struct.pack("<I", len(good) + 4096)           # pixel data offset past the end
    cases = {"empty.bmp": b"", "magic.bmp": b"XX" + good[2:], "truncated.bmp": good[:len(good) // 2],
             "header_only.bmp": good[:54], "huge.bmp": bytes(huge), "neg.bmp": bytes(neg),
             "offset.bmp": bytes(badoff), "short.pgm": b"P5\n5 3\n255\n" + bytes(4),
             "bad.pgm": b"P5\n-5 3\n255\n" + bytes(15)}
    for name, data in cases.items():
        p = tmp_path / name
        p.write_bytes(data)
        r = _run(io_check, p, tmp_path)
        assert r.returncode == 1 and _clean(r), (name, r.returncode, r.stderr[-4000:])
