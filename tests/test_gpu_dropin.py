"""The C++ drop-in (include/gnlm_main.h) driven like the reference's driver.

tests/dropin/dropin_main.cpp calls NLMeansProcessor::NL_Means with the
reference's signatures on a z-replicated padded 8-bit volume (3DNLM/nlm.cpp's
own construction) and on a padded 8-bit image; h is the reference's
PARAMETER = 10.  Tolerance: relative 1e-5 against the FP64 oracle (uint8
inputs, outputs in the 0-255 range)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from nlm_metrics import max_rel, report

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "dropin", "dropin_main")


def test_dropin_program(tmp_path):
    assert os.path.exists(EXE), "tests/dropin/dropin_main not built (run __graft_entry__.build())"
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    X, Y, D, SS, NS = 40, 36, 12, 7, 3          # tests/dropin/config.h
    img = np.fromfile(tmp_path / "img3.u8", np.uint8).reshape(Y, X)
    out3 = np.fromfile(tmp_path / "out3.f32", np.float32).reshape(D, Y, X)
    vol = np.repeat(img[None].astype(np.float64), D, axis=0)
    ref3 = oracle.nlm3d_brute(vol, (NS - 1) // 2, (SS - 1) // 2, 10.0)
    assert max_rel(out3, ref3) <= 1e-5, report(out3, ref3)
    img2 = np.fromfile(tmp_path / "img2.u8", np.uint8).reshape(32, 32)
    out2 = np.fromfile(tmp_path / "out2.f32", np.float32).reshape(32, 32)
    ref2 = oracle.nlm2d_brute(img2.astype(np.float64), 1, 2, 10.0)
    assert max_rel(out2, ref2) <= 1e-5, report(out2, ref2)
