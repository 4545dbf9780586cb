"""GPU: the packed FP16 envelope-width decode the join's LB_Keogh and DTW suffix terms use
(csrc/fp16_env.hpp, one mixed-precision FMA per bound) equals the plain conversion and addition
bit for bit, for every FP16 pattern in both halves of a word and sixteen sample values (the
fp16_env_check tool, built with the library)."""
import os
import subprocess

import pytest

from conftest import PKG_DIR

pytestmark = pytest.mark.gpu


def test_packed_width_decode_is_exact():
    tool = os.path.join(PKG_DIR, "bin", "fp16_env_check")
    assert os.path.exists(tool), "build the package first (python <package>/build.py)"
    r = subprocess.run([tool], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "fp16_env_check: 0 mismatches of 2097152 decodes" in r.stdout, r.stdout
