"""The conversion kernels' fast Montgomery forms (csd_mont_fast.hpp) return the bits of the general
forms in csd_arith.hpp (red_words, csd_mont_mul, csd_double_red128; hal/mont.rs:38-76) within
their stated operand ranges. The header is host-and-device code, so the check compiles it with g++
and compares the two on random and boundary operands (quotients near +-2^63, |x| near 2^125 and
2^63 - 2^31), for moduli from 2^20 to 2^58."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fast_montgomery_forms_match(tmp_path):
    exe = tmp_path / "mont_fast_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(HERE, "native", "mont_fast_check.cpp")],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe), "400000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout
