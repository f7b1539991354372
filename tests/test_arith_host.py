"""The product's modular arithmetic (csd_arith.hpp, compiled for the host) agrees with the oracle's
Montgomery and Barrett restatements on millions of random and extreme inputs."""
import os
import subprocess

from _pkg import ROOT
from oracle import cpu_backend


def test_arith_matches_oracle(tmp_path):
    cpu_backend.build()
    exe = tmp_path / "arith_harness"
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "arith_harness.cpp"),
                    os.path.join(ROOT, "oracle", "liboracle.so"), "-Wl,-rpath," + os.path.join(ROOT, "oracle"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "400000"], stdout=subprocess.PIPE, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    assert out.stdout.startswith("ok"), out.stdout
