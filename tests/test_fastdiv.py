"""The step kernel's division by invariant integers (epirust_amd/csrc/fastdiv.hpp) against
the hardware division, compiled for the host with g++ (the same header the kernel uses)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_fast_div_matches_division(tmp_path):
    exe = str(tmp_path / "fastdiv_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "epirust_amd", "csrc"),
                           os.path.join(HERE, "native", "fastdiv_check.cpp"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
