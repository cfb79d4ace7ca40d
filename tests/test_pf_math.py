"""CPU accuracy of the closures' FP64 exp / log / log1p (parflow_amd/csrc/
pf_math.h): the header is compiled for the host with g++ and every function
is compared with long double references over the argument ranges the van
Genuchten closures use (exp on [-50, 50], log on [1e-12, 1e12], log1p on
[1e-20, 1e12]).  The device build replaces the host divisions by the
reciprocal-plus-Newton quotient, which the GPU closure tests cover."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pf_math_ulp_error(tmp_path):
    exe = str(tmp_path / "pf_math_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "parflow_amd", "csrc"),
                    os.path.join(HERE, "native", "pf_math_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    err = {l.split()[0]: float(l.split()[1]) for l in out.strip().splitlines()}
    assert err["exp"] < 1.0, err
    assert err["log"] < 1.5, err
    assert err["log1p"] < 2.0, err
