"""rd_div_rc, the FP64 division by a host-computed reciprocal that the guide
uses for the reference's double-precision q = V2K*(-2.0)*vdotn/sqrt(nlen2)
(guide.cl:104-128): equal bit for bit to IEEE double division on random,
exact-product, root and near-power-of-two cases.  Also that a float
quotient of two floats equals the float rounding of their double quotient,
which lets the guide compute the reference's double d = 2.0*vdotn/nlen2 as
a float division, and that a float quotient with a normal result equals
the float rounding of the numerator times the divisor's double reciprocal
(the monochromator's GAUSS exponents, monochromator.cl:5-8)
(tests/native/div_rc_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("divrc") / "div_rc_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-march=x86-64-v3",
                    os.path.join(HERE, "native", "div_rc_check.cpp"), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_div_rc_matches_ieee_division(checker, seed):
    r = subprocess.run([checker, "4000000", str(seed)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0 float_mismatches 0" in r.stdout
