"""CPU test of the kernels' modular arithmetic (cheddar-fhe_amd/csrc/ck_arith.h:
signed lazy Montgomery, exact FP64 products in both quotient forms, the
reductions and canonical/centred conversions) against exact 128-bit integer
arithmetic: builds and runs tests/cpp/test_device_arith (no GPU runtime)."""
import os
import subprocess

from conftest import REPO

CPP = os.path.join(REPO, "tests", "cpp")


def test_device_arith():
    subprocess.run(["make", "-s", "-C", CPP, "test_device_arith"], check=True, timeout=300)
    r = subprocess.run([os.path.join(CPP, "test_device_arith")], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout
