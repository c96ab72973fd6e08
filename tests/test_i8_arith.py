"""CPU test of the exact int8 matrix-core arithmetic (cheddar-fhe_amd/csrc/
ck_i8_arith.h: balanced byte digits, recombination of the four int32
partial sums, the FP64 reduction, and the operand tables of the 4-step NTT
stages and of BaseConv) against 128-bit integer arithmetic, with the matrix
cores emulated in the lane layout measured on gfx950: builds and runs
tests/cpp/test_i8_arith (no GPU runtime)."""
import os
import subprocess

from conftest import REPO

CPP = os.path.join(REPO, "tests", "cpp")


def test_i8_arith():
    subprocess.run(["make", "-s", "-C", CPP, "test_i8_arith"], check=True, timeout=300)
    r = subprocess.run([os.path.join(CPP, "test_i8_arith")], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout
