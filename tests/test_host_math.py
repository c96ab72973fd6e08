"""CPU test of the C++ host arithmetic behind encoding (special FFT pair,
double -> residue conversion, CRT reconstruction): builds and runs
tests/cpp/test_host_math (no GPU runtime involved)."""
import os
import subprocess

from conftest import REPO

CPP = os.path.join(REPO, "tests", "cpp")


def test_host_math():
    subprocess.run(["make", "-s", "-C", CPP, "test_host_math"], check=True, timeout=300)
    r = subprocess.run([os.path.join(CPP, "test_host_math")], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK: 0 failure(s)" in r.stdout
