"""Shared test setup: markers, import paths and parameter fixtures.

The oracle (oracle/) is the checker; the product is the HIP library in
cheddar-fhe_amd/. GPU tests are marked `gpu` and call the product through
its C ABI (ctypes); CPU tests never touch a GPU.
"""
import json
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "cheddar-fhe_amd")
ORACLE = os.path.join(REPO, "oracle")
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)

# seeds: base 0x4348454444415200 ("CHEDDAR\0") + config index
SEED_BASE = 0x4348454444415200


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_params(name):
    for d in (os.path.join(PKG, "params"), os.path.join(GOLDEN, "params")):
        p = os.path.join(d, name)
        if os.path.exists(p):
            with open(p) as f:
                return json.load(f)
    raise FileNotFoundError(name)


@pytest.fixture(scope="session")
def rns24():
    return load_params("rns24_aux8.json")


@pytest.fixture(scope="session")
def boot40():
    return load_params("bootparam_40.json")


@pytest.fixture(scope="session")
def oracle_rns24(rns24):
    import ckks_oracle as O
    O.set_threads(min(8, os.cpu_count() or 1))
    return O.Param.from_json(rns24)


@pytest.fixture(scope="session")
def oracle_boot40(boot40):
    import ckks_oracle as O
    O.set_threads(min(8, os.cpu_count() or 1))
    return O.Param.from_json(boot40)
