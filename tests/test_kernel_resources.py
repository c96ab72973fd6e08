"""Build-time guard (CPU): every HIP kernel compiles for gfx950 without
private-memory spills and at the occupancy its launch geometry is sized
for. A spill turns register arrays into scratch traffic and has slowed a
transform pass by 2x, so it is caught here rather than in a benchmark."""
import os
import re
import subprocess

import pytest

from conftest import PKG

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CSRC = os.path.join(PKG, "csrc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "--cuda-device-only", "-c", "-o", os.devnull,
         "-Rpass-analysis=kernel-resource-usage"]

# minimum waves per SIMD for the transform passes (LDS allows five); the
# column pass's instantiations for BaseConv over 10..16 sources (mangled
# with template argument Li1E) and 17..24 sources (Li2E) are built for four
# and three
MIN_OCCUPANCY = {"ntt_fwd_cols": 5, "ntt_fwd_rows": 5, "intt_rows": 5, "intt_cols": 5,
                 "ntt8_cols": 3, "ntt8_rows": 3}
WIDE_OCCUPANCY = {"Li1E": 4, "Li2E": 3}


def extra_flags(src):
    """the per-file flags the build adds (cheddar-fhe_amd/build.py EXTRA)"""
    import importlib.util
    spec = importlib.util.spec_from_file_location("ck_build", os.path.join(PKG, "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.EXTRA.get(src, [])


def resources(src):
    r = subprocess.run([HIPCC] + FLAGS + extra_flags(src) + [os.path.join(CSRC, src)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    out, cur = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            out[cur] = {}
            continue
        m = re.search(r"(ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]|VGPRs): (\d+)",
                      line)
        if m and cur:
            out[cur][m.group(1).split()[0]] = int(m.group(2))
    return out


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("src", ["ck_ntt.hip", "ck_elem.hip", "ck_bconv.hip", "ck_i8.hip"])
def test_no_spills_and_occupancy(src):
    res = resources(src)
    assert res, "no kernels reported"
    for fn, r in res.items():
        assert r.get("ScratchSize", 0) == 0, (fn, r)
        for k, occ in MIN_OCCUPANCY.items():
            if k in fn:
                if k == "ntt_fwd_cols":
                    for tag, o in WIDE_OCCUPANCY.items():
                        if tag in fn:
                            occ = o
                assert r.get("Occupancy", 0) >= occ, (fn, r)
