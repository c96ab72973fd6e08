"""Build libcheddar_amd.so (the HIP core for gfx950) in-tree.

Every translation unit is compiled by hipcc for gfx950 only; packed FP32
instructions are disabled (-packed-fp32-ops) and FP contraction is off so
the exact FP64 modular products keep the operation order they are written
in. Objects go to cheddar-fhe_amd/_build/, the library next to this file.

Usage: python cheddar-fhe_amd/build.py [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcheddar_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I" + CSRC,
]


# per-file flags: the int8 matrix-core kernels keep their MFMA results in
# ordinary VGPRs (the default allocation parks them in accumulation
# registers and copies every result back with v_accvgpr_read)
EXTRA = {
    "ck_i8.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip") or f.endswith(".cpp"))


def compile_one(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + EXTRA.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined",
           "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_boot_bench(verbose)
    return LIB


BOOT_BENCH = os.path.join(HERE, "boot_bench")
PLAN_BENCH = os.path.join(HERE, "plan_bench")


def build_host_program(name, out, verbose=False):
    """A host program of host/ (C++ over the facade headers and the C ABI),
    linked against the library just built."""
    repo = os.path.dirname(HERE)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(repo, "include"),
           "-I" + os.path.join(HERE, "host", "include"),
           os.path.join(HERE, "host", name), "-o", out + ".tmp",
           "-L" + HERE, "-lcheddar_amd", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{name} build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {out}")


def build_boot_bench(verbose=False):
    """The bootstrapping benchmark (host/boot_bench.cpp) and the batched-plan
    benchmark (host/plan_bench.cpp, no torch in the process)."""
    build_host_program("boot_bench.cpp", BOOT_BENCH, verbose)
    build_host_program("plan_bench.cpp", PLAN_BENCH, verbose)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
