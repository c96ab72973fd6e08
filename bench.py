"""Benchmark: bit-exact HMult + relinearize + rescale throughput on one MI355X.

Workload (the plain GPU key-switching configuration): N = 2^16, L = 24 q
limbs (the largest primes below 2^31 with q = 1 mod 2^17), alpha = 8
special primes, dnum = 3; two ciphertexts of 2 polynomials with uniform
residues and a uniform evaluation key, all resident in HBM before timing.
One step = one Context::HMult with the merged RelinearizeRescale: Tensor ->
ModUp (INTT, BaseConv, NTT) -> key inner product (+ P * (d0, d1)) ->
ModDownAndRescale of both polynomials (INTT, BaseConv, NTT + epilogue).

Prints one JSON line (see the "roofline" and "cpu_baseline" objects).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "cheddar-fhe_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
LEVEL = 23
PARAMS = "rns24_aux8.json"
PROFILE_STEPS = 20


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def uniform_poly(torch, ctx, pidx, gen):
    """Uniform residues in [0, q_i) for the listed global prime indices."""
    n = ctx.n
    out = torch.empty((len(pidx), n), dtype=torch.int32, device="cuda")
    for t, p in enumerate(pidx):
        q = ctx.prime(p)
        out[t] = torch.randint(0, q, (n,), dtype=torch.int64, device="cuda",
                               generator=gen).to(torch.int32)
    return out


def alg_limbs(ctx, level, name):
    """(limbs read, limbs written) that one launch of kernel `name` ("kernel/
    output limbs") must move at least, for HMult+relin+rescale at `level`:
    every operand read once and every result written once, 4 bytes a
    coefficient."""
    kern, n = name.rsplit("/", 1)
    n = int(n)
    num_q, alpha, beta = ctx.num_q(level), ctx.alpha, ctx.modup_beta(level)
    next_q = ctx.num_q(level - 1)
    if kern == "intt_rows+tensor":       # a1 (*) a2 -> INTT pass 1
        return 2 * n, n
    if kern in ("intt_rows", "intt_cols", "ntt_fwd_rows", "ntt_fwd_cols", "permute"):
        return n, n
    if kern in ("ntt_fwd_cols+bconv", "bconv_mfma"):
        if n == beta * (num_q + alpha) - num_q:   # ModUp: the q limbs, once
            return num_q, n
        polys = n // next_q                        # ModDown(+Rescale)
        return polys * (alpha + num_q - next_q), n
    if kern == "ntt_fwd_rows+epilogue":            # + the kept q limbs
        return 2 * n, n
    if kern == "tensor":
        return 4 * num_q, 3 * num_q
    if kern in ("key_accum", "ntt_fwd_rows+key"):  # ModUp out + ct1, ct2 + key
        ext = num_q + alpha
        return (beta * ext - num_q) + 4 * num_q + 2 * beta * ext, 2 * ext
    raise ValueError(f"no byte model for kernel {name}")


def kernel_key(name):
    """'ntt_fwd_cols+bconv/72' -> 'ntt_fwd_cols/72': the kernel symbol and
    its limb count, the key rocprof summaries use."""
    kern, n = name.rsplit("/", 1)
    if kern == "ntt_fwd_rows+key":
        return f"key_rows/{n}"
    return f"{kern.split('+')[0]}/{n}"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(REPO, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    k = rec.get("kernels", {}).get(kernel_key(kernel))
    return None if k is None else k.get("hbm_bytes")


# ---- rocprofv3 output summaries --------------------------------------------
# Launch grids -> limb counts: column passes run 8 workgroups per limb,
# row passes and element-wise kernels 16 / 64 per limb (grid y = limbs).
def _short(sym):
    s = sym.replace("void ", "").replace("ck::(anonymous namespace)::", "")
    s = s.replace("ck::", "").split("(")[0]
    s = s.split("<")[0]
    return s[:-7] if s.endswith("_kernel") else s


_KERNELS = ("ntt_fwd_cols", "intt_cols", "ntt_fwd_rows", "intt_rows", "key_accum",
            "key_accum_table", "bs_fused", "bs_fused_table", "tensor", "permute",
            "binary", "unary", "const", "accum", "imag_unit", "crt_lift", "bconv_mfma",
            "bconv_mfma_table", "key_rows", "key_rows_table", "gs_accum")
# int8 matrix-core kernels (batched plans): labelled by workgroup count, since
# their workgroups per limb depend on the plan's blocks per workgroup
_I8_KERNELS = ("ntt8_cols", "ntt8_rows", "bconv_i8", "bconv_valu")


def _limbs(short, grid_x, grid_y, wg_x):
    """limbs of one launch from its grid (None for kernels not of this
    library)"""
    wgs = grid_x // max(1, wg_x)
    if short in _I8_KERNELS:
        return f"{wgs}wg"
    if short == "copy16":  # ck_copy: 16 bytes per thread, labelled by size
        return f"{(wgs * wg_x * 16) >> 20}MiB"
    if short not in _KERNELS:
        return None
    if short in ("ntt_fwd_cols", "intt_cols"):
        return wgs // 16
    if short in ("bs_fused", "bs_fused_table"):
        return wgs // 64
    return grid_y


def _read_csv(path):
    import csv
    with open(path) as f:
        return list(csv.DictReader(f))


def rocprof_summary(trace_csv):
    """Markdown table of a --kernel-trace CSV: per library kernel and limb
    count, launches and mean/median duration."""
    groups = {}
    for r in _read_csv(trace_csv):
        short = _short(r["Kernel_Name"])
        n = _limbs(short, int(r["Grid_Size_X"]), int(r["Grid_Size_Y"]),
                   int(r["Workgroup_Size_X"]))
        if n is None:
            continue
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        groups.setdefault(f"{short}/{n}", []).append(dur)
    lines = ["| kernel/limbs | launches | mean us | median us | min us |",
             "|---|---|---|---|---|"]
    for k, v in sorted(groups.items(), key=lambda kv: -sum(kv[1])):
        v.sort()
        lines.append(f"| {k} | {len(v)} | {sum(v) / len(v):.2f} | {v[len(v) // 2]:.2f} "
                     f"| {v[0]:.2f} |")
    return "\n".join(lines)


def pmc_traffic(fetch_csv, write_csv, scale=None):
    """Per kernel/limbs HBM bytes per launch from FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch), each scaled so that the 256 MiB device copy of
    --copy-calibration reads 256 MiB and writes 256 MiB; for runs without
    that copy (boot_bench), `scale` gives the (fetch, write) factors of an
    earlier calibration."""
    def per_kernel(path, counter):
        acc, copy = {}, None
        for r in _read_csv(path):
            if r["Counter_Name"] != counter:
                continue
            v = float(r["Counter_Value"])
            if "copyBuffer" in r["Kernel_Name"] and int(r["Grid_Size"]) == 131072:
                copy = v
                continue
            short = _short(r["Kernel_Name"])
            wg = int(r["Workgroup_Size"])
            grid = int(r["Grid_Size"])
            if short in ("ntt_fwd_cols", "intt_cols"):
                n = grid // wg // 16
            elif short in ("ntt_fwd_rows", "intt_rows", "key_rows"):
                n = grid // wg // 16
            elif short in ("key_accum", "tensor", "permute"):
                n = grid // wg // 64
            elif short in ("gs_accum", "permute_accum"):
                n = grid // 32768  # two words per thread
            elif short == "bs_fused":
                n = grid // wg // 64  # limbs rounded up to a multiple of 8
            else:
                continue
            acc.setdefault(f"{short}/{n}", []).append(v)
        return {k: sum(v) / len(v) for k, v in acc.items()}, copy
    fetch, fcopy = per_kernel(fetch_csv, "FETCH_SIZE")
    write, wcopy = per_kernel(write_csv, "WRITE_SIZE")
    if scale:
        fs, ws = scale
    elif not fcopy or not wcopy:
        raise SystemExit("no 256 MiB calibration copy in the counter files (give --pmc-scale)")
    else:
        fs, ws = 262144.0 / fcopy, 262144.0 / wcopy
    out = {"fetch_scale": fs, "write_scale": ws, "kernels": {}}
    for k in sorted(set(fetch) & set(write)):
        out["kernels"][k] = {"fetch_kib": fetch[k], "write_kib": write[k],
                             "hbm_bytes": int((fetch[k] * fs + write[k] * ws) * 1024)}
    return out


def copy_calibration(torch):
    """One plain device-to-device copy of 256 MiB: 256 MiB read + 256 MiB
    written, the reference point for scaling the HBM counters."""
    src = torch.ones(64 << 20, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log("copy calibration: 268435456 bytes copied")


def device_copy_bandwidth(torch, ctx, reps=10):
    """A plain device-to-device copy of 1 GiB (1 GiB read + 1 GiB written)
    by the library's copy kernel (ck_copy: one 16-byte vector per thread,
    non-temporal), timed over `reps` copies with HIP events on the current
    stream after one warm-up: the HBM rate a streaming kernel reaches on this
    GPU, reported beside the published peak; the same for torch's copy_."""
    src = torch.ones(256 << 20, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    moved = 2 * src.numel() * 4

    def rate(fn):
        fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        return moved / (ms * 1e-3) / 1e9, ms

    gbs, ms = rate(lambda: ctx.copy(dst, src, stream=stream))
    ok = bool(torch.equal(dst, src))
    torch_gbs, _ = rate(lambda: dst.copy_(src))
    del src, dst
    if not ok:
        raise SystemExit("ck_copy: the copy differs from its source")
    return {"gb_s": gbs, "bytes_per_copy": moved, "ms_per_copy": ms,
            "frac_of_peak": gbs / HBM_PEAK_GBS, "kernel": "ck_copy",
            "torch_copy_gb_s": torch_gbs}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """threads for the all-cores leg: the cores this process may run on,
    at most 16 (the host CPU share of one GPU on the measurement boxes)"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(prm, seconds=10.0):
    """The CPU oracle's HMult at the same shape on the host cores, once on
    one thread and once on every host thread (bounded samples: as many
    HMults as fit in about `seconds` each, at least 2)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import ckks_oracle as O

    P = O.Param.from_json(prm)
    pidx = P.level_primes(LEVEL)
    seed = 0x4348454444415200 + 1
    ct1 = [P.uniform(pidx, seed + i) for i in range(2)]
    ct2 = [P.uniform(pidx, seed + 2 + i) for i in range(2)]
    allp = list(range(P.num_primes))
    key = (np.stack([P.uniform(allp, seed + 10 + g) for g in range(P.dnum)]),
           np.stack([P.uniform(allp, seed + 20 + g) for g in range(P.dnum)]))

    def rate(threads):
        O.set_threads(threads)
        P.hmult(LEVEL, ct1, ct2, key)  # warm-up
        t0 = time.perf_counter()
        n = 0
        while True:
            P.hmult(LEVEL, ct1, ct2, key)
            n += 1
            el = time.perf_counter() - t0
            if (el >= seconds and n >= 2) or el >= 3 * seconds:
                return n / el, n, el

    cores = host_threads()
    v1, n1, e1 = rate(1)
    vn, nn, en = rate(cores)
    return {"value": vn, "unit": "HMult/s", "cores": cores, "kind": "port",
            "sample": f"{nn} HMult+relin+rescale at L=24 on {cores} host threads "
                      f"({en:.1f} s); {n1} on 1 thread ({e1:.1f} s)",
            "value_1_thread": v1, "nproc": os.cpu_count(), "cpu_model": cpu_model()}


BATCH = 64
BATCH_ROTS = 16


def run_batched(torch, ck, steps, warmup, autotune=True, bconv="auto", check=4):
    """Batched key switching (config 2): BATCH ciphertexts with limb counts
    uniform in [8, 24] (seeded) each go through HMult+relin+rescale (one
    plan, one launch per step across the batch), then BATCH_ROTS hoisted
    rotations by distinct offsets in [1, 2^15) of each product (a second
    plan sharing one set of rotation keys). Returns whole-batch HMult/s and
    HRot/s. Outside the timed region, `check` seeded items of each plan are
    compared word for word with the single-ciphertext entry points
    (ck_hmult, ck_hrot) on the same inputs."""
    import random
    prm = ck.load_param_json(PARAMS)
    ctx = ck.Context.from_json(prm)
    ctx.set_bconv_mfma(None if bconv == "auto" else bconv == "mfma")
    ctx.set_plan_autotune(autotune)
    torch.cuda.set_device(0)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0x4348454444415202)
    rng = random.Random(0x4348454444415202)
    levels = [rng.randint(8, 24) - 1 for _ in range(BATCH)]
    rots = rng.sample(range(1, 1 << 15), BATCH_ROTS)
    allp = list(range(ctx.num_primes))

    def qprimes(level):
        nm, nt = ctx.level_np(level)
        return list(range(ctx.max_ter - nt, ctx.max_ter + nm))

    def key():
        return ([uniform_poly(torch, ctx, allp, gen) for _ in range(ctx.dnum)],
                [uniform_poly(torch, ctx, allp, gen) for _ in range(ctx.dnum)])

    ct1 = [[uniform_poly(torch, ctx, qprimes(l), gen) for _ in range(2)] for l in levels]
    ct2 = [[uniform_poly(torch, ctx, qprimes(l), gen) for _ in range(2)] for l in levels]
    prod = [[torch.empty((ctx.num_q(l - 1), ctx.n), dtype=torch.int32, device="cuda")
             for _ in range(2)] for l in levels]
    rot_out = [[[torch.empty((ctx.num_q(l - 1), ctx.n), dtype=torch.int32, device="cuda")
                 for _ in range(2)] for _ in rots] for l in levels]
    mkey = key()
    rkeys = [key() for _ in rots]
    t0 = time.perf_counter()
    p_mult = ck.Plan.hmult(ctx, levels, prod, ct1, ct2, mkey)
    p_rot = ck.Plan.hrot_hoisted(ctx, [l - 1 for l in levels], prod, rkeys, rots, rot_out)
    log(f"plans built in {time.perf_counter() - t0:.1f} s: hmult {p_mult.info()}, "
        f"hoisted {p_rot.info()}")
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        p_mult.run(stream)
        p_rot.run(stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(stream)
    for _ in range(steps):
        p_mult.run(stream)
    ev[1].record(stream)
    for _ in range(steps):
        p_rot.run(stream)
    ev[2].record(stream)
    torch.cuda.synchronize()
    t_mult = ev[0].elapsed_time(ev[1]) / steps
    t_rot = ev[1].elapsed_time(ev[2]) / steps
    hmult_s = BATCH / (t_mult * 1e-3)
    hrot_s = BATCH * BATCH_ROTS / (t_rot * 1e-3)
    log(f"batched HMult: {t_mult:.3f} ms per {BATCH} ({hmult_s:.0f}/s); hoisted HRot: "
        f"{t_rot:.3f} ms per {BATCH * BATCH_ROTS} ({hrot_s:.0f}/s)")
    # sampled items against the single-ciphertext entry points
    sample = rng.sample(range(BATCH), check)
    ok = True
    for i in sample:
        lv = levels[i]
        o = [torch.empty_like(t) for t in prod[i]]
        ctx.hmult(o, ct1[i], ct2[i], mkey, lv, rescale=True, stream=stream)
        r = rng.randrange(BATCH_ROTS)
        orot = [torch.empty_like(t) for t in rot_out[i][r]]
        ctx.hrot(orot, prod[i], rkeys[r], rots[r], lv - 1, stream=stream)
        torch.cuda.synchronize()
        ok = ok and all(torch.equal(x, y) for x, y in zip(o, prod[i]))
        ok = ok and all(torch.equal(x, y) for x, y in zip(orot, rot_out[i][r]))
    log(f"batched outputs of items {sample} equal to ck_hmult / ck_hrot: {ok}")
    if not ok:
        raise RuntimeError("batched plan output differs from the single-ciphertext path")
    # per-kernel durations of one run of each plan (library dispatch events)
    for name, plan in (("hmult", p_mult), ("hoisted", p_rot)):
        ctx.profile(True)
        plan.run(stream)
        torch.cuda.synchronize()
        rep = ctx.profile_report()
        ctx.profile(False)
        tot = sum(v[0] for v in rep.values())
        for k, (ms, calls) in sorted(rep.items(), key=lambda kv: -kv[1][0]):
            log(f"  {name} plan: {k:32s} {calls} x {1e3 * ms / calls:8.1f} us "
                f"({100 * ms / tot:4.1f}%)")
    return {"workload": f"{BATCH} ciphertexts at 8..24 limbs: HMult+relin+rescale plan, "
                        f"then {BATCH_ROTS} hoisted HRot each",
            "hmult_per_s": hmult_s, "hrot_per_s": hrot_s,
            "ms_hmult_batch": t_mult, "ms_hrot_batch": t_rot, "steps": steps,
            "levels": levels, "rotations": rots, "autotune": autotune,
            "hmult_plan": p_mult.info(), "hrot_plan": p_rot.info(),
            "hmult_plan_candidates_us": {
                ("fused", "fp64", "int8")[f] + ("+key_rows" if kr else "+key_accum") +
                (f"+{ch}chunks" if ch > 1 else ""): us
                for f, kr, ch, us in p_mult.tuning()},
            "hrot_plan_bconv_us": {("fused", "fp64", "int8")[f]: us
                                   for f, _, _, us in p_rot.tuning()},
            "checked_items": sample, "equal_to_single_ops": ok}


def bench_batched(a, torch, ck):
    """--workload batched: the batched configuration as its own JSON line"""
    r = run_batched(torch, ck, a.steps, a.warmup, not a.no_autotune, a.bconv)
    res = {
        "metric": "hmult_relin_rescale_per_s", "value": r["hmult_per_s"], "unit": "HMult/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": r["ms_hmult_batch"] + r["ms_hrot_batch"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic uniform residues and keys (seeded)",
        "config": {"workload": "batched: " + r["workload"], "log_n": 16,
                   "parallelism": "single GPU"},
        "batched": r,
    }
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 256 MiB device copy (HBM counter "
                         "calibration for rocprofv3 --pmc runs)")
    ap.add_argument("--rocprof-summary", metavar="KERNEL_TRACE_CSV",
                    help="print a per-kernel summary of a rocprofv3 --kernel-trace "
                         "CSV and exit")
    ap.add_argument("--bconv", choices=("auto", "fused", "mfma"), default="auto",
                    help="BaseConv fused into the forward NTT column pass, on the FP64 "
                         "matrix cores as its own step, or auto (matrix cores in batched "
                         "plans only)")
    ap.add_argument("--workload", choices=("hmult", "batched"), default="hmult",
                    help="hmult: one HMult+relin+rescale per step at L=24 (default); "
                         "batched: 64 ciphertexts at 8..24 limbs through one HMult "
                         "plan, then 16 hoisted rotations each through one plan")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched workload: use --bconv instead of timing both BaseConv "
                         "forms per plan")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-plan measurement in the default run")
    ap.add_argument("--no-boot", action="store_true",
                    help="skip the bootstrapping measurement (bootparam_40, full slots)")
    ap.add_argument("--boot-runs", type=int, default=3)
    ap.add_argument("--pmc-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print per-kernel HBM traffic (JSON) from two rocprofv3 --pmc "
                         "runs (FETCH_SIZE, WRITE_SIZE) and exit")
    ap.add_argument("--pmc-scale", nargs=2, type=float, metavar=("FETCH", "WRITE"),
                    help="with --pmc-traffic: counter scale factors of an earlier "
                         "calibration, for runs without the calibration copy")
    a = ap.parse_args()
    if a.rocprof_summary:
        print(rocprof_summary(a.rocprof_summary))
        return
    if a.pmc_traffic:
        print(json.dumps(pmc_traffic(*a.pmc_traffic, scale=a.pmc_scale), indent=1))
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported "
              "(the key-switching path has no multi-GPU form here)", file=sys.stderr)
        sys.exit(2)

    import torch
    import cheddar_amd as ck

    if a.workload == "batched":
        return bench_batched(a, torch, ck)

    prm = ck.load_param_json(PARAMS)
    ctx = ck.Context.from_json(prm)
    ctx.set_bconv_mfma(None if a.bconv == "auto" else a.bconv == "mfma")
    torch.cuda.set_device(0)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0x4348454444415201)
    nm, nt = ctx.level_np(LEVEL)
    qp = list(range(ctx.max_ter - nt, ctx.max_ter + nm))  # [ter reversed | main]
    allp = list(range(ctx.num_primes))
    log(f"context: L={ctx.L} alpha={ctx.alpha} dnum={ctx.dnum} level={LEVEL}")
    ct1 = [uniform_poly(torch, ctx, qp, gen) for _ in range(2)]
    ct2 = [uniform_poly(torch, ctx, qp, gen) for _ in range(2)]
    key = ([uniform_poly(torch, ctx, allp, gen) for _ in range(ctx.dnum)],
           [uniform_poly(torch, ctx, allp, gen) for _ in range(ctx.dnum)])
    out = [torch.empty((ctx.num_q(LEVEL - 1), ctx.n), dtype=torch.int32, device="cuda")
           for _ in range(2)]
    stream = torch.cuda.current_stream()

    def step():
        ctx.hmult(out, ct1, ct2, key, LEVEL, rescale=True, stream=stream)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    log(f"warmup done ({a.warmup} steps)")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(a.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / a.steps
    log(f"timed {a.steps} steps: {ms * 1e3:.1f} us/step (host wall "
        f"{wall * 1e6 / a.steps:.1f} us/step)")

    # Per-kernel durations, measured live: HIP events recorded by the
    # library on `stream` around every launch of a further run of steps.
    ctx.profile(True)
    for _ in range(PROFILE_STEPS):
        step()
    torch.cuda.synchronize()
    rep = ctx.profile_report()
    ctx.profile(False)
    per_step = {k: (v[0] * 1e3 / PROFILE_STEPS, v[1] // PROFILE_STEPS, v[0] * 1e3 / v[1])
                for k, v in rep.items()}
    total = sum(v[0] for v in per_step.values())
    for k, (us, calls, avg) in sorted(per_step.items(), key=lambda kv: -kv[1][0]):
        log(f"  {k:28s} {calls} x {avg:7.2f} us = {us:7.2f} us/step "
            f"({100 * us / total:4.1f}%)")
    log(f"  kernel sum {total:.1f} us/step")
    dom = max(per_step, key=lambda k: per_step[k][0])
    launch_us = per_step[dom][2]
    limbs_in, limbs_out = alg_limbs(ctx, LEVEL, dom)
    alg_bytes = (limbs_in + limbs_out) * ctx.n * 4
    achieved = alg_bytes / (launch_us * 1e-6) / 1e9
    log(f"dominant kernel {dom}: {launch_us:.2f} us/launch, {alg_bytes / 1e6:.1f} MB "
        f"algorithmic -> {achieved:.0f} GB/s")
    traffic = measured_traffic(dom)
    copy_bw = device_copy_bandwidth(torch, ctx)
    log(f"device copy: {copy_bw['gb_s']:.0f} GB/s ({copy_bw['frac_of_peak']:.2f} of the "
        f"{HBM_PEAK_GBS:.0f} GB/s peak); dominant kernel at "
        f"{achieved / copy_bw['gb_s']:.2f} of the copy rate")
    if a.copy_calibration:
        copy_calibration(torch)

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(prm)
        log(f"cpu baseline: {cpu['value']:.3f} HMult/s")

    config3 = bench_config3(torch, ck, ctx)
    for k in ("bconv", "ntt"):
        config3[f"{k}_frac_of_device_copy"] = (config3[f"{k}_hbm_frac"] * HBM_PEAK_GBS /
                                              copy_bw["gb_s"])
    batched = None
    if not a.no_batched:
        batched = run_batched(torch, ck, steps=10, warmup=2)
        batched["hmult_speedup_vs_plain"] = batched["hmult_per_s"] / (1e3 / ms)
    boot = None if a.no_boot else bench_boot(a.boot_runs)

    res = {
        "metric": "hmult_relin_rescale_per_s",
        "value": 1e3 / ms,
        "unit": "HMult/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic uniform residues and evaluation key (seeded)",
        "config": {"workload": "HMult+relin+rescale, N=2^16, L=24, alpha=8, dnum=3, "
                               "plain HIP form", "bconv": a.bconv, "log_n": 16, "L": ctx.L, "alpha": ctx.alpha,
                   "dnum": ctx.dnum, "level": LEVEL, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "launch_us": launch_us,
                     "alg_bytes_per_launch": alg_bytes,
                     "frac_of_device_copy": achieved / copy_bw["gb_s"]},
        "device_copy": copy_bw,
        "cpu_baseline": cpu,
        "config3": config3,
        "batched": batched,
        "bootstrap": boot,
    }
    print(json.dumps(res), flush=True)


INT8_MFMA_PEAK_TOPS = 5000.0  # MI355X dense int8 matrix peak


def bench_config3(torch, ck, ctx, reps=10):
    """The config-3 contractions at their stated shape, 64 polynomials x 32
    limbs, each as one batched plan in every form: BaseConv 8 -> 32 limbs
    (the special primes to all primes) on the int8 matrix cores, on the FP64
    matrix cores and on the vector unit; the forward NTT of all 32 limbs of
    each polynomial as the int8 4-step transform and as radix-16 FP64
    butterflies on the vector unit, and the autotuned transform plan. All
    forms' outputs are compared on the GPU (they must be identical). Inputs
    resident in HBM; HIP events on the launch stream."""
    npoly, src_p, dst_p = 64, list(range(24, 32)), list(range(32))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0x4348454444415203)
    srcs = []
    for _ in range(npoly):
        v = torch.empty((len(src_p), ctx.n), dtype=torch.int32, device="cuda")
        for i, pi in enumerate(src_p):
            h = (ctx.prime(pi) - 1) // 2
            v[i] = torch.randint(-h, h + 1, (ctx.n,), generator=gen, device="cuda",
                                 dtype=torch.int64).to(torch.int32)
        srcs.append(v)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps  # ms

    b_bytes = npoly * (len(src_p) + len(dst_p)) * ctx.n * 4
    # matrix-core work as executed: int8 16x16x32 steps over K = 8 sources x
    # 4 digits, four byte planes; FP64 16x16x4 over {hi, lo} table halves
    b_int8_ops = npoly * len(dst_p) * ctx.n * 32 * 4 * 2
    b_f64_flop = npoly * len(src_p) * len(dst_p) * ctx.n * 2 * 2
    bconv, outs = {}, {}
    for form in ("int8", "fp64", "valu"):
        ctx.set_plan_forms(bconv=form)
        d = [torch.empty((len(dst_p), ctx.n), dtype=torch.int32, device="cuda")
             for _ in range(npoly)]
        plan = ck.Plan.bconv(ctx, src_p, dst_p, srcs, d)
        t = timed(lambda: plan.run(stream))
        outs[form] = d
        rec = {"ms": t, "hbm_gb_s": b_bytes / (t * 1e-3) / 1e9,
               "hbm_frac": b_bytes / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if form == "int8":
            rec["int8_mfma_tops"] = b_int8_ops / (t * 1e-3) / 1e12
            rec["int8_mfma_frac"] = rec["int8_mfma_tops"] / INT8_MFMA_PEAK_TOPS
        if form == "fp64":
            rec["fp64_mfma_tflop_s"] = b_f64_flop / (t * 1e-3) / 1e12
            rec["fp64_mfma_frac"] = rec["fp64_mfma_tflop_s"] / FP64_MFMA_PEAK_TFLOPS
        bconv[form] = rec
    bconv_equal = all(torch.equal(x, y) for f in ("fp64", "valu")
                      for x, y in zip(outs[f], outs["int8"]))
    ctx.set_plan_forms(bconv="int8")
    del outs

    polys = [uniform_poly(torch, ctx, dst_p, gen) for _ in range(npoly)]
    n_bytes = npoly * len(dst_p) * ctx.n * 8
    # int8 transform: four 16-point stages, each 4 planes of K = 16 x 4 digits
    n_int8_ops = npoly * len(dst_p) * ctx.n * 4 * 4 * 64 * 2
    ntt, nouts = {}, {}
    for form in ("int8", "valu", "auto"):
        ctx.set_plan_forms(ntt=form)
        o = [torch.empty_like(x) for x in polys]
        plan = ck.Plan.ntt(ctx, polys, o, (ctx.L, 0, ctx.alpha))
        t = timed(lambda: plan.run(stream))
        nouts[form] = o
        rec = {"ms": t, "us_per_limb": 1e3 * t / (npoly * len(dst_p)),
               "hbm_gb_s": n_bytes / (t * 1e-3) / 1e9,
               "hbm_frac": n_bytes / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
        if form == "int8":
            rec["int8_mfma_tops"] = n_int8_ops / (t * 1e-3) / 1e12
            rec["int8_mfma_frac"] = rec["int8_mfma_tops"] / INT8_MFMA_PEAK_TOPS
        if form == "auto":
            rec["chosen_form"] = "int8" if plan.info()["form"] == 1 else "valu"
            rec["candidates"] = [{"form": "int8" if f == 1 else "valu", "blocks": b,
                                  "chunk": c, "us": us} for f, b, c, us in plan.tuning()]
        ntt[form] = rec
    ntt_equal = all(torch.equal(x, y) for f in ("valu", "auto")
                    for x, y in zip(nouts[f], nouts["int8"]))
    ctx.set_plan_forms(ntt="int8")
    best_b = min(("int8", "fp64", "valu"), key=lambda f: bconv[f]["ms"])
    # the autotuned plan is a product form too (the fastest candidate kept)
    best_n = min(("int8", "valu", "auto"), key=lambda f: ntt[f]["ms"])
    res = {
        "shape": "64 polynomials x 32 limbs, N=2^16",
        "bconv_8_to_32": bconv, "bconv_fastest": best_b, "bconv_forms_equal": bconv_equal,
        "ntt_forward": ntt, "ntt_fastest": best_n, "ntt_forms_equal": ntt_equal,
        # the fastest form of each, under the summary keys
        "bconv_8_to_32_ms": bconv[best_b]["ms"], "bconv_hbm_frac": bconv[best_b]["hbm_frac"],
        "ntt_fwd_ms": ntt[best_n]["ms"], "ntt_us_per_limb": ntt[best_n]["us_per_limb"],
        "ntt_hbm_frac": ntt[best_n]["hbm_frac"],
    }
    log("config 3: BaseConv 8->32 x 64 polys: " +
        ", ".join(f"{f} {bconv[f]['ms']:.3f} ms ({bconv[f]['hbm_frac']:.2f} HBM)"
                  for f in bconv) + f"; forms equal: {bconv_equal}")
    log("config 3: forward NTT 64 x 32 limbs: " +
        ", ".join(f"{f} {ntt[f]['ms']:.3f} ms ({ntt[f]['hbm_frac']:.2f} HBM)" for f in ntt) +
        f"; forms equal: {ntt_equal}")
    if not (bconv_equal and ntt_equal):
        raise RuntimeError("config-3 forms disagree")
    return res


def bench_boot(runs):
    """Bootstrap latency and precision: cheddar-fhe_amd/boot_bench (the C++
    facade's BootContext) on bootparam_40.json, full slots: one ciphertext
    with a message uniform in the unit box (ms per bootstrap over `runs`
    after a warm-up), and 8 ciphertexts with messages uniform in the unit
    disk in one captured graph over 4 lane contexts, each lane a stream of
    its own (ms per bootstrap, precision over all 8, outputs equal to eager
    bootstraps of the same inputs)."""
    exe = os.path.join(REPO, "cheddar-fhe_amd", "boot_bench")
    prm = os.path.join(REPO, "tests", "golden", "params", "bootparam_40.json")
    if not os.path.exists(exe):
        raise RuntimeError("cheddar-fhe_amd/boot_bench is not built (run build())")
    log(f"bootstrap: {os.path.basename(prm)}, {runs} runs ...")
    r = subprocess.run([exe, prm, str(runs)], capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"boot_bench failed ({r.returncode}): {r.stdout[-2000:]} "
                           f"{r.stderr[-2000:]}")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    kern = out.pop("kernels", {})
    total = sum(v[0] for v in kern.values())
    by_name = {}
    for k, (kms, calls) in kern.items():
        base = k.split("/")[0]
        t, c = by_name.get(base, (0.0, 0))
        by_name[base] = (t + kms, c + calls)
    for k, (kms, calls) in sorted(by_name.items(), key=lambda kv: -kv[1][0]):
        log(f"  boot {k:28s} {calls:5d} launches {kms:8.3f} ms ({100 * kms / total:4.1f}%)")
    out["kernels_by_name"] = {k: [round(v[0], 4), v[1]] for k, v in by_name.items()}
    out["launches"] = sum(v[1] for v in by_name.values())
    log(f"bootstrap: {out['ms_per_bootstrap']:.2f} ms, precision {out['precision_bits_max']:.2f} "
        f"bits (max error), {out['precision_bits_mean']:.2f} bits (mean); kernel sum "
        f"{total:.2f} ms; {out['graph_batch']} in one captured graph over "
        f"{out['graph_lanes']} lanes: {out['graph_ms_per_bootstrap']:.2f} ms per bootstrap "
        f"(equal to eager: {out['graph_equal_to_eager']}) "
        f"({out['graph_precision_bits_max']:.2f} bits max, "
        f"{out['graph_precision_bits_mean']:.2f} mean, unit-disk messages); min-KS schedule: "
        f"{out['minks_ms_per_bootstrap']:.2f} ms ({out['minks_precision_bits_max']:.2f} bits)")
    out["kernel_ms"] = total
    out["config"] = ("bootparam_40.json as shipped: N=2^16, 2^15 slots, dense secret "
                     "(Hamming weight 32768) with sparse-secret encapsulation (32)")
    return out


if __name__ == "__main__":
    main()
