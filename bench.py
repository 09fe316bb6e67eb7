#!/usr/bin/env python3
"""Benchmark of the MI355X block-GEMM path (one JSON line on stdout).

Default workload: FP64 block GEMM C = A B^T at M = N = K = 128 over 4096 independent
problems (one per workgroup), entries sign-random with log10|x| ~ U[-4, 4].  It is timed
on the same inputs as native FP64 MFMA and as split-precision FP64: int8 digits on int8
MFMA (split_i8), FP16 digits on FP16 MFMA (split: per-chunk exponents; split_fullk:
exponents over the whole K).  `value` is the TFLOPS (2MNK per problem, FP64-equivalent)
of the path named by --path (default: the fastest path whose error against the
compensated Dot2 oracle is at most 1e-13 componentwise, |C - C*| <= 1e-13 (|A||B|), on a
sample of problems).

Inputs are resident in HBM before timing.  W untimed warm-up steps, then K timed steps
bracketed by device synchronisation; one step is one library call over the whole batch.
The defaults (W = 100, K = 200: about 0.1 s of timed work) measure sustained operation:
under load the GPU's power management first lowers the clock (launches 10-40 of a cold
start run ~15 % slower than the first few) and then settles, and the settled rate is
what a long-running caller sees.
The roofline object describes the dominant kernel of the reported path; the CPU
baseline times the oracle's OpenMP FP64 triple loop (the reference's CPU check loop)
on the host cores, and `cpu_baselines` the reference's CPU loops of the other workloads.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64_mfma_tflops": 78.6, "f16_mfma_tflops": 2500.0, "hbm_gbs": 8000.0}
SEED = 0x4B414D49

METRIC = "Batched block-GEMM TFLOPS & % MFMA peak, FP64 native vs split-FP16; max rel err"


def log_uniform_signed(rng, shape, lo=-4.0, hi=4.0):
    mag = 10.0 ** rng.uniform(lo, hi, size=shape)
    sign = np.where(rng.random(shape) < 0.5, -1.0, 1.0)
    return sign * mag


def time_calls(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps  # ms per step


def cpu_baseline_f64(m, n, k, count, budget_s=10.0):
    """The oracle's OpenMP FP64 triple loop over the same workload shape, repeated until
    about `budget_s` seconds of CPU time have run (a bounded sample)."""
    import oracle
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    rng = np.random.default_rng(SEED + 99)
    A = log_uniform_signed(rng, (count, m, k))
    B = log_uniform_signed(rng, (count, n, k))
    oracle.gemm_f64(A[:8], B[:8], nthreads=cores)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.gemm_f64(A, B, nthreads=cores)
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or reps >= 200:
            break
    tflops = 2.0 * m * n * k * count * reps / dt / 1e12
    return {"value": tflops, "unit": "TFLOPS", "cores": cores, "kind": "port",
            "sample": f"{reps} passes over all {count} problems {m}x{n}x{k} FP64 (oracle "
                      f"triple loop, OpenMP over problems, {cores} threads), {dt:.1f} s",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


def cpu_model():
    """The host CPU's model name from /proc/cpuinfo (None where it is not readable)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


RECORD_KEYS = {"native": "w3_native", "split": "w3_split", "split_i8": "w3_split_i8"}


def _timed_passes(fn, budget_s):
    fn()  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or reps >= 200:
            return reps, dt


def cpu_baselines_other(budget_s=2.5):
    """The reference's CPU loops for the other workloads, each timed for about `budget_s`
    seconds on the host cores (OpenMP over problems / block rows): FP64 64^3 x 1024 (the
    triple loop, F64-1D:294-304), FP16 128^3 x 8192 (the FP16 check loop with double
    accumulation, F16-1D:285-296), the 20000 uneven FP64 problems (BAT-1D:304-318 over
    problems of their own sizes) and the 8192^2 block SpGEMM with the SPA (symbolic with
    the column fill, and numeric, timed separately, SPG-1D:285-370)."""
    import oracle
    from kami_amd import workloads as W
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    out = {"cores": cores, "kind": "port"}
    rng = np.random.default_rng(SEED + 98)
    A = rng.uniform(-1, 1, (1024, 64, 64))
    B = rng.uniform(-1, 1, (1024, 64, 64))
    reps, dt = _timed_passes(lambda: oracle.gemm_f64(A, B, nthreads=cores), budget_s)
    out["w1_f64_64cubed_x1024"] = {"tflops": 2.0 * 64 ** 3 * 1024 * reps / dt / 1e12,
                                   "sample": f"{reps} passes over all 1024 problems, {dt:.1f} s"}
    cnt = 8192
    Ah = rng.standard_normal((cnt, 128, 128)).astype(np.float16).view(np.uint16)
    Bh = rng.standard_normal((cnt, 128, 128)).astype(np.float16).view(np.uint16)
    reps, dt = _timed_passes(lambda: oracle.gemm_u16(Ah, Bh, nthreads=cores), budget_s)
    out["w2_f16_128cubed_x8192"] = {"tflops": 2.0 * 128 ** 3 * cnt * reps / dt / 1e12,
                                    "sample": f"{reps} passes over all {cnt} problems, {dt:.1f} s"}
    del Ah, Bh
    sizes = W.uneven_sizes(np.random.default_rng(SEED + 3), 20000)
    na = int((sizes[:, 0] * sizes[:, 2]).sum())
    nb = int((sizes[:, 1] * sizes[:, 2]).sum())
    Au, Bu = rng.uniform(-1, 1, na), rng.uniform(-1, 1, nb)
    reps, dt = _timed_passes(lambda: oracle.batched_uneven_f64(sizes, Au, Bu, nthreads=cores),
                             budget_s)
    flop = 2.0 * float((sizes[:, 0] * sizes[:, 1] * sizes[:, 2]).sum())
    out["w4_batched_uneven_x20000_f64"] = {"tflops": flop * reps / dt / 1e12,
                                           "sample": f"{reps} passes over all 20000 problems, {dt:.1f} s"}
    rp, ci, vals = W.spgemm_input(np.random.default_rng(SEED + 4))
    mb, bs = 256, 32
    v16 = vals.view(np.uint16)

    def symbolic():
        rpc, _ = oracle.spgemm_symbolic(mb, mb, rp, ci, rp, ci)
        oracle.spgemm_fill(mb, mb, rp, ci, rp, ci, rpc)
    reps, dt = _timed_passes(symbolic, budget_s / 2)
    rpc, nnzc = oracle.spgemm_symbolic(mb, mb, rp, ci, rp, ci)
    sym_ms = dt / reps * 1e3
    reps, dt = _timed_passes(lambda: oracle.spgemm_numeric(mb, mb, bs, rp, ci, v16, rp, ci, v16, rpc),
                             budget_s / 2)
    nnzcub = oracle.spgemm_nnzcub(ci, rp)
    out["w5_block_spgemm"] = {"symbolic_ms": sym_ms, "numeric_ms": dt / reps * 1e3,
                              "numeric_tflops": 2.0 * nnzcub * bs ** 3 * reps / dt / 1e12,
                              "nnzb_C": int(nnzc)}
    return out


def measured_traffic(path, cfg, m, count):
    """HBM bytes per launch of the path's kernel from the PMC record in profiles/traffic.json
    (FETCH_SIZE and WRITE_SIZE passes scaled by a 1 GiB copy), when that record was taken
    on the same instance and workload; otherwise None."""
    rec = None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f).get(RECORD_KEYS.get(path, ""))
    except (OSError, ValueError):
        return None
    if not rec or m != 128 or count != 4096:
        return None
    if any(cfg.get(k) != v for k, v in rec.get("config", {}).items()):
        return None
    return rec["traffic_bytes"]


def measured_lds(path, cfg, m, count):
    """LDS traffic of the path's kernel from the SQ_INSTS_LDS record in profiles/lds.json
    (same instance and workload only): bytes per FLOP where every LDS access of the
    kernel has one width, else the instruction count alone; None without a record."""
    try:
        with open(os.path.join(ROOT, "profiles", "lds.json")) as f:
            rec = json.load(f).get(RECORD_KEYS.get(path, ""))
    except (OSError, ValueError):
        return None
    if not rec or m != 128 or count != 4096:
        return None
    if any(cfg.get(k) != v for k, v in rec.get("config", {}).items()):
        return None
    out = {"instructions_per_launch": rec["SQ_INSTS_LDS"],
           "bank_conflict_cycles": rec["SQ_LDS_BANK_CONFLICT"], "source": "rocprofv3 --pmc SQ_INSTS_LDS"}
    if rec.get("bytes_per_instruction"):
        out["bytes_per_flop"] = rec["SQ_INSTS_LDS"] * rec["bytes_per_instruction"] / rec["flop_per_launch"]
    elif rec.get("lds_bytes_per_launch"):  # mixed widths: bytes counted from the kernel's mix
        out["bytes_per_flop"] = rec["lds_bytes_per_launch"] / rec["flop_per_launch"]
    return out


def rel_errors(C, A, B, sample):
    """Max errors of C against the compensated (Dot2) oracle on the first `sample`
    problems: componentwise |dC| / (|A||B|) and normwise max|dC| / max|C*| per problem."""
    import oracle
    exact, absC = oracle.gemm_dot2(A[:sample], B[:sample])
    d = np.abs(C[:sample] - exact)
    comp = float(np.max(d / np.where(absC > 0, absC, 1.0)))
    norm = float(np.max(np.max(d, axis=(1, 2)) / np.maximum(np.max(np.abs(exact), axis=(1, 2)), 1e-300)))
    return {"componentwise": comp, "normwise": norm}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--path", choices=["auto", "native", "split_i8", "split", "split_fullk"],
                    default="auto")
    ap.add_argument("--m", type=int, default=128)
    ap.add_argument("--problems", type=int, default=4096)
    ap.add_argument("--err-sample", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true",
                    help="only the headline workload (skip the other workloads)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard the problems across GPUs, one process each)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import kami_amd as K

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    m = n = k = args.m
    count = args.problems
    rng = np.random.default_rng(SEED + 2)
    A = log_uniform_signed(rng, (count, m, k))
    B = log_uniform_signed(rng, (count, n, k))
    Ad = torch.from_numpy(A).to(dev)
    Bd = torch.from_numpy(B).to(dev)
    flops = 2.0 * m * n * k * count
    bytes_alg = 8.0 * (m * k + n * k + m * n) * count

    results = {}
    outs = {}
    # split_i8: int8 digits, persistent, exponents over the whole K; split: FP16 digits,
    # persistent with per-chunk exponents; split_fullk: FP16 digits, one workgroup per
    # problem with exponents over the whole K
    paths = [("native", K.PREC_F64, 0), ("split_i8", K.PREC_F64_SPLIT_I8, 0),
             ("split", K.PREC_F64_SPLIT_F16, 0), ("split_fullk", K.PREC_F64_SPLIT_F16, 1)]
    for name, prec, pipe in paths:
        try:
            cfg = K.query(m, n, k, prec, pipeline=pipe)
        except K.KamiError as e:
            results[name] = {"available": False, "reason": str(e)}
            continue
        Cd = torch.empty((count, m, n), dtype=torch.float64, device=dev)
        fn = lambda: K.block_gemm(Ad, Bd, Cd, precision=prec, pipeline=pipe,  # noqa: E731
                                  stream=stream)
        ms = time_calls(fn, args.steps, args.warmup, stream)
        C = Cd.cpu().numpy()
        outs[name] = C
        err = rel_errors(C, A, B, min(args.err_sample, count))
        tflops = flops / (ms * 1e-3) / 1e12
        results[name] = {"available": True, "ms": ms, "tflops": tflops,
                         "frac_f64_peak": tflops / PEAK["f64_mfma_tflops"],
                         "hbm_gbs": bytes_alg / (ms * 1e-3) / 1e9, "max_rel_err": err,
                         "config": cfg}
        print(f"[bench] {name}: {ms:.4f} ms/step  {tflops:.2f} TFLOPS  err={err}",
              file=sys.stderr, flush=True)

    ok = [p for p in ("split_i8", "split", "split_fullk", "native")
          if results.get(p, {}).get("available")
          and (p == "native" or results[p]["max_rel_err"]["componentwise"] <= 1e-13)]
    if args.path == "auto":
        chosen = max(ok, key=lambda p: results[p]["tflops"])
    else:
        chosen = args.path
        if not results.get(chosen, {}).get("available"):
            print(f"bench.py: path {chosen} unavailable", file=sys.stderr)
            sys.exit(1)
    r = results[chosen]
    ms = r["ms"]
    traffic = measured_traffic(chosen, r["config"], m, count)
    lds = measured_lds(chosen, r["config"], m, count)
    if chosen == "native":
        roof = {"bound": "mfma", "achieved": r["tflops"], "peak": PEAK["f64_mfma_tflops"],
                "unit": "TFLOP/s", "frac": r["tflops"] / PEAK["f64_mfma_tflops"],
                "traffic": traffic}
    else:
        gbs = r["hbm_gbs"]
        roof = {"bound": "hbm", "achieved": gbs, "peak": PEAK["hbm_gbs"], "unit": "GB/s",
                "frac": gbs / PEAK["hbm_gbs"], "traffic": traffic}
    cpu = None if args.no_cpu else cpu_baseline_f64(m, n, k, count)
    dtypes = {"native": "f64", "split_i8": "i8-digits/i32-acc -> f64",
              "split": "f16-digits/f32-acc -> f64", "split_fullk": "f16-digits/f32-acc -> f64"}
    line = {
        "metric": METRIC,
        "value": r["tflops"],
        "unit": "TFLOPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtypes[chosen],
        "data": "synthetic",
        "config": {"workload": f"FP64 block GEMM {m}x{n}x{k}, {count} problems, log-uniform "
                               f"|x| in [1e-4, 1e4] with random signs; path={chosen}",
                   "m": m, "n": n, "k": k, "problems": count,
                   "partition": r["config"]["partition"], "ranks": r["config"]["ranks"]},
        "roofline": roof,
        "lds": lds,
        "cpu_baseline": cpu,
        "paths": {p: ({kk: vv for kk, vv in v.items() if kk != "config"} if v.get("available")
                      else v) for p, v in results.items()},
        "max_rel_err": r["max_rel_err"],
    }
    if not args.quick:
        from kami_amd import workloads as W
        others = {}
        for name, fn in [("w1_f64_64cubed_x1024", W.run_w1), ("w2_f16_block_gemm", W.run_w2),
                         ("w3r_resident_f64_64cubed_x4096", W.run_w3r),
                         ("w4_batched_uneven_x20000", W.run_w4), ("w5_block_spgemm", W.run_w5),
                         ("w5r_spgemm_reference_measurement", W.run_w5r),
                         ("w6_block_spmm", W.run_w6)]:
            t0 = time.perf_counter()
            try:
                others[name] = fn()
            except Exception as e:  # reported, not hidden: the line says which failed
                others[name] = {"error": f"{type(e).__name__}: {e}"}
            print(f"[bench] {name} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
        line["workloads"] = others
        if not args.no_cpu:
            try:
                line["cpu_baselines"] = cpu_baselines_other()
            except Exception as e:  # reported, not hidden
                line["cpu_baselines"] = {"error": f"{type(e).__name__}: {e}"}
    line["targets"] = targets(results, line.get("workloads", {}))
    print(json.dumps(line))


def targets(results, workloads):
    """The two stated targets in one small object (last in the line): split-precision
    FP64 against native FP64 at the same shapes, streaming (w3, 128^3 x 4096) and
    resident-repeat (w3r, 64^3 x 4096), with each split path's componentwise error; and
    the best FP16 resident-repeat fraction of the FP16 MFMA peak per w2 shape."""
    t = {"bound_componentwise": 1e-13}
    nat = results.get("native", {})
    for p in ("split_i8", "split"):
        r = results.get(p, {})
        if r.get("available") and nat.get("available"):
            t[f"w3_{p}_over_native"] = r["tflops"] / nat["tflops"]
            t[f"w3_{p}_componentwise"] = r["max_rel_err"]["componentwise"]
    w3r = workloads.get("w3r_resident_f64_64cubed_x4096", {})
    for p in ("split_i8", "split_f16"):
        if f"{p}_over_native" in w3r:
            t[f"w3r_{p}_over_native"] = w3r[f"{p}_over_native"]
            t[f"w3r_{p}_componentwise"] = w3r[p]["max_rel_err_vs_torch_fp64"]["componentwise"]
    w2 = workloads.get("w2_f16_block_gemm", {})
    best = {}
    for shape, parts in w2.items() if isinstance(w2, dict) else []:
        fr = [(v["resident_frac_f16_peak"], k) for k, v in parts.items() if v.get("available", True)
              and "resident_frac_f16_peak" in v]
        if fr:
            f, part = max(fr)
            best[shape] = [round(f, 3), part]
    if best:
        t["w2_best_resident_frac_f16_peak"] = best
    return t


if __name__ == "__main__":
    main()
