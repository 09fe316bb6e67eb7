"""Instance sweep for the block GEMM (the reference's bash sweeps, made a module).

For every compiled (shape, partition, P) instance of a precision, times two modes:
  streaming  -- `count` distinct problems read from and written to HBM once;
  resident   -- the reference's method: the product is repeated `iters` times inside
                the kernel on register/LDS-resident operands (NUM_ITER, F64-1D:140),
                TFLOPS = 2 M N K * count * iters / t (F64-1D:279-282).
Each line is the reference's block-GEMM CSV schema
`[hemeng_log],dim,M,N,K,NUM_RANK,NUM_ALLOC,TFLOPS,threads` (F64-1D:287) followed by
mode, problems, iters, fraction of MFMA peak, max error vs the FP64 oracle and the
pipeline (1: one workgroup per problem, 2: persistent with deep prefetch).

    python -m kami_amd.sweep --prec f64 --sizes 64 128 --count 4096
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

import kami_amd as K

PEAK_TFLOPS = {K.PREC_F64: 78.6, K.PREC_F16_F32: 2500.0, K.PREC_BF16_F32: 2500.0,
               K.PREC_F64_SPLIT_F16: 78.6, K.PREC_F64_SPLIT_BF16: 78.6}
PRECS = {"f64": K.PREC_F64, "f16": K.PREC_F16_F32, "bf16": K.PREC_BF16_F32,
         "split16": K.PREC_F64_SPLIT_F16, "splitbf16": K.PREC_F64_SPLIT_BF16}


def instances(prec, shape):
    m, n, k = shape
    out = []
    for part in ("1d", "2d", "3d"):
        for r in (1, 2, 4, 8, 16):
            for pipe in (1, 2):
                try:
                    cfg = K.query(m, n, k, prec, partition=part, ranks=r, pipeline=pipe)
                except K.KamiError:
                    continue
                if (part, r, pipe) not in [(o[0], o[1], o[3]) for o in out]:
                    out.append((part, r, cfg, pipe))
    return out


def make_inputs(prec, count, shape, seed, dev):
    import torch
    m, n, k = shape
    rng = np.random.default_rng(seed)
    if prec in (K.PREC_F16_F32, K.PREC_BF16_F32):
        dt = torch.float16 if prec == K.PREC_F16_F32 else torch.bfloat16
        A = torch.from_numpy(rng.standard_normal((count, m, k)).astype(np.float32)).to(dev).to(dt)
        B = torch.from_numpy(rng.standard_normal((count, n, k)).astype(np.float32)).to(dev).to(dt)
    else:
        A = torch.from_numpy(rng.uniform(-1, 1, (count, m, k))).to(dev)
        B = torch.from_numpy(rng.uniform(-1, 1, (count, n, k))).to(dev)
    return A, B


def time_ms(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def max_err(prec, A, B, C, sample=8):
    """max |C - C_fp64| / max |C_fp64| over the first `sample` problems."""
    import torch
    a = A[:sample].double()
    b = B[:sample].double()
    ref = torch.bmm(a, b.transpose(1, 2).contiguous())
    return float((C[:sample].double() - ref).abs().max() / ref.abs().max())


def sweep(prec, shapes, count, target_tflop, steps, warmup, modes, out=sys.stdout):
    import torch
    dev = torch.device("cuda:0")
    best = {}
    for shape in shapes:
        m, n, k = shape
        A, B = make_inputs(prec, count, shape, 0x4B414D49 + m + k, dev)
        out_dt = torch.float32 if prec in (K.PREC_F16_F32, K.PREC_BF16_F32) else torch.float64
        C = torch.empty((count, m, n), dtype=out_dt, device=dev)
        # resident mode: at most 2048 workgroups, repeats sized to ~target_tflop per call
        cres = min(count, 2048)
        it_res = max(1, int(round(target_tflop * 1e12 / (2.0 * m * n * k * cres))))
        for (part, r, cfg, pipe) in instances(prec, shape):
            for mode in modes:
                if mode == "resident" and pipe == 2:
                    continue  # the persistent form runs iters = 1 only
                it = it_res if mode == "resident" else 1
                cnt = cres if mode == "resident" else count
                fn = lambda: K.block_gemm(A[:cnt], B[:cnt], C[:cnt], precision=prec,  # noqa
                                          partition=part, ranks=r, iters=it, pipeline=pipe)
                ms = time_ms(fn, steps, warmup)
                tflops = 2.0 * m * n * k * cnt * it / (ms * 1e-3) / 1e12
                err = max_err(prec, A, B, C) if it == 1 else float("nan")
                frac = tflops / PEAK_TFLOPS[prec]
                dim = {1: "1d", 2: "2d", 3: "3d"}[cfg["partition"]]
                waves = cfg["waves_m"] * cfg["waves_n"] * cfg["waves_k"]
                print(f"[hemeng_log],{dim},{m},{n},{k},{r},{waves},{tflops:.3f},{cfg['threads']},"
                      f"{mode},{cnt},{it},{frac:.4f},{err:.3e},pipeline{pipe}", file=out, flush=True)
                key = (shape, mode)
                if key not in best or tflops > best[key][2]:
                    best[key] = (part, r, tflops, pipe)
    return best


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--prec", choices=sorted(PRECS), default="f64")
    ap.add_argument("--sizes", type=int, nargs="+", default=[64, 128])
    ap.add_argument("--lowrank", type=int, nargs="*", default=[],
                    help="K values of 256 x 256 x K low-rank shapes")
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--target-tflop", type=float, default=2.0,
                    help="resident mode: FLOP per call (in 1e12) that sets the repeat count")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--modes", nargs="+", default=["streaming", "resident"])
    args = ap.parse_args(argv)
    prec = PRECS[args.prec]
    shapes = [(s, s, s) for s in args.sizes] + [(256, 256, kk) for kk in args.lowrank]
    best = sweep(prec, shapes, args.count, args.target_tflop, args.steps, args.warmup, args.modes)
    summary = {f"{s[0]}x{s[1]}x{s[2]}/{mode}": {"partition": p, "ranks": r, "tflops": t,
                                                "pipeline": pp}
               for (s, mode), (p, r, t, pp) in best.items()}
    print("[best] " + json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
