"""Small drivers for profiling one kernel under rocprofv3.

    python -m kami_amd.profiling copy --mib 1024         # plain device copy (calibration)
    python -m kami_amd.profiling add --mib 512           # C = A + B: the w3 traffic mix
    python -m kami_amd.profiling w3 --path native        # FP64 128^3 x 4096 native
    python -m kami_amd.profiling w3 --path split         # same through FP16 slices
    python -m kami_amd.profiling w3 --path split_i8      # same through int8 digits
    python -m kami_amd.profiling w1 | w3r | w4 --prec f16 | w5 | w5r | w6
    python -m kami_amd.profiling w2 --shape 64x64x64 --part 3d --iters 200  # FP16 resident repeat

Each run launches the kernel `--reps` times after one warm-up, so per-dispatch counters
(FETCH_SIZE, WRITE_SIZE in KiB) can be divided by the algorithmic bytes of one launch.
The copy reads and writes exactly `--mib` MiB per launch: its counters give the scale
factor between counter KiB and bytes moved.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy", "add", "w1", "w2", "w3", "w3r", "w4", "w5", "w5r", "w6"])
    ap.add_argument("--shape", default="64x64x64")
    ap.add_argument("--part", choices=["1d", "2d", "3d"], default="2d")
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--count", type=int, default=2048)
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--path", choices=["native", "split", "split_i8"], default="native")
    ap.add_argument("--prec", choices=["f16", "f64"], default="f16")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)

    import torch

    import kami_amd as K
    from kami_amd import workloads as W

    dev = torch.device("cuda:0")
    if args.what == "copy":
        n = args.mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(args.reps + 1):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy: {args.reps + 1} launches of {args.mib} MiB read + {args.mib} MiB written")
        return
    if args.what == "add":
        # two reads and one write of --mib MiB each: the bytes of one w3 launch at 512
        n = args.mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device=dev)
        b = torch.ones_like(a)
        c = torch.empty_like(a)
        for _ in range(args.reps + 1):
            torch.add(a, b, out=c)
        torch.cuda.synchronize()
        print(f"add: {args.reps + 1} launches of 2 x {args.mib} MiB read + {args.mib} MiB written")
        return
    if args.what == "w2":
        # FP16 block GEMM in the resident-repeat mode (the reference's NUM_ITER loop):
        # N(0,1) inputs, `iters` products per launch on `count` problems
        m, n, k = (int(x) for x in args.shape.split("x"))
        g = torch.Generator(device=dev)
        g.manual_seed(W.SEED + 1)
        A = torch.randn((args.count, m, k), generator=g, device=dev).half()
        B = torch.randn((args.count, n, k), generator=g, device=dev).half()
        C = torch.empty((args.count, m, n), dtype=torch.float32, device=dev)
        for _ in range(args.reps + 1):
            K.block_gemm(A, B, C, partition=args.part, iters=args.iters)
        torch.cuda.synchronize()
        flop = 2.0 * m * n * k * args.count * args.iters
        print(f"w2 {args.shape} {args.part}: {args.reps + 1} launches, {flop:.4g} FLOP each")
        return
    if args.what == "w3":
        rng = np.random.default_rng(W.SEED + 2)
        m = n = k = 128
        count = 4096
        A = torch.from_numpy(W.log_uniform_signed(rng, (count, m, k))).to(dev)
        B = torch.from_numpy(W.log_uniform_signed(rng, (count, n, k))).to(dev)
        C = torch.empty((count, m, n), dtype=torch.float64, device=dev)
        prec = {"native": K.PREC_F64, "split": K.PREC_F64_SPLIT_F16,
                "split_i8": K.PREC_F64_SPLIT_I8}[args.path]
        for _ in range(args.reps + 1):
            K.block_gemm(A, B, C, precision=prec)
        torch.cuda.synchronize()
        byts = 8 * (m * k + n * k + m * n) * count
        print(f"w3 {args.path}: {args.reps + 1} launches, {byts} algorithmic bytes each")
        return
    if args.what == "w1":
        W.run_w1(steps=args.reps, warmup=1)
    elif args.what == "w3r":
        W.run_w3r(steps=args.reps, warmup=1)
    elif args.what == "w4":
        W.run_w4(steps=args.reps, warmup=1, precisions=(args.prec,))
    elif args.what == "w5":
        W.run_w5(steps=args.reps, warmup=1)
    elif args.what == "w5r":
        print(W.run_w5r(steps=args.reps, warmup=1))
    elif args.what == "w6":
        W.run_w6(steps=args.reps, warmup=1)
    torch.cuda.synchronize()
    print(f"{args.what}: done", file=sys.stderr)


if __name__ == "__main__":
    main()
