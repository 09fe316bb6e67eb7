"""Synthetic workloads of the block-GEMM path and their GPU measurements.

Five workloads (seed 0x4B414D49 + index, numpy PCG64 on the host or torch's generator
on the device):
  w1  FP64 block GEMM, M = N = K = 64, 1024 problems, U[-1, 1], 2D partition;
  w2  FP16 in / FP32 accumulate, M = N = K in {32, 64, 128, 256} and 256 x 256 x {16, 32},
      8192 problems per shape, N(0, 1) cast to FP16, every partition timed;
  w3  FP64 through int8 / FP16 digits, M = N = K = 128, 4096 problems, sign-random
      entries with log10|x| ~ U[-4, 4], against native FP64 (timed by bench.py);
  w4  batched uneven sizes: 20000 problems, M, N, K log-uniform in [16, 256] rounded to
      multiples of 16, U[-1, 1], FP16->FP32 and FP64, one persistent launch per
      execution in a hipGraph, autotuned per size bucket;
  w5  block SpGEMM C = A * A, 8192 x 8192 with 32 x 32 blocks at 2 % block density,
      N(0, 1) values in FP16, FP32 accumulation.
Timing: inputs resident in HBM, warm-up calls, then torch.cuda events on the stream the
library launches on.  FLOP counts are algorithmic: 2 M N K per problem (FP64-equivalent
for split precision) and 2 * nnzCub * bs^3 for SpGEMM (the reference's accounting,
src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:513).
"""
from __future__ import annotations

import numpy as np

import kami_amd as K

SEED = 0x4B414D49
PEAK = {"f64": 78.6, "f16": 2500.0, "hbm_gbs": 8000.0}


def _torch():
    import torch
    return torch


def time_ms(fn, steps=10, warmup=2):
    torch = _torch()
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


def graph_time_ms(fn, steps=10, warmup=2):
    """Milliseconds per call of `fn` (launches on the current stream only), with `steps`
    calls captured into one graph and replayed: host overhead between calls drops out."""
    torch = _torch()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            for _ in range(steps):
                fn()
    torch.cuda.current_stream().wait_stream(s)
    g.replay()  # first replay: uploads the graph
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    g.replay()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def log_uniform_signed(rng, shape, lo=-4.0, hi=4.0):
    return np.where(rng.random(shape) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(lo, hi, shape)


# ------------------------------------------------------------------ w1

def run_w1(steps=20, warmup=3):
    torch = _torch()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(SEED + 0)
    m = n = k = 64
    count = 1024
    A = torch.from_numpy(rng.uniform(-1, 1, (count, m, k))).to(dev)
    B = torch.from_numpy(rng.uniform(-1, 1, (count, n, k))).to(dev)
    C = torch.empty((count, m, n), dtype=torch.float64, device=dev)
    ms = time_ms(lambda: K.block_gemm(A, B, C, partition="2d"), steps, warmup)
    flops = 2.0 * m * n * k * count
    byts = 8.0 * (m * k + n * k + m * n) * count
    ref = torch.bmm(A[:64], B[:64].transpose(1, 2).contiguous())
    err = float((C[:64] - ref).abs().max() / ref.abs().max())
    tf = flops / ms / 1e9
    return {"workload": "FP64 64x64x64 x1024, U[-1,1], 2D", "ms": ms, "tflops": tf,
            "frac_f64_peak": tf / PEAK["f64"], "hbm_gbs": byts / ms / 1e6,
            "note": "100 MB of operands fit the 256 MB last-level cache: MALL-warm",
            "max_rel_err_vs_fp64": err}


# ------------------------------------------------------------------ w2

W2_SHAPES = [(32, 32, 32), (64, 64, 64), (128, 128, 128), (256, 256, 256), (256, 256, 16),
             (256, 256, 32)]


def run_w2(count=8192, steps=10, warmup=2, res_count=16384, res_iters=1000):
    """FP16 block GEMM at the cfg2 shapes, every partition: streaming over `count`
    problems, and the reference's resident-repeat measurement, `res_count` problems (its
    BLOCKS_PER_GRID) each accumulating its product `res_iters` times inside the kernel
    (NUM_ITER; src/block_gemm/KAMI/1d/fp16/block_gemm_1d_half_mma.cu:37,41,130,256)."""
    torch = _torch()
    dev = torch.device("cuda:0")
    out = {}
    for (m, n, k) in W2_SHAPES:
        g = torch.Generator(device=dev)
        g.manual_seed(SEED + 1 + m * 1000 + k)
        A = torch.randn((count, m, k), generator=g, device=dev, dtype=torch.float32).half()
        B = torch.randn((count, n, k), generator=g, device=dev, dtype=torch.float32).half()
        C = torch.empty((count, m, n), dtype=torch.float32, device=dev)
        ref = torch.bmm(A[:16].double(), B[:16].double().transpose(1, 2).contiguous())
        flops = 2.0 * m * n * k * count
        byts = (2.0 * (m * k + n * k) + 4.0 * m * n) * count
        reps = (res_count + count - 1) // count  # the resident problems: the inputs repeated
        Ar = A.repeat(reps, 1, 1)[:res_count]
        Br = B.repeat(reps, 1, 1)[:res_count]
        Cr = torch.empty((res_count, m, n), dtype=torch.float32, device=dev)
        shape_out = {}
        for part in ("1d", "2d", "3d"):
            try:
                cfg = K.query(m, n, k, K.PREC_F16_F32, partition=part)
            except K.KamiError:
                shape_out[part] = {"available": False}
                continue
            ms = time_ms(lambda: K.block_gemm(A, B, C, partition=part), steps, warmup)
            err = float((C[:16].double() - ref).abs().max() / ref.abs().max())
            msr = time_ms(lambda: K.block_gemm(Ar, Br, Cr, partition=part, iters=res_iters), 3, 1)
            tf = flops / ms / 1e9
            tfr = 2.0 * m * n * k * res_count * res_iters / msr / 1e9
            shape_out[part] = {"ranks": cfg["ranks"], "streaming_tflops": tf,
                               "streaming_hbm_gbs": byts / ms / 1e6,
                               "resident_tflops": tfr, "resident_frac_f16_peak": tfr / PEAK["f16"],
                               "max_rel_err_vs_fp64": err}
        out[f"{m}x{n}x{k}"] = shape_out
        del A, B, C, Ar, Br, Cr
    return out


# ------------------------------------------------------------------ w3r

def run_w3r(count=4096, iters=200, steps=20, warmup=10):
    """Resident-repeat FP64 block GEMM at 64^3 (the reference's NUM_ITER mode): operands
    read once, `iters` products accumulated inside the kernel.  Native FP64 MFMA against
    FP64 through FP16 / BF16 slices (digits made once; the slice products and their FP64
    recombination repeated), log-uniform signed inputs as in the split workload.  Errors are
    against torch's FP64 product (the error of that reference is itself ~1e-16 |A||B|)."""
    torch = _torch()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(SEED + 5)
    s = 64
    A = torch.from_numpy(log_uniform_signed(rng, (count, s, s))).to(dev)
    B = torch.from_numpy(log_uniform_signed(rng, (count, s, s))).to(dev)
    C = torch.empty((count, s, s), dtype=torch.float64, device=dev)
    ref = torch.bmm(A[:32], B[:32].transpose(1, 2).contiguous())
    flops = 2.0 * s ** 3 * count * iters
    out = {"workload": f"FP64 block GEMM 64x64x64 x{count}, log-uniform |x| in [1e-4, 1e4], "
                       f"resident-repeat iters={iters}"}
    absAB = torch.bmm(A[:32].abs(), B[:32].abs().transpose(1, 2).contiguous())
    for name, prec in [("native", K.PREC_F64), ("split_i8", K.PREC_F64_SPLIT_I8),
                       ("split_f16", K.PREC_F64_SPLIT_F16), ("split_bf16", K.PREC_F64_SPLIT_BF16)]:
        cfg = K.query(s, s, s, prec, iters=iters)
        ms = time_ms(lambda: K.block_gemm(A, B, C, precision=prec, iters=iters), steps, warmup)
        d = ((C[:32] / iters) - ref).abs()
        tf = flops / ms / 1e9
        out[name] = {"ms": ms, "tflops": tf, "frac_f64_peak": tf / PEAK["f64"],
                     "max_rel_err_vs_torch_fp64": {
                         "normwise": float(d.max() / ref.abs().max()),
                         "componentwise": float((d / absAB).max())},
                     "partition": cfg["partition"], "threads": cfg["threads"],
                     "slices": cfg["slices"]}
    for name in ("split_i8", "split_f16"):
        out[f"{name}_over_native"] = out[name]["tflops"] / out["native"]["tflops"]
    return out


# ------------------------------------------------------------------ w4

def uneven_sizes(rng, count, lo=16, hi=256):
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 3)))
    return np.clip(np.round(x / 16) * 16, lo, hi).astype(np.int64)


def run_w4(count=20000, steps=10, warmup=2, precisions=("f16", "f64")):
    torch = _torch()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(SEED + 3)
    sizes = uneven_sizes(rng, count)
    out = {"problems": count}
    for pname in precisions:
        in_dt = torch.float16 if pname == "f16" else torch.float64
        out_dt = torch.float32 if pname == "f16" else torch.float64
        na = int((sizes[:, 0] * sizes[:, 2]).sum())
        nb = int((sizes[:, 1] * sizes[:, 2]).sum())
        nc = int((sizes[:, 0] * sizes[:, 1]).sum())
        g = torch.Generator(device=dev)
        g.manual_seed(SEED + 3)
        Abuf = (torch.rand(na, generator=g, device=dev, dtype=torch.float64) * 2 - 1).to(in_dt)
        Bbuf = (torch.rand(nb, generator=g, device=dev, dtype=torch.float64) * 2 - 1).to(in_dt)
        Cbuf = torch.empty(nc, dtype=out_dt, device=dev)
        probs = []
        oa = ob = oc = 0
        for (m, n, k) in sizes:
            m, n, k = int(m), int(n), int(k)
            probs.append((Abuf[oa:oa + m * k].view(m, k), Bbuf[ob:ob + n * k].view(n, k),
                          Cbuf[oc:oc + m * n].view(m, n)))
            oa += m * k
            ob += n * k
            oc += m * n
        flops = 2.0 * float((sizes[:, 0] * sizes[:, 1] * sizes[:, 2]).sum())
        elt = 2 if pname == "f16" else 8
        byts = elt * (na + nb) + (4 if pname == "f16" else 8) * nc
        with K.BatchedPlan(probs, autotune=True, use_graph=True) as plan:
            info = plan.info()
            ms = time_ms(plan.execute, steps, warmup)
        errs = []
        for idx in range(0, count, max(1, count // 32)):
            A, B, C = probs[idx]
            ref = A.double() @ B.double().t()
            errs.append(float((C.double() - ref).abs().max() / ref.abs().max()))
        tf = flops / ms / 1e9
        peak = PEAK["f16"] if pname == "f16" else PEAK["f64"]
        out[pname] = {"ms": ms, "tflops": tf, "frac_mfma_peak": tf / peak,
                      "hbm_gbs": byts / ms / 1e6, "gflop": flops / 1e9,
                      "work_items": info["work_items"], "buckets": info["buckets"],
                      "workgroups": info["workgroups"], "max_rel_err_vs_fp64": max(errs)}
        del probs, Abuf, Bbuf, Cbuf
    return out


# ------------------------------------------------------------------ w5

def spgemm_input(rng, mb=256, bs=32, density=0.02):
    mask = rng.random((mb, mb)) < density
    rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    ci = np.nonzero(mask)[1].astype(np.int32)
    vals = rng.standard_normal((len(ci), bs, bs)).astype(np.float16)
    return rp, ci, vals


def run_w5(steps=10, warmup=2):
    torch = _torch()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(SEED + 4)
    mb, bs = 256, 32
    rp, ci, vals = spgemm_input(rng, mb, bs)
    A = K.BsrMatrix(mb, mb, bs, torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev),
                    torch.from_numpy(vals.view(np.int16)).to(dev).view(torch.float16))
    # nnzCub = sum over A's blocks (i, k) of nnz(B row k), the reference's FLOP unit
    nnzcub = int(sum(int(rp[c + 1] - rp[c]) for c in ci))
    ws = K.spgemm_workspace(A, A)  # the rows' column bitmaps, reused by the fill
    rowptr, nnz = K.spgemm_symbolic(A, A, workspace=ws)
    colidx = K.spgemm_fill(A, A, rowptr, nnz, workspace=ws)

    def symbolic():
        r, z = K.spgemm_symbolic(A, A, workspace=ws)
        K.spgemm_fill(A, A, r, z, workspace=ws)

    ms_sym = time_ms(symbolic, steps, warmup)
    Cv = torch.empty((nnz, bs, bs), dtype=torch.float32, device=dev)

    def numeric():
        a = A._c()
        import ctypes
        K.check(K.lib().kami_bsr_spgemm_numeric(
            ctypes.byref(a), ctypes.byref(a), ctypes.c_void_p(rowptr.data_ptr()),
            ctypes.c_void_p(colidx.data_ptr()), nnz, ctypes.c_void_p(Cv.data_ptr()),
            K.PREC_F16_F32, 1.0, K._stream_ptr(None)), "numeric")

    # the numeric calls replayed from a captured graph (the ~10 us kernel is shorter than a
    # Python call through ctypes, so eager timing measures the host); the eager figure
    # is kept beside it
    ms_num = graph_time_ms(numeric, 50, warmup)
    ms_num_eager = time_ms(numeric, steps, warmup)
    flops = 2.0 * nnzcub * bs ** 3
    byts = 2.0 * len(ci) * bs * bs + 4.0 * nnz * bs * bs  # A read once (A = B), C written
    return {"workload": "BSR C=A*A, 8192^2, 32x32 blocks, 2% density, FP16->FP32",
            "nnzb_A": int(len(ci)), "nnzb_C": int(nnz), "nnzCub": nnzcub,
            "symbolic_ms": ms_sym, "numeric_ms": ms_num, "numeric_tflops": flops / ms_num / 1e9,
            "numeric_hbm_gbs": byts / ms_num / 1e6, "numeric_ms_eager": ms_num_eager,
            "symbolic_ms_host_program": symbolic_ms_cli()}


def symbolic_ms_cli(blocks=256, bs=32, sparsity=0.98, reps=20):
    """The symbolic phase timed by the command-line driver, the way the reference's host
    program times it (count, scan, the total read back, C's column array allocated,
    fill; SPG-1D:285-370), on the reference's generator at the w5 size and density:
    milliseconds per call over `reps` calls, or None when the driver is not built."""
    import os
    import re
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kami_bench")
    if not os.path.exists(cli):
        return None
    r = subprocess.run([cli, "spgemm", "--blocks", str(blocks), "--bs", str(bs), "--sparsity",
                        str(sparsity), "--sym-reps", str(reps)], capture_output=True, text=True,
                       timeout=300)
    m = re.search(r"Symbolic time: ([0-9.e+-]+) ms", r.stdout)
    return float(m.group(1)) if r.returncode == 0 and m else None


# ------------------------------------------------------------------ w5r

def reference_bsr(blocks=8, bs=16, sparsity=0.5):
    """The reference programs' BSR input (src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu,
    the generator at :372-397): srand(0), then for every block (i, j) in row order a
    block is stored when (float) rand() / (float) RAND_MAX > sparsity, and its bs x bs
    values are rand() % 3 + 1, drawn right after.  The C library's own rand() is called
    through ctypes, so the pattern is the reference's on this host's glibc."""
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(0)
    rand_max = np.float32(2147483647)
    rp, ci, vals = [0], [], []
    for _ in range(blocks):
        for j in range(blocks):
            if np.float32(libc.rand()) / rand_max > np.float32(sparsity):
                ci.append(j)
                vals.append([libc.rand() % 3 + 1 for _ in range(bs * bs)])
        rp.append(len(ci))
    return (np.array(rp, np.int32), np.array(ci, np.int32),
            np.array(vals, np.float64).reshape(-1, bs, bs))


def run_w5r(copies=16384, iters=1000, blocks=8, bs=16, steps=3, warmup=1):
    """The reference's SpGEMM measurement: every one of its 16384 CTAs multiplies the same
    BSR matrix by itself NUM_ITER = 1000 times inside the kernel and the FLOP count is
    2 nnzCub bs^3 x 16384 x 1000 (src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:39,43,
    512-514).  Here the matrix (the reference's default: 8 x 8 blocks of 16 at 50 %
    density, its generator) is laid out `copies` times along the diagonal of one
    block-diagonal BSR matrix, whose product is the copies' products; the numeric phase
    repeats each block product `iters` times on register-held operands
    (kami_bsr_spgemm_numeric_repeat)."""
    torch = _torch()
    dev = torch.device("cuda:0")
    rp1, ci1, v1 = reference_bsr(blocks, bs)
    nnz1 = len(ci1)
    nnzcub1 = int(sum(int(rp1[c + 1] - rp1[c]) for c in ci1))
    mb = blocks * copies
    rp = np.concatenate([rp1[:-1] + c * nnz1 for c in range(copies)] + [[nnz1 * copies]]).astype(np.int32)
    ci = (np.tile(ci1, copies) + np.repeat(np.arange(copies, dtype=np.int32) * blocks, nnz1)).astype(np.int32)
    vals = torch.from_numpy(v1.astype(np.float16)).to(dev).repeat(copies, 1, 1)
    A = K.BsrMatrix(mb, mb, bs, torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev), vals)
    rowptr, nnz = K.spgemm_symbolic(A, A)
    colidx = K.spgemm_fill(A, A, rowptr, nnz)
    out = {}
    ms = time_ms(lambda: out.__setitem__("C", K.spgemm_numeric(A, A, rowptr, colidx, iters=iters)),
                 steps, warmup)
    # the first copy against its product in FP64 (integer values: exact)
    Cv = out["C"][: int(rowptr[blocks].item())].double().cpu().numpy() / iters
    rpc = rowptr[: blocks + 1].cpu().numpy()
    cic = colidx[: int(rpc[-1])].cpu().numpy()
    err = 0.0
    for i in range(blocks):
        for q in range(rpc[i], rpc[i + 1]):
            ref = np.zeros((bs, bs))
            for a in range(rp1[i], rp1[i + 1]):
                k = ci1[a]
                for b in range(rp1[k], rp1[k + 1]):
                    if ci1[b] == cic[q]:
                        ref += v1[a] @ v1[b]
            err = max(err, float(np.max(np.abs(Cv[q] - ref))))
    flops = 2.0 * nnzcub1 * bs ** 3 * copies * iters
    return {"workload": f"reference SpGEMM measurement: {blocks * bs}^2 BSR ({blocks}x{blocks} blocks of "
                        f"{bs}, 50% density, the reference's generator) x {copies} copies x {iters} "
                        "in-kernel repetitions, FP16->FP32",
            "nnzCub_per_matrix": nnzcub1, "nnzb_C": int(nnz), "ms": ms,
            "tflops": flops / ms / 1e9, "max_abs_err_first_copy": err}


# ------------------------------------------------------------------ w6

def run_w6(steps=20, warmup=3, n=256):
    """Block SpMM C = A B^T with the w5 matrix (8192^2, 32x32 blocks, 2 % of blocks, FP16
    N(0,1) values) and a dense B of n rows x 8192 (the reference's N x K layout), FP16
    inputs with FP32 C and FP64 inputs and C."""
    torch = _torch()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(SEED + 4)
    mb, bs = 256, 32
    rp, ci, vals = spgemm_input(rng, mb, bs)
    nnzb, kdim = len(ci), mb * bs
    Bh = rng.standard_normal((n, kdim)).astype(np.float16)
    Bd = Bh.astype(np.float64)
    ref = np.zeros((kdim, n))  # block by block in FP64
    for i in range(mb):
        for q in range(rp[i], rp[i + 1]):
            ref[i * bs:(i + 1) * bs] += vals[q].astype(np.float64) @ Bd[:, ci[q] * bs:(ci[q] + 1) * bs].T
    out = {"workload": f"BSR {kdim}^2 (32x32 blocks, 2% density) x dense {n} x {kdim}",
           "nnzb_A": int(nnzb), "n": n}
    flops = 2.0 * nnzb * bs * bs * n
    for name, prec, dt in [("f16", K.PREC_F16_F32, torch.float16), ("f64", K.PREC_F64, torch.float64)]:
        A = K.BsrMatrix(mb, mb, bs, torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev),
                        torch.from_numpy(vals.astype(np.float64)).to(dev).to(dt))
        B = torch.from_numpy(Bh.astype(np.float64)).to(dev).to(dt)
        C = K.bsr_spmm(A, B, precision=prec)
        # one call is a ~10 µs kernel, about the cost of the Python wrapper's checks, so the
        # calls are replayed from a captured graph: the time is the GPU's
        ms = graph_time_ms(lambda: K.bsr_spmm(A, B, C, precision=prec), steps, warmup)
        err = float(np.max(np.abs(C.double().cpu().numpy() - ref)) / np.max(np.abs(ref)))
        ein, eout = (2, 4) if name == "f16" else (8, 8)
        byts = ein * (nnzb * bs * bs + n * kdim) + eout * kdim * n  # A, B read once, C written
        out[name] = {"ms": ms, "tflops": flops / ms / 1e9, "hbm_gbs": byts / ms / 1e6,
                     "max_rel_err_vs_fp64": err}
    return out

