"""kami_batched_gemm_static: the batched GEMM without a plan and without allocation.

Its work queue and counter live in a fixed device array shared by every call on the
device, so the GPU cases run it twice in a row, the second call on fewer and smaller
problems than the first: stale queue entries or a counter left from the first call would
show as tiles of the first batch written again or tiles of the second one skipped.  Each
call is compared bit for bit with a plan created from the same problems (no autotuning,
eager execution): the same kernel on the same queue.  The host-side checks (size limit,
rejected options) run without a device.
"""
import ctypes

import numpy as np
import pytest

import kami_amd as K
from kami_amd import _abi

MAX_ENTRIES = 32768  # queue entries the static array holds


def problem_array(shapes, ptrs=None):
    arr = (_abi.Problem * len(shapes))()
    for i, (m, n, k) in enumerate(shapes):
        p = arr[i]
        p.m, p.n, p.k = m, n, k
        p.lda = p.ldb = p.ldc = 0
        a, b, c = ptrs[i] if ptrs else (0x10000, 0x20000, 0x30000)
        p.A, p.B, p.C = a, b, c
    return arr


def call_static(arr, count, prec, b_layout=_abi.B_NK, alpha=1.0, beta=0.0, autotune=0,
                use_graph=0, stream=None):
    opts = _abi.BatchedOpts()
    opts.b_layout = b_layout
    opts.autotune = autotune
    opts.use_graph = use_graph
    opts.alpha, opts.beta = alpha, beta
    return K.lib().kami_batched_gemm_static(arr, count, prec, ctypes.byref(opts),
                                            ctypes.c_void_p(stream or 0))


def test_limits_and_options_checked_on_the_host():
    # more problems than queue entries, and more entries (four 128 x 128 tiles per
    # 256 x 256 FP64 problem) than the array holds: refused before any device call
    arr = problem_array([(16, 16, 16)] * (MAX_ENTRIES + 1))
    assert call_static(arr, MAX_ENTRIES + 1, K.PREC_F64) == _abi.ENOSPC
    n = MAX_ENTRIES // 4 + 1
    arr = problem_array([(256, 256, 16)] * n)
    assert call_static(arr, n, K.PREC_F64) == _abi.ENOSPC
    # autotuning and graphs need memory of their own
    arr = problem_array([(64, 64, 64)])
    assert call_static(arr, 1, K.PREC_F64, autotune=1) == _abi.EINVAL
    assert call_static(arr, 1, K.PREC_F64, use_graph=1) == _abi.EINVAL
    # the problem checks of plan creation
    assert call_static(problem_array([(24, 16, 16)]), 1, K.PREC_F64) == _abi.EINVAL
    assert call_static(problem_array([(16, 16, 16)]), 0, K.PREC_F64) == _abi.EINVAL
    assert call_static(problem_array([(16, 16, 16)]), 1, K.PREC_F64_SPLIT_F16) == _abi.ENOTSUP


def _uneven(rng, count, lo=16, hi=256):
    def one():
        return int(np.clip(np.round(np.exp(rng.uniform(np.log(lo), np.log(hi))) / 16) * 16, lo, hi))
    return [(one(), one(), one()) for _ in range(count)]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f16"])
def test_two_calls_match_a_plan(prec, dev):
    torch = pytest.importorskip("torch")
    print(f"\n[case] static batched {prec}", flush=True)
    dtype = torch.float64 if prec == "f64" else torch.float16
    out_dtype = torch.float64 if prec == "f64" else torch.float32
    code = K.PREC_F64 if prec == "f64" else K.PREC_F16_F32
    rng = np.random.default_rng(31 if prec == "f64" else 32)
    stream = torch.cuda.current_stream().cuda_stream

    def batch(shapes):
        probs = []
        for (m, n, k) in shapes:
            A = torch.from_numpy(rng.uniform(-1, 1, (m, k))).to(dev).to(dtype)
            B = torch.from_numpy(rng.uniform(-1, 1, (n, k))).to(dev).to(dtype)
            probs.append((A, B, torch.full((m, n), float("nan"), dtype=out_dtype, device=dev)))
        return probs

    first = batch(_uneven(rng, 300))
    second = batch(_uneven(rng, 40, hi=128))  # fewer and smaller problems
    dc = K.lib().kami_device_check()
    assert dc == _abi.OK, (K.lib().kami_status_string(dc), K.lib().kami_last_hip_error())
    for probs, alpha in ((first, 1.0), (second, -0.5)):
        shapes = [(A.shape[0], B.shape[0], A.shape[1]) for (A, B, _) in probs]
        arr = problem_array(shapes, [(A.data_ptr(), B.data_ptr(), C.data_ptr()) for (A, B, C) in probs])
        st = call_static(arr, len(probs), code, alpha=alpha, stream=stream)
        assert st == _abi.OK, (K.lib().kami_status_string(st), K.lib().kami_last_hip_error())
        torch.cuda.synchronize()
        got = [C.clone() for (_, _, C) in probs]
        for (_, _, C) in probs:
            C.fill_(float("nan"))
        with K.BatchedPlan(probs, use_graph=False, autotune=False, alpha=alpha) as plan:
            plan.execute()
            torch.cuda.synchronize()
        for i, ((_, _, C), g) in enumerate(zip(probs, got)):
            assert torch.equal(g, C), f"problem {i} of {len(probs)} differs from the plan's result"
    # the first batch's outputs stay as the first call left them: compare once more
    # against a fresh plan run on the first batch
    kept = [C.clone() for (_, _, C) in first]
    with K.BatchedPlan(first, use_graph=False, autotune=False) as plan:
        plan.execute()
        torch.cuda.synchronize()
    for (_, _, C), g in zip(first, kept):
        assert torch.equal(g, C)
