"""The benchmark workloads at their full sizes, checked through properties that do not
depend on size: integer known answers compared bit for bit with the oracle, closed forms
(all-ones A gives C(i, j) = sum_k B(j, k), exact for small integers in FP32 and FP64),
and native FP64 against split precision within the 1e-13 bound on every problem, with
the compensated Dot2 oracle on a sample.

Workloads: FP64 64^3 x 1024 (2D); FP16 {32, 64, 128, 256}^3 and 256^2 x {16, 32} x 8192
problems, every compiled partition; split FP64 128^3 x 4096 with log-uniform entries;
20000 uneven problems (M, N, K log-uniform in [16, 256], multiples of 16) in FP64 and
FP16 through one graph-captured plan; block SpMM of an 8192^2 BSR matrix (32 x 32 blocks,
2 % density) with a 256 x 8192 B.  (The 8192^2 SpGEMM is in test_gpu_spgemm.py.)
"""
import numpy as np
import pytest

import kami_amd as K
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def test_fp64_64cubed_x1024_2d_known_answer(dev):
    count, s = 1024, 64
    n = count * s * s
    A = oracle.randr_mod3(n).reshape(count, s, s) - 1.0
    B = oracle.randr_mod3_from(n, n).reshape(count, s, s) * 3.0 - 2.0
    C = K.block_gemm(torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev),
                     partition="2d").cpu().numpy()
    np.testing.assert_array_equal(C, oracle.gemm_f64(A, B))


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 64, 64), (128, 128, 128),
                                   (256, 256, 256), (256, 256, 16), (256, 256, 32)],
                         ids=lambda s: "x".join(map(str, s)))
def test_fp16_8192_problems_every_partition_closed_form(shape, dev):
    m, n, k = shape
    count = 8192
    g = torch.Generator(device=dev)
    g.manual_seed(m * 1000 + k)
    A = torch.ones((count, m, k), dtype=torch.float16, device=dev)
    B = torch.randint(0, 3, (count, n, k), generator=g, device=dev).to(torch.float16)
    want = B.float().sum(dim=2)[:, None, :]  # at most 2 k <= 512: exact in FP32
    C = torch.empty((count, m, n), dtype=torch.float32, device=dev)
    ran = 0
    for part in ("1d", "2d", "3d"):
        try:
            K.query(m, n, k, K.PREC_F16_F32, partition=part)
        except K.KamiError:
            continue
        C.fill_(float("nan"))
        K.block_gemm(A, B, C, partition=part)
        assert torch.equal(C, want.expand(count, m, n)), part
        ran += 1
    assert ran >= 1


def log_uniform(rng, shape):
    return np.where(rng.random(shape) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-4, 4, shape)


@pytest.mark.parametrize("prec", [K.PREC_F64_SPLIT_I8, K.PREC_F64_SPLIT_F16, K.PREC_F64_SPLIT_BF16],
                         ids=["i8", "f16", "bf16"])
def test_split_128cubed_x4096_agrees_with_native_on_every_problem(prec, dev):
    """The benchmark's split workload at full size (FP64 128^3 x 4096, log-uniform).
    Int8 and FP16 digits (their default 8 and 7 digits) meet 1e-13 componentwise,
    |C - C*| <= 1e-13 (|A||B|): against native FP64 on every problem and against the Dot2
    oracle on a sample.  BF16 (7 digits of 8 bits) meets 1e-13 normwise only."""
    rng = np.random.default_rng(0x4B414D49 + 2)
    count, s = 4096, 128
    A = log_uniform(rng, (count, s, s))
    B = log_uniform(rng, (count, s, s))
    Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    Cn = K.block_gemm(Ad, Bd)
    Cs = K.block_gemm(Ad, Bd, precision=prec)
    componentwise = prec != K.PREC_F64_SPLIT_BF16
    if componentwise:
        absAB = torch.bmm(Ad.abs(), Bd.abs().transpose(1, 2).contiguous())
        diff = ((Cs - Cn).abs() / absAB).amax(dim=(1, 2))
    else:
        diff = (Cs - Cn).abs().amax(dim=(1, 2)) / Cn.abs().amax(dim=(1, 2))
    print(f"  max difference to native FP64: {float(diff.max()):.3e}", flush=True)
    assert float(diff.max()) <= 1e-13
    idx = np.arange(0, count, count // 16)
    exact, absC = oracle.gemm_dot2(A[idx], B[idx])
    for C in (Cs, Cn):
        got = C[torch.from_numpy(idx).to(dev)].cpu().numpy()
        if componentwise:
            err = np.max(np.abs(got - exact) / absC)
        else:
            err = np.max(np.max(np.abs(got - exact), axis=(1, 2)) / np.max(np.abs(exact), axis=(1, 2)))
        assert float(err) <= 1e-13


def uneven_sizes(rng, count, lo=16, hi=256):
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), size=(count, 3)))
    return np.clip(np.round(x / 16) * 16, lo, hi).astype(np.int64)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float16], ids=["f64", "f16"])
def test_batched_20000_uneven_graph_closed_form(dtype, dev):
    rng = np.random.default_rng(0x4B414D49 + 3)
    sizes = uneven_sizes(rng, 20000)
    out_dt = torch.float64 if dtype == torch.float64 else torch.float32
    na = int((sizes[:, 0] * sizes[:, 2]).sum())
    nb = int((sizes[:, 1] * sizes[:, 2]).sum())
    nc = int((sizes[:, 0] * sizes[:, 1]).sum())
    Abuf = torch.ones(na, dtype=dtype, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    Bbuf = torch.randint(0, 3, (nb,), generator=g, device=dev).to(dtype)
    Cbuf = torch.full((nc,), float("nan"), dtype=out_dt, device=dev)
    probs, offs = [], []
    oa = ob = oc = 0
    for (m, n, k) in sizes.tolist():
        probs.append((Abuf[oa:oa + m * k].view(m, k), Bbuf[ob:ob + n * k].view(n, k),
                      Cbuf[oc:oc + m * n].view(m, n)))
        offs.append((ob, oc))
        oa, ob, oc = oa + m * k, ob + n * k, oc + m * n
    Bh = Bbuf.double().cpu().numpy()
    with K.BatchedPlan(probs, use_graph=True) as plan:
        for replay in range(2):
            plan.execute()
            torch.cuda.synchronize()
            Ch = Cbuf.double().cpu().numpy()
            for (m, n, k), (ob_, oc_) in zip(sizes.tolist(), offs):
                rows = Bh[ob_:ob_ + n * k].reshape(n, k).sum(axis=1)
                got = Ch[oc_:oc_ + m * n].reshape(m, n)
                if not np.array_equal(got, np.broadcast_to(rows, (m, n))):
                    raise AssertionError(f"replay {replay}: problem {m}x{n}x{k} at C offset {oc_}")
            Cbuf.fill_(float("nan"))


@pytest.mark.parametrize("prec,dt", [(K.PREC_F16_F32, torch.float16), (K.PREC_F64, torch.float64)],
                         ids=["f16", "f64"])
def test_spmm_8192_bs32_n256_known_answer(prec, dt, dev):
    rng = np.random.default_rng(0x4B414D49 + 4)
    mb, bs, n = 256, 32, 256
    mask = rng.random((mb, mb)) < 0.02
    rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    ci = np.nonzero(mask)[1].astype(np.int32)
    vals = rng.integers(-2, 3, (len(ci), bs, bs)).astype(np.float64)
    Bnk = rng.integers(-2, 3, (n, mb * bs)).astype(np.float64)
    A = K.BsrMatrix(mb, mb, bs, torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev),
                    torch.from_numpy(vals).to(dt).to(dev))
    C = K.bsr_spmm(A, torch.from_numpy(Bnk).to(dt).to(dev), precision=prec)
    np.testing.assert_array_equal(C.cpu().double().numpy(),
                                  oracle.bsr_spmm(mb, mb, bs, rp, ci, vals, Bnk))
