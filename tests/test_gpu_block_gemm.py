"""GPU parity of kami_block_gemm against the CPU oracle (calls go through the C ABI).

Integer-valued inputs make every summation order exact, so those cases are bit-exact:
the reference's own known-answer inputs (ramp A[i]=i, B[i]=i; rand_r(i) % 3; A=1 with
B in {0,1,2}).  Random inputs are checked with the rigorous forward bound
|fl(AB) - AB| <= gamma_K |A||B| for FP64 and with the 1e-2 relative tolerance (vs the
FP64 oracle, relative to |A||B|) for FP16/BF16 inputs with FP32 accumulation.
"""
import ctypes

import numpy as np
import pytest

import kami_amd as K
import oracle
from kami_amd import _abi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SQUARE = {K.PREC_F64: [16, 32, 64, 128], K.PREC_F16_F32: [16, 32, 64, 128, 192, 208, 224, 256],
          K.PREC_BF16_F32: [16, 32, 64, 128, 192, 208, 224, 256]}
LOWRANK = [(s, s, k) for s in (32, 64, 128, 192, 256) for k in (16, 32) if s != k]
# K beyond one stage, folded through LDS
KFOLD = {K.PREC_F64: [(64, 64, 512), (128, 128, 512)],
         K.PREC_F16_F32: [(128, 128, 512), (256, 256, 512)],
         K.PREC_BF16_F32: [(128, 128, 512), (256, 256, 512)]}


def instances(prec, shapes):
    out = []
    for (m, n, k) in shapes:
        for part in ("1d", "2d", "3d"):
            for r in (1, 2, 4, 8, 12, 13, 14, 16):
                try:
                    K.query(m, n, k, prec, partition=part, ranks=r)
                except K.KamiError:
                    continue
                out.append((m, n, k, part, r))
    return out


F64_CASES = instances(K.PREC_F64, [(s, s, s) for s in SQUARE[K.PREC_F64]] + KFOLD[K.PREC_F64])
F16_CASES = instances(K.PREC_F16_F32, [(s, s, s) for s in SQUARE[K.PREC_F16_F32]] + LOWRANK +
                      KFOLD[K.PREC_F16_F32])
BF16_CASES = instances(K.PREC_BF16_F32, [(s, s, s) for s in SQUARE[K.PREC_BF16_F32]] + LOWRANK +
                       KFOLD[K.PREC_BF16_F32])


def ids(cases):
    return [f"{m}x{n}x{k}-{p}-P{r}" for (m, n, k, p, r) in cases]


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def ramp_closed_form(m, n, k):
    i = np.arange(m, dtype=np.float64)[:, None]
    j = np.arange(n, dtype=np.float64)[None, :]
    K_ = float(k)
    return i * j * K_**3 + (i + j) * K_**2 * (K_ - 1) / 2 + (K_ - 1) * K_ * (2 * K_ - 1) / 6


def to_dev(x, dev, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(x))
    if dtype is not None:
        t = t.view(dtype)
    return t.to(dev)


# ------------------------------------------------------------------ FP64

@pytest.mark.parametrize("case", F64_CASES, ids=ids(F64_CASES))
def test_f64_ramp_kat_bitexact(case, dev):
    m, n, k, part, r = case
    A = np.tile(oracle.ramp(m * k).reshape(1, m, k), (3, 1, 1))
    B = np.tile(oracle.ramp(n * k).reshape(1, n, k), (3, 1, 1))
    C = K.block_gemm(to_dev(A, dev), to_dev(B, dev), partition=part, ranks=r).cpu().numpy()
    ref = ramp_closed_form(m, n, k)
    for p in range(3):
        np.testing.assert_array_equal(C[p], ref)


@pytest.mark.parametrize("case", F64_CASES, ids=ids(F64_CASES))
def test_f64_randr_batched_bitexact(case, dev):
    m, n, k, part, r = case
    count = 37
    A = oracle.randr_mod3(count * m * k).reshape(count, m, k)
    B = oracle.randr_mod3(count * n * k)[::-1].copy().reshape(count, n, k)
    C = K.block_gemm(to_dev(A, dev), to_dev(B, dev), partition=part, ranks=r).cpu().numpy()
    np.testing.assert_array_equal(C, oracle.gemm_f64(A, B))


def test_f64_uniform_64cubed_1024_problems_2d(dev):
    """The 64^3 x 1024 FP64 workload with the 2D partition, U[-1,1] inputs."""
    rng = np.random.default_rng(0x4B414D49)
    A = rng.uniform(-1, 1, size=(1024, 64, 64))
    B = rng.uniform(-1, 1, size=(1024, 64, 64))
    Ad, Bd = to_dev(A, dev), to_dev(B, dev)
    C1 = K.block_gemm(Ad, Bd, partition="2d").cpu().numpy()
    C2 = K.block_gemm(Ad, Bd, partition="2d").cpu().numpy()
    np.testing.assert_array_equal(C1, C2)  # deterministic
    exact, absC = oracle.gemm_dot2(A, B)
    u = 2.0**-53
    gamma = 64 * u / (1 - 64 * u)
    assert np.all(np.abs(C1 - exact) <= gamma * absC)
    ref = oracle.gemm_f64(A, B)
    rel = np.max(np.abs(C1 - ref)) / np.max(np.abs(ref))
    assert rel < 1e-14


def test_partitions_bitwise_identical_on_integers(dev):
    count, s = 16, 64
    A = oracle.randr_mod3(count * s * s).reshape(count, s, s) - 1.0
    B = oracle.randr_mod3(2 * count * s * s)[count * s * s:].reshape(count, s, s) - 1.0
    Ad, Bd = to_dev(A, dev), to_dev(B, dev)
    outs = [K.block_gemm(Ad, Bd, partition=p).cpu().numpy() for p in ("1d", "2d", "3d")]
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], outs[2])


@pytest.mark.parametrize("prec,s", [(K.PREC_F64, 64), (K.PREC_F64, 128), (K.PREC_F16_F32, 128)])
@pytest.mark.parametrize("part", ["1d", "2d", "3d"])
def test_b_layout_kn_matches_nk(prec, s, part, dev):
    count = 5
    rng = np.random.default_rng(s)
    A = rng.integers(-3, 4, size=(count, s, s)).astype(np.float64)
    Bnk = rng.integers(-3, 4, size=(count, s, s)).astype(np.float64)
    Bkn = np.ascontiguousarray(Bnk.transpose(0, 2, 1))
    if prec == K.PREC_F16_F32:
        A, Bnk, Bkn = (x.astype(np.float16) for x in (A, Bnk, Bkn))
    C1 = K.block_gemm(to_dev(A, dev), to_dev(Bnk, dev), partition=part).cpu().numpy()
    C2 = K.block_gemm(to_dev(A, dev), to_dev(Bkn, dev), partition=part, b_layout="kn").cpu().numpy()
    np.testing.assert_array_equal(C1, C2)
    ref = np.einsum("pik,pjk->pij", A.astype(np.float64), Bnk.astype(np.float64))
    np.testing.assert_array_equal(C1, ref)


@pytest.mark.parametrize("part", ["1d", "2d", "3d"])
def test_alpha_beta_iters(part, dev):
    count, s = 4, 64
    rng = np.random.default_rng(1)
    A = rng.integers(-2, 3, size=(count, s, s)).astype(np.float64)
    B = rng.integers(-2, 3, size=(count, s, s)).astype(np.float64)
    C0 = rng.integers(-9, 10, size=(count, s, s)).astype(np.float64)
    Cd = to_dev(C0, dev)
    K.block_gemm(to_dev(A, dev), to_dev(B, dev), Cd, partition=part, alpha=2.0, beta=-1.0, iters=3)
    ref = 6.0 * np.einsum("pik,pjk->pij", A, B) - C0
    np.testing.assert_array_equal(Cd.cpu().numpy(), ref)


def test_padded_leading_dimensions_and_strides(dev):
    m = n = k = 64
    count = 3
    lda, ldb, ldc = k + 6, k + 10, n + 3
    rng = np.random.default_rng(2)
    A = rng.integers(-3, 4, size=(count, m, lda)).astype(np.float64)
    B = rng.integers(-3, 4, size=(count, n, ldb)).astype(np.float64)
    C = np.full((count, m, ldc), 7.0)
    Ad, Bd, Cd = to_dev(A, dev), to_dev(B, dev), to_dev(C, dev)
    d = K.make_desc(m, n, k, K.PREC_F64, partition="2d", lda=lda, ldb=ldb, ldc=ldc)
    st = K.lib().kami_block_gemm(ctypes.byref(d), ctypes.c_void_p(Ad.data_ptr()),
                                 ctypes.c_void_p(Bd.data_ptr()), ctypes.c_void_p(Cd.data_ptr()),
                                 count, m * lda, n * ldb, m * ldc,
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == _abi.OK
    got = Cd.cpu().numpy()
    ref = np.einsum("pik,pjk->pij", A[:, :, :k], B[:, :, :k])
    np.testing.assert_array_equal(got[:, :, :n], ref)
    np.testing.assert_array_equal(got[:, :, n:], 7.0)  # padding untouched


# ------------------------------------------------------------------ FP16 / BF16

def _u16(x, bf16):
    if bf16:
        return torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    return x.astype(np.float16).view(np.uint16)


@pytest.mark.parametrize("bf16", [False, True], ids=["f16", "bf16"])
def test_half_integer_kat_exact_all_instances(bf16, dev):
    """A = 1 and B = rand() % 3 (the reference's FP16 self-check inputs,
    src/block_gemm/KAMI/1d/fp16/block_gemm_1d_half_mma.cu:208-217): sums <= 2K are exact
    in FP32, so every instance must reproduce the oracle bit for bit."""
    cases = BF16_CASES if bf16 else F16_CASES
    prec = K.PREC_BF16_F32 if bf16 else K.PREC_F16_F32
    tdt = torch.bfloat16 if bf16 else torch.float16
    for (m, n, k, part, r) in cases:
        print(f"  instance {m}x{n}x{k} {part} P={r}", flush=True)
        count = 6
        rng = np.random.default_rng(m + k)
        A = np.ones((count, m, k))
        A[1] = rng.integers(0, 2, size=(m, k))
        B = rng.integers(0, 3, size=(count, n, k)).astype(np.float64)
        Au, Bu = _u16(A, bf16), _u16(B, bf16)
        C = K.block_gemm(torch.from_numpy(Au.view(np.int16)).to(dev).view(tdt),
                         torch.from_numpy(Bu.view(np.int16)).to(dev).view(tdt),
                         precision=prec, partition=part, ranks=r).cpu().numpy()
        np.testing.assert_array_equal(C, oracle.gemm_u16(Au, Bu, is_bf16=bf16))


@pytest.mark.parametrize("bf16", [False, True], ids=["f16", "bf16"])
@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 64, 64), (128, 128, 128), (256, 256, 256),
                                   (256, 256, 16), (256, 256, 32)])
@pytest.mark.parametrize("part", ["1d", "2d", "3d"])
def test_half_normal_inputs_tolerance(bf16, shape, part, dev):
    m, n, k = shape
    prec = K.PREC_BF16_F32 if bf16 else K.PREC_F16_F32
    try:
        K.query(m, n, k, prec, partition=part)
    except K.KamiError as e:
        assert e.status == _abi.ENOTSUP
        pytest.skip(f"no {part} instance for {shape} (ENOTSUP by design)")
    tdt = torch.bfloat16 if bf16 else torch.float16
    count = 32
    rng = np.random.default_rng(m * 7 + k)
    A = rng.standard_normal((count, m, k))
    B = rng.standard_normal((count, n, k))
    Au, Bu = _u16(A, bf16), _u16(B, bf16)
    C = K.block_gemm(torch.from_numpy(Au.view(np.int16)).to(dev).view(tdt),
                     torch.from_numpy(Bu.view(np.int16)).to(dev).view(tdt),
                     precision=prec, partition=part).cpu().numpy().astype(np.float64)
    ref = oracle.gemm_u16(Au, Bu, is_bf16=bf16)
    Ab = np.vectorize(oracle.bf16_to_double if bf16 else oracle.half_to_double)
    mag = np.einsum("pik,pjk->pij", np.abs(Ab(Au[:2])), np.abs(Ab(Bu[:2])))
    err = np.abs(C - ref)
    assert np.all(err[:2] <= 1e-2 * mag)          # the stated tolerance, relative to |A||B|
    assert np.max(err) / np.max(np.abs(ref)) < 1e-5  # FP32 accumulation in practice


# every persistent 2D shape with r x r waves whose wave tile (s / r rows) holds a
# 16-row atom (32^3 with 4 x 4 waves cannot exist)
PIPE_CASES = [(prec, s, r) for prec, sizes in ((K.PREC_F64, (32, 64, 128)), (K.PREC_F16_F32, (64, 128)),
                                               (K.PREC_BF16_F32, (64, 128)))
              for s in sizes for r in (2, 4) if s // r >= 16]


@pytest.mark.parametrize("prec,s,r", PIPE_CASES)
def test_persistent_pipeline_matches_oracle(prec, s, r, dev):
    """The persistent 2D form walks several problems per workgroup and prefetches across
    problem boundaries: counts below, equal to and far above the grid size."""
    try:
        K.query(s, s, s, prec, partition="2d", ranks=r, pipeline=2)
    except K.KamiError:
        pytest.skip("no persistent instance")
    rng = np.random.default_rng(s + r)
    for count in (1, 7, 2500):
        A = rng.integers(-3, 4, size=(count, s, s)).astype(np.float64)
        B = rng.integers(-3, 4, size=(count, s, s)).astype(np.float64)
        C0 = rng.integers(-3, 4, size=(count, s, s)).astype(np.float64)
        if prec == K.PREC_F64:
            Ad, Bd = to_dev(A, dev), to_dev(B, dev)
            Cd = to_dev(C0, dev)
        else:
            tdt = torch.float16 if prec == K.PREC_F16_F32 else torch.bfloat16
            Ad, Bd = to_dev(A, dev).to(tdt), to_dev(B, dev).to(tdt)
            Cd = to_dev(C0, dev).float()
        K.block_gemm(Ad, Bd, Cd, precision=prec, partition="2d", ranks=r, pipeline=2,
                     alpha=2.0, beta=-1.0)
        ref = 2.0 * np.einsum("pik,pjk->pij", A, B) - C0
        np.testing.assert_array_equal(Cd.double().cpu().numpy(), ref)


@pytest.mark.parametrize("prec,s,part,ranks", [
    (K.PREC_F16_F32, 256, "1d", 8), (K.PREC_F16_F32, 256, "1d", 16), (K.PREC_BF16_F32, 256, "1d", 8),
    (K.PREC_F64, 128, "1d", 8), (K.PREC_F16_F32, 192, "2d", 4), (K.PREC_BF16_F32, 192, "2d", 4),
    (K.PREC_F16_F32, 192, "1d", 12)])
@pytest.mark.parametrize("layout", ["nk", "kn"])
def test_resident_forms_repeat_exactly(prec, s, part, ranks, layout, dev):
    """iters > 1 dispatches the resident form of a key (all of K in one LDS stage; in 1D
    A in registers); on integer inputs the result is exactly iters * alpha * A B^T + beta C."""
    cfg = K.query(s, s, s, prec, partition=part, ranks=ranks, iters=3, b_layout=layout)
    assert cfg["k_stage"] == s and cfg["waves_m"] == ranks
    if (s, ranks) != (192, 12):  # 192^3 1D with 12 waves holds all of K in either form
        assert K.query(s, s, s, prec, partition=part, ranks=ranks)["k_stage"] < s  # streaming
    count = 6
    rng = np.random.default_rng(s + ranks)
    A = rng.integers(-2, 3, size=(count, s, s)).astype(np.float64)
    Bnk = rng.integers(-2, 3, size=(count, s, s)).astype(np.float64)
    C0 = rng.integers(-9, 10, size=(count, s, s)).astype(np.float64)
    dt = {K.PREC_F16_F32: torch.float16, K.PREC_BF16_F32: torch.bfloat16, K.PREC_F64: torch.float64}[prec]
    out_dt = torch.float64 if prec == K.PREC_F64 else torch.float32
    Bin = Bnk if layout == "nk" else np.ascontiguousarray(Bnk.transpose(0, 2, 1))
    Cd = torch.from_numpy(C0).to(dev).to(out_dt)
    K.block_gemm(torch.from_numpy(A).to(dev).to(dt), torch.from_numpy(Bin).to(dev).to(dt), Cd,
                 precision=prec, partition=part, ranks=ranks, b_layout=layout, alpha=2.0,
                 beta=-1.0, iters=3)
    ref = 6.0 * np.einsum("pik,pjk->pij", A, Bnk) - C0
    np.testing.assert_array_equal(Cd.double().cpu().numpy(), ref)
