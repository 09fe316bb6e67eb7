"""GPU parity of split-precision FP64 (FP16 / BF16 digit slices on MFMA).

The oracle is the compensated Dot2 product (twice-working-precision accurate).  The
default FP16 form (seven 9-bit digits) meets 1e-13 componentwise on the log-uniform
workload, |C - C*| <= 1e-13 (|A||B|); every form meets it normwise per problem,
    max_ij |C - C*| / max_ij |C*|  <= 1e-13,
and integer-valued inputs (the reference's ramp and rand_r % 3 known answers) must be
reproduced exactly, since every digit, digit product and level sum is then exact.
"""
import numpy as np
import pytest

import kami_amd as K
import oracle
from kami_amd import _abi

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SPLITS = [(K.PREC_F64_SPLIT_F16, 6), (K.PREC_F64_SPLIT_F16, 7), (K.PREC_F64_SPLIT_BF16, 7)]
SIDS = ["f16-s6", "f16-s7", "bf16-s7"]
# 1: one workgroup per problem, exponents over the full K; 2: persistent, per-chunk exponents
PIPES = [1, 2]


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def run(A, B, prec, slices, dev, C=None, pipe=0, **kw):
    Ad = torch.from_numpy(np.ascontiguousarray(A)).to(dev)
    Bd = torch.from_numpy(np.ascontiguousarray(B)).to(dev)
    Cd = None if C is None else torch.from_numpy(np.ascontiguousarray(C)).to(dev)
    if prec in (K.PREC_F64_SPLIT_F16, K.PREC_F64_SPLIT_BF16):
        kw["pipeline"] = pipe
    out = K.block_gemm(Ad, Bd, Cd, precision=prec, slices=slices, **kw)
    return out.cpu().numpy()


def normwise(C, exact):
    d = np.max(np.abs(C - exact), axis=(1, 2))
    return float(np.max(d / np.maximum(np.max(np.abs(exact), axis=(1, 2)), 1e-300)))


def sizes_for(prec, slices):
    out = []
    for s in (64, 128):
        try:
            K.query(s, s, s, prec, slices=slices)
            out.append(s)
        except K.KamiError:
            pass
    return out


@pytest.mark.parametrize("prec,slices", SPLITS, ids=SIDS)
@pytest.mark.parametrize("pipe", PIPES)
def test_ramp_kat_exact(prec, slices, pipe, dev):
    for s in sizes_for(prec, slices):
        A = np.tile(oracle.ramp(s * s).reshape(1, s, s), (2, 1, 1))
        C = run(A, A, prec, slices, dev, pipe=pipe)
        np.testing.assert_array_equal(C, oracle.gemm_f64(A, A))


@pytest.mark.parametrize("prec,slices", SPLITS, ids=SIDS)
@pytest.mark.parametrize("pipe", PIPES)
def test_randr_kat_exact(prec, slices, pipe, dev):
    for s in sizes_for(prec, slices):
        count = 9
        A = oracle.randr_mod3(count * s * s).reshape(count, s, s) - 1.0
        B = oracle.randr_mod3(2 * count * s * s)[count * s * s:].reshape(count, s, s) * 3.0
        C = run(A, B, prec, slices, dev, pipe=pipe)
        np.testing.assert_array_equal(C, oracle.gemm_f64(A, B))


def log_uniform(rng, shape):
    return np.where(rng.random(shape) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-4, 4, shape)


@pytest.mark.parametrize("prec,slices", SPLITS, ids=SIDS)
@pytest.mark.parametrize("pipe", PIPES)
def test_log_uniform_normwise_1e13(prec, slices, pipe, dev):
    """The split-precision workload: sign-random entries, log10|x| ~ U[-4, 4]."""
    rng = np.random.default_rng(0x4B414D49 + 3)
    for s in sizes_for(prec, slices):
        count = 96
        A = log_uniform(rng, (count, s, s))
        B = log_uniform(rng, (count, s, s))
        C = run(A, B, prec, slices, dev, pipe=pipe)
        exact, absC = oracle.gemm_dot2(A, B)
        err = normwise(C, exact)
        comp = float(np.max(np.abs(C - exact) / absC))
        print(f"  {s}^3 normwise {err:.3e}  componentwise {comp:.3e}", flush=True)
        assert err <= 1e-13
        if (prec, slices) == (K.PREC_F64_SPLIT_F16, 7):
            # the default FP16 digit count meets the bound componentwise, |C - C*| <=
            # 1e-13 (|A||B|); six digits and BF16's seven meet it normwise only
            assert comp <= 1e-13
        Cn = run(A, B, K.PREC_F64, 0, dev)
        assert normwise(Cn, exact) <= 1e-14


@pytest.mark.parametrize("prec,slices", SPLITS, ids=SIDS)
@pytest.mark.parametrize("pipe", PIPES)
def test_adversarial_rows(prec, slices, pipe, dev):
    """Zero rows, a single huge entry among tiny ones, signed zeros, exact powers of two,
    entries 2^-40 below their row maximum, all-equal rows and cancelling pairs."""
    rng = np.random.default_rng(17)
    s = 128 if 128 in sizes_for(prec, slices) else 64
    count = 4
    A = rng.uniform(-1, 1, (count, s, s))
    B = rng.uniform(-1, 1, (count, s, s))
    A[0, 0, :] = 0.0                                  # zero row
    A[0, 1, :] = -0.0                                 # negative zeros
    A[0, 2, :] = 1e-6 * rng.uniform(-1, 1, s)
    A[0, 2, 5] = 1e6                                  # one dominant entry
    A[0, 3, :] = 2.0 ** rng.integers(-30, 30, s)      # powers of two
    A[1, 4, :] = 0.75                                 # all equal
    A[1, 5, :] = 2.0 ** -40 * rng.uniform(-1, 1, s)
    A[1, 5, 0] = 1.0                                  # entries 2^-40 below the row max
    B[2, 7, :] = 1.0
    B[2, 8, :] = -1.0                                 # cancelling columns
    B[3] *= 1e200                                     # large exponents
    A[3] *= 1e-200                                    # small exponents
    C = run(A, B, prec, slices, dev, pipe=pipe)
    exact, _ = oracle.gemm_dot2(A, B)
    assert normwise(C, exact) <= 1e-13
    assert np.all(C[0, 0] == 0.0) and np.all(C[0, 1] == 0.0)


@pytest.mark.parametrize("prec,slices", SPLITS[:1] + SPLITS[2:], ids=[SIDS[0], SIDS[2]])
@pytest.mark.parametrize("pipe", PIPES)
def test_alpha_beta(prec, slices, pipe, dev):
    s = 64
    rng = np.random.default_rng(3)
    A = rng.integers(-4, 5, (3, s, s)).astype(np.float64)
    B = rng.integers(-4, 5, (3, s, s)).astype(np.float64)
    C0 = rng.integers(-9, 10, (3, s, s)).astype(np.float64)
    C = run(A, B, prec, slices, dev, C=C0, pipe=pipe, alpha=-0.5, beta=2.0)
    np.testing.assert_array_equal(C, -0.5 * oracle.gemm_f64(A, B) + 2.0 * C0)


@pytest.mark.parametrize("prec,slices", SPLITS, ids=SIDS)
def test_persistent_walks_many_problems(prec, slices, dev):
    """More problems than resident workgroups, so each persistent workgroup handles
    several problems and prefetches across problem boundaries; integer inputs make
    every digit product exact, so the result must equal the FP64 product."""
    for s in sizes_for(prec, slices):
        count = 1100 if s == 64 else 601
        A = oracle.randr_mod3(count * s * s).reshape(count, s, s) - 1.0
        B = (oracle.randr_mod3(2 * count * s * s)[count * s * s:].reshape(count, s, s) - 1.0) * 5.0
        C = run(A, B, prec, slices, dev, pipe=2)
        np.testing.assert_array_equal(C, oracle.gemm_f64(A, B))


# resident-repeat forms (iters > 1): 32^3 and 64^3, the default slice count of each
# digit type
RESIDENT = [(K.PREC_F64_SPLIT_F16, 6), (K.PREC_F64_SPLIT_F16, 7), (K.PREC_F64_SPLIT_BF16, 7)]
RIDS = ["f16-s6", "f16-s7", "bf16-s7"]


@pytest.mark.parametrize("prec,slices", RESIDENT, ids=RIDS)
@pytest.mark.parametrize("iters", [2, 5])
@pytest.mark.parametrize("s", [32, 64])
def test_resident_repeat_kat_exact(prec, slices, iters, s, dev):
    """The resident form repeats the slice products `iters` times into one accumulator, as
    the reference's NUM_ITER loop does; on integer inputs every digit, level sum and FP64
    addition is exact, so C = iters * A B exactly (ramp and rand_r % 3 inputs)."""
    cfg = K.query(s, s, s, prec, slices=slices, iters=iters)
    assert cfg["k_stage"] == s and cfg["pipeline"] == 1 and cfg["prefetch"] == 0
    count = 3
    ramp = np.tile(oracle.ramp(s * s).reshape(1, s, s), (count, 1, 1))
    randr = oracle.randr_mod3(count * s * s).reshape(count, s, s) - 1.0
    for A, B in [(ramp, ramp), (randr, ramp), (randr, randr * 7.0)]:
        C = run(A, B, prec, slices, dev, iters=iters)
        np.testing.assert_array_equal(C, iters * oracle.gemm_f64(A, B))


@pytest.mark.parametrize("prec,slices", RESIDENT, ids=RIDS)
@pytest.mark.parametrize("s", [32, 64])
def test_resident_repeat_log_uniform_normwise_1e13(prec, slices, s, dev):
    rng = np.random.default_rng(29 + s)
    count, iters = 64, 3
    A = log_uniform(rng, (count, s, s))
    B = log_uniform(rng, (count, s, s))
    C = run(A, B, prec, slices, dev, iters=iters)
    exact, _ = oracle.gemm_dot2(A, B)
    assert normwise(C / iters, exact) <= 1e-13
    # the one-workgroup-per-problem form (iters = 1, 64^3 only) agrees far inside the
    # bound; the resident FP16 form slices with an 8-bit first digit, so the two differ
    # by a few ulps
    if s == 64:
        C1 = run(A, B, prec, slices, dev, pipe=1)
        assert normwise(C / iters, C1) <= 1e-14


@pytest.mark.parametrize("prec,slices", RESIDENT, ids=RIDS)
@pytest.mark.parametrize("s", [32, 64])
def test_resident_repeat_alpha_beta_and_adversarial_rows(prec, slices, s, dev):
    rng = np.random.default_rng(31 + s)
    iters = 2
    A = rng.integers(-4, 5, (2, s, s)).astype(np.float64)
    B = rng.integers(-4, 5, (2, s, s)).astype(np.float64)
    C0 = rng.integers(-9, 10, (2, s, s)).astype(np.float64)
    C = run(A, B, prec, slices, dev, C=C0, iters=iters, alpha=-0.5, beta=2.0)
    np.testing.assert_array_equal(C, -0.5 * iters * oracle.gemm_f64(A, B) + 2.0 * C0)
    A = rng.uniform(-1, 1, (2, s, s))
    B = rng.uniform(-1, 1, (2, s, s))
    A[0, 0, :] = 0.0
    A[0, 2, :] = 1e-6 * rng.uniform(-1, 1, s)
    A[0, 2, 5] = 1e6
    A[1, 5, :] = 2.0 ** -40 * rng.uniform(-1, 1, s)
    A[1, 5, 0] = 1.0
    B[1] *= 1e200
    A[1] *= 1e-200
    C = run(A, B, prec, slices, dev, iters=iters)
    exact, _ = oracle.gemm_dot2(A, B)
    assert normwise(C / iters, exact) <= 1e-13
    assert np.all(C[0, 0] == 0.0)


@pytest.mark.parametrize("prec,slices", RESIDENT, ids=RIDS)
@pytest.mark.parametrize("s,pipe,iters", [(64, 1, 1), (64, 2, 1), (128, 2, 1), (128, 1, 1),
                                          (64, 0, 3), (32, 0, 3)])
def test_padded_leading_dimensions_and_strides(prec, slices, s, pipe, iters, dev):
    """Rows of A and B padded past K, C rows past N and problems spaced past their spans,
    through the C ABI, for the one-workgroup, persistent and resident (iters = 3) forms;
    integer inputs make the result exact and the padding of C must stay untouched."""
    import ctypes
    rng = np.random.default_rng(s + pipe + iters)
    count = 3
    lda, ldb, ldc = s + 6, s + 10, s + 3
    sa, sb, sc = s * lda + 32, s * ldb + 16, s * ldc + 5
    A = rng.integers(-3, 4, size=count * sa).astype(np.float64)
    B = rng.integers(-3, 4, size=count * sb).astype(np.float64)
    C = np.full(count * sc, 7.0)
    Ad, Bd, Cd = (torch.from_numpy(x).to(dev) for x in (A, B, C))
    d = K.make_desc(s, s, s, prec, slices=slices, iters=iters, pipeline=pipe, lda=lda, ldb=ldb,
                    ldc=ldc)
    st = K.lib().kami_block_gemm(ctypes.byref(d), ctypes.c_void_p(Ad.data_ptr()),
                                 ctypes.c_void_p(Bd.data_ptr()), ctypes.c_void_p(Cd.data_ptr()),
                                 count, sa, sb, sc,
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == _abi.OK
    got = Cd.cpu().numpy()
    for p in range(count):
        Ap = A[p * sa: p * sa + s * lda].reshape(s, lda)[:, :s]
        Bp = B[p * sb: p * sb + s * ldb].reshape(s, ldb)[:, :s]
        Cp = got[p * sc: p * sc + s * ldc].reshape(s, ldc)
        np.testing.assert_array_equal(Cp[:, :s], iters * (Ap @ Bp.T))
        np.testing.assert_array_equal(Cp[:, s:], 7.0)
        np.testing.assert_array_equal(got[p * sc + s * ldc: (p + 1) * sc], 7.0)


def test_split_unsupported_requests_are_enotsup():
    for kw in [dict(b_layout="kn"), dict(partition="3d"), dict(iters=2), dict(slices=3)]:
        with pytest.raises(K.KamiError) as e:
            K.query(128, 128, 128, K.PREC_F64_SPLIT_F16, **kw)
        assert e.value.status == _abi.ENOTSUP
    with pytest.raises(K.KamiError):
        K.query(96, 96, 96, K.PREC_F64_SPLIT_F16)


SPLIT_GEMM_SHAPES = [(256, 384, 96), (192, 320, 64), (128, 128, 1024), (64, 448, 32)]


@pytest.mark.parametrize("prec,slices", SPLITS, ids=SIDS)
def test_split_gemm_any_size_kat_exact(prec, slices, dev):
    """kami_split_gemm cuts larger problems into tiles; integer inputs stay exact."""
    for (m, n, k) in SPLIT_GEMM_SHAPES:
        count = 2
        A = oracle.randr_mod3(count * m * k).reshape(count, m, k) - 1.0
        B = (oracle.randr_mod3(2 * count * n * k)[count * n * k:].reshape(count, n, k) - 1.0) * 7.0
        try:
            C = K.split_gemm(torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev),
                             precision=prec, slices=slices)
        except K.KamiError as e:
            assert e.status == _abi.ENOTSUP  # no 64-tile instance for this slice count
            continue
        np.testing.assert_array_equal(C.cpu().numpy(), oracle.gemm_f64(A, B))


@pytest.mark.parametrize("prec,slices", SPLITS, ids=SIDS)
def test_split_gemm_log_uniform_normwise(prec, slices, dev):
    rng = np.random.default_rng(0x4B414D49 + 8)
    for (m, n, k) in SPLIT_GEMM_SHAPES[:3]:
        A = log_uniform(rng, (1, m, k))
        B = log_uniform(rng, (1, n, k))
        C0 = rng.standard_normal((1, m, n))
        try:
            Cd = torch.from_numpy(C0).to(dev)
            K.split_gemm(torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev), Cd,
                         precision=prec, slices=slices, alpha=0.5, beta=-2.0)
        except K.KamiError as e:
            assert e.status == _abi.ENOTSUP
            continue
        exact, _ = oracle.gemm_dot2(A, B)
        ref = 0.5 * exact - 2.0 * C0
        err = normwise(Cd.cpu().numpy(), ref)
        print(f"  {m}x{n}x{k} normwise {err:.3e}", flush=True)
        assert err <= 1e-13
