"""GPU parity of split-precision FP64 through int8 digits (KAMI_PREC_F64_SPLIT_I8).

Each operand row is scaled by its exponent and truncated to a 62-bit fixed-point
integer, cut into eight balanced byte digits; the digit products run on int8 MFMA with
exact int32 level sums and are combined in FP64.  Checks:
  * integer-valued inputs (the reference's ramp and rand_r % 3 known answers) are
    reproduced exactly: their fixed-point values are exact and every level sum is exact;
  * on the split-precision workload (sign-random, log10|x| ~ U[-4, 4]) the error against
    the compensated Dot2 product is at most 1e-13 componentwise, |C - C*| <= 1e-13 (|A||B|),
    which bounds the normwise error too;
  * resident repetition (iters > 1), alpha / beta, padded leading dimensions and problem
    strides, adversarial rows, and more problems than resident workgroups.
"""
import ctypes

import numpy as np
import pytest

import kami_amd as K
import oracle
from kami_amd import _abi

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

I8 = K.PREC_F64_SPLIT_I8
SIZES = [64, 128]


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def run(A, B, dev, C=None, **kw):
    Ad = torch.from_numpy(np.ascontiguousarray(A)).to(dev)
    Bd = torch.from_numpy(np.ascontiguousarray(B)).to(dev)
    Cd = None if C is None else torch.from_numpy(np.ascontiguousarray(C)).to(dev)
    return K.block_gemm(Ad, Bd, Cd, precision=I8, **kw).cpu().numpy()


def errors(C, exact, absC):
    d = np.abs(C - exact)
    comp = float(np.max(d / np.where(absC > 0, absC, 1.0)))
    norm = float(np.max(np.max(d, axis=(1, 2)) /
                        np.maximum(np.max(np.abs(exact), axis=(1, 2)), 1e-300)))
    return comp, norm


def log_uniform(rng, shape):
    return np.where(rng.random(shape) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-4, 4, shape)


@pytest.mark.parametrize("s", SIZES)
def test_ramp_and_randr_kat_exact(s, dev):
    count = 5
    ramp = np.tile(oracle.ramp(s * s).reshape(1, s, s), (count, 1, 1))
    randr = oracle.randr_mod3(count * s * s).reshape(count, s, s) - 1.0
    randr3 = oracle.randr_mod3(2 * count * s * s)[count * s * s:].reshape(count, s, s) * 3.0
    for A, B in [(ramp, ramp), (randr, randr3), (randr3, ramp)]:
        np.testing.assert_array_equal(run(A, B, dev), oracle.gemm_f64(A, B))


@pytest.mark.parametrize("s", SIZES)
def test_log_uniform_componentwise_1e13(s, dev):
    """The split-precision workload of the benchmark, at both block sizes."""
    rng = np.random.default_rng(0x4B414D49 + 5)
    count = 64
    A = log_uniform(rng, (count, s, s))
    B = log_uniform(rng, (count, s, s))
    C = run(A, B, dev)
    exact, absC = oracle.gemm_dot2(A, B)
    comp, norm = errors(C, exact, absC)
    print(f"  {s}^3 componentwise {comp:.3e} normwise {norm:.3e}", flush=True)
    assert comp <= 1e-13
    assert norm <= 1e-15


@pytest.mark.parametrize("s", SIZES)
def test_uniform_inputs_componentwise(s, dev):
    rng = np.random.default_rng(11 + s)
    A = rng.uniform(-1, 1, (32, s, s))
    B = rng.uniform(-1, 1, (32, s, s))
    exact, absC = oracle.gemm_dot2(A, B)
    comp, _ = errors(run(A, B, dev), exact, absC)
    assert comp <= 1e-13


@pytest.mark.parametrize("s", SIZES)
def test_adversarial_rows(s, dev):
    """Zero and negative-zero rows, one dominant entry among tiny ones, exact powers of
    two, entries 2^-40 below their row maximum, all-equal rows, cancelling columns,
    subnormal entries in a normal row and in a tiny row, and exponents near +-660."""
    rng = np.random.default_rng(19 + s)
    count = 4
    A = rng.uniform(-1, 1, (count, s, s))
    B = rng.uniform(-1, 1, (count, s, s))
    A[0, 0, :] = 0.0
    A[0, 1, :] = -0.0
    A[0, 2, :] = 1e-6 * rng.uniform(-1, 1, s)
    A[0, 2, 5] = 1e6
    A[0, 3, :] = 2.0 ** rng.integers(-30, 30, s)
    A[1, 4, :] = 0.75
    A[1, 5, :] = 2.0 ** -40 * rng.uniform(-1, 1, s)
    A[1, 5, 0] = 1.0
    A[1, 6, :] = 5e-324 * rng.integers(1, 1000, s)       # a row of subnormals
    A[1, 7, 3] = 1e-310                                  # one subnormal in a normal row
    B[2, 7, :] = 1.0
    B[2, 8, :] = -1.0
    B[3] *= 1e200
    A[3] *= 1e-200
    C = run(A, B, dev)
    exact, absC = oracle.gemm_dot2(A, B)
    _, norm = errors(C, exact, absC)
    assert norm <= 1e-13
    assert np.all(C[0, 0] == 0.0) and np.all(C[0, 1] == 0.0)
    # a row of subnormals times O(1) values: its digits are exact (10-bit subnormals), so
    # the result is the subnormal sum rounded once, within two units of 2^-1074 of the
    # same sum formed in the normal range and scaled back
    ref = np.ldexp(np.ldexp(A[1, 6], 1000) @ B[1].T, -1000)
    assert np.max(np.abs(C[1, 6] - ref)) <= 2.0 ** -1073


@pytest.mark.parametrize("s", SIZES)
@pytest.mark.parametrize("iters", [2, 5])
def test_resident_repeat_kat_exact(s, iters, dev):
    """iters > 1 repeats the digit products on the sliced operands, as the reference's
    NUM_ITER loop does: C = iters * A B exactly on integer inputs."""
    cfg = K.query(s, s, s, I8, iters=iters)
    assert cfg["slices"] == 8 and cfg["partition"] == K.PART_1D
    count = 3
    ramp = np.tile(oracle.ramp(s * s).reshape(1, s, s), (count, 1, 1))
    randr = oracle.randr_mod3(count * s * s).reshape(count, s, s) - 1.0
    for A, B in [(ramp, ramp), (randr, randr * 7.0)]:
        np.testing.assert_array_equal(run(A, B, dev, iters=iters), iters * oracle.gemm_f64(A, B))


@pytest.mark.parametrize("s", SIZES)
def test_resident_repeat_past_one_int32_run(s, dev):
    """The repetitions sum the level sums in int32 for at most 2^31 / (8 K 2^14) of them
    (255 at K = 64, 127 at K = 128), then combine in FP64 and start again: iters past
    that bound still give iters * A B exactly on integer inputs, and on random inputs
    (digits over the whole byte range, the largest level sums) the same as one product
    times iters to within the final rounding."""
    iters = 300
    count = 3
    rng = np.random.default_rng(41 + s)
    ramp = np.tile(oracle.ramp(s * s).reshape(1, s, s), (count, 1, 1))
    randr = oracle.randr_mod3(count * s * s).reshape(count, s, s) - 1.0
    np.testing.assert_array_equal(run(ramp, randr, dev, iters=iters), iters * oracle.gemm_f64(ramp, randr))
    A = rng.uniform(-1, 1, (count, s, s))
    B = rng.uniform(-1, 1, (count, s, s))
    C1 = run(A, B, dev)
    Cn = run(A, B, dev, iters=iters)
    exact, absC = oracle.gemm_dot2(A, B)
    comp, _ = errors(Cn / iters, exact, absC)
    assert comp <= 1e-13
    assert np.max(np.abs(Cn / iters - C1) / np.where(absC > 0, absC, 1.0)) <= 1e-15


@pytest.mark.parametrize("s", SIZES)
def test_resident_repeat_log_uniform(s, dev):
    rng = np.random.default_rng(37 + s)
    A = log_uniform(rng, (16, s, s))
    B = log_uniform(rng, (16, s, s))
    C1 = run(A, B, dev)
    C3 = run(A, B, dev, iters=3)
    exact, absC = oracle.gemm_dot2(A, B)
    comp, _ = errors(C3 / 3.0, exact, absC)
    assert comp <= 1e-13
    assert np.max(np.abs(C3 / 3.0 - C1) / np.where(absC > 0, absC, 1.0)) <= 1e-15


@pytest.mark.parametrize("s", SIZES)
def test_alpha_beta(s, dev):
    rng = np.random.default_rng(3 + s)
    A = rng.integers(-4, 5, (3, s, s)).astype(np.float64)
    B = rng.integers(-4, 5, (3, s, s)).astype(np.float64)
    C0 = rng.integers(-9, 10, (3, s, s)).astype(np.float64)
    C = run(A, B, dev, C=C0, alpha=-0.5, beta=2.0)
    np.testing.assert_array_equal(C, -0.5 * oracle.gemm_f64(A, B) + 2.0 * C0)
    C = run(A, B, dev, C=C0, alpha=2.0, beta=-1.0, iters=2)
    np.testing.assert_array_equal(C, 4.0 * oracle.gemm_f64(A, B) - C0)


@pytest.mark.parametrize("s", SIZES)
def test_persistent_walks_many_problems(s, dev):
    """Far more problems than resident workgroups: every workgroup slices, multiplies and
    prefetches across many problem boundaries (B rows through LDS, A rows through
    registers); integer inputs make the result exact."""
    count = 2300 if s == 64 else 1100
    A = oracle.randr_mod3(count * s * s).reshape(count, s, s) - 1.0
    B = (oracle.randr_mod3(2 * count * s * s)[count * s * s:].reshape(count, s, s) - 1.0) * 5.0
    np.testing.assert_array_equal(run(A, B, dev), oracle.gemm_f64(A, B))


@pytest.mark.parametrize("s,iters", [(64, 1), (128, 1), (64, 3), (128, 2)])
def test_padded_leading_dimensions_and_strides(s, iters, dev):
    """Padded rows of A, B and C and problems spaced past their spans through the C ABI;
    the padding of C must stay untouched."""
    rng = np.random.default_rng(s + iters)
    count = 3
    lda, ldb, ldc = s + 6, s + 10, s + 3
    sa, sb, sc = s * lda + 32, s * ldb + 16, s * ldc + 5
    A = rng.integers(-3, 4, size=count * sa).astype(np.float64)
    B = rng.integers(-3, 4, size=count * sb).astype(np.float64)
    C = np.full(count * sc, 7.0)
    Ad, Bd, Cd = (torch.from_numpy(x).to(dev) for x in (A, B, C))
    d = K.make_desc(s, s, s, I8, iters=iters, lda=lda, ldb=ldb, ldc=ldc)
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


def test_agrees_with_native_fp64(dev):
    """Against native FP64 MFMA on the benchmark's distribution (1024 problems of 128^3):
    both are within 1e-13 of each other relative to |A||B| on every entry."""
    rng = np.random.default_rng(0x4B414D49 + 6)
    count, s = 1024, 128
    A = torch.from_numpy(log_uniform(rng, (count, s, s))).to(dev)
    B = torch.from_numpy(log_uniform(rng, (count, s, s))).to(dev)
    Ci = K.block_gemm(A, B, precision=I8)
    Cn = K.block_gemm(A, B, precision=K.PREC_F64)
    absC = torch.bmm(A.abs(), B.abs().transpose(1, 2).contiguous())
    rel = ((Ci - Cn).abs() / absC).max().item()
    print(f"  max |C_i8 - C_f64| / (|A||B|) = {rel:.3e}", flush=True)
    assert rel <= 1e-13
