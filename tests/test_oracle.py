"""Pins the CPU oracle against facts that do not depend on it (CPU only).

Each oracle function is checked against at least one of: a closed form, a brute-force
computation by another method (numpy integer/float arithmetic, exact rationals), an
independent library conversion (numpy float16, torch bfloat16), the reference's own
known-answer inputs, or a published constant.  A plausible slip in the oracle (a
transposed operand, a dropped term, an off-by-one in the scan) fails one of these.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ----------------------------------------------------------------- dense GEMM

def ramp_closed_form(m, n, k):
    """C_ij = sum_k (iK + k)(jK + k) for A[i] = i, B[i] = i with B stored N x K."""
    i = np.arange(m, dtype=np.float64)[:, None]
    j = np.arange(n, dtype=np.float64)[None, :]
    K = float(k)
    return i * j * K**3 + (i + j) * K**2 * (K - 1) / 2 + (K - 1) * K * (2 * K - 1) / 6


@pytest.mark.parametrize("s", [16, 32, 64, 128])
def test_ramp_kat_closed_form(s):
    A = oracle.ramp(s * s).reshape(1, s, s)
    B = oracle.ramp(s * s).reshape(1, s, s)
    C = oracle.gemm_f64(A, B)
    np.testing.assert_array_equal(C[0], ramp_closed_form(s, s, s))


def test_gemm_f64_matches_integer_matmul_both_layouts():
    rng = np.random.default_rng(7)
    A = rng.integers(-50, 50, size=(3, 24, 40)).astype(np.float64)
    Bnk = rng.integers(-50, 50, size=(3, 16, 40)).astype(np.float64)
    ref = np.einsum("pik,pjk->pij", A.astype(np.int64), Bnk.astype(np.int64)).astype(np.float64)
    np.testing.assert_array_equal(oracle.gemm_f64(A, Bnk), ref)
    Bkn = np.ascontiguousarray(Bnk.transpose(0, 2, 1))
    np.testing.assert_array_equal(oracle.gemm_f64(A, Bkn, b_kn=True), ref)


def test_randr_mod3_kat_exact():
    """The reference's batched inputs (rand_r(i) % 3) are integers in {0,1,2}; products
    are exact in any summation order, so the oracle must equal integer matmul."""
    m = n = k = 32
    count = 4
    A = oracle.randr_mod3(count * m * k).reshape(count, m, k)
    B = oracle.randr_mod3(count * n * k).reshape(count, n, k)
    assert set(np.unique(A)) <= {0.0, 1.0, 2.0} and len(np.unique(A)) == 3
    ref = np.einsum("pik,pjk->pij", A.astype(np.int64), B.astype(np.int64))
    np.testing.assert_array_equal(oracle.gemm_f64(A, B), ref.astype(np.float64))


def test_randr_mod3_values_from_libc():
    """rand_r with seed i is the C library's; check the first values against ctypes libc."""
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    x = oracle.randr_mod3(64)
    for i in range(64):
        s = ctypes.c_uint(i)
        assert x[i] == libc.rand_r(ctypes.byref(s)) % 3


def test_gemm_f64_random_vs_numpy():
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 1, size=(2, 64, 64))
    B = rng.uniform(-1, 1, size=(2, 64, 64))
    C = oracle.gemm_f64(A, B)
    ref = np.einsum("pik,pjk->pij", A, B)
    bound = 64 * np.finfo(np.float64).eps * np.einsum("pik,pjk->pij", abs(A), abs(B))
    assert np.all(abs(C - ref) <= bound)


def test_dot2_is_exact_to_double_rounding():
    """Dot2 must agree with exact rational arithmetic to (about) one rounding."""
    rng = np.random.default_rng(11)
    k = 48
    # ill-conditioned dot products: large cancelling terms plus a small remainder
    a = rng.uniform(-1, 1, size=(1, 4, k)) * 10.0 ** rng.integers(-8, 8, size=(1, 4, k))
    b = rng.uniform(-1, 1, size=(1, 3, k)) * 10.0 ** rng.integers(-8, 8, size=(1, 3, k))
    C, absC = oracle.gemm_dot2(a, b)
    for i in range(4):
        for j in range(3):
            exact = sum(Fraction(float(a[0, i, t])) * Fraction(float(b[0, j, t])) for t in range(k))
            err = abs(Fraction(float(C[0, i, j])) - exact)
            # |fl(dot2) - exact| <= u|exact| + gamma_{n}^2 |a||b| (Ogita-Rump-Oishi Prop. 5.5)
            u = 2.0 ** -53
            gam = k * u / (1 - k * u)
            assert err <= u * abs(exact) + 2 * gam * gam * float(absC[0, i, j])
            assert absC[0, i, j] == pytest.approx(np.sum(abs(a[0, i] * b[0, j])), rel=1e-15)


def test_slab_sum_matches_sum():
    rng = np.random.default_rng(5)
    S = rng.integers(-9, 9, size=(3, 8, 8)).astype(np.float64)
    np.testing.assert_array_equal(oracle.slab_sum(S), S.sum(axis=0))


# ----------------------------------------------------------------- formats

def test_half_to_double_all_patterns():
    bits = np.arange(65536, dtype=np.uint16)
    ref = bits.view(np.float16).astype(np.float64)
    got = np.array([oracle.half_to_double(b) for b in range(0, 65536, 7)])
    r = ref[::7]
    finite = np.isfinite(r)
    np.testing.assert_array_equal(got[finite], r[finite])
    assert np.all(np.isnan(got[np.isnan(r)]))


def test_double_to_half_matches_numpy():
    rng = np.random.default_rng(1)
    xs = np.concatenate([
        rng.standard_normal(3000),
        rng.uniform(-70000, 70000, 500),
        rng.uniform(-1e-6, 1e-6, 500),           # subnormal range
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 2.0**-24, 2.0**-25, 3 * 2.0**-26]),
        (np.arange(2048) + 0.5) * 2.0**-10,       # exact ties: round to even
    ])
    got = np.array([oracle.double_to_half(x) for x in xs], dtype=np.uint16)
    ref = xs.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(got, ref)


def test_bf16_conversions_match_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.standard_normal(2000), rng.uniform(-1e30, 1e30, 200),
                         np.array([0.0, 1.0, -3.5, 1 + 2.0**-8, 1 + 3 * 2.0**-9])])
    ref = torch.from_numpy(xs.astype(np.float32)).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = np.array([oracle.double_to_bf16(float(np.float32(x))) for x in xs], dtype=np.uint16)
    np.testing.assert_array_equal(got, ref)
    back = torch.from_numpy(ref.view(np.int16)).view(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(np.array([oracle.bf16_to_double(b) for b in ref]), back)


def test_gemm_u16_matches_widened_numpy():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((2, 32, 32)).astype(np.float16)
    B = rng.standard_normal((2, 16, 32)).astype(np.float16)
    C = oracle.gemm_u16(A.view(np.uint16), B.view(np.uint16))
    ref = np.einsum("pik,pjk->pij", A.astype(np.float64), B.astype(np.float64))
    np.testing.assert_allclose(C, ref, rtol=0, atol=1e-12)


# ----------------------------------------------------------------- sparse

def test_glibc_rand_known_answer_drives_bsr_generator():
    """glibc rand() after srand(0) (== srand(1)) is the published sequence
    1804289383, 846930886, 1681692777.  The first draw decides block (0, 0), the next two
    are its first two values (rand() % 3 + 1)."""
    rp, ci, vals = oracle.bsr_generate(8, 8, 0.5, 16)
    assert 1804289383 / 2147483647 > 0.5
    assert ci[0] == 0
    assert oracle.half_to_double(vals[0, 0, 0]) == 846930886 % 3 + 1
    assert oracle.half_to_double(vals[0, 0, 1]) == 1681692777 % 3 + 1
    assert rp[0] == 0 and np.all(np.diff(rp) >= 0) and rp[-1] == len(ci)
    for i in range(8):
        assert np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0)
    v = np.vectorize(oracle.half_to_double)(vals)
    assert set(np.unique(v)) <= {1.0, 2.0, 3.0}


def test_spmm_generator_follows_glibc_sequence():
    """The SpMM program's generator: block (0, 0) is kept by the first draw, its values are
    rand() % 2 + 1 of the next draws (846930886, 1681692777, 1714636915, 1957747793), and B
    continues the same sequence."""
    rp, ci, vals, B = oracle.spmm_generate(4, 4, 0.5, 16, 32)
    assert ci[0] == 0
    assert list(vals[0, 0, :4]) == [846930886 % 2 + 1, 1681692777 % 2 + 1,
                                    1714636915 % 2 + 1, 1957747793 % 2 + 1]
    assert set(np.unique(vals)) <= {1.0, 2.0} and set(np.unique(B)) <= {1.0, 2.0}
    assert B.shape == (32, 64)
    rp2, ci2, _ = oracle.bsr_generate(4, 4, 0.5, 16)
    np.testing.assert_array_equal(rp, rp2)  # same keep decisions as the SpGEMM generator
    np.testing.assert_array_equal(ci, ci2)


@pytest.mark.parametrize("b_kn", [False, True])
def test_bsr_spmm_oracle_vs_dense_product(b_kn):
    """The restated check loop equals the dense product of the expanded matrix, in both B
    layouts, on integer values (exact in any order) and within rounding on random ones."""
    rng = np.random.default_rng(21)
    for mb, nb, bs, n, dens in [(3, 5, 16, 48, 0.4), (6, 4, 32, 16, 0.7), (1, 1, 16, 16, 1.0),
                                (4, 3, 16, 32, 0.0)]:
        mask = rng.random((mb, nb)) < dens
        rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
        ci = np.nonzero(mask)[1].astype(np.int32)
        for integer in (True, False):
            vals = (rng.integers(-3, 4, (len(ci), bs, bs)) if integer
                    else rng.standard_normal((len(ci), bs, bs))).astype(np.float64)
            Bnk = (rng.integers(-3, 4, (n, nb * bs)) if integer
                   else rng.standard_normal((n, nb * bs))).astype(np.float64)
            dense = np.zeros((mb * bs, nb * bs))
            for i in range(mb):
                for q in range(rp[i], rp[i + 1]):
                    dense[i * bs:(i + 1) * bs, ci[q] * bs:(ci[q] + 1) * bs] = vals[q]
            Bin = np.ascontiguousarray(Bnk.T) if b_kn else Bnk
            C = oracle.bsr_spmm(mb, nb, bs, rp, ci, vals, Bin, b_kn=b_kn)
            if integer:
                np.testing.assert_array_equal(C, dense @ Bnk.T)
            else:
                np.testing.assert_allclose(C, dense @ Bnk.T, rtol=1e-12, atol=1e-12)


def test_exclusive_scan_vs_cumsum():
    rng = np.random.default_rng(9)
    for n in [0, 1, 2, 7, 257]:
        x = rng.integers(1, 100, size=n).astype(np.int32)
        y = oracle.exclusive_scan(x)
        if n >= 2:
            np.testing.assert_array_equal(y, np.concatenate([[0], np.cumsum(x)[:-1]]))
        else:
            # the reference returns early for length 0 and 1 and leaves the input as is
            # (src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:272-273)
            np.testing.assert_array_equal(y, x)


def _random_pattern(rng, mb, nb, p):
    mask = rng.random((mb, nb)) < p
    rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    ci = np.nonzero(mask)[1].astype(np.int32)
    return mask, rp, ci


@pytest.mark.parametrize("mb,kb,nb,p", [(1, 1, 1, 1.0), (9, 7, 11, 0.3), (64, 64, 64, 0.05),
                                        (40, 30, 50, 0.0)])
def test_spgemm_symbolic_and_fill_vs_boolean_product(mb, kb, nb, p):
    rng = np.random.default_rng(mb * 1000 + nb)
    ma, rpA, ciA = _random_pattern(rng, mb, kb, p)
    mbm, rpB, ciB = _random_pattern(rng, kb, nb, p)
    prod = (ma.astype(np.int64) @ mbm.astype(np.int64)) > 0
    rpC, nnz = oracle.spgemm_symbolic(mb, nb, rpA, ciA, rpB, ciB)
    np.testing.assert_array_equal(rpC, np.concatenate([[0], np.cumsum(prod.sum(1))]))
    assert nnz == prod.sum()
    ciC = oracle.spgemm_fill(mb, nb, rpA, ciA, rpB, ciB, rpC)
    np.testing.assert_array_equal(ciC, np.nonzero(prod)[1])
    assert oracle.spgemm_nnzcub(ciA, rpB) == int((ma.astype(np.int64) @ mbm.sum(1)).sum())


def test_spgemm_numeric_vs_dense_product():
    rng = np.random.default_rng(21)
    bs, mb, kb, nb = 16, 6, 5, 7
    ma, rpA, ciA = _random_pattern(rng, mb, kb, 0.5)
    mbm, rpB, ciB = _random_pattern(rng, kb, nb, 0.5)
    vA = rng.standard_normal((len(ciA), bs, bs)).astype(np.float16)
    vB = rng.standard_normal((len(ciB), bs, bs)).astype(np.float16)
    dA = np.zeros((mb * bs, kb * bs))
    dB = np.zeros((kb * bs, nb * bs))
    for i in range(mb):
        for t in range(rpA[i], rpA[i + 1]):
            dA[i * bs:(i + 1) * bs, ciA[t] * bs:(ciA[t] + 1) * bs] = vA[t]
    for i in range(kb):
        for t in range(rpB[i], rpB[i + 1]):
            dB[i * bs:(i + 1) * bs, ciB[t] * bs:(ciB[t] + 1) * bs] = vB[t]
    dC = dA @ dB
    rpC, _ = oracle.spgemm_symbolic(mb, nb, rpA, ciA, rpB, ciB)
    ciC = oracle.spgemm_fill(mb, nb, rpA, ciA, rpB, ciB, rpC)
    vC = oracle.spgemm_numeric(mb, nb, bs, rpA, ciA, vA.view(np.uint16), rpB, ciB,
                               vB.view(np.uint16), rpC)
    for i in range(mb):
        for t in range(rpC[i], rpC[i + 1]):
            j = ciC[t]
            np.testing.assert_allclose(vC[t], dC[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs],
                                       rtol=0, atol=1e-11)
    vC64 = oracle.spgemm_numeric(mb, nb, bs, rpA, ciA, vA.astype(np.float64), rpB, ciB,
                                 vB.astype(np.float64), rpC)
    np.testing.assert_allclose(vC64, vC, rtol=0, atol=1e-11)


def test_golden_fixtures_match_oracle():
    """The committed fixtures (tests/golden/make_golden.py) still match the oracle."""
    path = os.path.join(GOLDEN, "spgemm_ref_default.npz")
    g = np.load(path, allow_pickle=False)
    rp, ci, vals = oracle.bsr_generate(8, 8, 0.5, 16)
    np.testing.assert_array_equal(rp, g["rowptr"])
    np.testing.assert_array_equal(ci, g["colidx"])
    np.testing.assert_array_equal(vals, g["vals"])
    rpC, _ = oracle.spgemm_symbolic(8, 8, rp, ci, rp, ci)
    np.testing.assert_array_equal(rpC, g["rowptrC"])
    np.testing.assert_array_equal(oracle.spgemm_fill(8, 8, rp, ci, rp, ci, rpC), g["colidxC"])
    with open(os.path.join(GOLDEN, "ramp_kat.json")) as f:
        kat = json.load(f)
    for entry in kat:
        s = entry["size"]
        A = oracle.ramp(s * s).reshape(1, s, s)
        C = oracle.gemm_f64(A, A)
        for (i, j, v) in entry["samples"]:
            assert C[0, i, j] == v


@pytest.mark.parametrize("beta,slices,bf16", [(9, 6, False), (8, 7, True), (9, 2, False)])
def test_split_slices_reconstruct_within_bound(beta, slices, bf16):
    """The digits and exponents give X back within 2^(e - beta slices - 1) per element,
    each chunk exponent bounds its chunk (|x| < 2^e, and some |x| >= 2^(e-1)), and the
    digits obey |d_0| <= 2^beta, |d_s| <= 2^(beta-1) -- checked from the outputs alone."""
    rng = np.random.default_rng(7)
    rows, k = 9, 70  # a partial last chunk
    X = np.sign(rng.standard_normal((rows, k))) * 10.0 ** rng.uniform(-12, 12, (rows, k))
    X[3] = 0.0                       # all-zero row
    X[4, :32] = 0.0                  # one all-zero chunk
    X[5, 7] = -0.0
    X[6, :5] = [2.0 ** -1074, 3 * 2.0 ** -1070, 1e300, -1e-300, 0.5]
    digits, e = oracle.split_slices(X, beta, slices, bf16)
    if bf16:
        vals = (digits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    else:
        vals = digits.view(np.float16).astype(np.float64)
    assert np.all(vals == np.rint(vals))
    assert np.all(np.abs(vals[0]) <= 2.0 ** beta) and np.all(np.abs(vals[1:]) <= 2.0 ** (beta - 1))
    ec = np.repeat(e, 32, axis=1)[:, :k]
    for i in range(rows):  # exact rational arithmetic: no rounding in the check itself
        for j in range(k):
            recon = sum(Fraction(vals[s, i, j]) * Fraction(2) ** int(ec[i, j] - beta * (s + 1))
                        for s in range(slices))
            assert abs(recon - Fraction(X[i, j])) <= Fraction(2) ** int(ec[i, j] - beta * slices - 1)
    for i in range(rows):
        for c in range((k + 31) // 32):
            chunk = np.abs(X[i, 32 * c:32 * c + 32])
            if chunk.max() == 0:
                assert e[i, c] == 0 and np.all(vals[:, i, 32 * c:32 * c + 32] == 0)
            else:
                assert chunk.max() < 2.0 ** e[i, c] and chunk.max() >= 2.0 ** (e[i, c] - 1)


def test_split_slices_exact_on_short_mantissas():
    """Values with at most beta * slices significant bits below the chunk maximum come
    back exactly; a one-slice split of powers of two is the power itself."""
    X = np.array([[3.0, -5.0, 255.0, 1.0, 0.25, -0.125] + [0.0] * 26])
    d, e = oracle.split_slices(X, 9, 2)
    vals = d.view(np.float16).astype(np.float64)
    recon = sum(np.ldexp(vals[s], e[0, 0] - 9 * (s + 1)) for s in range(2))
    assert np.array_equal(recon, X) and e[0, 0] == 8
    P = np.array([[2.0 ** 40, -2.0 ** 40]])
    d1, e1 = oracle.split_slices(P, 9, 1)
    assert e1[0, 0] == 41 and list(d1[0, 0].view(np.float16).astype(float)) == [256.0, -256.0]


def test_randr_mod3_from_offset_is_a_window():
    assert np.array_equal(oracle.randr_mod3_from(100, 37), oracle.randr_mod3(137)[37:])
