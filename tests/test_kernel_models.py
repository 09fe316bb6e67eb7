"""CPU models of two index and exactness arguments the kernels rely on.

* The block SpMM numbers its workgroups so that XCD x (workgroup L runs on XCD L mod 8)
  takes the contiguous run [x q + min(x, r), ...) of the tile-major work order,
  q = total / 8, r = total mod 8 (spmm.hip): the map must be a bijection of [0, total).
* The split kernels sum digit products in FP32 (split_f64.hip): a level sum
  T_g = sum_{s+t=g} sum_k d_s(a_k) d_t(b_k) is exact when its bound is at most 2^24, and a
  level that may exceed it is allowed only when its worst rounding stays below 2^-50 of
  2^(e_i + f_j).  The bound is checked here against brute force on small cases, and the
  level choices of the resident FP16 form (an 8-bit first digit, so every level sum runs
  over the whole K = 64) against the full-width first digit, where levels 4 and 5 would
  need a drain per 32-wide chunk.
"""
import itertools

import numpy as np
import pytest


def xcd_order(L, total):
    x, q8, r8 = L % 8, total // 8, total % 8
    return x * q8 + min(x, r8) + L // 8


@pytest.mark.parametrize("total", list(range(1, 70)) + [255, 256, 257, 1024, 1031, 4093])
def test_xcd_order_is_a_bijection_with_contiguous_runs_per_xcd(total):
    P = [xcd_order(L, total) for L in range(total)]
    assert sorted(P) == list(range(total))
    for x in range(min(8, total)):
        run = sorted(P[L] for L in range(x, total, 8))
        assert run == list(range(run[0], run[0] + len(run)))


def level_bound(beta, S, g, kc, b1):
    tot = 0
    for s in range(1, S + 1):
        t = g - s
        if 1 <= t <= S:
            bs = 2 ** b1 if s == 1 else 2 ** (beta - 1)
            bt = 2 ** b1 if t == 1 else 2 ** (beta - 1)
            tot += kc * bs * bt
    return tot


def level_ok(beta, S, g, kc, b1):
    b = level_bound(beta, S, g, kc, b1)
    if b <= 2 ** 24:
        return True
    return b * 2.0 ** -(24 + beta * g - 2 * (beta - b1)) <= 2.0 ** -50


def test_level_bound_is_attained_by_extreme_digits():
    """Brute force over digit magnitudes on a small case: the largest |T_g| over all digit
    choices with |d_1| <= 2^b1 and |d_s| <= 2^(beta-1) equals the bound."""
    beta, S, b1, kc = 3, 3, 2, 2
    lim = [2 ** b1] + [2 ** (beta - 1)] * (S - 1)
    for g in range(2, 2 * S + 1):
        best = 0
        # per-k digit vectors at +-max magnitude (extremes of a multilinear sum)
        choices = list(itertools.product(*[(-m, m) for m in lim]))
        for da in choices:
            for db in choices:
                T = sum(da[s] * db[g - 2 - s] for s in range(S) if 0 <= g - 2 - s < S)
                best = max(best, abs(T) * kc)
        assert best == level_bound(beta, S, g, kc, b1)


def test_streaming_form_levels_fit_fp32_per_32_chunk():
    # FP16 digits (beta 9, 6 slices, levels <= 7) and BF16 (beta 8, 7 slices, levels <= 8)
    for beta, S, G in [(9, 6, 7), (9, 7, 8), (8, 7, 8)]:
        assert all(level_ok(beta, S, g, 32, beta) for g in range(2, G + 1))
    # and the two smallest levels may be added in FP32 first
    for beta, S, G in [(9, 6, 7), (9, 7, 8), (8, 7, 8)]:
        assert level_bound(beta, S, G - 1, 32, beta) * 2.0 ** (-24 - beta * (G - 1)) <= 2.0 ** -50


def test_resident_form_levels_run_over_the_whole_k():
    # FP16 resident: 8-bit first digit, K = 64 (and 32): every level over the whole K
    for K in (32, 64):
        assert all(level_ok(9, 6, g, K, 8) for g in range(2, 8))
        assert all(level_ok(8, 7, g, K, 8) for g in range(2, 9))
    # with a 9-bit first digit, levels 4 and 5 of K = 64 could round beyond 2^-50
    assert [g for g in range(2, 8) if not level_ok(9, 6, g, 64, 9)] == [4, 5]


def test_fp32_level_sum_exact_up_to_2_pow_24():
    """Integers of magnitude up to 2^24 are exact in FP32, one more is not."""
    assert float(np.float32(2 ** 24)) == 2 ** 24
    assert float(np.float32(2 ** 24 - 1)) == 2 ** 24 - 1
    assert float(np.float32(2 ** 24 + 1)) != 2 ** 24 + 1


# ---- int8 digits of the split path (split_i8.hip) ----------------------------------

def _ldexp_bits(x, e):
    return np.frombuffer(np.array([np.ldexp(x, e)], dtype=np.float64).tobytes(), np.uint64)[0]


def digit_word_model(x, neg_e):
    """The kernel's digit_word step for step with 32-bit words: scale by 2^-E, place the
    significand's leading bit at bit 62 (v_alignbit + bit insert), shift right by
    min(1023 - exponent, 63), two's complement by the sign, add 0x80 to every byte with
    carries, flip each byte's top bit."""
    M32 = 0xFFFFFFFF
    bits = int(_ldexp_bits(x, neg_e))
    h, l = bits >> 32, bits & M32
    ex = (h >> 20) & 0x7FF
    r = min((1023 - ex) & M32, 63)
    align = ((h << 32 | l) >> 22) & M32
    mh = (align & 0x3FFFFFFF) | 0x40000000
    X = ((mh << 32) | ((l << 10) & M32)) >> r
    s = 0xFFFFFFFFFFFFFFFF if h >> 31 else 0
    Y = ((X ^ s) + 0x8080808080808080 + (s & 1)) & 0xFFFFFFFFFFFFFFFF
    return Y ^ 0x8080808080808080


def balanced_digits(word):
    return [((word >> (8 * b)) & 0xFF) - 256 * (((word >> (8 * b)) & 0xFF) >= 128) for b in range(8)]


def row_exponent_model(row):
    """|x| < 2^E from the largest upper word of |x|, completed with all-ones low bits."""
    hi = max(int(np.frombuffer(np.float64(v).tobytes(), np.uint64)[0]) >> 32 & 0x7FFFFFFF for v in row)
    return int(np.frexp(np.frombuffer(np.uint64(hi << 32 | 0xFFFFFFFF).tobytes(), np.float64)[0])[1])


def test_int8_digits_reconstruct_the_truncated_fixed_point_value():
    """For every x of a row with |x| < 2^E the eight balanced byte digits d_b in
    [-128, 127] satisfy sum_b d_b 2^(8b) = trunc(|x| 2^(62 - E)) sign(x) exactly (Python
    integers and fractions as the exact reference); the row maximum lands in [2^61, 2^62)."""
    from fractions import Fraction
    rng = np.random.default_rng(5)
    rows = [np.where(rng.random(64) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-4, 4, 64),
            rng.uniform(-1, 1, 64), np.array([0.0, -0.0, 1.0, -1.0, 2.0 ** -40, 3.0, 0.75, 1e-310]),
            np.array([5e-324 * k for k in (1, 7, 999, 1000)]),       # a row of subnormals
            np.array([1.7e308, -1e300, 1.0]), np.array([2.0 ** -1022, 2.0 ** -1060])]
    for row in rows:
        E = row_exponent_model(row)
        assert all(abs(Fraction(float(v))) < Fraction(2) ** E for v in row)
        big = max(abs(Fraction(float(v))) for v in row)
        for v in row:
            d = balanced_digits(digit_word_model(float(v), -E))
            assert all(-128 <= q <= 127 for q in d)
            X = sum(q << (8 * b) for b, q in enumerate(d))
            fx = Fraction(float(v))
            want = int(abs(fx) * Fraction(2) ** (62 - E))  # truncation toward zero
            assert X == (want if fx >= 0 else -want), (v, E)
        if big > 0 and big.numerator.bit_length() < 60:
            top = int(big * Fraction(2) ** (62 - E))
            assert 2 ** 61 <= top < 2 ** 62 or row_exponent_model(row) < -1021


def v_perm(s0, s1, sel):
    data = (s0 << 32) | s1
    out = 0
    for i in range(4):
        b = (sel >> (8 * i)) & 0xFF
        assert b < 8
        out |= ((data >> (8 * b)) & 0xFF) << (8 * i)
    return out


def test_int8_transpose_selectors_gather_one_byte_plane_per_word():
    """transpose4's v_perm_b32 selectors: out[b] holds byte b of w0, w1, w2, w3."""
    rng = np.random.default_rng(6)
    for _ in range(50):
        w = [int(x) for x in rng.integers(0, 2 ** 63, 4, dtype=np.uint64)]
        out = []
        for half in range(2):
            a, b, c, d = ((x >> (32 * half)) & 0xFFFFFFFF for x in w)
            ab_lo, ab_hi = v_perm(b, a, 0x05010400), v_perm(b, a, 0x07030602)
            cd_lo, cd_hi = v_perm(d, c, 0x05010400), v_perm(d, c, 0x07030602)
            out += [v_perm(cd_lo, ab_lo, 0x05040100), v_perm(cd_lo, ab_lo, 0x07060302),
                    v_perm(cd_hi, ab_hi, 0x05040100), v_perm(cd_hi, ab_hi, 0x07060302)]
        for byte in range(8):
            assert out[byte] == sum(((w[q] >> (8 * byte)) & 0xFF) << (8 * q) for q in range(4))


def test_int8_level_sums_fit_int32_and_the_horner_scale():
    """Level l sums at most min(l, 14 - l) + 1 <= 8 digit products of magnitude <= 2^14 per
    k, so over K <= 128 every level is below 2^31 (exact in int32).  The kept levels
    combine as C = 2^(e + f - 124) sum_l T_l 2^(8 (14 - l)) = 2^(e + f - 12) sum_l T_l 2^(-8 l)."""
    for K in (64, 128):
        for l in range(15):
            terms = sum(1 for p in range(8) for q in range(8) if p + q == l)
            assert terms <= 8 and terms * K * 128 * 128 < 2 ** 31
    # Horner scale on a random exact example
    rng = np.random.default_rng(7)
    da = rng.integers(-128, 128, (8, 5))
    db = rng.integers(-128, 128, (8, 5))
    X = [sum(int(da[p, k]) << (8 * (7 - p)) for p in range(8)) for k in range(5)]
    Y = [sum(int(db[q, k]) << (8 * (7 - q)) for q in range(8)) for k in range(5)]
    exact = sum(x * y for x, y in zip(X, Y))
    T = [sum(int(da[p, k]) * int(db[l - p, k]) for k in range(5) for p in range(8) if 0 <= l - p < 8)
         for l in range(15)]
    assert sum(t << (8 * (14 - l)) for l, t in enumerate(T)) == exact


def _int8_split_model(A, B, levels):
    """The int8 split's arithmetic in exact integers: per-row exponents over K, 62-bit
    fixed point truncated toward zero, balanced byte digits, level sums of the digit
    products for l < levels in int64 (exact), Horner in FP64 and the 2^(e + f - 12) scale."""
    def digits(X):
        E = np.frexp(np.max(np.abs(X), axis=-1))[1]
        y = np.ldexp(np.abs(X), (30 - E)[..., None])
        hi = np.floor(y)
        lo = np.floor(np.ldexp(y - hi, 32))
        V = (hi.astype(np.int64) << 32) + lo.astype(np.int64)  # trunc(|x| 2^(62 - E))
        V = np.where(X < 0, -V, V)
        d = []
        for _ in range(8):
            r = ((V + 128) & 255) - 128
            d.append(r)
            V = (V - r) >> 8
        assert np.all(V == 0)
        return E, d[::-1]
    Ea, da = digits(A)
    Eb, db = digits(B)
    T = [np.zeros(A.shape[:-1] + (B.shape[-2],), np.int64) for _ in range(levels)]
    for p in range(8):
        for q in range(8):
            if p + q < levels:
                T[p + q] += np.matmul(da[p], np.swapaxes(db[q], -1, -2))
    acc = T[levels - 1].astype(np.float64)
    for l in range(levels - 2, -1, -1):
        acc = acc / 256.0 + T[l].astype(np.float64)
    return np.ldexp(acc, Ea[..., :, None] + Eb[..., None, :] - 12)


def test_int8_level_cutoff_meets_the_componentwise_bound_with_eight_levels_only():
    """On the split-precision workload (sign-random, log10|x| ~ U[-4, 4], 128^3), keeping
    the levels l < 8 (36 digit products) is within 1e-13 |A||B| of the Dot2 product
    componentwise; dropping level 7 (28 products) is not."""
    import oracle
    rng = np.random.default_rng(0x4B414D49 + 2)
    shape = (6, 128, 128)
    A = np.where(rng.random(shape) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-4, 4, shape)
    B = np.where(rng.random(shape) < 0.5, -1.0, 1.0) * 10.0 ** rng.uniform(-4, 4, shape)
    exact, absC = oracle.gemm_dot2(A, B)
    comp = {}
    for L in (8, 7):
        d = np.abs(_int8_split_model(A, B, L) - exact)
        comp[L] = float(np.max(d / np.where(absC > 0, absC, 1.0)))
    assert comp[8] <= 1e-14
    assert comp[7] > 1e-13
