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
