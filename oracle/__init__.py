"""CPU oracle for the block-GEMM path -- TEST INFRASTRUCTURE ONLY.

Python binding of ``oracle/libkami_oracle.so`` (built from ``kami_oracle.c`` by
``oracle/Makefile``), a plain-C restatement of the reference's host-side algorithms:
the FP64 validation triple loop (src/block_gemm/KAMI/1d/fp64/block_gemm_1d_double_mma.cu:
294-304), the OpenMP batched loop (src/batched_gemm/KAMI/1d/batched_gemm_1d_double_mma.cu:
304-318), the FP16 loop with double casts (src/block_gemm/KAMI/1d/fp16/
block_gemm_1d_half_mma.cu:285-296), the 3D slab sum (src/block_gemm/KAMI/3d/fp64/
block_gemm_3d_double_mma.cu:324-333), the SpGEMM SPA symbolic/numeric phases, exclusive
scan and BSR generator (src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:269-397) and the
column fill (src/sparse/spgemm/2d/block_spgemm_2d_half_mma.cu:88-116), plus a compensated
Dot2 product (Ogita-Rump-Oishi 2005) as the exact reference for split-precision errors.

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py may load
this package.  Pinning of each function: see tests/test_oracle.py (every function is
checked against closed forms, brute force or an independent library routine).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkami_oracle.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_c = ctypes
_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "oracle_half_to_double": (_c.c_double, [_c.c_uint16]),
            "oracle_double_to_half": (_c.c_uint16, [_c.c_double]),
            "oracle_bf16_to_double": (_c.c_double, [_c.c_uint16]),
            "oracle_double_to_bf16": (_c.c_uint16, [_c.c_double]),
            "oracle_gemm_f64": (None, [_c.c_int, _c.c_int, _c.c_int, _f64p, _c.c_int64, _f64p,
                                       _c.c_int64, _c.c_int, _f64p, _c.c_int64]),
            "oracle_batched_gemm_f64": (None, [_c.c_int64, _c.c_int, _c.c_int, _c.c_int, _f64p,
                                               _c.c_int64, _f64p, _c.c_int64, _f64p, _c.c_int64,
                                               _c.c_int, _c.c_int]),
            "oracle_batched_uneven_f64": (None, [_c.c_int64, _i32p, _f64p, _f64p, _f64p, _c.c_int]),
            "oracle_batched_gemm_u16": (None, [_c.c_int64, _c.c_int, _c.c_int, _c.c_int, _u16p,
                                               _c.c_int64, _u16p, _c.c_int64, _f64p, _c.c_int64,
                                               _c.c_int, _c.c_int, _c.c_int]),
            "oracle_batched_gemm_dot2_f64": (None, [_c.c_int64, _c.c_int, _c.c_int, _c.c_int,
                                                    _f64p, _c.c_int64, _f64p, _c.c_int64, _f64p,
                                                    _c.c_int64, _f64p, _c.c_int, _c.c_int]),
            "oracle_slab_sum_f64": (None, [_c.c_int, _c.c_int, _c.c_int, _f64p]),
            "oracle_fill_ramp": (None, [_f64p, _c.c_int64]),
            "oracle_fill_randr_mod3": (None, [_f64p, _c.c_int64]),
            "oracle_bsr_generate": (_c.c_int, [_c.c_int, _c.c_int, _c.c_float, _c.c_int, _i32p,
                                               _i32p, _u16p]),
            "oracle_exclusive_scan_i32": (None, [_i32p, _c.c_int]),
            "oracle_spgemm_symbolic": (_c.c_int64, [_c.c_int, _c.c_int, _i32p, _i32p, _i32p,
                                                    _i32p, _i32p]),
            "oracle_spgemm_fill": (None, [_c.c_int, _c.c_int, _i32p, _i32p, _i32p, _i32p, _i32p,
                                          _i32p]),
            "oracle_spgemm_numeric_u16": (None, [_c.c_int, _c.c_int, _c.c_int, _i32p, _i32p, _u16p,
                                                 _i32p, _i32p, _u16p, _i32p, _f64p, _c.c_int]),
            "oracle_spgemm_numeric_f64": (None, [_c.c_int, _c.c_int, _c.c_int, _i32p, _i32p, _f64p,
                                                 _i32p, _i32p, _f64p, _i32p, _f64p]),
            "oracle_spgemm_nnzcub": (_c.c_int64, [_c.c_int, _i32p, _i32p]),
            "oracle_bsr_spmm_f64": (None, [_c.c_int, _c.c_int, _c.c_int, _i32p, _i32p, _f64p,
                                           _f64p, _c.c_int, _c.c_int64, _c.c_int, _f64p,
                                           _c.c_int64]),
            "oracle_spmm_generate": (_c.c_int, [_c.c_int, _c.c_int, _c.c_float, _c.c_int, _i32p,
                                                _i32p, _f64p, _c.c_int, _f64p]),
            "oracle_max_threads": (_c.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _L = L
    return _L


def _c64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def gemm_f64(A, B, b_kn=False, nthreads=0):
    """C[p] = A[p] @ op(B[p]) for (count, M, K) / (count, N, K) (or (count, K, N)) arrays."""
    A, B = _c64(A), _c64(B)
    count, m, k = A.shape
    n = B.shape[2] if b_kn else B.shape[1]
    C = np.empty((count, m, n), dtype=np.float64)
    lib().oracle_batched_gemm_f64(count, m, n, k, A, m * k, B, n * k, C, m * n, int(b_kn), nthreads)
    return C


def batched_uneven_f64(sizes, A, B, nthreads=0):
    """C of every problem p = sizes[p] = (m, n, k): A (m x k), B (n x k) and C (m x n)
    packed one after another in problem order (the batched workload's layout); the
    reference's per-problem FP64 loop with OpenMP over the batch
    (src/batched_gemm/KAMI/1d/batched_gemm_1d_double_mma.cu:304-318)."""
    mnk = np.ascontiguousarray(sizes, dtype=np.int32).reshape(-1)
    count = mnk.size // 3
    nc = int(np.sum(mnk[0::3].astype(np.int64) * mnk[1::3]))
    C = np.empty(nc, dtype=np.float64)
    lib().oracle_batched_uneven_f64(count, mnk, _c64(A), _c64(B), C, nthreads)
    return C


def gemm_u16(A, B, is_bf16=False, b_kn=False, nthreads=0):
    """FP16/BF16 bit patterns (uint16) in, double out, double accumulation."""
    A = np.ascontiguousarray(A, dtype=np.uint16)
    B = np.ascontiguousarray(B, dtype=np.uint16)
    count, m, k = A.shape
    n = B.shape[2] if b_kn else B.shape[1]
    C = np.empty((count, m, n), dtype=np.float64)
    lib().oracle_batched_gemm_u16(count, m, n, k, A, m * k, B, n * k, C, m * n, int(b_kn),
                                  int(is_bf16), nthreads)
    return C


def gemm_dot2(A, B, b_kn=False, nthreads=0):
    """Compensated (twice-working-precision) product and the |A||B| magnitudes."""
    A, B = _c64(A), _c64(B)
    count, m, k = A.shape
    n = B.shape[2] if b_kn else B.shape[1]
    C = np.empty((count, m, n), dtype=np.float64)
    absC = np.empty((count, m, n), dtype=np.float64)
    lib().oracle_batched_gemm_dot2_f64(count, m, n, k, A, m * k, B, n * k, C, m * n, absC,
                                       int(b_kn), nthreads)
    return C, absC


def slab_sum(slabs):
    """Sum P partial (M, N) slabs stored contiguously (P, M, N) -> (M, N)."""
    S = _c64(slabs).copy()
    p, m, n = S.shape
    lib().oracle_slab_sum_f64(m, n, p, S)
    return S[0]


def ramp(n):
    x = np.empty(n, dtype=np.float64)
    lib().oracle_fill_ramp(x, n)
    return x


def randr_mod3(n):
    x = np.empty(n, dtype=np.float64)
    lib().oracle_fill_randr_mod3(x, n)
    return x


def randr_mod3_from(n, offset):
    """rand_r(seed = offset + i) % 3 for i < n (the per-element generator above, from an
    offset)."""
    return randr_mod3(n + offset)[offset:]


def split_slices(X, beta, slices, bf16=False):
    """Digit slices of the split-precision path, restated from their definition: per row
    and 32-wide chunk the exponent e of the largest |x| (|x| < 2^e, 0 for an all-zero
    chunk); r = x 2^(beta - e); `slices` times d = rint(r) (ties to even), r = (r - d) 2^beta.
    Returns (digits uint16 (slices, rows, k) as binary16 or bfloat16 bits, exps int32
    (rows, ceil(k / 32))).  Pinned in tests/test_oracle.py by reconstruction within
    2^(e - beta slices - 1), the digit bounds and exactness on short-mantissa inputs."""
    X = np.asarray(X, dtype=np.float64)
    rows, k = X.shape
    nch = (k + 31) // 32
    pad = np.zeros((rows, nch * 32))
    pad[:, :k] = X
    ex = np.frexp(pad)[1]
    ex = np.where(pad == 0, -4096, ex)
    e = ex.reshape(rows, nch, 32).max(axis=2)
    e = np.where(e == -4096, 0, e).astype(np.int32)
    r = np.ldexp(X, beta - np.repeat(e, 32, axis=1)[:, :k])
    digits = np.empty((slices, rows, k), dtype=np.uint16)
    for s in range(slices):
        q = np.rint(r)
        if bf16:
            digits[s] = (q.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
        else:
            digits[s] = q.astype(np.float16).view(np.uint16)
        r = (r - q) * float(2 ** beta)
    return digits, e


def half_to_double(h):
    return lib().oracle_half_to_double(int(h))


def double_to_half(x):
    return lib().oracle_double_to_half(float(x))


def bf16_to_double(b):
    return lib().oracle_bf16_to_double(int(b))


def double_to_bf16(x):
    return lib().oracle_double_to_bf16(float(x))


def bsr_generate(rows, cols, sparsity, bs):
    """The reference's random BSR generator: (rowptr, colidx, vals uint16 (nnzb, bs, bs))."""
    rp = np.zeros(rows + 1, dtype=np.int32)
    ci = np.zeros(rows * cols, dtype=np.int32)
    vals = np.zeros(rows * cols * bs * bs, dtype=np.uint16)
    nnzb = lib().oracle_bsr_generate(rows, cols, sparsity, bs, rp, ci, vals)
    return rp, ci[:nnzb].copy(), vals[: nnzb * bs * bs].reshape(nnzb, bs, bs).copy()


def exclusive_scan(x):
    y = np.ascontiguousarray(x, dtype=np.int32).copy()
    lib().oracle_exclusive_scan_i32(y, len(y))
    return y


def spgemm_symbolic(mb, nb, rpA, ciA, rpB, ciB):
    rpC = np.zeros(mb + 1, dtype=np.int32)
    nnz = lib().oracle_spgemm_symbolic(mb, nb, np.ascontiguousarray(rpA, np.int32),
                                       np.ascontiguousarray(ciA, np.int32),
                                       np.ascontiguousarray(rpB, np.int32),
                                       np.ascontiguousarray(ciB, np.int32), rpC)
    return rpC, int(nnz)


def spgemm_fill(mb, nb, rpA, ciA, rpB, ciB, rpC):
    ciC = np.zeros(max(int(rpC[-1]), 1), dtype=np.int32)
    lib().oracle_spgemm_fill(mb, nb, np.ascontiguousarray(rpA, np.int32),
                             np.ascontiguousarray(ciA, np.int32),
                             np.ascontiguousarray(rpB, np.int32),
                             np.ascontiguousarray(ciB, np.int32),
                             np.ascontiguousarray(rpC, np.int32), ciC)
    return ciC[: int(rpC[-1])]


def spgemm_numeric(mb, nb, bs, rpA, ciA, vA, rpB, ciB, vB, rpC, is_bf16=False):
    """vA, vB: (nnzb, bs, bs) uint16 (FP16/BF16 bits) or float64; returns (nnzbC, bs, bs)."""
    nnzc = int(rpC[-1])
    vC = np.zeros(max(nnzc, 1) * bs * bs, dtype=np.float64)
    args = [np.ascontiguousarray(rpA, np.int32), np.ascontiguousarray(ciA, np.int32)]
    if np.asarray(vA).dtype == np.float64:
        lib().oracle_spgemm_numeric_f64(mb, nb, bs, *args, _c64(vA).ravel(),
                                        np.ascontiguousarray(rpB, np.int32),
                                        np.ascontiguousarray(ciB, np.int32), _c64(vB).ravel(),
                                        np.ascontiguousarray(rpC, np.int32), vC)
    else:
        lib().oracle_spgemm_numeric_u16(mb, nb, bs, *args,
                                        np.ascontiguousarray(vA, np.uint16).ravel(),
                                        np.ascontiguousarray(rpB, np.int32),
                                        np.ascontiguousarray(ciB, np.int32),
                                        np.ascontiguousarray(vB, np.uint16).ravel(),
                                        np.ascontiguousarray(rpC, np.int32), vC, int(is_bf16))
    return vC[: nnzc * bs * bs].reshape(nnzc, bs, bs)


def spgemm_nnzcub(ciA, rpB):
    ciA = np.ascontiguousarray(ciA, np.int32)
    return int(lib().oracle_spgemm_nnzcub(len(ciA), ciA, np.ascontiguousarray(rpB, np.int32)))


def max_threads():
    return int(lib().oracle_max_threads())


def bsr_spmm(mb, nb, bs, rp, ci, vals, B, b_kn=False):
    """C (mb*bs, n) = A_bsr * op(B) in double; vals (nnzb, bs, bs) float64, B (n, K) or
    (K, n) for b_kn."""
    B = _c64(B)
    n = B.shape[1] if b_kn else B.shape[0]
    C = np.zeros((mb * bs, n), dtype=np.float64)
    lib().oracle_bsr_spmm_f64(mb, nb, bs, np.ascontiguousarray(rp, np.int32),
                              np.ascontiguousarray(ci, np.int32), _c64(vals).ravel(), B.ravel(),
                              n, B.shape[1], int(bool(b_kn)), C.ravel(), n)
    return C


def spmm_generate(rows, cols, sparsity, bs, n):
    """The reference SpMM program's inputs: (rowptr, colidx, vals (nnzb, bs, bs), B (n, K))."""
    rp = np.zeros(rows + 1, dtype=np.int32)
    ci = np.zeros(rows * cols, dtype=np.int32)
    vals = np.zeros(rows * cols * bs * bs, dtype=np.float64)
    B = np.zeros((n, cols * bs), dtype=np.float64)
    nnzb = lib().oracle_spmm_generate(rows, cols, sparsity, bs, rp, ci, vals, n, B.ravel())
    return rp, ci[:nnzb].copy(), vals[: nnzb * bs * bs].reshape(nnzb, bs, bs).copy(), B
