"""kami_amd -- MI355X (gfx950) block-level GEMM, batched GEMM and block SpGEMM.

Python host side over the C ABI of ``libkami.so`` (``include/kami.h``).  Device
memory and streams come from PyTorch-ROCm tensors; every computation runs in the
library's HIP kernels.
"""
from __future__ import annotations

import ctypes
import os

from . import _abi
from ._abi import (B_KN, B_NK, GEN_RAMP, GEN_RAND_R3, PART_1D, PART_2D, PART_3D, PART_AUTO,
                   PREC_BF16_F32, PREC_F16_F32, PREC_F64, PREC_F64_SPLIT_BF16,
                   PREC_F64_SPLIT_F16, PREC_F64_SPLIT_I8, KamiError, check, lib)

__all__ = [
    "block_gemm", "query", "make_desc", "KamiError", "lib", "version", "device_ok",
    "PREC_F64", "PREC_F16_F32", "PREC_BF16_F32", "PREC_F64_SPLIT_F16", "PREC_F64_SPLIT_BF16",
    "PREC_F64_SPLIT_I8",
    "PART_AUTO", "PART_1D", "PART_2D", "PART_3D", "B_NK", "B_KN", "GEN_RAMP", "GEN_RAND_R3",
]

_PART_NAMES = {"auto": PART_AUTO, "1d": PART_1D, "2d": PART_2D, "3d": PART_3D}
_LAYOUT_NAMES = {"nk": B_NK, "kn": B_KN}


def _part(p):
    return _PART_NAMES[p.lower()] if isinstance(p, str) else int(p)


def _layout(b):
    return _LAYOUT_NAMES[b.lower()] if isinstance(b, str) else int(b)


def version() -> str:
    return lib().kami_version().decode()


def device_ok() -> bool:
    return lib().kami_device_check() == _abi.OK


def make_desc(m, n, k, precision, *, partition="auto", ranks=0, b_layout="nk", iters=1,
              alpha=1.0, beta=0.0, lda=0, ldb=0, ldc=0, slices=0, pipeline=0) -> _abi.GemmDesc:
    d = _abi.GemmDesc()
    check(lib().kami_gemm_desc_init(ctypes.byref(d), m, n, k, precision), "kami_gemm_desc_init")
    d.partition = _part(partition)
    d.ranks = ranks
    d.b_layout = _layout(b_layout)
    d.iters = iters
    d.alpha = alpha
    d.beta = beta
    d.lda, d.ldb, d.ldc = lda, ldb, ldc
    d.slices = slices
    d.pipeline = pipeline
    return d


def query(m, n, k, precision, **kw) -> dict:
    """The compiled instance a descriptor resolves to (no device access)."""
    d = make_desc(m, n, k, precision, **kw)
    cfg = _abi.GemmConfig()
    check(lib().kami_block_gemm_query(ctypes.byref(d), ctypes.byref(cfg)), "kami_block_gemm_query")
    return cfg.as_dict()


def _precision_of(A, precision):
    import torch
    if precision is not None:
        return precision
    if A.dtype == torch.float64:
        return PREC_F64
    if A.dtype == torch.float16:
        return PREC_F16_F32
    if A.dtype == torch.bfloat16:
        return PREC_BF16_F32
    raise TypeError(f"unsupported input dtype {A.dtype}")


_IN_DTYPE = {}  # precision -> torch dtype of A and B, filled on first use


def _in_dtype(prec):
    import torch
    if not _IN_DTYPE:
        _IN_DTYPE.update({PREC_F64: torch.float64, PREC_F64_SPLIT_F16: torch.float64,
                          PREC_F64_SPLIT_BF16: torch.float64, PREC_F64_SPLIT_I8: torch.float64,
                          PREC_F16_F32: torch.float16,
                          PREC_BF16_F32: torch.bfloat16})
    if prec not in _IN_DTYPE:
        raise ValueError(f"unknown precision {prec}")
    return _IN_DTYPE[prec]


def _out_dtype(prec):
    import torch
    return torch.float32 if prec in (PREC_F16_F32, PREC_BF16_F32) else torch.float64


def _on_gpu(t, name):
    """Raise ValueError unless tensor t lives in device memory: the library hands its
    pointer to a kernel, and a host pointer there faults the GPU."""
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a GPU tensor, got one on {t.device}")


def _expect(t, name, shape, dtype, device):
    """Raise ValueError unless tensor t has exactly this shape, dtype and device, is
    contiguous and lives on a GPU: the library receives raw device pointers and would
    read or write past a tensor that does not match."""
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    _on_gpu(t, name)


def _stream_ptr(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def block_gemm(A, B, C=None, *, precision=None, partition="auto", ranks=0, b_layout="nk",
               iters=1, alpha=1.0, beta=0.0, slices=0, pipeline=0, stream=None):
    """C = alpha * A @ op(B) + beta * C for a batch of problems, one per workgroup.

    A: (count, M, K) or (M, K); B: (count, N, K) ("nk", the reference layout) or
    (count, K, N) ("kn"); C: (count, M, N), float64 for FP64 and split precision,
    float32 for FP16/BF16 inputs.  Tensors must be contiguous device tensors.
    """
    import torch
    squeeze = A.dim() == 2
    if squeeze:
        A, B = A.unsqueeze(0), B.unsqueeze(0)
        if C is not None:
            C = C.unsqueeze(0)
    if A.dim() != 3 or B.dim() != 3:
        raise ValueError("block_gemm expects A (count, M, K) and B (count, N, K) or (count, K, N)")
    count, M, K = A.shape
    lay = _layout(b_layout)
    N = B.shape[1] if lay == B_NK else B.shape[2]
    prec = _precision_of(A, precision)
    out_dtype = _out_dtype(prec)
    if C is None:
        C = torch.empty((count, M, N), dtype=out_dtype, device=A.device)
    _expect(A, "A", (count, M, K), _in_dtype(prec), A.device)
    _expect(B, "B", (count, N, K) if lay == B_NK else (count, K, N), _in_dtype(prec), A.device)
    _expect(C, "C", (count, M, N), out_dtype, A.device)
    d = make_desc(M, N, K, prec, partition=partition, ranks=ranks, b_layout=lay, iters=iters,
                  alpha=alpha, beta=beta, slices=slices, pipeline=pipeline)
    st = lib().kami_block_gemm(ctypes.byref(d), ctypes.c_void_p(A.data_ptr()),
                               ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(C.data_ptr()),
                               count, 0, 0, 0, _stream_ptr(stream))
    check(st, f"kami_block_gemm({M}x{N}x{K}, prec={prec}, part={partition}, ranks={ranks})")
    return C[0] if squeeze else C


def split_gemm(A, B, C=None, *, precision=PREC_F64_SPLIT_F16, slices=0, alpha=1.0, beta=0.0,
               stream=None):
    """FP64 C = alpha * A @ B^T + beta * C through FP16/BF16 slices at any size whose M and
    N are multiples of 128 (or 64) and K a multiple of 32 (kami_split_gemm).

    A: (count, M, K) or (M, K) float64; B: (count, N, K) or (N, K); C like A @ B^T.
    """
    import torch
    squeeze = A.dim() == 2
    if squeeze:
        A, B = A.unsqueeze(0), B.unsqueeze(0)
        if C is not None:
            C = C.unsqueeze(0)
    if A.dim() != 3 or B.dim() != 3:
        raise ValueError("split_gemm expects A (count, M, K) and B (count, N, K)")
    count, M, K = A.shape
    N = B.shape[1]
    if C is None:
        C = torch.empty((count, M, N), dtype=torch.float64, device=A.device)
    _expect(A, "A", (count, M, K), torch.float64, A.device)
    _expect(B, "B", (count, N, K), torch.float64, A.device)
    _expect(C, "C", (count, M, N), torch.float64, A.device)
    d = make_desc(M, N, K, precision, alpha=alpha, beta=beta, slices=slices)
    check(lib().kami_split_gemm(ctypes.byref(d), ctypes.c_void_p(A.data_ptr()),
                                ctypes.c_void_p(B.data_ptr()), ctypes.c_void_p(C.data_ptr()),
                                count, 0, 0, 0, _stream_ptr(stream)),
          f"kami_split_gemm({M}x{N}x{K}, prec={precision})")
    return C[0] if squeeze else C


__all__ += ["split_gemm"]


def split_slices(X, *, precision=PREC_F64_SPLIT_F16, slices=6, stream=None):
    """Digit slices of an FP64 matrix as the split-precision path forms them
    (kami_split_slices): returns (digits, exps) with digits a uint16 tensor of shape
    (slices, rows, k) holding binary16 (FP16 precision) or bfloat16 (BF16) bits and exps
    an int32 tensor (rows, ceil(k / 32)), so that
    X[i, j] ~= sum_s digits[s, i, j] * 2^(exps[i, j // 32] - BETA * (s + 1)).
    X: (rows, k) float64 device tensor whose rows are contiguous."""
    import torch
    if X.dim() != 2 or X.dtype != torch.float64 or X.stride(1) != 1:
        raise ValueError("split_slices expects a (rows, k) float64 tensor with unit column stride")
    _on_gpu(X, "X")
    rows, k = X.shape
    digits = torch.empty((slices, rows, k), dtype=torch.int16, device=X.device)
    exps = torch.empty((rows, (k + 31) // 32), dtype=torch.int32, device=X.device)
    check(lib().kami_split_slices(ctypes.c_void_p(X.data_ptr()), rows, k, X.stride(0),
                                  precision, slices, ctypes.c_void_p(digits.data_ptr()),
                                  ctypes.c_void_p(exps.data_ptr()), _stream_ptr(stream)),
          f"kami_split_slices({rows}x{k}, prec={precision}, slices={slices})")
    return digits.view(torch.uint16), exps


def generate(kind, count, *, precision=PREC_F64, offset=0, device="cuda", stream=None):
    """A known-answer input of the reference's programs generated on the device
    (kami_generate): kind GEN_RAMP (x[i] = i + offset) or GEN_RAND_R3 (rand_r(i + offset)
    % 3).  Returns a float64 tensor, or float16 / bfloat16 for the FP16 / BF16 precisions."""
    import torch
    dt = {PREC_F64: torch.float64, PREC_F16_F32: torch.float16,
          PREC_BF16_F32: torch.bfloat16}.get(precision)
    if dt is None:
        raise KamiError(_abi.ENOTSUP, f"generate: precision {precision}")
    out = torch.empty(count, dtype=dt, device=device)
    _on_gpu(out, "generate's output")
    check(lib().kami_generate(kind, precision, ctypes.c_void_p(out.data_ptr()), count, offset,
                              _stream_ptr(stream)), f"kami_generate(kind={kind})")
    return out


__all__ += ["split_slices", "generate"]


class BatchedPlan:
    """A batched GEMM plan over problems of uneven size (kami_batched_plan_*).

    problems: sequence of (A, B, C) contiguous 2-D device tensors, A (M, K), B (N, K)
    ("nk") or (K, N) ("kn"), C (M, N); M, N, K multiples of 16.  The tensors must stay
    alive while the plan exists.  execute() is one persistent launch, replayed through a
    hipGraph when use_graph is set.
    """

    def __init__(self, problems, *, precision=None, b_layout="nk", autotune=False,
                 use_graph=True, alpha=1.0, beta=0.0, stream=None):
        if not problems:
            raise ValueError("empty problem list")
        A0 = problems[0][0]
        prec = _precision_of(A0, precision)
        lay = _layout(b_layout)
        arr = (_abi.Problem * len(problems))()
        in_dt, out_dt = _in_dtype(prec), _out_dtype(prec)
        for i, (A, B, C) in enumerate(problems):
            if A.dim() != 2 or B.dim() != 2 or C.dim() != 2:
                raise ValueError(f"problem {i}: A, B and C must be 2-D")
            m, k = A.shape
            n = B.shape[0] if lay == B_NK else B.shape[1]
            _expect(A, f"problem {i} A", (m, k), in_dt, A0.device)
            _expect(B, f"problem {i} B", (n, k) if lay == B_NK else (k, n), in_dt, A0.device)
            _expect(C, f"problem {i} C", (m, n), out_dt, A0.device)
            p = arr[i]
            p.m, p.n, p.k = m, n, k
            p.lda = p.ldb = p.ldc = 0
            p.A, p.B, p.C = A.data_ptr(), B.data_ptr(), C.data_ptr()
        opts = _abi.BatchedOpts()
        opts.b_layout = lay
        opts.autotune = int(autotune)
        opts.use_graph = int(use_graph)
        opts.alpha, opts.beta = alpha, beta
        handle = ctypes.c_void_p()
        check(lib().kami_batched_plan_create(arr, len(problems), prec, ctypes.byref(opts),
                                             _stream_ptr(stream), ctypes.byref(handle)),
              "kami_batched_plan_create")
        self._h = handle
        self._keep = problems
        self.precision = prec

    def execute(self, stream=None):
        check(lib().kami_batched_execute(self._h, _stream_ptr(stream)), "kami_batched_execute")

    def info(self) -> dict:
        info = _abi.PlanInfo()
        check(lib().kami_batched_plan_info(self._h, ctypes.byref(info)), "kami_batched_plan_info")
        return {"problems": info.problems, "work_items": info.work_items, "buckets": info.buckets,
                "workgroups": info.workgroups, "flops": info.flops,
                "variant_of_bucket": list(info.variant_of_bucket)[: min(16, info.buckets)]}

    def close(self):
        if getattr(self, "_h", None):
            lib().kami_batched_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


__all__ += ["BatchedPlan"]


class BsrMatrix:
    """Block-sparse row matrix on the device: mb x nb blocks of bs x bs (row-major).

    rowptr: int32 (mb + 1,), colidx: int32 (nnzb,) ascending per row, vals: (nnzb, bs, bs).
    """

    def __init__(self, mb, nb, bs, rowptr, colidx, vals):
        self.mb, self.nb, self.bs = int(mb), int(nb), int(bs)
        self.rowptr, self.colidx, self.vals = rowptr, colidx, vals

    @property
    def nnzb(self):
        return int(self.colidx.numel())

    def _key(self):
        """What a successful check and the C descriptor depend on: the arrays' identities
        and storage, and the block sizes (reassigning or reallocating any of them, or
        changing a size, makes a new key)."""
        return (self.mb, self.nb, self.bs, id(self.rowptr), id(self.colidx), id(self.vals),
                self.rowptr.data_ptr(), self.colidx.data_ptr(), self.vals.data_ptr(),
                self.colidx.numel())

    def check(self, name="matrix"):
        """Raise ValueError unless the arrays match the sizes: rowptr int32 (mb + 1,),
        colidx int32 (nnzb,), vals (nnzb, bs, bs), all contiguous on one device.  A
        passing check is remembered for the same arrays."""
        key = self._key()
        if getattr(self, "_checked", None) == key:
            return
        self._check(name)
        self._checked = key

    def _check(self, name):
        import torch
        if self.rowptr.dtype != torch.int32 or tuple(self.rowptr.shape) != (self.mb + 1,):
            raise ValueError(f"{name}: rowptr must be int32 of shape ({self.mb + 1},)")
        if self.colidx.dtype != torch.int32 or self.colidx.dim() != 1:
            raise ValueError(f"{name}: colidx must be a 1-D int32 tensor")
        if self.nnzb and tuple(self.vals.shape) != (self.nnzb, self.bs, self.bs):
            raise ValueError(f"{name}: vals must have shape ({self.nnzb}, {self.bs}, {self.bs})")
        for t in (self.rowptr, self.colidx, self.vals):
            if not t.is_contiguous():
                raise ValueError(f"{name}: rowptr, colidx and vals must be contiguous")
            if t.device != self.rowptr.device:
                raise ValueError(f"{name}: rowptr, colidx and vals must be on one device")
        _on_gpu(self.rowptr, f"{name}: rowptr")

    def _c(self):
        """The C descriptor (kami_bsr), rebuilt only when the arrays change."""
        key = self._key()
        cached = getattr(self, "_desc", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        b = _abi.Bsr()
        b.mb, b.nb, b.bs, b.nnzb = self.mb, self.nb, self.bs, self.nnzb
        b.rowptr = self.rowptr.data_ptr()
        b.colidx = self.colidx.data_ptr() if self.nnzb else None
        b.vals = self.vals.data_ptr() if self.nnzb else None
        self._desc = (key, b)
        return b


def _check_pair(A: BsrMatrix, B: BsrMatrix):
    A.check("A")
    B.check("B")
    if A.nb != B.mb or A.bs != B.bs:
        raise ValueError(f"A ({A.mb}x{A.nb} blocks of {A.bs}) and B ({B.mb}x{B.nb} blocks of "
                         f"{B.bs}) do not multiply")
    if A.rowptr.device != B.rowptr.device:
        raise ValueError("A and B must be on one device")


def spgemm_workspace(A: BsrMatrix, B: BsrMatrix):
    """Device workspace for spgemm_symbolic / spgemm_fill of C = A * B: the rows' column
    bitmaps, which the symbolic call leaves for the fill (uint8 tensor)."""
    import torch
    _check_pair(A, B)
    n = ctypes.c_size_t()
    check(lib().kami_bsr_spgemm_workspace_size(A.mb, B.nb, ctypes.byref(n)),
          "kami_bsr_spgemm_workspace_size")
    return torch.empty(max(int(n.value), 1), dtype=torch.uint8, device=A.rowptr.device)


def _workspace_args(ws, device):
    if ws is None:
        return None, 0
    import torch
    if ws.dtype != torch.uint8 or ws.dim() != 1 or not ws.is_contiguous() or ws.device != device:
        raise ValueError("workspace must be a contiguous 1-D uint8 tensor on the matrices' device")
    return ctypes.c_void_p(ws.data_ptr()), ws.numel()


def spgemm_symbolic(A: BsrMatrix, B: BsrMatrix, stream=None, workspace=None):
    """Row pointer of C = A * B (exclusive scan of per-row block counts) and nnzb(C).
    Synchronises the stream to read nnzb(C) back, as the reference does
    (src/sparse/spgemm/2d/block_spgemm_2d_half_mma.cu:808).  With `workspace`
    (spgemm_workspace) the rows' column bitmaps stay there for spgemm_fill."""
    import torch
    _check_pair(A, B)
    rowptr = torch.empty(A.mb + 1, dtype=torch.int32, device=A.rowptr.device)
    nnz = ctypes.c_int64()
    a, b = A._c(), B._c()
    ws, ws_bytes = _workspace_args(workspace, A.rowptr.device)
    check(lib().kami_bsr_spgemm_symbolic(ctypes.byref(a), ctypes.byref(b),
                                         ctypes.c_void_p(rowptr.data_ptr()), ctypes.byref(nnz),
                                         ws, ws_bytes, _stream_ptr(stream)), "kami_bsr_spgemm_symbolic")
    return rowptr, int(nnz.value)


def spgemm_fill(A: BsrMatrix, B: BsrMatrix, rowptrC, nnzbC, stream=None, workspace=None):
    """Ascending block columns of every row of C (from the bitmaps the symbolic call left
    in `workspace`, when given)."""
    import torch
    _check_pair(A, B)
    if rowptrC.dtype != torch.int32 or tuple(rowptrC.shape) != (A.mb + 1,) or \
            not rowptrC.is_contiguous() or rowptrC.device != A.rowptr.device:
        raise ValueError(f"rowptrC must be a contiguous int32 tensor of shape ({A.mb + 1},)")
    colidx = torch.empty(max(nnzbC, 1), dtype=torch.int32, device=A.rowptr.device)
    a, b = A._c(), B._c()
    ws, ws_bytes = _workspace_args(workspace, A.rowptr.device)
    check(lib().kami_bsr_spgemm_fill(ctypes.byref(a), ctypes.byref(b),
                                     ctypes.c_void_p(rowptrC.data_ptr()),
                                     ctypes.c_void_p(colidx.data_ptr()), ws, ws_bytes,
                                     _stream_ptr(stream)), "kami_bsr_spgemm_fill")
    return colidx[:nnzbC]


def spgemm_numeric(A: BsrMatrix, B: BsrMatrix, rowptrC, colidxC, *, precision=None, alpha=1.0,
                   iters=1, stream=None):
    """Values of C = alpha * iters * A * B on the symbolic structure: (nnzbC, bs, bs),
    float32 for FP16/BF16 inputs, float64 for FP64.  iters > 1 repeats the block products
    inside the kernel (kami_bsr_spgemm_numeric_repeat)."""
    import torch
    _check_pair(A, B)
    prec = _precision_of(A.vals, precision)
    if prec not in (PREC_F64, PREC_F16_F32, PREC_BF16_F32):
        raise ValueError(f"spgemm_numeric: precision {prec} is not supported")
    in_dt = _in_dtype(prec)
    if (A.nnzb and A.vals.dtype != in_dt) or (B.nnzb and B.vals.dtype != in_dt):
        raise ValueError(f"A and B values must be {in_dt}")
    if rowptrC.dtype != torch.int32 or tuple(rowptrC.shape) != (A.mb + 1,) or \
            not rowptrC.is_contiguous() or rowptrC.device != A.rowptr.device:
        raise ValueError(f"rowptrC must be a contiguous int32 tensor of shape ({A.mb + 1},)")
    if colidxC.dtype != torch.int32 or colidxC.dim() != 1 or not colidxC.is_contiguous() or \
            colidxC.device != A.rowptr.device:
        raise ValueError("colidxC must be a contiguous 1-D int32 tensor")
    out_dt = torch.float64 if prec == PREC_F64 else torch.float32
    nnz = int(colidxC.numel())
    vals = torch.empty((max(nnz, 1), A.bs, A.bs), dtype=out_dt, device=A.rowptr.device)
    a, b = A._c(), B._c()
    if iters == 1:
        check(lib().kami_bsr_spgemm_numeric(ctypes.byref(a), ctypes.byref(b),
                                            ctypes.c_void_p(rowptrC.data_ptr()),
                                            ctypes.c_void_p(colidxC.data_ptr() if nnz else 0), nnz,
                                            ctypes.c_void_p(vals.data_ptr()), prec, alpha,
                                            _stream_ptr(stream)), "kami_bsr_spgemm_numeric")
    else:
        check(lib().kami_bsr_spgemm_numeric_repeat(ctypes.byref(a), ctypes.byref(b),
                                                   ctypes.c_void_p(rowptrC.data_ptr()),
                                                   ctypes.c_void_p(colidxC.data_ptr() if nnz else 0),
                                                   nnz, ctypes.c_void_p(vals.data_ptr()), prec,
                                                   alpha, int(iters), _stream_ptr(stream)),
              "kami_bsr_spgemm_numeric_repeat")
    return vals[:nnz]


def bsr_spgemm(A: BsrMatrix, B: BsrMatrix, *, precision=None, alpha=1.0, stream=None):
    """C = alpha * A * B as a BsrMatrix (symbolic, fill and numeric phases)."""
    ws = spgemm_workspace(A, B)
    rowptr, nnz = spgemm_symbolic(A, B, stream, workspace=ws)
    colidx = spgemm_fill(A, B, rowptr, nnz, stream, workspace=ws)
    vals = spgemm_numeric(A, B, rowptr, colidx, precision=precision, alpha=alpha, stream=stream)
    return BsrMatrix(A.mb, B.nb, A.bs, rowptr, colidx, vals)


def bsr_spmm(A: BsrMatrix, B, C=None, *, b_layout="nk", precision=None, alpha=1.0, beta=0.0,
             stream=None):
    """C = alpha * A * op(B) + beta * C with A block-sparse and B dense.

    b_layout "nk": B is (n, K) row-major and C(i, j) = sum_k A(i, k) B[j, k] (the
    reference's layout); "kn": B is (K, n).  K = A.nb * A.bs.  C is (A.mb * A.bs, n),
    float32 for FP16/BF16 inputs, float64 for FP64."""
    import torch
    A.check("A")
    prec = _precision_of(A.vals if A.nnzb else B, precision)
    if prec not in (PREC_F64, PREC_F16_F32, PREC_BF16_F32):
        raise ValueError(f"bsr_spmm: precision {prec} is not supported")
    kn = b_layout == "kn"
    if B.dim() != 2 or B.stride(1) != 1:
        raise ValueError("bsr_spmm expects a 2-D B with unit column stride")
    n = int(B.shape[1] if kn else B.shape[0])
    kdim = A.nb * A.bs
    if int(B.shape[0] if kn else B.shape[1]) != kdim:
        raise ValueError(f"B must have {kdim} {'rows' if kn else 'columns'} (A.nb * A.bs)")
    in_dt = _in_dtype(prec)
    if B.dtype != in_dt or (A.nnzb and A.vals.dtype != in_dt):
        raise ValueError(f"A and B values must be {in_dt}")
    out_dt = torch.float64 if prec == PREC_F64 else torch.float32
    if C is None:
        C = torch.zeros((A.mb * A.bs, n), dtype=out_dt, device=B.device)
    if C.dim() != 2 or tuple(C.shape) != (A.mb * A.bs, n) or C.dtype != out_dt or C.stride(1) != 1:
        raise ValueError(f"C must be {out_dt} of shape {(A.mb * A.bs, n)} with unit column stride")
    if C.device != B.device or A.rowptr.device != B.device:
        raise ValueError("A, B and C must be on one device")
    _on_gpu(B, "B")
    a = A._c()
    check(lib().kami_bsr_spmm(ctypes.byref(a), ctypes.c_void_p(B.data_ptr()), n, B.stride(0),
                              B_KN if kn else B_NK, ctypes.c_void_p(C.data_ptr()), C.stride(0),
                              prec, alpha, beta, _stream_ptr(stream)), "kami_bsr_spmm")
    return C


__all__ += ["BsrMatrix", "spgemm_workspace", "spgemm_symbolic", "spgemm_fill", "spgemm_numeric",
            "bsr_spgemm", "bsr_spmm"]


def tune_save(path):
    """Write the recorded block-GEMM and batched tuning choices to `path`."""
    check(lib().kami_tune_save(os.fsencode(path)), "kami_tune_save")


def tune_load(path):
    """Merge the tuning choices in `path` (all or nothing)."""
    check(lib().kami_tune_load(os.fsencode(path)), "kami_tune_load")


def tune_clear():
    check(lib().kami_tune_clear(), "kami_tune_clear")


__all__ += ["tune_save", "tune_load", "tune_clear"]
