"""The Python wrappers reject operands that do not match each other before any library
call: the C ABI takes raw device pointers, so a mismatched tensor would make a kernel
read or write past it.  CPU tensors only; no kernel is launched."""
import pytest
import torch

import kami_amd as K


def t(*shape, dtype=torch.float64):
    return torch.zeros(shape, dtype=dtype)


@pytest.mark.parametrize("case", [
    "b_count", "b_k", "b_dtype", "c_dtype", "c_shape", "a_dtype_for_prec", "noncontig",
])
def test_block_gemm_rejects_mismatch(case):
    A, B, C = t(4, 64, 64), t(4, 64, 64), t(4, 64, 64)
    kw = {}
    if case == "b_count":
        B = t(1, 64, 64)
    elif case == "b_k":
        B = t(4, 64, 32)
    elif case == "b_dtype":
        B = t(4, 64, 64, dtype=torch.float32)
    elif case == "c_dtype":
        C = t(4, 64, 64, dtype=torch.float32)
    elif case == "c_shape":
        C = t(4, 64, 32)
    elif case == "a_dtype_for_prec":
        kw["precision"] = K.PREC_F16_F32
        C = t(4, 64, 64, dtype=torch.float32)
    elif case == "noncontig":
        B = t(4, 64, 64).transpose(1, 2)
    with pytest.raises(ValueError):
        K.block_gemm(A, B, C, **kw)


def test_block_gemm_kn_layout_shape():
    A = t(2, 64, 32)
    with pytest.raises(ValueError):
        K.block_gemm(A, t(2, 64, 32), t(2, 64, 64), b_layout="kn")  # (count, K, N) expected


def test_split_gemm_rejects_mismatch():
    with pytest.raises(ValueError):
        K.split_gemm(t(2, 128, 128), t(4096, 128, 128))
    with pytest.raises(ValueError):
        K.split_gemm(t(2, 128, 128), t(2, 128, 64))
    with pytest.raises(ValueError):
        K.split_gemm(t(2, 128, 128), t(2, 128, 128), t(2, 128, 128, dtype=torch.float32))


def test_batched_plan_rejects_mismatch():
    good = (t(32, 16), t(48, 16), t(32, 48))
    with pytest.raises(ValueError):
        K.BatchedPlan([good, (t(32, 16), t(48, 32), t(32, 48))])  # B's K
    with pytest.raises(ValueError):
        K.BatchedPlan([good, (t(32, 16), t(48, 16), t(32, 16))])  # C's columns
    with pytest.raises(ValueError):
        K.BatchedPlan([good, (t(32, 16), t(48, 16), t(32, 48, dtype=torch.float32))])
    with pytest.raises(ValueError):
        K.BatchedPlan([(t(32, 16, dtype=torch.float16), t(48, 16, dtype=torch.float16),
                        t(32, 48, dtype=torch.float64))])  # FP16 writes FP32 C


def bsr(mb, nb, bs, nnzb, dtype=torch.float16):
    rp = torch.zeros(mb + 1, dtype=torch.int32)
    rp[1:] = nnzb
    return K.BsrMatrix(mb, nb, bs, rp, torch.zeros(nnzb, dtype=torch.int32),
                       torch.zeros((nnzb, bs, bs), dtype=dtype))


def test_bsr_spmm_rejects_mismatch():
    A = bsr(4, 4, 32, 1)
    with pytest.raises(ValueError):
        K.bsr_spmm(A, t(64, 100, dtype=torch.float16))  # K = 128 columns expected
    with pytest.raises(ValueError):
        K.bsr_spmm(A, t(100, 64, dtype=torch.float16), b_layout="kn")  # K rows
    with pytest.raises(ValueError):
        K.bsr_spmm(A, t(64, 128, dtype=torch.float16), t(128, 64, dtype=torch.float16))
    with pytest.raises(ValueError):
        K.bsr_spmm(A, t(64, 128, dtype=torch.float16), t(64, 64, dtype=torch.float32))
    with pytest.raises(ValueError):
        K.bsr_spmm(A, t(64, 128, dtype=torch.float64))  # values FP16, B FP64


def test_spgemm_rejects_mismatch():
    with pytest.raises(ValueError):
        K.spgemm_symbolic(bsr(4, 4, 32, 1), bsr(5, 4, 32, 1))
    with pytest.raises(ValueError):
        K.spgemm_symbolic(bsr(4, 4, 32, 1), bsr(4, 4, 16, 1))
    bad = bsr(4, 4, 32, 2)
    bad.vals = torch.zeros((1, 32, 32), dtype=torch.float16)
    with pytest.raises(ValueError):
        K.spgemm_symbolic(bad, bsr(4, 4, 32, 1))
    A = bsr(4, 4, 32, 1)
    with pytest.raises(ValueError):
        K.spgemm_fill(A, A, torch.zeros(3, dtype=torch.int32), 1)
    with pytest.raises(ValueError):
        K.spgemm_numeric(A, A, torch.zeros(5, dtype=torch.int32),
                         torch.zeros(1, dtype=torch.int64))
    with pytest.raises(ValueError):
        K.spgemm_numeric(A, bsr(4, 4, 32, 1, dtype=torch.float64),
                         torch.zeros(5, dtype=torch.int32), torch.zeros(1, dtype=torch.int32))


def test_host_tensors_are_refused_before_any_library_call():
    """Operands that match in every other way but live in host memory: the library would
    hand their pointers to a kernel, so the wrappers refuse them."""
    with pytest.raises(ValueError, match="GPU tensor"):
        K.block_gemm(t(2, 64, 64), t(2, 64, 64), t(2, 64, 64))
    with pytest.raises(ValueError, match="GPU tensor"):
        K.split_gemm(t(2, 128, 128), t(2, 128, 128))
    with pytest.raises(ValueError, match="GPU tensor"):
        K.split_slices(t(4, 64))
    with pytest.raises(ValueError, match="GPU tensor"):
        K.generate(K.GEN_RAMP, 16, device="cpu")
    with pytest.raises(ValueError, match="GPU tensor"):
        K.BatchedPlan([(t(32, 16), t(48, 16), t(32, 48))])
