"""GPU parity of block SpGEMM: symbolic structure bit-exact, numeric values vs the oracle.

Inputs: the reference's own generator (srand(0), 8 x 8 blocks of 16 x 16, sparsity 0.5,
values rand() % 3 + 1; src/sparse/spgemm/1d/block_spgemm_1d_half_mma.cu:372-397) whose
integer products are exact in FP32, and the large workload (8192 x 8192, 32 x 32 blocks,
2 % block density, N(0,1) values in FP16) at full size.  Edge cases: empty rows, an empty
matrix, rectangular block grids, dense patterns, bs 16 and 32, FP16 / BF16 / FP64 values.
"""
import os

import numpy as np
import pytest

import kami_amd as K
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def to_bsr(mb, nb, bs, rp, ci, vals, dev, dtype=torch.float16):
    rpt = torch.from_numpy(np.ascontiguousarray(rp, np.int32)).to(dev)
    cit = torch.from_numpy(np.ascontiguousarray(ci, np.int32)).to(dev)
    if vals.dtype == np.uint16:
        vt = torch.from_numpy(vals.view(np.int16)).to(dev).view(dtype)
    else:
        vt = torch.from_numpy(np.ascontiguousarray(vals)).to(dev).to(dtype)
    return K.BsrMatrix(mb, nb, bs, rpt, cit, vt)


def random_pattern(rng, mb, nb, p):
    mask = rng.random((mb, nb)) < p
    rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    ci = np.nonzero(mask)[1].astype(np.int32)
    return rp, ci


def check_structure(A, B, rpA, ciA, rpB, ciB):
    """Structure bit-exact against the oracle, without a workspace (the fill rebuilds
    the bitmaps) and with one (the fill expands the bitmaps the symbolic call stored)."""
    rp_ref, nnz_ref = oracle.spgemm_symbolic(A.mb, B.nb, rpA, ciA, rpB, ciB)
    ci_ref = oracle.spgemm_fill(A.mb, B.nb, rpA, ciA, rpB, ciB, rp_ref)
    for ws in (None, K.spgemm_workspace(A, B)):
        rowptr, nnz = K.spgemm_symbolic(A, B, workspace=ws)
        colidx = K.spgemm_fill(A, B, rowptr, nnz, workspace=ws)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(rowptr.cpu().numpy(), rp_ref)
        assert nnz == nnz_ref
        np.testing.assert_array_equal(colidx.cpu().numpy(), ci_ref)
    return rowptr, colidx, rp_ref, nnz


def test_reference_default_input_exact(dev):
    g = np.load(os.path.join(GOLDEN, "spgemm_ref_default.npz"), allow_pickle=False)
    rp, ci, vals = g["rowptr"], g["colidx"], g["vals"]
    A = to_bsr(8, 8, 16, rp, ci, vals, dev)
    rowptr, colidx, rp_ref, nnz = check_structure(A, A, rp, ci, rp, ci)
    np.testing.assert_array_equal(rowptr.cpu().numpy(), g["rowptrC"])
    np.testing.assert_array_equal(colidx.cpu().numpy(), g["colidxC"])
    C = K.spgemm_numeric(A, A, rowptr, colidx).cpu().numpy()
    ref = oracle.spgemm_numeric(8, 8, 16, rp, ci, vals, rp, ci, vals, rp_ref)
    np.testing.assert_array_equal(C.astype(np.float64), ref)


@pytest.mark.parametrize("bs,dt", [(16, torch.float16), (16, torch.bfloat16), (32, torch.float16),
                                   (16, torch.float64), (32, torch.float64)],
                         ids=["f16-16", "bf16-16", "f16-32", "f64-16", "f64-32"])
def test_in_kernel_repeat_is_iters_times_the_product(bs, dt, dev):
    """kami_bsr_spgemm_numeric_repeat (the reference's NUM_ITER in-kernel repeat): on the
    reference generator's integer values every repetition is exact, so C(iters) equals
    iters * C(1) and iters * the oracle."""
    rp, ci, vals, _ = oracle.spmm_generate(8, 8, 0.5, bs, 16)
    A = to_bsr(8, 8, bs, rp, ci, vals, dev, dtype=dt)
    rowptr, colidx, rp_ref, nnz = check_structure(A, A, rp, ci, rp, ci)
    ref = oracle.spgemm_numeric(8, 8, bs, rp, ci, vals, rp, ci, vals, rp_ref)
    one = K.spgemm_numeric(A, A, rowptr, colidx).cpu().double().numpy()
    np.testing.assert_array_equal(one, ref)
    for iters in (2, 7):
        got = K.spgemm_numeric(A, A, rowptr, colidx, iters=iters, alpha=0.5).cpu().double().numpy()
        np.testing.assert_array_equal(got, 0.5 * iters * ref)


def test_large_2pct_32x32_blocks(dev):
    """C = A * A, 8192 x 8192 with 32 x 32 blocks at 2 % block density, N(0,1) FP16."""
    rng = np.random.default_rng(0x4B414D49 + 4)
    mb = 256
    rp, ci = random_pattern(rng, mb, mb, 0.02)
    vals = rng.standard_normal((len(ci), 32, 32)).astype(np.float16)
    A = to_bsr(mb, mb, 32, rp, ci, vals.view(np.uint16), dev)
    rowptr, colidx, rp_ref, nnz = check_structure(A, A, rp, ci, rp, ci)
    C = K.spgemm_numeric(A, A, rowptr, colidx).cpu().numpy().astype(np.float64)
    ref = oracle.spgemm_numeric(mb, mb, 32, rp, ci, vals.view(np.uint16), rp, ci,
                                vals.view(np.uint16), rp_ref)
    absv = np.abs(vals.astype(np.float64))
    mag = oracle.spgemm_numeric(mb, mb, 32, rp, ci, absv, rp, ci, absv, rp_ref)
    err = np.abs(C - ref)
    assert np.all(err <= 1e-2 * mag)
    assert np.max(err) / np.max(np.abs(ref)) < 1e-5
    print(f"  nnzb(A)={len(ci)} nnzb(C)={nnz} nnzCub={oracle.spgemm_nnzcub(ci, rp)}", flush=True)


@pytest.mark.parametrize("bs", [16, 32])
@pytest.mark.parametrize("dt", ["f16", "bf16", "f64"])
def test_rectangular_and_empty_rows(bs, dt, dev):
    rng = np.random.default_rng(bs + len(dt))
    mb, kb, nb = 37, 23, 51
    rpA, ciA = random_pattern(rng, mb, kb, 0.15)
    rpB, ciB = random_pattern(rng, kb, nb, 0.15)
    # force empty rows in A (B has some by chance at this density)
    maskA = np.zeros((mb, kb), bool)
    for i in range(mb):
        maskA[i, ciA[rpA[i]:rpA[i + 1]]] = True
    maskA[[0, 7, 36]] = False
    rpA = np.concatenate([[0], np.cumsum(maskA.sum(1))]).astype(np.int32)
    ciA = np.nonzero(maskA)[1].astype(np.int32)
    tdt = {"f16": torch.float16, "bf16": torch.bfloat16, "f64": torch.float64}[dt]
    vA = rng.integers(-2, 3, (len(ciA), bs, bs)).astype(np.float64)
    vB = rng.integers(-2, 3, (len(ciB), bs, bs)).astype(np.float64)
    A = to_bsr(mb, kb, bs, rpA, ciA, vA, dev, tdt)
    B = to_bsr(kb, nb, bs, rpB, ciB, vB, dev, tdt)
    rowptr, colidx, rp_ref, nnz = check_structure(A, B, rpA, ciA, rpB, ciB)
    C = K.spgemm_numeric(A, B, rowptr, colidx).cpu().numpy().astype(np.float64)
    ref = oracle.spgemm_numeric(mb, nb, bs, rpA, ciA, vA, rpB, ciB, vB, rp_ref)
    np.testing.assert_array_equal(C, ref)  # small integers: exact in every precision


def test_empty_and_dense_patterns(dev):
    # A without blocks: C is empty
    rpA = np.zeros(5, np.int32)
    A = to_bsr(4, 6, 16, rpA, np.zeros(0, np.int32), np.zeros((0, 16, 16)), dev)
    rpB, ciB = random_pattern(np.random.default_rng(1), 6, 9, 0.5)
    B = to_bsr(6, 9, 16, rpB, ciB, np.ones((len(ciB), 16, 16)), dev)
    rowptr, nnz = K.spgemm_symbolic(A, B)
    assert nnz == 0 and np.all(rowptr.cpu().numpy() == 0)
    # fully dense block patterns, one block row with many products
    rng = np.random.default_rng(2)
    rp, ci = random_pattern(rng, 12, 12, 1.0)
    v = rng.integers(-1, 2, (len(ci), 32, 32)).astype(np.float64)
    D = to_bsr(12, 12, 32, rp, ci, v, dev)
    rowptr, colidx, rp_ref, nnz = check_structure(D, D, rp, ci, rp, ci)
    assert nnz == 144
    C = K.spgemm_numeric(D, D, rowptr, colidx, alpha=-2.0).cpu().numpy().astype(np.float64)
    ref = oracle.spgemm_numeric(12, 12, 32, rp, ci, v, rp, ci, v, rp_ref)
    np.testing.assert_array_equal(C, -2.0 * ref)


def test_wide_bitmap_structure(dev):
    """Many block columns (several scan chunks of bitmap words) with sparse rows."""
    rng = np.random.default_rng(3)
    mb, kb, nb = 40, 30, 20000
    rpA, ciA = random_pattern(rng, mb, kb, 0.3)
    rpB, ciB = random_pattern(rng, kb, nb, 0.01)
    A = to_bsr(mb, kb, 16, rpA, ciA, np.ones((len(ciA), 16, 16)), dev)
    B = to_bsr(kb, nb, 16, rpB, ciB, np.ones((len(ciB), 16, 16)), dev)
    check_structure(A, B, rpA, ciA, rpB, ciB)


def test_many_rows_multi_chunk_scan(dev):
    """More than 1024 block rows: the scan crosses chunk boundaries."""
    rng = np.random.default_rng(4)
    mb = 3000
    rpA, ciA = random_pattern(rng, mb, 50, 0.05)
    rpB, ciB = random_pattern(rng, 50, 60, 0.1)
    A = to_bsr(mb, 50, 16, rpA, ciA, np.ones((len(ciA), 16, 16)), dev)
    B = to_bsr(50, 60, 16, rpB, ciB, np.ones((len(ciB), 16, 16)), dev)
    check_structure(A, B, rpA, ciA, rpB, ciB)


@pytest.mark.parametrize("mb", [1, 1023, 1024, 4095, 4096, 5000])
def test_scan_at_chunk_edges_and_repeated_calls(mb, dev):
    """Row counts on both sides of the scan's 1024-entry chunk edges (the scan covers mb
    counts and the total) and a single row; each product symbolically computed ten more
    times: every call gives the oracle's row pointers and total (the pinned total word
    and its sequence number are reused from call to call)."""
    rng = np.random.default_rng(8 + mb)
    rpA, ciA = random_pattern(rng, mb, 40, 0.05)
    rpB, ciB = random_pattern(rng, 40, 64, 0.1)
    A = to_bsr(mb, 40, 16, rpA, ciA, np.ones((len(ciA), 16, 16)), dev)
    B = to_bsr(40, 64, 16, rpB, ciB, np.ones((len(ciB), 16, 16)), dev)
    rowptr, colidx, rp_ref, nnz_ref = check_structure(A, B, rpA, ciA, rpB, ciB)
    for _ in range(10):
        rowptr, nnz = K.spgemm_symbolic(A, B)
        assert nnz == nnz_ref
        np.testing.assert_array_equal(rowptr.cpu().numpy(), rp_ref)


def test_long_rows_cross_the_pair_chunks(dev):
    """Rows of A with more than 256 blocks and rows whose (A block, B block) pairs number
    several thousand: the symbolic kernels take A's entries 256 at a time and spread the
    pairs over the workgroup; rows of every length from empty to dense."""
    rng = np.random.default_rng(6)
    mb, kb, nb = 24, 700, 300
    dens = np.linspace(0.0, 1.0, mb)[:, None]          # row i of A at density i / (mb - 1)
    maskA = rng.random((mb, kb)) < dens
    rpA = np.concatenate([[0], np.cumsum(maskA.sum(1))]).astype(np.int32)
    ciA = np.nonzero(maskA)[1].astype(np.int32)
    rpB, ciB = random_pattern(rng, kb, nb, 0.05)
    A = to_bsr(mb, kb, 16, rpA, ciA, np.ones((len(ciA), 16, 16)), dev)
    B = to_bsr(kb, nb, 16, rpB, ciB, np.ones((len(ciB), 16, 16)), dev)
    assert (rpA[1:] - rpA[:-1]).max() > 600
    check_structure(A, B, rpA, ciA, rpB, ciB)


def test_workspace_smaller_than_the_bitmaps_is_not_used(dev):
    """A workspace below the bitmap size is ignored: the fill rebuilds the bitmaps, and
    a stale larger workspace of another product is never read when too small."""
    rng = np.random.default_rng(7)
    rpA, ciA = random_pattern(rng, 50, 40, 0.2)
    rpB, ciB = random_pattern(rng, 40, 70, 0.2)
    A = to_bsr(50, 40, 16, rpA, ciA, np.ones((len(ciA), 16, 16)), dev)
    B = to_bsr(40, 70, 16, rpB, ciB, np.ones((len(ciB), 16, 16)), dev)
    full = K.spgemm_workspace(A, B)
    small = torch.full((full.numel() - 4,), 0xFF, dtype=torch.uint8, device=dev)
    rowptr, nnz = K.spgemm_symbolic(A, B, workspace=small)
    colidx = K.spgemm_fill(A, B, rowptr, nnz, workspace=small)
    rp_ref, nnz_ref = oracle.spgemm_symbolic(50, 70, rpA, ciA, rpB, ciB)
    assert nnz == nnz_ref
    np.testing.assert_array_equal(colidx.cpu().numpy(),
                                  oracle.spgemm_fill(50, 70, rpA, ciA, rpB, ciB, rp_ref))
    assert torch.all(small == 0xFF)
