"""GPU parity of block SpMM (kami_bsr_spmm) against the oracle's restated check loop.

Integer-valued blocks and B (the reference program's rand() % 2 + 1 generator, and small
signed integers) are exact in FP32 and FP64 accumulation, so those results must be
bit-identical; N(0, 1) FP16 values are held to 1e-2 of |A||B| against FP64.
"""
import numpy as np
import pytest

import kami_amd as K
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

PRECS = {"f16": (K.PREC_F16_F32, torch.float16), "bf16": (K.PREC_BF16_F32, torch.bfloat16),
         "f64": (K.PREC_F64, torch.float64)}


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def random_bsr(rng, mb, nb, bs, density, values):
    mask = rng.random((mb, nb)) < density
    rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    ci = np.nonzero(mask)[1].astype(np.int32)
    return rp, ci, values((len(ci), bs, bs))


def to_dev(rp, ci, vals, mb, nb, bs, dt, dev):
    return K.BsrMatrix(mb, nb, bs, torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev),
                       torch.from_numpy(vals).to(dt).to(dev))


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("bs", [16, 32])
@pytest.mark.parametrize("layout", ["nk", "kn"])
def test_reference_generator_exact(prec, bs, layout, dev):
    code, dt = PRECS[prec]
    blocks, n = (16 if bs == 16 else 8), 128
    rp, ci, vals, Bnk = oracle.spmm_generate(blocks, blocks, 0.5, bs, n)
    A = to_dev(rp, ci, vals, blocks, blocks, bs, dt, dev)
    Bin = np.ascontiguousarray(Bnk.T) if layout == "kn" else Bnk
    Bd = torch.from_numpy(Bin).to(dt).to(dev)
    C = K.bsr_spmm(A, Bd, b_layout=layout, precision=code).cpu().double().numpy()
    ref = oracle.bsr_spmm(blocks, blocks, bs, rp, ci, vals, Bin, b_kn=layout == "kn")
    np.testing.assert_array_equal(C, ref)


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("bs", [16, 32])
def test_ragged_shapes_and_empty_rows_exact(prec, bs, dev):
    """n not a multiple of the 64-column tile, block rows without blocks, one block row,
    a matrix with no stored blocks (C = beta C), signed integer values."""
    code, dt = PRECS[prec]
    rng = np.random.default_rng(bs)
    ints = lambda shape: rng.integers(-3, 4, shape).astype(np.float64)  # noqa: E731
    for mb, nb, n, dens in [(7, 5, 80, 0.3), (1, 9, 16, 0.6), (12, 3, 208, 0.0), (5, 5, 48, 1.0)]:
        rp, ci, vals = random_bsr(rng, mb, nb, bs, dens, ints)
        Bnk = ints((n, nb * bs))
        C0 = ints((mb * bs, n))
        A = to_dev(rp, ci, vals, mb, nb, bs, dt, dev)
        out_dt = torch.float64 if prec == "f64" else torch.float32
        Cd = torch.from_numpy(C0).to(out_dt).to(dev)
        K.bsr_spmm(A, torch.from_numpy(Bnk).to(dt).to(dev), Cd, precision=code, alpha=2.0,
                   beta=-1.0)
        ref = 2.0 * oracle.bsr_spmm(mb, nb, bs, rp, ci, vals, Bnk) - C0
        np.testing.assert_array_equal(Cd.cpu().double().numpy(), ref)


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("layout", ["nk", "kn"])
def test_long_block_rows_exact(prec, layout, dev):
    """Block rows of 200, 64, 65 and 129 stored blocks: the kernel reads a row's column
    indices 64 at a time, so these rows cross that chunk boundary 0 to 3 times."""
    code, dt = PRECS[prec]
    rng = np.random.default_rng(77)
    mb, nb, bs, n = 5, 200, 16, 32
    mask = np.zeros((mb, nb), dtype=bool)
    mask[0, :] = True
    for i, cnt in [(1, 64), (2, 65), (4, 129)]:
        mask[i, np.sort(rng.choice(nb, cnt, replace=False))] = True
    rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    ci = np.nonzero(mask)[1].astype(np.int32)
    vals = rng.integers(-2, 3, (len(ci), bs, bs)).astype(np.float64)
    Bnk = rng.integers(-2, 3, (n, nb * bs)).astype(np.float64)
    Bin = np.ascontiguousarray(Bnk.T) if layout == "kn" else Bnk
    A = to_dev(rp, ci, vals, mb, nb, bs, dt, dev)
    C = K.bsr_spmm(A, torch.from_numpy(Bin).to(dt).to(dev), b_layout=layout,
                   precision=code).cpu().double().numpy()
    np.testing.assert_array_equal(C, oracle.bsr_spmm(mb, nb, bs, rp, ci, vals, Bin,
                                                     b_kn=layout == "kn"))


def test_padded_leading_dimensions(dev):
    rng = np.random.default_rng(5)
    mb, nb, bs, n = 6, 4, 32, 96
    rp, ci, vals = random_bsr(rng, mb, nb, bs, 0.5, lambda s: rng.integers(-2, 3, s).astype(np.float64))
    Bbig = rng.integers(-2, 3, (n, nb * bs + 24)).astype(np.float64)
    Bd = torch.from_numpy(Bbig).to(torch.float16).to(dev)[:, : nb * bs]  # ldb = K + 24
    Cbig = torch.zeros((mb * bs, n + 32), dtype=torch.float32, device=dev)
    A = to_dev(rp, ci, vals, mb, nb, bs, torch.float16, dev)
    K.bsr_spmm(A, Bd, Cbig[:, :n], precision=K.PREC_F16_F32)
    ref = oracle.bsr_spmm(mb, nb, bs, rp, ci, vals, np.ascontiguousarray(Bbig[:, : nb * bs]))
    np.testing.assert_array_equal(Cbig[:, :n].cpu().double().numpy(), ref)
    assert torch.all(Cbig[:, n:] == 0)


@pytest.mark.parametrize("prec", ["f16", "bf16"])
def test_normal_values_within_tolerance(prec, dev):
    """N(0, 1) values cast to the input type: |C - C*| <= 1e-2 (|A||B|) against the FP64
    product of the cast values."""
    code, dt = PRECS[prec]
    rng = np.random.default_rng(99)
    mb, nb, bs, n = 64, 64, 32, 256
    rp, ci, vals = random_bsr(rng, mb, nb, bs, 0.05, rng.standard_normal)
    Bnk = rng.standard_normal((n, nb * bs))
    vals_c = torch.from_numpy(vals).to(dt).double().numpy()
    B_c = torch.from_numpy(Bnk).to(dt).double().numpy()
    A = to_dev(rp, ci, vals, mb, nb, bs, dt, dev)
    C = K.bsr_spmm(A, torch.from_numpy(Bnk).to(dt).to(dev), precision=code).cpu().double().numpy()
    ref = oracle.bsr_spmm(mb, nb, bs, rp, ci, vals_c, B_c)
    absref = oracle.bsr_spmm(mb, nb, bs, rp, ci, np.abs(vals_c), np.abs(B_c))
    assert np.all(np.abs(C - ref) <= 1e-2 * absref + 1e-30)


def test_matrix_market_input_through_spgemm_and_spmm(tmp_path, dev):
    """A Matrix Market file read into BSR drives both sparse entry points; integer values
    make the FP64 results exact."""
    import scipy.io
    import scipy.sparse as sp

    from kami_amd import mtx

    rng = np.random.default_rng(11)
    M = sp.random(300, 300, density=0.01, random_state=rng, format="coo",
                  data_rvs=lambda k: rng.integers(-3, 4, k).astype(np.float64))
    f = tmp_path / "m.mtx"
    scipy.io.mmwrite(str(f), M)
    mb, nb, rp, ci, blocks = mtx.load_matrix_market_bsr(f, 16)
    dense = np.zeros((mb * 16, nb * 16))
    dense[:300, :300] = M.toarray()
    A = to_dev(rp, ci, blocks, mb, nb, 16, torch.float64, dev)
    C = K.bsr_spgemm(A, A, precision=K.PREC_F64)
    got = np.zeros_like(dense)
    rpc, cic, vc = C.rowptr.cpu().numpy(), C.colidx.cpu().numpy(), C.vals.cpu().numpy()
    for i in range(mb):
        for q in range(rpc[i], rpc[i + 1]):
            got[i * 16:(i + 1) * 16, cic[q] * 16:(cic[q] + 1) * 16] = vc[q]
    np.testing.assert_array_equal(got, dense @ dense)
    Bnk = rng.integers(-2, 3, (48, nb * 16)).astype(np.float64)
    Cs = K.bsr_spmm(A, torch.from_numpy(Bnk).to(dev), precision=K.PREC_F64).cpu().numpy()
    np.testing.assert_array_equal(Cs, dense @ Bnk.T)
