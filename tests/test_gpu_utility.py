"""GPU parity of the standalone slicing (kami_split_slices) and the device-side
known-answer generators (kami_generate) against the oracle: bit-exact."""
import numpy as np
import pytest

import kami_amd as K
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def _edge_matrix(rng, rows, k):
    X = np.sign(rng.standard_normal((rows, k))) * 10.0 ** rng.uniform(-12, 12, (rows, k))
    if rows > 6:
        X[1] = 0.0
        X[2, :32] = 0.0
        X[3, 5] = -0.0
        X[4, :6] = [2.0 ** -1074, 3 * 2.0 ** -1070, 1e300, -1e-300, 0.5, -2.0 ** 40]
        X[5] = np.round(X[5] * 0) + rng.integers(-5, 6, k)  # small integers
    return X


@pytest.mark.parametrize("prec,beta,bf16", [(K.PREC_F64_SPLIT_F16, 9, False),
                                            (K.PREC_F64_SPLIT_BF16, 8, True)])
@pytest.mark.parametrize("rows,k,pad", [(9, 70, 0), (33, 256, 8), (3, 1000, 24), (128, 128, 0)])
def test_split_slices_bit_exact(prec, beta, bf16, rows, k, pad, dev):
    rng = np.random.default_rng(rows * 1000 + k)
    X = _edge_matrix(rng, rows, k)
    Xd = torch.zeros((rows, k + pad), dtype=torch.float64, device=dev)
    Xd[:, :k] = torch.from_numpy(X).to(dev)
    for slices in (1, 3, 7):
        d, e = K.split_slices(Xd[:, :k], precision=prec, slices=slices)
        rd, re = oracle.split_slices(X, beta, slices, bf16)
        np.testing.assert_array_equal(e.cpu().numpy(), re)
        np.testing.assert_array_equal(d.cpu().numpy(), rd)


def test_split_slices_many_rows(dev):
    """More rows than one grid dimension holds (rows are strided over the grid)."""
    rng = np.random.default_rng(3)
    X = rng.standard_normal((70000, 40))
    d, e = K.split_slices(torch.from_numpy(X).to(dev), slices=2)
    rd, re = oracle.split_slices(X, 9, 2)
    np.testing.assert_array_equal(e.cpu().numpy(), re)
    np.testing.assert_array_equal(d.cpu().numpy(), rd)


@pytest.mark.parametrize("prec", [K.PREC_F64, K.PREC_F16_F32, K.PREC_BF16_F32])
@pytest.mark.parametrize("offset", [0, 12345])
def test_generate_matches_oracle(prec, offset, dev):
    n = 100003
    conv = {K.PREC_F64: lambda v: v,
            K.PREC_F16_F32: lambda v: np.array([oracle.double_to_half(x) for x in v], np.uint16),
            K.PREC_BF16_F32: lambda v: np.array([oracle.double_to_bf16(x) for x in v], np.uint16)}[prec]
    for kind, ref in [(K.GEN_RAMP, oracle.ramp(n + offset)[offset:]),
                      (K.GEN_RAND_R3, oracle.randr_mod3_from(n, offset))]:
        got = K.generate(kind, n, precision=prec, offset=offset, device=dev).cpu()
        if prec == K.PREC_F64:
            np.testing.assert_array_equal(got.numpy(), ref)
        else:
            sample = np.r_[0:2000, n - 2000:n]  # the conversion oracle is per value
            np.testing.assert_array_equal(got.view(torch.int16).numpy().view(np.uint16)[sample],
                                          conv(ref[sample]))


def test_generated_inputs_drive_the_batched_known_answer(dev):
    """rand_r(i) % 3 inputs made on the device give the same block-GEMM result as the
    oracle's host-generated ones (integer values: exact)."""
    m = n = k = 64
    count = 32
    A = K.generate(K.GEN_RAND_R3, count * m * k, device=dev).view(count, m, k)
    B = K.generate(K.GEN_RAND_R3, count * n * k, device=dev).view(count, n, k)
    C = K.block_gemm(A, B, precision=K.PREC_F64).cpu().numpy()
    Ah = oracle.randr_mod3(count * m * k).reshape(count, m, k)
    Bh = oracle.randr_mod3(count * n * k).reshape(count, n, k)
    np.testing.assert_array_equal(C, oracle.gemm_f64(Ah, Bh))
