"""GPU parity of the batched (uneven-size, persistent, work-queue) GEMM.

Sizes are M, N, K multiples of 16 in [16, 256] (the uneven-size workload).  Integer
inputs make each problem exact, so every problem is compared bit for bit with the
oracle's triple loop; random inputs use the FP64 forward bound gamma_K |A||B| and the
1e-2 (relative to |A||B|) bound for FP16/BF16 inputs.  hipGraph replays are checked
against the eager launch after eager work and synchronisation between replays.
"""
import numpy as np
import pytest

import kami_amd as K
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[case] {request.node.nodeid}", flush=True)


def uneven_sizes(rng, count, lo=16, hi=256):
    def one():
        return int(np.clip(np.round(np.exp(rng.uniform(np.log(lo), np.log(hi))) / 16) * 16, lo, hi))
    return [(one(), one(), one()) for _ in range(count)]


def make_problems(sizes, dtype, dev, rng, kind="int", b_layout="nk"):
    host, dev_probs = [], []
    out_dtype = torch.float64 if dtype == torch.float64 else torch.float32
    for (m, n, k) in sizes:
        if kind == "int":
            A = rng.integers(-2, 3, (m, k)).astype(np.float64)
            B = rng.integers(-2, 3, (n, k)).astype(np.float64)
        else:
            A = rng.uniform(-1, 1, (m, k)) if dtype == torch.float64 else rng.standard_normal((m, k))
            B = rng.uniform(-1, 1, (n, k)) if dtype == torch.float64 else rng.standard_normal((n, k))
        At = torch.from_numpy(A).to(dev).to(dtype)
        Bt = torch.from_numpy(B).to(dev).to(dtype)
        if b_layout == "kn":
            Bt = Bt.t().contiguous()
        Ct = torch.full((m, n), float("nan"), dtype=out_dtype, device=dev)
        host.append((At.double().cpu().numpy(), Bt.double().cpu().numpy()))
        dev_probs.append((At, Bt, Ct))
    return host, dev_probs


def reference(host, b_layout="nk"):
    out = []
    for A, B in host:
        Bn = B if b_layout == "nk" else np.ascontiguousarray(B.T)
        out.append(oracle.gemm_f64(A[None], Bn[None])[0])
    return out


@pytest.mark.parametrize("dtype", [torch.float64, torch.float16, torch.bfloat16],
                         ids=["f64", "f16", "bf16"])
@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_uneven_integer_exact(dtype, use_graph, dev):
    rng = np.random.default_rng(11)
    sizes = uneven_sizes(rng, 150) + [(16, 16, 16), (256, 256, 256), (16, 256, 16), (256, 16, 256)]
    host, probs = make_problems(sizes, dtype, dev, rng)
    with K.BatchedPlan(probs, use_graph=use_graph) as plan:
        plan.execute()
        torch.cuda.synchronize()
        info = plan.info()
        assert info["problems"] == len(sizes) and info["work_items"] >= len(sizes)
    for (At, Bt, Ct), ref in zip(probs, reference(host)):
        np.testing.assert_array_equal(Ct.double().cpu().numpy(), ref)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float16], ids=["f64", "f16"])
def test_uneven_random_tolerance(dtype, dev):
    rng = np.random.default_rng(12)
    sizes = uneven_sizes(rng, 120)
    host, probs = make_problems(sizes, dtype, dev, rng, kind="random")
    with K.BatchedPlan(probs) as plan:
        plan.execute()
        torch.cuda.synchronize()
    u = 2.0 ** -53
    for (At, Bt, Ct), (A, B), ref in zip(probs, host, reference(host)):
        mag = np.abs(A) @ np.abs(B).T
        err = np.abs(Ct.double().cpu().numpy() - ref)
        if dtype == torch.float64:
            k = A.shape[1]
            assert np.all(err <= k * u / (1 - k * u) * mag * 1.01)
        else:
            assert np.all(err <= 1e-2 * mag)


def test_kn_layout_alpha_beta(dev):
    rng = np.random.default_rng(13)
    sizes = uneven_sizes(rng, 60)
    host, probs = make_problems(sizes, torch.float64, dev, rng, b_layout="kn")
    c0 = []
    for (_, _, Ct) in probs:
        c = torch.from_numpy(rng.integers(-5, 6, tuple(Ct.shape)).astype(np.float64)).to(dev)
        Ct.copy_(c)
        c0.append(c.cpu().numpy())
    with K.BatchedPlan(probs, b_layout="kn", alpha=2.0, beta=-1.0) as plan:
        plan.execute()
        torch.cuda.synchronize()
    for (At, Bt, Ct), ref, c in zip(probs, reference(host, "kn"), c0):
        np.testing.assert_array_equal(Ct.cpu().numpy(), 2.0 * ref - c)


def test_graph_replays_match_eager_with_work_between(dev):
    rng = np.random.default_rng(14)
    sizes = uneven_sizes(rng, 100)
    host, probs = make_problems(sizes, torch.float64, dev, rng, kind="random")
    with K.BatchedPlan(probs, use_graph=False) as eager:
        eager.execute()
        torch.cuda.synchronize()
        expect = [Ct.clone() for (_, _, Ct) in probs]
    with K.BatchedPlan(probs, use_graph=True) as g:
        for rep in range(4):
            for (_, _, Ct) in probs:
                Ct.fill_(float("nan"))              # eager work between replays
            junk = torch.randn(1 << 16, device=dev).sum()  # unrelated eager kernels
            torch.cuda.synchronize()
            g.execute()
            torch.cuda.synchronize()
            for (_, _, Ct), e in zip(probs, expect):
                assert torch.equal(Ct, e), f"replay {rep} differs from the eager result"
            del junk
        # the graph reads live buffers: new inputs give the new product
        probs[0][0].mul_(2.0)
        g.execute()
        torch.cuda.synchronize()
        assert torch.equal(probs[0][2], 2.0 * expect[0])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float16], ids=["f64", "f16"])
def test_autotuned_plan_is_exact_and_reports_buckets(dtype, dev):
    rng = np.random.default_rng(15)
    sizes = uneven_sizes(rng, 200)
    host, probs = make_problems(sizes, dtype, dev, rng)
    with K.BatchedPlan(probs, autotune=True) as plan:
        info = plan.info()
        assert 1 <= info["buckets"] <= 125
        assert all(0 <= v < 6 for v in info["variant_of_bucket"])
        for (_, _, Ct) in probs:
            Ct.fill_(float("nan"))
        plan.execute()
        torch.cuda.synchronize()
    for (At, Bt, Ct), ref in zip(probs, reference(host)):
        np.testing.assert_array_equal(Ct.double().cpu().numpy(), ref)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float16], ids=["f64", "f16"])
@pytest.mark.parametrize("autotune", [False, True], ids=["priors", "tuned"])
def test_short_queue_of_one_stage_tiles(dtype, autotune, dev):
    """A short queue (few tiles per workgroup, so two entries per pop) of mixed K where
    most tiles have a single K stage: the cursor runs past the fetched tiles again right
    after a boundary repair, which must leave it blocked until the next boundary (a
    regression test: the repair used to loop on a blocked cursor)."""
    rng = np.random.default_rng(2024)
    sizes = [(int(rng.choice([16, 32, 48, 64])), int(rng.choice([16, 32, 64])),
              int(rng.choice([16, 16, 16, 16, 32, 64, 256]))) for _ in range(1500)]
    host, probs = make_problems(sizes, dtype, dev, rng)
    with K.BatchedPlan(probs, autotune=autotune, use_graph=False) as plan:
        assert plan.info()["work_items"] < 6 * plan.info()["workgroups"]  # two per pop
        plan.execute()
        torch.cuda.synchronize()
    for (A, B, C), ref in zip(probs, reference(host)):
        np.testing.assert_array_equal(C.double().cpu().numpy(), ref)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float16], ids=["f64", "f16"])
def test_large_and_tiny_problems_in_one_plan(dtype, dev):
    """Problems far beyond 256 (hundreds of K stages, tens of tiles each) next to 16^3
    ones and ragged edges in one plan; integer inputs keep every result exact (FP16:
    |sums| <= 4 K stay exact in FP32)."""
    rng = np.random.default_rng(77)
    sizes = [(1024, 768, 2048), (528, 1040, 1200), (16, 16, 16), (48, 1008, 4096),
             (1008, 16, 512)] + [(16, 16, 16)] * 40 + [(272, 304, 336)] * 3
    host, probs = make_problems(sizes, dtype, dev, rng)
    with K.BatchedPlan(probs, autotune=False, use_graph=True) as plan:
        plan.execute()
        torch.cuda.synchronize()
    for (A, B, C), ref in zip(probs, reference(host)):
        np.testing.assert_array_equal(C.double().cpu().numpy(), ref)


def test_hundred_thousand_tiny_problems(dev):
    """A long queue of the smallest problems (16^3, one tile and one K stage each)."""
    count, s = 100000, 16
    rng = np.random.default_rng(78)
    A = torch.from_numpy(rng.integers(-2, 3, (count, s, s)).astype(np.float64)).to(dev)
    B = torch.from_numpy(rng.integers(-2, 3, (count, s, s)).astype(np.float64)).to(dev)
    C = torch.full((count, s, s), float("nan"), dtype=torch.float64, device=dev)
    probs = [(A[i], B[i], C[i]) for i in range(count)]
    with K.BatchedPlan(probs, precision=K.PREC_F64, autotune=False, use_graph=False) as plan:
        assert plan.info()["work_items"] == count
        plan.execute()
        torch.cuda.synchronize()
    ref = np.einsum("pik,pjk->pij", A.cpu().numpy(), B.cpu().numpy())
    np.testing.assert_array_equal(C.cpu().numpy(), ref)
