"""C-ABI checks that run without a GPU: exports, validation, instance selection."""
import ctypes

import pytest

import kami_amd as K
from kami_amd import _abi


def test_library_exports_every_header_function():
    names = _abi.header_functions()
    assert len(names) >= 15
    lib = K.lib()
    for n in names:
        assert hasattr(lib, n), f"libkami.so does not export {n}"


def test_status_strings_and_version():
    lib = K.lib()
    for s in range(7):
        assert lib.kami_status_string(s).decode().startswith("KAMI_")
    assert "gfx950" in K.version()


def test_struct_sizes_match_header_layout():
    assert ctypes.sizeof(_abi.GemmDesc) == 80
    assert ctypes.sizeof(_abi.GemmConfig) == 48
    assert ctypes.sizeof(_abi.Problem) == 64
    assert ctypes.sizeof(_abi.BatchedOpts) == 32
    assert ctypes.sizeof(_abi.Bsr) == 40


def _no_device():
    """True where no gfx950 device is current: only then may a test hand a valid request
    with made-up pointers to the library, which stops it at the device check.  With a
    device present the request would launch on those addresses."""
    return K.lib().kami_device_check() == _abi.ENODEV


def _st(d, A=1 << 20, B=1 << 20, C=1 << 20, count=1):
    return K.lib().kami_block_gemm(ctypes.byref(d), ctypes.c_void_p(A), ctypes.c_void_p(B),
                                   ctypes.c_void_p(C), count, 0, 0, 0, None)


@pytest.mark.parametrize("field,value", [("m", 0), ("n", -1), ("k", 0), ("precision", 9),
                                         ("partition", 7), ("ranks", -1), ("iters", 0),
                                         ("b_layout", 3), ("lda", 8), ("ldb", -1)])
def test_invalid_descriptor_is_einval(field, value):
    d = K.make_desc(64, 64, 64, K.PREC_F64)
    setattr(d, field, value)
    assert _st(d) == _abi.EINVAL


def test_invalid_pointers_and_counts():
    d = K.make_desc(64, 64, 64, K.PREC_F64)
    assert K.lib().kami_block_gemm(ctypes.byref(d), None, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                   1, 0, 0, 0, None) == _abi.EINVAL
    assert _st(d, A=(1 << 20) + 8) == _abi.EINVAL          # A not 16-byte aligned
    assert _st(d, count=-1) == _abi.EINVAL
    assert _st(d, count=0) == _abi.OK                       # empty batch is a no-op
    d.lda = 65                                              # 65 * 8 bytes: not a 16B multiple
    assert _st(d) == _abi.EINVAL
    d.lda = 68
    d.ldc = 32                                              # ldc < n
    assert _st(d) == _abi.EINVAL


def test_unsupported_shape_is_enotsup():
    d = K.make_desc(48, 48, 48, K.PREC_F64)
    cfg = _abi.GemmConfig()
    assert K.lib().kami_block_gemm_query(ctypes.byref(d), ctypes.byref(cfg)) == _abi.ENOTSUP
    d = K.make_desc(256, 256, 256, K.PREC_F16_F32, partition="3d", ranks=4)  # 64 waves
    assert K.lib().kami_block_gemm_query(ctypes.byref(d), ctypes.byref(cfg)) == _abi.ENOTSUP


@pytest.mark.parametrize("prec", [K.PREC_F16_F32, K.PREC_BF16_F32])
@pytest.mark.parametrize("shape", [(192, 192, 192), (256, 256, 256), (256, 256, 16), (256, 256, 32)])
def test_3d_at_wide_shapes_resolves_to_2x2x2_waves(prec, shape):
    """3D at M = N >= 192 runs 2 x 2 x 2 waves over column passes of the output."""
    cfg = K.query(*shape, prec, partition="3d")
    assert (cfg["waves_m"], cfg["waves_n"], cfg["waves_k"], cfg["threads"]) == (2, 2, 2, 512)
    assert cfg["partition"] == K.PART_3D and cfg["lds_bytes"] <= 160 * 1024


@pytest.mark.parametrize("prec,s,part,waves", [
    (K.PREC_F64, 64, "1d", (4, 1, 1)), (K.PREC_F64, 64, "2d", (2, 2, 1)),
    (K.PREC_F64, 64, "3d", (2, 2, 2)), (K.PREC_F64, 128, "1d", (8, 1, 1)),
    (K.PREC_F16_F32, 128, "1d", (4, 1, 1)), (K.PREC_F16_F32, 256, "2d", (4, 4, 1)),
    (K.PREC_BF16_F32, 64, "3d", (2, 2, 2)),
])
def test_query_default_ranks_follow_reference_tables(prec, s, part, waves):
    cfg = K.query(s, s, s, prec, partition=part)
    assert (cfg["waves_m"], cfg["waves_n"], cfg["waves_k"]) == waves
    assert cfg["threads"] == 64 * waves[0] * waves[1] * waves[2]
    assert 0 < cfg["lds_bytes"] <= 160 * 1024


def test_every_instance_fits_lds_and_workgroup_limits():
    for prec, sizes in [(K.PREC_F64, [16, 32, 64, 128]),
                        (K.PREC_F16_F32, [32, 64, 128, 256]),
                        (K.PREC_BF16_F32, [32, 64, 128, 256])]:
        for s in sizes:
            for part in ("1d", "2d", "3d"):
                try:
                    cfg = K.query(s, s, s, prec, partition=part)
                except K.KamiError as e:
                    assert e.status == _abi.ENOTSUP
                    continue
                assert cfg["threads"] <= 1024 and cfg["lds_bytes"] <= 160 * 1024


def test_device_check_without_gpu_reports_nodev():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    assert K.lib().kami_device_check() == _abi.ENODEV


def test_split_precision_query_and_enotsup():
    cfg = K.query(128, 128, 128, K.PREC_F64_SPLIT_F16)
    assert cfg["slices"] == 7 and cfg["threads"] == 512 and cfg["lds_bytes"] <= 160 * 1024
    assert K.query(128, 128, 128, K.PREC_F64_SPLIT_BF16)["slices"] == 7
    for kw in [dict(b_layout="kn"), dict(partition="3d"), dict(iters=2), dict(slices=3)]:
        with pytest.raises(K.KamiError) as e:
            K.query(128, 128, 128, K.PREC_F64_SPLIT_F16, **kw)
        assert e.value.status == _abi.ENOTSUP
    # 64^3 with iters > 1 resolves to the resident-repeat form (whole K on chip); the
    # persistent form takes iters = 1 only
    for prec, slices in [(K.PREC_F64_SPLIT_F16, 7), (K.PREC_F64_SPLIT_BF16, 7)]:
        cfg = K.query(64, 64, 64, prec, iters=3)
        assert (cfg["k_stage"], cfg["pipeline"], cfg["prefetch"], cfg["slices"]) == (64, 1, 0, slices)
        assert cfg["threads"] == 512 and cfg["lds_bytes"] <= 160 * 1024
        assert K.query(64, 64, 64, prec, iters=1)["k_stage"] == 32
        with pytest.raises(K.KamiError) as e:
            K.query(64, 64, 64, prec, iters=3, pipeline=2)
        assert e.value.status == _abi.ENOTSUP
        # 32^3 exists only as a resident form
        assert K.query(32, 32, 32, prec, iters=2)["k_stage"] == 32
        with pytest.raises(K.KamiError) as e:
            K.query(32, 32, 32, prec, iters=1)
        assert e.value.status == _abi.ENOTSUP


def test_split_i8_query_and_enotsup():
    """The int8-digit form: 1D (a wave per 16 rows of C), one persistent form for any
    iters, eight digits, whole-K exponents; K x N B, 2D / 3D partitions, other digit
    counts, shapes without an instance and kami_split_gemm are ENOTSUP."""
    for s, waves in [(128, 8), (64, 4)]:
        for iters in (1, 4):
            cfg = K.query(s, s, s, K.PREC_F64_SPLIT_I8, iters=iters)
            assert (cfg["partition"], cfg["ranks"], cfg["slices"], cfg["pipeline"]) == \
                (K.PART_1D, waves, 8, 2)
            assert cfg["threads"] == 64 * waves and cfg["k_stage"] == s
            assert cfg["lds_bytes"] <= 160 * 1024
        assert K.query(s, s, s, K.PREC_F64_SPLIT_I8, partition="1d")["ranks"] == waves
    for kw in [dict(b_layout="kn"), dict(partition="2d"), dict(partition="3d"), dict(slices=7),
               dict(pipeline=1)]:
        with pytest.raises(K.KamiError) as e:
            K.query(128, 128, 128, K.PREC_F64_SPLIT_I8, **kw)
        assert e.value.status == _abi.ENOTSUP
    with pytest.raises(K.KamiError) as e:
        K.query(32, 32, 32, K.PREC_F64_SPLIT_I8)
    assert e.value.status == _abi.ENOTSUP
    d = K.make_desc(256, 256, 128, K.PREC_F64_SPLIT_I8)
    p = ctypes.c_void_p(1 << 20)
    assert K.lib().kami_split_gemm(ctypes.byref(d), p, p, p, 1, 0, 0, 0, None) == _abi.ENOTSUP
    # rows addressed with 32-bit byte offsets: a problem spanning 2^31 bytes is refused
    d = K.make_desc(128, 128, 128, K.PREC_F64_SPLIT_I8, lda=1 << 22)
    assert K.lib().kami_block_gemm(ctypes.byref(d), p, p, p, 1, 0, 0, 0, None) == _abi.ENOTSUP


def test_spgemm_numeric_repeat_rejects_iters_below_one():
    bsr = _abi.Bsr()
    assert K.lib().kami_bsr_spgemm_numeric_repeat(ctypes.byref(bsr), ctypes.byref(bsr), None,
                                                  None, 0, None, K.PREC_F16_F32, 1.0, 0,
                                                  None) == _abi.EINVAL


def test_batched_plan_validation_without_device():
    arr = (_abi.Problem * 2)()
    for p in arr:
        p.m, p.n, p.k = 32, 32, 32
        p.A = p.B = p.C = 1 << 20
    opts = _abi.BatchedOpts()
    opts.alpha = 1.0
    h = ctypes.c_void_p()
    lib = K.lib()
    arr[1].k = 24  # not a multiple of 16
    assert lib.kami_batched_plan_create(arr, 2, K.PREC_F64, ctypes.byref(opts), None,
                                        ctypes.byref(h)) == _abi.EINVAL
    arr[1].k = 32
    arr[0].A = None
    assert lib.kami_batched_plan_create(arr, 2, K.PREC_F64, ctypes.byref(opts), None,
                                        ctypes.byref(h)) == _abi.EINVAL
    arr[0].A = 1 << 20
    opts.autotune, opts.beta = 1, 1.0  # tuning overwrites C: refused with beta != 0
    assert lib.kami_batched_plan_create(arr, 2, K.PREC_F64, ctypes.byref(opts), None,
                                        ctypes.byref(h)) == _abi.EINVAL
    opts.autotune, opts.beta = 0, 0.0
    assert lib.kami_batched_plan_create(arr, 2, K.PREC_F64_SPLIT_F16, ctypes.byref(opts), None,
                                        ctypes.byref(h)) == _abi.ENOTSUP
    assert lib.kami_batched_plan_create(arr, 0, K.PREC_F64, ctypes.byref(opts), None,
                                        ctypes.byref(h)) == _abi.EINVAL
    # A tile's operand pieces are addressed by 32-bit byte offsets: 32 rows of A 2^24
    # doubles apart span 31 * 2^27 bytes > 2^31 - 16.
    arr[0].lda = 1 << 24
    assert lib.kami_batched_plan_create(arr, 2, K.PREC_F64, ctypes.byref(opts), None,
                                        ctypes.byref(h)) == _abi.ENOTSUP
    if _no_device():
        arr[0].lda = 1 << 20  # 31 * 2^23 bytes fits: the request reaches the device check
        assert lib.kami_batched_plan_create(arr, 2, K.PREC_F64, ctypes.byref(opts), None,
                                            ctypes.byref(h)) == _abi.ENODEV
    arr[0].lda = 0
    opts.b_layout = K.B_KN
    arr[1].ldb = 1 << 26  # K x N B: 31 rows of 2^26 halves = 31 * 2^27 bytes
    assert lib.kami_batched_plan_create(arr, 2, K.PREC_F16_F32, ctypes.byref(opts), None,
                                        ctypes.byref(h)) == _abi.ENOTSUP


def test_spmm_validation_without_device():
    lib = K.lib()
    rp = (ctypes.c_int32 * 3)(0, 1, 1)
    ci = (ctypes.c_int32 * 1)(0)
    A = _abi.Bsr()
    A.mb, A.nb, A.bs, A.nnzb = 2, 2, 16, 1
    A.rowptr = ctypes.addressof(rp)
    A.colidx = ctypes.addressof(ci)
    A.vals = 1 << 20
    B, C = 1 << 21, 1 << 22

    def call(n=32, ldb=0, layout=K.B_NK, ldc=0, prec=K.PREC_F16_F32, b=B, c=C):
        return lib.kami_bsr_spmm(ctypes.byref(A), b, n, ldb, layout, c, ldc, prec, 1.0, 0.0, None)

    assert call(n=24) == _abi.EINVAL          # n not a multiple of 16
    assert call(n=0) == _abi.EINVAL
    assert call(b=None) == _abi.EINVAL
    assert call(c=None) == _abi.EINVAL
    assert call(layout=7) == _abi.EINVAL
    assert call(ldb=16) == _abi.EINVAL        # n x K layout needs ldb >= K = 32
    assert call(layout=K.B_KN, ldb=16) == _abi.EINVAL
    assert call(ldc=16) == _abi.EINVAL
    assert call(ldb=33) == _abi.EINVAL        # 66-byte rows break 16-byte alignment
    assert call(prec=K.PREC_F64_SPLIT_F16) == _abi.ENOTSUP
    A.bs = 8
    assert call() == _abi.ENOTSUP
    A.bs = 16
    if _no_device():
        assert call() == _abi.ENODEV          # valid request, no GPU here
    # one workgroup per (block row, 64 columns): 2^24 of them is beyond one dispatch
    A.mb = 1 << 22
    assert call(n=256) == _abi.ENOTSUP        # 2^22 x 4 = 2^24
    if _no_device():
        assert call(n=192) == _abi.ENODEV     # 2^22 x 3 fits
    A.mb = 2


def test_split_gemm_validation_without_device():
    lib = K.lib()

    def call(m, n, k, prec=K.PREC_F64_SPLIT_F16, **kw):
        d = K.make_desc(m, n, k, prec, **kw)
        return lib.kami_split_gemm(ctypes.byref(d), 1 << 20, 1 << 21, 1 << 22, 1, 0, 0, 0, None)

    assert call(256, 256, 256, K.PREC_F64) == _abi.ENOTSUP       # not a split precision
    assert call(96, 128, 64) == _abi.ENOTSUP                      # M not a multiple of 64
    assert call(128, 128, 48) == _abi.ENOTSUP                     # K not a multiple of 32
    assert call(128, 128, 64, b_layout=K.B_KN) == _abi.ENOTSUP
    assert call(128, 128, 64, partition="3d") == _abi.ENOTSUP
    assert call(128, 128, 64, slices=3) == _abi.ENOTSUP           # no such instance
    assert call(0, 128, 64) == _abi.EINVAL
    if _no_device():
        for shape in [(256, 384, 96), (192, 64, 32), (4096, 4096, 4096)]:
            assert call(*shape) == _abi.ENODEV                    # valid; no GPU here


def test_split_slices_and_generate_validation_without_device():
    lib = K.lib()
    X, D, E = 1 << 20, 1 << 21, 1 << 22

    def sl(x=X, rows=4, k=64, ldx=0, prec=K.PREC_F64_SPLIT_F16, slices=6, d=D, e=E):
        return lib.kami_split_slices(x, rows, k, ldx, prec, slices, d, e, None)

    assert sl(x=None) == _abi.EINVAL
    assert sl(d=None) == _abi.EINVAL
    assert sl(rows=-1) == _abi.EINVAL
    assert sl(slices=0) == _abi.EINVAL
    assert sl(slices=9) == _abi.EINVAL
    assert sl(ldx=32) == _abi.EINVAL           # ldx < k
    assert sl(prec=K.PREC_F64) == _abi.ENOTSUP
    assert sl(rows=0) == _abi.OK               # nothing to do
    if _no_device():
        assert sl() == _abi.ENODEV

    def gen(kind=K.GEN_RAMP, prec=K.PREC_F64, dst=X, count=16, offset=0):
        return lib.kami_generate(kind, prec, dst, count, offset, None)

    assert gen(kind=5) == _abi.EINVAL
    assert gen(dst=None) == _abi.EINVAL
    assert gen(count=-1) == _abi.EINVAL
    assert gen(offset=-3) == _abi.EINVAL
    assert gen(dst=X + 4) == _abi.EINVAL       # doubles need 8-byte alignment
    assert gen(prec=K.PREC_F64_SPLIT_F16) == _abi.ENOTSUP
    assert gen(count=0) == _abi.OK
    if _no_device():
        assert gen() == _abi.ENODEV


def test_launch_limits_and_overlapping_outputs_without_device():
    """Requests whose grid would reach 2^32 work-items come back ENOTSUP, and problems
    whose C regions overlap (0 < stride_c < span of one C) EINVAL, before any device
    call."""
    lib = K.lib()
    big = 0x7fffffff
    d = K.make_desc(64, 64, 64, K.PREC_F64, partition="2d", pipeline=1)
    cfg = K.query(64, 64, 64, K.PREC_F64, partition="2d", pipeline=1)
    assert cfg["pipeline"] == 1 and big * cfg["threads"] >= 1 << 32
    p = ctypes.c_void_p(1 << 20)
    assert lib.kami_block_gemm(ctypes.byref(d), p, p, p, big, 0, 0, 0, None) == _abi.ENOTSUP
    ds = K.make_desc(128, 128, 128, K.PREC_F64_SPLIT_F16, pipeline=1)
    assert lib.kami_block_gemm(ctypes.byref(ds), p, p, p, big, 0, 0, 0, None) == _abi.ENOTSUP
    assert lib.kami_block_gemm(ctypes.byref(d), p, p, p, 2, 0, 0, 64 * 64 - 16, None) == _abi.EINVAL
    digits = ctypes.c_void_p(1 << 21)
    exps = ctypes.c_void_p(1 << 22)
    assert lib.kami_split_slices(p, 1, 1 << 32, 0, K.PREC_F64_SPLIT_F16, 6, digits, exps,
                                 None) == _abi.ENOTSUP
