"""ctypes binding of ``include/kami.h`` (the library's C ABI).

Struct layouts mirror the header field for field.  Loading fails loudly when
``libkami.so`` has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkami.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "kami.h")

c_int32, c_int64, c_double, c_void_p, c_size_t = (
    ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t)

# kami_status
OK, EINVAL, ENOTSUP, ENOMEM, EHIP, ENOSPC, ENODEV = range(7)
STATUS_NAMES = {OK: "KAMI_OK", EINVAL: "KAMI_EINVAL", ENOTSUP: "KAMI_ENOTSUP",
                ENOMEM: "KAMI_ENOMEM", EHIP: "KAMI_EHIP", ENOSPC: "KAMI_ENOSPC",
                ENODEV: "KAMI_ENODEV"}

# kami_precision
(PREC_F64, PREC_F16_F32, PREC_BF16_F32, PREC_F64_SPLIT_F16, PREC_F64_SPLIT_BF16,
 PREC_F64_SPLIT_I8) = range(6)
# kami_partition
PART_AUTO, PART_1D, PART_2D, PART_3D = range(4)
# kami_b_layout
B_NK, B_KN = 0, 1
# kami_generator
GEN_RAMP, GEN_RAND_R3 = 0, 1


class GemmDesc(ctypes.Structure):
    _fields_ = [("m", c_int32), ("n", c_int32), ("k", c_int32), ("precision", c_int32),
                ("partition", c_int32), ("ranks", c_int32), ("b_layout", c_int32),
                ("iters", c_int32), ("lda", c_int64), ("ldb", c_int64), ("ldc", c_int64),
                ("alpha", c_double), ("beta", c_double), ("slices", c_int32),
                ("pipeline", c_int32)]


class GemmConfig(ctypes.Structure):
    _fields_ = [("partition", c_int32), ("ranks", c_int32), ("waves_m", c_int32),
                ("waves_n", c_int32), ("waves_k", c_int32), ("threads", c_int32),
                ("k_stage", c_int32), ("lds_bytes", c_int32), ("slices", c_int32),
                ("instance", c_int32), ("pipeline", c_int32), ("prefetch", c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Problem(ctypes.Structure):
    _fields_ = [("m", c_int32), ("n", c_int32), ("k", c_int32), ("reserved", c_int32),
                ("lda", c_int64), ("ldb", c_int64), ("ldc", c_int64),
                ("A", c_void_p), ("B", c_void_p), ("C", c_void_p)]


class BatchedOpts(ctypes.Structure):
    _fields_ = [("b_layout", c_int32), ("autotune", c_int32), ("use_graph", c_int32),
                ("reserved", c_int32), ("alpha", c_double), ("beta", c_double)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("problems", c_int64), ("work_items", c_int64), ("buckets", c_int32),
                ("workgroups", c_int32), ("flops", c_double),
                ("variant_of_bucket", c_int32 * 16)]


class Bsr(ctypes.Structure):
    _fields_ = [("mb", c_int32), ("nb", c_int32), ("bs", c_int32), ("nnzb", c_int32),
                ("rowptr", c_void_p), ("colidx", c_void_p), ("vals", c_void_p)]


_PROTOTYPES = {
    "kami_status_string": (ctypes.c_char_p, [c_int32]),
    "kami_last_hip_error": (c_int32, []),
    "kami_version": (ctypes.c_char_p, []),
    "kami_device_check": (c_int32, []),
    "kami_gemm_desc_init": (c_int32, [ctypes.POINTER(GemmDesc), c_int32, c_int32, c_int32, c_int32]),
    "kami_block_gemm_query": (c_int32, [ctypes.POINTER(GemmDesc), ctypes.POINTER(GemmConfig)]),
    "kami_block_gemm": (c_int32, [ctypes.POINTER(GemmDesc), c_void_p, c_void_p, c_void_p,
                                  c_int64, c_int64, c_int64, c_int64, c_void_p]),
    "kami_split_gemm": (c_int32, [ctypes.POINTER(GemmDesc), c_void_p, c_void_p, c_void_p,
                                  c_int64, c_int64, c_int64, c_int64, c_void_p]),
    "kami_block_gemm_autotune": (c_int32, [ctypes.POINTER(GemmDesc), c_void_p, c_void_p, c_void_p,
                                           c_int64, c_void_p, ctypes.POINTER(GemmConfig)]),
    "kami_batched_plan_create": (c_int32, [ctypes.POINTER(Problem), c_int64, c_int32,
                                           ctypes.POINTER(BatchedOpts), c_void_p,
                                           ctypes.POINTER(c_void_p)]),
    "kami_batched_execute": (c_int32, [c_void_p, c_void_p]),
    "kami_batched_plan_info": (c_int32, [c_void_p, ctypes.POINTER(PlanInfo)]),
    "kami_batched_plan_destroy": (c_int32, [c_void_p]),
    "kami_batched_gemm_static": (c_int32, [ctypes.POINTER(Problem), c_int64, c_int32,
                                           ctypes.POINTER(BatchedOpts), c_void_p]),
    "kami_bsr_spgemm_workspace_size": (c_int32, [c_int32, c_int32, ctypes.POINTER(c_size_t)]),
    "kami_bsr_spgemm_symbolic": (c_int32, [ctypes.POINTER(Bsr), ctypes.POINTER(Bsr), c_void_p,
                                           ctypes.POINTER(c_int64), c_void_p, c_size_t, c_void_p]),
    "kami_bsr_spgemm_fill": (c_int32, [ctypes.POINTER(Bsr), ctypes.POINTER(Bsr), c_void_p,
                                       c_void_p, c_void_p, c_size_t, c_void_p]),
    "kami_bsr_spgemm_numeric": (c_int32, [ctypes.POINTER(Bsr), ctypes.POINTER(Bsr), c_void_p,
                                          c_void_p, c_int64, c_void_p, c_int32, c_double,
                                          c_void_p]),
    "kami_bsr_spgemm_numeric_repeat": (c_int32, [ctypes.POINTER(Bsr), ctypes.POINTER(Bsr),
                                                 c_void_p, c_void_p, c_int64, c_void_p, c_int32,
                                                 c_double, c_int32, c_void_p]),
    "kami_tune_save": (c_int32, [ctypes.c_char_p]),
    "kami_tune_load": (c_int32, [ctypes.c_char_p]),
    "kami_tune_clear": (c_int32, []),
    "kami_bsr_spmm": (c_int32, [ctypes.POINTER(Bsr), c_void_p, c_int32, c_int64, c_int32,
                                c_void_p, c_int64, c_int32, c_double, c_double, c_void_p]),
    "kami_split_slices": (c_int32, [c_void_p, c_int64, c_int64, c_int64, c_int32, c_int32,
                                    c_void_p, c_void_p, c_void_p]),
    "kami_generate": (c_int32, [c_int32, c_int32, c_void_p, c_int64, c_int64, c_void_p]),
}


class KamiError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        name = STATUS_NAMES.get(status, str(status))
        extra = ""
        if status == EHIP and _LIB is not None:
            extra = f" (hipError {_LIB.kami_last_hip_error()})"
        super().__init__(f"{what}: {name}{extra}" if what else f"{name}{extra}")


_LIB = None


def lib():
    """The loaded library; raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C kami_amd/csrc` "
                "(or __graft_entry__.build()); there is no fallback implementation")
        # One HIP runtime per process: torch's bundled libamdhip64.so carries the soname
        # libamdhip64.so.7 that libkami.so needs, so once torch is loaded libkami binds to
        # that same runtime.  Loaded the other way round, torch would load its own copy
        # by file name, and of two runtimes in one process the second to start finds no
        # device (hipErrorNoDevice).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(status: int, what: str = "") -> None:
    if status != OK:
        raise KamiError(status, what)


def header_functions(path: str = HEADER_PATH) -> list[str]:
    """Names of the functions declared in include/kami.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kami_[a-z0-9_]+)\s*\(", text)))
