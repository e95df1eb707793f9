"""ctypes binding of libsptrsv_hip.so (the C ABI declared in include/sptrsv_hip.h).

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsptrsv_hip.so")

# constants (SpTRSV_cuda/common.h:25-33 values)
SUBSTITUTION_FORWARD = 0
SUBSTITUTION_BACKWARD = 1
OPT_WARP_NNZ = 1
OPT_WARP_RHS = 2
OPT_WARP_AUTO = 3
COL_PARTITION = 0
NNZ_PARTITION = 1
F64 = 0
F32 = 1
PLAN_VALIDATE = 1
PLAN_MIXED = 2
ENGINE_GROUP = 0
ENGINE_LANE = 1
ENGINE_ROW = 2
REMOTE_RECEIVER = 1
REMOTE_DIRECT = 2

STATUS = {
    0: "OK",
    -1: "NOT_SQUARE",
    -2: "INVALID_ARG",
    -3: "MISSING_DIAG",
    -4: "ZERO_DIAG",
    -5: "NOT_TRIANGULAR",
    -6: "UNSORTED",
    -7: "INDEX",
    -8: "HIP",
    -9: "TIMEOUT",
    -10: "NO_DEVICE",
    -11: "UNSUPPORTED",
    -12: "NO_MEMORY",
}

# every symbol include/sptrsv_hip.h declares
EXPORTS = [
    "sptrsv_status_string", "sptrsv_version",
    "sptrsv_syncfree_hip", "sptrsv_syncfree_hip_f32", "sptrsv_syncfree_hip_csr",
    "sptrsv_set_verbose", "sptrsv_set_bench_repeat",
    "sptrsv_device_count", "sptrsv_set_device", "sptrsv_device_name",
    "sptrsv_device_cu_count", "sptrsv_malloc", "sptrsv_free",
    "sptrsv_memcpy_htod", "sptrsv_memcpy_dtoh", "sptrsv_memcpy_dtod",
    "sptrsv_memset", "sptrsv_synchronize", "sptrsv_stream_create",
    "sptrsv_stream_destroy", "sptrsv_stream_synchronize", "sptrsv_event_create",
    "sptrsv_event_destroy", "sptrsv_event_record", "sptrsv_event_elapsed",
    "sptrsv_exclusive_scan", "sptrsv_csr2csc", "sptrsv_analyse",
    "sptrsv_validate", "sptrsv_equal_col_partition",
    "sptrsv_equal_nnz_partition", "sptrsv_levels",
    "sptrsv_plan_create", "sptrsv_batch_plan_create", "sptrsv_plan_destroy",
    "sptrsv_plan_set_config", "sptrsv_plan_get_config", "sptrsv_plan_solve",
    "sptrsv_plan_solve_ir", "sptrsv_plan_autotune", "sptrsv_graph_create",
    "sptrsv_graph_launch", "sptrsv_graph_destroy", "sptrsv_plan_check",
    "sptrsv_plan_reset", "sptrsv_plan_set_spin_cap", "sptrsv_plan_info",
    "sptrsv_plan_prepare", "sptrsv_plan_levels", "sptrsv_exclusive_scan_noalloc",
]


class SptrsvError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {STATUS.get(status, status)} ({status})")


class Config(ctypes.Structure):
    _fields_ = [
        ("lanes_per_column", ctypes.c_int),
        ("rhs_lanes", ctypes.c_int),
        ("rhs_per_lane", ctypes.c_int),
        ("block", ctypes.c_int),
        ("window", ctypes.c_int),
        ("partition", ctypes.c_int),
        ("grid", ctypes.c_int),
        ("engine", ctypes.c_int),
        ("remote", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("nnz", ctypes.c_int),
        ("substitution", ctypes.c_int),
        ("value_type", ctypes.c_int),
        ("nsys", ctypes.c_int),
        ("levels", ctypes.c_int),
        ("scratch_bytes", ctypes.c_longlong),
        ("analysis_ms", ctypes.c_double),
        ("prepare_ms", ctypes.c_double),
    ]


_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
IP = ctypes.POINTER(ctypes.c_int)
DP = ctypes.POINTER(ctypes.c_double)
FP = ctypes.POINTER(ctypes.c_float)
SZ = ctypes.c_size_t


def _sig(lib):
    def s(name, argtypes, restype=I):
        f = getattr(lib, name)
        f.argtypes = argtypes
        f.restype = restype

    s("sptrsv_status_string", [I], ctypes.c_char_p)
    s("sptrsv_version", [], ctypes.c_char_p)
    drop = [P, P, P, I, I, I, I, I, I, P, P, P, DP]
    s("sptrsv_syncfree_hip", drop)
    s("sptrsv_syncfree_hip_f32", drop)
    s("sptrsv_syncfree_hip_csr", drop)
    s("sptrsv_set_verbose", [I], None)
    s("sptrsv_set_bench_repeat", [I], None)
    s("sptrsv_device_count", [IP])
    s("sptrsv_set_device", [I])
    s("sptrsv_device_name", [ctypes.c_char_p, I])
    s("sptrsv_device_cu_count", [IP])
    s("sptrsv_malloc", [ctypes.POINTER(P), SZ])
    s("sptrsv_free", [P])
    s("sptrsv_memcpy_htod", [P, P, SZ])
    s("sptrsv_memcpy_dtoh", [P, P, SZ])
    s("sptrsv_memcpy_dtod", [P, P, SZ])
    s("sptrsv_memset", [P, I, SZ])
    s("sptrsv_synchronize", [])
    s("sptrsv_stream_create", [ctypes.POINTER(P)])
    s("sptrsv_stream_destroy", [P])
    s("sptrsv_stream_synchronize", [P])
    s("sptrsv_event_create", [ctypes.POINTER(P)])
    s("sptrsv_event_destroy", [P])
    s("sptrsv_event_record", [P, P])
    s("sptrsv_event_elapsed", [P, P, FP])
    s("sptrsv_exclusive_scan", [P, I, P])
    s("sptrsv_exclusive_scan_noalloc", [P, I, P])
    s("sptrsv_csr2csc", [I, I, I, P, P, P, I, P, P, P, P])
    s("sptrsv_analyse", [P, I, I, P, P])
    s("sptrsv_validate", [P, P, P, I, I, I, I, P])
    s("sptrsv_equal_col_partition", [I, I, P, P])
    s("sptrsv_equal_nnz_partition", [I, I, I, P, P, P])
    s("sptrsv_levels", [P, P, I, I, IP, IP, IP, IP])
    s("sptrsv_plan_create", [ctypes.POINTER(P), I, I, P, P, P, I, I, I, P])
    s("sptrsv_batch_plan_create", [ctypes.POINTER(P), I, P, I, I, P, P, P, I, I, I, P])
    s("sptrsv_plan_destroy", [P])
    s("sptrsv_plan_set_config", [P, I, ctypes.POINTER(Config)])
    s("sptrsv_plan_get_config", [P, I, ctypes.POINTER(Config)])
    s("sptrsv_plan_solve", [P, P, P, I, P])
    s("sptrsv_plan_solve_ir", [P, P, P, I, I, P])
    s("sptrsv_plan_autotune", [P, I, I, ctypes.POINTER(Config), FP])
    s("sptrsv_graph_create", [ctypes.POINTER(P), P, P, P, I, I, I, P])
    s("sptrsv_graph_launch", [P, P])
    s("sptrsv_graph_destroy", [P])
    s("sptrsv_plan_check", [P, P])
    s("sptrsv_plan_reset", [P, P])
    s("sptrsv_plan_set_spin_cap", [P, ctypes.c_longlong])
    s("sptrsv_plan_info", [P, ctypes.POINTER(PlanInfo)])
    s("sptrsv_plan_prepare", [P, I, I, P])
    s("sptrsv_plan_levels", [P, IP, IP, IP, IP, P])


def lib():
    """The loaded library; raises if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python benchmark-sptrsv_amd/build_hip.py` (no fallback exists)")
        l = ctypes.CDLL(LIB_PATH)
        _sig(l)
        _lib = l
    return _lib


def check(status, what=""):
    if status != 0:
        raise SptrsvError(status, what)
    return status
