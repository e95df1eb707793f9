"""ctypes access to the CPU oracle (liboracle.so) and, where it was built, the
reference's own serial routines (_ref/libsptrsv_ref_f64.so, _ref/libsptrsv_ref_f32.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = {np.float64: os.path.join(HERE, "_ref", "libsptrsv_ref_f64.so"),
          np.float32: os.path.join(HERE, "_ref", "libsptrsv_ref_f32.so")}

_o = None
_r = {}


def build():
    """Compile liboracle.so (and _ref/ when the reference checkout exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _o = ctypes.CDLL(ORACLE_SO)
        _o.oracle_rel_err_l1.restype = ctypes.c_double
        _o.oracle_rel_err_l1.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong]
    return _o


def ref(dtype=np.float64):
    """The reference's own routines compiled by the Makefile, or None."""
    dtype = np.dtype(dtype).type
    if dtype not in _r:
        path = REF_SO[dtype]
        _r[dtype] = ctypes.CDLL(path) if os.path.exists(path) else None
    return _r[dtype]


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def exclusive_scan(a):
    a = _i32(a).copy()
    lib().oracle_exclusive_scan(_p(a), len(a))
    return a


def transpose(m, n, rowPtr, colIdx, val):
    """Stable CSR -> CSC (tranpose.h:6-43)."""
    val = np.ascontiguousarray(val)
    nnz = len(colIdx)
    outIdx = np.zeros(max(nnz, 1), np.int32)
    outPtr = np.zeros(n + 1, np.int32)
    outVal = np.zeros(max(nnz, 1), val.dtype)
    fn = lib().oracle_transpose_f64 if val.dtype == np.float64 else lib().oracle_transpose_f32
    fn(m, n, nnz, _p(_i32(rowPtr)), _p(_i32(colIdx)), _p(val), _p(outIdx), _p(outPtr), _p(outVal))
    return outPtr, outIdx[:nnz], outVal[:nnz]


def analyser(rowIdx, m):
    out = np.zeros(m, np.int32)
    lib().oracle_analyser(_p(_i32(rowIdx)), m, len(rowIdx), _p(out))
    return out


def solve(colPtr, rowIdx, val, b, substitution=0, rhs=1, in_degree=None):
    """Serial sync-free executor (serialref.h:25-108); returns x like b."""
    val = np.ascontiguousarray(val)
    dt = val.dtype
    n = len(colPtr) - 1
    b = np.ascontiguousarray(b, dt)
    if in_degree is None:
        in_degree = analyser(rowIdx, n)
    x = np.zeros_like(b)
    fn = lib().oracle_executor_f64 if dt == np.float64 else lib().oracle_executor_f32
    st = fn(_p(_i32(colPtr)), _p(_i32(rowIdx)), _p(val), _p(_i32(in_degree)), n, n,
            substitution, rhs, _p(b), _p(x))
    if st != 0:
        raise RuntimeError(f"oracle executor status {st}")
    return x


def spmv(colPtr, rowIdx, val, x, rhs=1):
    n = len(colPtr) - 1
    x = np.ascontiguousarray(x, np.float64)
    b = np.zeros_like(x)
    lib().oracle_spmv_csc(_p(_i32(colPtr)), _p(_i32(rowIdx)), _p(np.ascontiguousarray(val, np.float64)),
                          n, n, rhs, _p(x), _p(b))
    return b


def rel_err_l1(x, x_ref):
    x = np.ascontiguousarray(x, np.float64).ravel()
    x_ref = np.ascontiguousarray(x_ref, np.float64).ravel()
    return lib().oracle_rel_err_l1(_p(x), _p(x_ref), len(x))


def findlevel(colPtr, rowIdx):
    n = len(colPtr) - 1
    out = [ctypes.c_int() for _ in range(4)]
    st = lib().oracle_findlevel_csc(_p(_i32(colPtr)), _p(_i32(rowIdx)), n, n, len(rowIdx),
                                    *[ctypes.byref(o) for o in out])
    if st != 0:
        raise RuntimeError(f"oracle findlevel status {st}")
    return tuple(o.value for o in out)


def equal_col_partition(nparts, cols):
    c = np.zeros(nparts, np.int32)
    d = np.zeros(nparts, np.int32)
    lib().oracle_equal_col_partition(nparts, cols, _p(c), _p(d))
    return c, d


def equal_nnz_partition(nparts, colPtr):
    colPtr = _i32(colPtr)
    c = np.zeros(nparts, np.int32)
    d = np.zeros(nparts, np.int32)
    lib().oracle_equal_nnz_partition(nparts, len(colPtr) - 1, int(colPtr[-1]), _p(colPtr),
                                     _p(c), _p(d))
    return c, d


def refine_f32(colPtr, rowIdx, val, b, iters, substitution=0, rhs=1):
    n = len(colPtr) - 1
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b)
    st = lib().oracle_refine_f32(_p(_i32(colPtr)), _p(_i32(rowIdx)),
                                 _p(np.ascontiguousarray(val, np.float64)), n, substitution,
                                 rhs, _p(b), _p(x), iters)
    if st != 0:
        raise RuntimeError(f"oracle refine status {st}")
    return x


# ------------------------------------------------ the reference's own code --
def ref_solve(colPtr, rowIdx, val, b, substitution=0, rhs=1, in_degree=None):
    val = np.ascontiguousarray(val)
    r = ref(val.dtype)
    if r is None:
        return None
    n = len(colPtr) - 1
    if in_degree is None:
        indeg = np.zeros(n, np.int32)
        r.ref_analyser(_p(_i32(rowIdx)), n, n, len(rowIdx), _p(indeg))
    else:
        indeg = _i32(in_degree)
    b = np.ascontiguousarray(b, val.dtype)
    x = np.zeros_like(b)
    r.ref_executor(_p(_i32(colPtr)), _p(_i32(rowIdx)), _p(val), _p(indeg), n, n, substitution,
                   rhs, _p(b), _p(x))
    return x


def ref_analyser(rowIdx, m):
    r = ref(np.float64)
    if r is None:
        return None
    out = np.zeros(m, np.int32)
    r.ref_analyser(_p(_i32(rowIdx)), m, m, len(rowIdx), _p(out))
    return out


def ref_transpose(m, n, rowPtr, colIdx, val):
    val = np.ascontiguousarray(val)
    r = ref(val.dtype)
    if r is None:
        return None
    nnz = len(colIdx)
    outIdx = np.zeros(max(nnz, 1), np.int32)
    outPtr = np.zeros(n + 1, np.int32)
    outVal = np.zeros(max(nnz, 1), val.dtype)
    r.ref_transpose(m, n, nnz, _p(_i32(rowPtr)), _p(_i32(colIdx)), _p(val), _p(outIdx),
                    _p(outPtr), _p(outVal))
    return outPtr, outIdx[:nnz], outVal[:nnz]


def ref_exclusive_scan(a):
    r = ref(np.float64)
    if r is None:
        return None
    a = _i32(a).copy()
    r.ref_exclusive_scan(_p(a), len(a))
    return a


def ref_findlevel(colPtr, rowIdx, val):
    val = np.ascontiguousarray(val, np.float64)
    r = ref(np.float64)
    if r is None:
        return None
    n = len(colPtr) - 1
    out = [ctypes.c_int() for _ in range(4)]
    st = r.ref_findlevel_csc(_p(_i32(colPtr)), _p(_i32(rowIdx)), _p(val), n, n, len(rowIdx),
                             *[ctypes.byref(o) for o in out])
    if st != 0:
        return None
    return tuple(o.value for o in out)
