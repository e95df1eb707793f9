"""Host-side mirror of the reference's solver interface over the gfx950 C ABI.

`sptrsv_syncfree` has the argument list of sptrsv_syncfree_cuda
(SpTRSV_cuda/sptrsv_syncfree_cuda.h:409-421): CSC arrays, m, n, nnz,
substitution, rhs, opt, x (overwritten), b, x_ref; it returns
(status, gflops).  `Plan` is the analyse-once / solve-many form on
device-resident arrays.  Everything runs in libsptrsv_hip.so; nothing here
computes a solve.
"""
import atexit
import ctypes
import sys
import weakref

import numpy as np

from ._lib import (COL_PARTITION, ENGINE_GROUP, ENGINE_LANE, ENGINE_ROW, EXPORTS, F32, F64,
                   REMOTE_DIRECT, REMOTE_RECEIVER,
                   LIB_PATH, NNZ_PARTITION,
                   OPT_WARP_AUTO, OPT_WARP_NNZ, OPT_WARP_RHS, PLAN_MIXED,
                   PLAN_VALIDATE, STATUS, SUBSTITUTION_BACKWARD,
                   SUBSTITUTION_FORWARD, Config, PlanInfo, SptrsvError, check,
                   lib)

__all__ = [
    "COL_PARTITION", "NNZ_PARTITION", "ENGINE_GROUP", "ENGINE_LANE", "ENGINE_ROW", "OPT_WARP_AUTO", "OPT_WARP_NNZ",
    "REMOTE_RECEIVER", "REMOTE_DIRECT",
    "OPT_WARP_RHS", "SUBSTITUTION_BACKWARD", "SUBSTITUTION_FORWARD", "F32",
    "F64", "PLAN_MIXED", "PLAN_VALIDATE", "STATUS", "EXPORTS", "LIB_PATH",
    "Config", "SptrsvError", "lib", "check", "DeviceArray", "Plan", "Graph",
    "Stream", "Event", "sptrsv_syncfree", "sptrsv_syncfree_f32",
    "sptrsv_syncfree_csr", "csr2csc", "analyse", "validate", "exclusive_scan",
    "equal_col_partition", "equal_nnz_partition", "levels", "device_count",
    "device_name",
]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# Live library objects, released at interpreter exit in dependency order
# (graphs, plans, events, streams, buffers) while the HIP runtime is still up;
# finalizers that run later, during interpreter teardown, do nothing.
_LIVE = {k: weakref.WeakSet() for k in ("graph", "plan", "event", "stream", "array")}


@atexit.register
def _release_all():
    for kind in ("graph", "plan", "event", "stream", "array"):
        for obj in list(_LIVE[kind]):
            try:
                obj._release()
            except Exception:
                pass


def _finalize(obj):
    if sys.is_finalizing():
        return
    try:
        obj._release()
    except Exception:
        pass


def device_count():
    n = ctypes.c_int(0)
    st = lib().sptrsv_device_count(ctypes.byref(n))
    return n.value if st == 0 else 0


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().sptrsv_device_name(buf, 256), "sptrsv_device_name")
    return buf.value.decode()


class DeviceArray:
    """A device buffer allocated by the library (hipMalloc)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().sptrsv_malloc(ctypes.byref(p), self.nbytes), "sptrsv_malloc")
        self.ptr = p
        _LIVE["array"].add(self)

    @classmethod
    def from_host(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        if d.nbytes:
            check(lib().sptrsv_memcpy_htod(d.ptr, _ptr(arr), d.nbytes), "htod")
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().sptrsv_memcpy_dtoh(_ptr(out), self.ptr, self.nbytes), "dtoh")
        return out

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes == self.nbytes
        if self.nbytes:
            check(lib().sptrsv_memcpy_htod(self.ptr, _ptr(arr), self.nbytes), "htod")

    def zero(self):
        if self.nbytes:
            check(lib().sptrsv_memset(self.ptr, 0, self.nbytes), "memset")

    def free(self):
        if self.ptr:
            lib().sptrsv_free(self.ptr)
            self.ptr = None

    _release = free

    def __del__(self):
        _finalize(self)


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().sptrsv_stream_create(ctypes.byref(p)), "stream_create")
        self.ptr = p
        _LIVE["stream"].add(self)

    def synchronize(self):
        check(lib().sptrsv_stream_synchronize(self.ptr), "stream_synchronize")

    def _release(self):
        if self.ptr:
            lib().sptrsv_stream_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        _finalize(self)


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        check(lib().sptrsv_event_create(ctypes.byref(p)), "event_create")
        self.ptr = p
        _LIVE["event"].add(self)

    def record(self, stream=None):
        check(lib().sptrsv_event_record(self.ptr, stream.ptr if stream else None), "event_record")

    def elapsed_ms(self, later):
        ms = ctypes.c_float(0.0)
        check(lib().sptrsv_event_elapsed(self.ptr, later.ptr, ctypes.byref(ms)), "event_elapsed")
        return ms.value

    def _release(self):
        if self.ptr:
            lib().sptrsv_event_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        _finalize(self)


def _sp(stream):
    return stream.ptr if stream is not None else None


class Plan:
    """Analyse once, solve many times (device-resident CSC factor)."""

    def __init__(self, colPtr, rowIdx, val, n, nnz, substitution=SUBSTITUTION_FORWARD,
                 flags=PLAN_VALIDATE, stream=None, sys_ptr=None):
        """colPtr/rowIdx/val are DeviceArray (kept alive by the plan)."""
        self._keep = (colPtr, rowIdx, val)
        self.n, self.nnz = n, nnz
        self.dtype = val.dtype
        vt = F64 if val.dtype == np.float64 else F32
        p = ctypes.c_void_p()
        if sys_ptr is None:
            st = lib().sptrsv_plan_create(ctypes.byref(p), n, nnz, colPtr.ptr, rowIdx.ptr,
                                          val.ptr, vt, substitution, flags, _sp(stream))
        else:
            sp = np.ascontiguousarray(sys_ptr, dtype=np.int32)
            self._sys = sp
            st = lib().sptrsv_batch_plan_create(ctypes.byref(p), len(sp) - 1, _ptr(sp), n, nnz,
                                                colPtr.ptr, rowIdx.ptr, val.ptr, vt,
                                                substitution, flags, _sp(stream))
        check(st, "plan_create")
        self.ptr = p
        _LIVE["plan"].add(self)

    @classmethod
    def from_host(cls, colPtr, rowIdx, val, substitution=SUBSTITUTION_FORWARD,
                  flags=PLAN_VALIDATE, sys_ptr=None):
        n = len(colPtr) - 1
        d = (DeviceArray.from_host(np.asarray(colPtr, np.int32)),
             DeviceArray.from_host(np.asarray(rowIdx, np.int32)),
             DeviceArray.from_host(val))
        return cls(*d, n=n, nnz=len(rowIdx), substitution=substitution, flags=flags,
                   sys_ptr=sys_ptr)

    def set_config(self, rhs=1, **kw):
        c = Config(**kw)
        check(lib().sptrsv_plan_set_config(self.ptr, rhs, ctypes.byref(c)), "set_config")

    def get_config(self, rhs=1):
        c = Config()
        check(lib().sptrsv_plan_get_config(self.ptr, rhs, ctypes.byref(c)), "get_config")
        return c.as_dict()

    def solve(self, b, x, rhs=1, stream=None):
        check(lib().sptrsv_plan_solve(self.ptr, b.ptr, x.ptr, rhs, _sp(stream)), "solve")

    def solve_ir(self, b, x, rhs=1, iters=2, stream=None):
        check(lib().sptrsv_plan_solve_ir(self.ptr, b.ptr, x.ptr, rhs, iters, _sp(stream)),
              "solve_ir")

    def autotune(self, rhs=1, mode=0):
        c = Config()
        ms = ctypes.c_float(0.0)
        check(lib().sptrsv_plan_autotune(self.ptr, rhs, mode, ctypes.byref(c), ctypes.byref(ms)),
              "autotune")
        return c.as_dict(), ms.value

    def check(self, stream=None):
        return lib().sptrsv_plan_check(self.ptr, _sp(stream))

    def reset(self, stream=None):
        check(lib().sptrsv_plan_reset(self.ptr, _sp(stream)), "reset")

    def set_spin_cap(self, cap):
        check(lib().sptrsv_plan_set_spin_cap(self.ptr, int(cap)), "set_spin_cap")

    def info(self):
        i = PlanInfo()
        check(lib().sptrsv_plan_info(self.ptr, ctypes.byref(i)), "plan_info")
        return {f: getattr(i, f) for f, _ in i._fields_}

    def prepare(self, rhs=1, mode=0, stream=None):
        """Run the per-configuration analysis now (instead of at the first solve)."""
        check(lib().sptrsv_plan_prepare(self.ptr, rhs, mode, _sp(stream)), "prepare")

    def levels(self, stream=None):
        """(nlevel, min, avg, max) rows per level, computed on the device."""
        out = [ctypes.c_int() for _ in range(4)]
        check(lib().sptrsv_plan_levels(self.ptr, *[ctypes.byref(o) for o in out],
                                       _sp(stream)), "plan_levels")
        return tuple(o.value for o in out)

    def graph(self, b, x, rhs=1, mode=0, iters=0, stream=None):
        return Graph(self, b, x, rhs, mode, iters, stream)

    def destroy(self):
        if getattr(self, "ptr", None):
            check(lib().sptrsv_plan_destroy(self.ptr), "plan_destroy")
            self.ptr = None

    _release = destroy

    def __del__(self):
        _finalize(self)


class Graph:
    def __init__(self, plan, b, x, rhs, mode, iters, stream):
        p = ctypes.c_void_p()
        check(lib().sptrsv_graph_create(ctypes.byref(p), plan.ptr, b.ptr, x.ptr, rhs, mode,
                                        iters, _sp(stream)), "graph_create")
        self.ptr = p
        self._keep = (plan, b, x)
        _LIVE["graph"].add(self)

    def launch(self, stream=None):
        check(lib().sptrsv_graph_launch(self.ptr, _sp(stream)), "graph_launch")

    def destroy(self):
        if getattr(self, "ptr", None):
            lib().sptrsv_graph_destroy(self.ptr)
            self.ptr = None

    _release = destroy

    def __del__(self):
        _finalize(self)


def _dropin(fn, colPtr, rowIdx, val, m, n, nnz, substitution, rhs, opt, x, b, x_ref, vdt):
    colPtr = np.ascontiguousarray(colPtr, np.int32)
    rowIdx = np.ascontiguousarray(rowIdx, np.int32)
    val = np.ascontiguousarray(val, vdt)
    b = np.ascontiguousarray(b, vdt)
    assert x.dtype == vdt and x.flags.c_contiguous
    xr = None if x_ref is None else np.ascontiguousarray(x_ref, vdt)
    g = ctypes.c_double(0.0)
    st = fn(_ptr(colPtr), _ptr(rowIdx), _ptr(val), m, n, nnz, substitution, rhs, opt,
            _ptr(x), _ptr(b), None if xr is None else _ptr(xr), ctypes.byref(g))
    return st, g.value


def sptrsv_syncfree(cscColPtr, cscRowIdx, cscVal, m, n, nnz, substitution, rhs, opt, x, b,
                    x_ref=None):
    """Drop-in for sptrsv_syncfree_cuda (FP64): returns (status, gflops)."""
    return _dropin(lib().sptrsv_syncfree_hip, cscColPtr, cscRowIdx, cscVal, m, n, nnz,
                   substitution, rhs, opt, x, b, x_ref, np.float64)


def sptrsv_syncfree_f32(cscColPtr, cscRowIdx, cscVal, m, n, nnz, substitution, rhs, opt, x,
                        b, x_ref=None):
    """The same entry built for VALUE_TYPE=float."""
    return _dropin(lib().sptrsv_syncfree_hip_f32, cscColPtr, cscRowIdx, cscVal, m, n, nnz,
                   substitution, rhs, opt, x, b, x_ref, np.float32)


def sptrsv_syncfree_csr(csrRowPtr, csrColIdx, csrVal, m, n, nnz, substitution, rhs, opt, x,
                        b, x_ref=None):
    """CSR in (transposed on the device), FP64."""
    return _dropin(lib().sptrsv_syncfree_hip_csr, csrRowPtr, csrColIdx, csrVal, m, n, nnz,
                   substitution, rhs, opt, x, b, x_ref, np.float64)


def set_verbose(v):
    lib().sptrsv_set_verbose(int(v))


def set_bench_repeat(r):
    lib().sptrsv_set_bench_repeat(int(r))


def csr2csc(m, n, rowPtr, colIdx, val=None):
    """Device transpose of host CSR arrays; returns host (colPtr, rowIdx, val)."""
    rowPtr = np.ascontiguousarray(rowPtr, np.int32)
    colIdx = np.ascontiguousarray(colIdx, np.int32)
    nnz = len(colIdx)
    dP, dI = DeviceArray.from_host(rowPtr), DeviceArray.from_host(colIdx)
    oP, oI = DeviceArray((n + 1,), np.int32), DeviceArray((max(nnz, 1),), np.int32)
    if val is not None:
        val = np.ascontiguousarray(val)
        dV = DeviceArray.from_host(val)
        oV = DeviceArray((max(nnz, 1),), val.dtype)
        vt = F64 if val.dtype == np.float64 else F32
        vp, ovp = dV.ptr, oV.ptr
    else:
        vt, vp, ovp = F64, None, None
    check(lib().sptrsv_csr2csc(m, n, nnz, dP.ptr, dI.ptr, vp, vt, oP.ptr, oI.ptr, ovp, None),
          "csr2csc")
    outv = oV.to_host()[:nnz] if val is not None else None
    return oP.to_host(), oI.to_host()[:nnz], outv


def analyse(rowIdx, m):
    rowIdx = np.ascontiguousarray(rowIdx, np.int32)
    dI = DeviceArray.from_host(rowIdx)
    out = DeviceArray((max(m, 1),), np.int32)
    check(lib().sptrsv_analyse(dI.ptr, m, len(rowIdx), out.ptr, None), "analyse")
    lib().sptrsv_synchronize()
    return out.to_host()[:m]


def validate(colPtr, rowIdx, val, substitution=SUBSTITUTION_FORWARD):
    n = len(colPtr) - 1
    dP = DeviceArray.from_host(np.asarray(colPtr, np.int32))
    dI = DeviceArray.from_host(np.asarray(rowIdx, np.int32))
    dV = DeviceArray.from_host(val)
    vt = F64 if np.asarray(val).dtype == np.float64 else F32
    return lib().sptrsv_validate(dP.ptr, dI.ptr, dV.ptr, vt, n, len(rowIdx), substitution, None)


def exclusive_scan(a):
    a = np.ascontiguousarray(a, np.int32)
    d = DeviceArray.from_host(a)
    check(lib().sptrsv_exclusive_scan(d.ptr, len(a), None), "exclusive_scan")
    return d.to_host()


def equal_col_partition(nparts, cols):
    c = np.zeros(nparts, np.int32)
    d = np.zeros(nparts, np.int32)
    check(lib().sptrsv_equal_col_partition(nparts, cols, _ptr(c), _ptr(d)), "col_partition")
    return c, d


def equal_nnz_partition(nparts, colPtr):
    colPtr = np.ascontiguousarray(colPtr, np.int32)
    c = np.zeros(nparts, np.int32)
    d = np.zeros(nparts, np.int32)
    check(lib().sptrsv_equal_nnz_partition(nparts, len(colPtr) - 1, int(colPtr[-1]),
                                           _ptr(colPtr), _ptr(c), _ptr(d)), "nnz_partition")
    return c, d


def levels(colPtr, rowIdx):
    colPtr = np.ascontiguousarray(colPtr, np.int32)
    rowIdx = np.ascontiguousarray(rowIdx, np.int32)
    out = [ctypes.c_int() for _ in range(4)]
    check(lib().sptrsv_levels(_ptr(colPtr), _ptr(rowIdx), len(colPtr) - 1, len(rowIdx),
                              *[ctypes.byref(o) for o in out]), "levels")
    return tuple(o.value for o in out)
