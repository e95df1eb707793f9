"""Helpers shared by the GPU tests: run the library on host data and compare
with the oracle."""
import numpy as np
import pytest

import sptrsv_amd as S
import sptrsv_oracle as orc


def require_gpu():
    if S.device_count() < 1:
        pytest.skip("no HIP device")


def plan_solve(f, b, rhs=1, subst=0, cfg=None, dtype=np.float64, plan=None):
    """Solve through the plan API; returns (x, plan)."""
    if plan is None:
        plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val.astype(dtype), substitution=subst,
                                sys_ptr=getattr(f, "sys_ptr", None))
    if cfg:
        plan.set_config(rhs, **cfg)
    db = S.DeviceArray.from_host(np.ascontiguousarray(b, dtype))
    dx = S.DeviceArray(np.shape(b), dtype)
    plan.solve(db, dx, rhs)
    st = plan.check()
    assert st == 0, S.STATUS.get(st, st)
    return dx.to_host(), plan


def oracle_x(f, b, rhs=1, subst=0, dtype=np.float64):
    x = orc.solve(f.colPtr, f.rowIdx, f.val.astype(dtype), np.ascontiguousarray(b, dtype).ravel(),
                  substitution=subst, rhs=rhs)
    return x.reshape(np.shape(b))


def rel_errs(x, ref):
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    l1 = np.abs(x - ref).sum() / max(np.abs(ref).sum(), 1e-300)
    linf = np.abs(x - ref).max() / max(np.abs(ref).max(), 1e-300)
    return l1, linf


def assert_close(x, ref, tol):
    l1, linf = rel_errs(x, ref)
    assert l1 <= tol and linf <= tol, f"rel err L1 {l1:.3e} Linf {linf:.3e} > {tol:.1e}"
