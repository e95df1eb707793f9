"""Row engine (pull over the CSR, solutions as their own arrival flags)
against the oracle.

Tolerances as in test_gpu_parity.py: FP64 1e-12 and FP32 2e-5 relative to
the serial oracle, refinement 1e-10 against the FP64 oracle.  The row sums
have a fixed order (per lane in entry order, then a fixed butterfly), so two
solves of the same system are bit-identical, which is checked too.
"""
import ctypes

import numpy as np
import pytest

import sptrsv_amd as S
from sptrsv_amd import matrices as mx
from gpu_util import assert_close, oracle_x, plan_solve, rel_errs, require_gpu

pytestmark = pytest.mark.gpu

ROW = S.ENGINE_ROW
CFGS = [dict(engine=ROW, lanes_per_column=G, block=blk)
        for G, blk in [(4, 256), (16, 1024), (64, 256)]]


def _id(c):
    return "G{lanes_per_column}-B{block}".format(**c)


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


def _factors():
    return {
        "lap2d": mx.laplacian2d_lower(70, 50),
        "lap3d": mx.laplacian3d_lower(12, 10, 9),
        "powerlaw": mx.power_law_lower(20_000, window=2000, seed=1),
        "bidiag": mx.bidiagonal(3000),
        "batch": mx.laplacian2d_batch(16, lo=8, hi=48, seed=2),
    }


@pytest.mark.parametrize("cfg", CFGS, ids=_id)
@pytest.mark.parametrize("name", ["lap2d", "lap3d", "powerlaw", "bidiag", "batch"])
def test_row_engine_forward(name, cfg):
    f = _factors()[name]
    x0 = mx.x_uniform(f.n, seed=3)
    b = f.spmv(x0)
    x, plan = plan_solve(f, b, cfg=cfg)
    assert_close(x, oracle_x(f, b), 1e-12)
    # fixed summation order: a second solve is bit-identical
    x2, _ = plan_solve(f, b, plan=plan)
    assert np.array_equal(x, x2)


@pytest.mark.parametrize("cfg", CFGS, ids=_id)
def test_row_engine_backward(cfg):
    f = mx.upper_from_lower(mx.power_law_lower(8000, window=900, seed=4))
    b = f.spmv(mx.x_uniform(f.n, seed=5))
    x, _ = plan_solve(f, b, subst=1, cfg=cfg)
    assert_close(x, oracle_x(f, b, subst=1), 1e-12)


def test_row_engine_integer_factor_exact():
    # unit diagonal, integer entries, x = 1: every partial sum is an exact
    # integer, so any summation order gives x exactly
    f = mx.laplacian2d_lower(40, 30, diag=1.0)
    b = f.spmv(np.ones(f.n))
    x, _ = plan_solve(f, b, cfg=CFGS[1])
    assert np.array_equal(x, np.ones(f.n))


@pytest.mark.parametrize("rhs", [2, 5])
def test_row_engine_multi_rhs(rhs):
    f = mx.power_law_lower(6000, window=800, seed=6)
    X = mx.x_normal(f.n, rhs, seed=7)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=rhs, cfg=CFGS[0])
    assert_close(x, oracle_x(f, B, rhs=rhs), 1e-12)


def test_row_engine_fp32_plan():
    f = mx.power_law_lower(10_000, window=1000, seed=8)
    b = f.spmv(mx.x_uniform(f.n, seed=9))
    x, _ = plan_solve(f, b, dtype=np.float32, cfg=CFGS[2])
    assert_close(x, oracle_x(f, b, dtype=np.float32), 2e-5)


def test_row_engine_refinement():
    f = mx.power_law_lower(40_000, window=4000, seed=7)
    X = mx.x_uniform(f.n, seed=0)
    B = f.spmv(X)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, flags=S.PLAN_VALIDATE | S.PLAN_MIXED)
    plan.set_config(1, **CFGS[2])
    db = S.DeviceArray.from_host(B)
    dx = S.DeviceArray(B.shape, np.float64)
    x64 = oracle_x(f, B)
    errs = []
    for iters in (0, 2, 3):
        plan.solve_ir(db, dx, iters=iters)
        assert plan.check() == 0
        errs.append(rel_errs(dx.to_host(), x64)[0])
    assert errs[0] > 1e-9
    assert errs[1] <= 1e-10 and errs[2] <= 1e-10, errs


def test_row_engine_graph_replays():
    f = mx.power_law_lower(20_000, window=2000, seed=11)
    b = f.spmv(mx.x_uniform(f.n, seed=12))
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    plan.set_config(1, **CFGS[0])
    db = S.DeviceArray.from_host(b)
    dx = S.DeviceArray(b.shape, np.float64)
    plan.solve(db, dx)
    assert plan.check() == 0
    ref = dx.to_host()
    assert_close(ref, oracle_x(f, b), 1e-12)
    stream = S.Stream()
    g = plan.graph(db, dx, stream=stream)
    for _ in range(3):
        dx.zero()
        g.launch(stream)
        stream.synchronize()
        assert plan.check(stream) == 0
        assert np.array_equal(dx.to_host(), ref)


def test_row_engine_spin_cap_timeout_then_recovery():
    f = mx.bidiagonal(50_000)
    B = f.spmv(mx.x_uniform(f.n, seed=1))
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    db = S.DeviceArray.from_host(B)
    dx = S.DeviceArray(B.shape, np.float64)
    plan.set_config(1, **CFGS[0])
    plan.set_spin_cap(1)  # 1 us: the chain's later rows wait longer
    plan.solve(db, dx)
    assert plan.check() == -9  # SPTRSV_ERR_TIMEOUT; the plan state is reset
    plan.set_spin_cap(0)
    plan.solve(db, dx)
    assert plan.check() == 0
    assert_close(dx.to_host(), oracle_x(f, B), 1e-12)


def test_row_engine_config_checks():
    f = mx.laplacian2d_lower(8, 8)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    for bad in (dict(engine=ROW, lanes_per_column=8), dict(engine=ROW, rhs_per_lane=2),
                dict(engine=ROW, block=100)):
        with pytest.raises(S.SptrsvError):
            plan.set_config(1, **bad)
    plan.set_config(1, engine=ROW)
    c = plan.get_config(1)
    assert c["engine"] == ROW and c["lanes_per_column"] == 64 and c["block"] == 1024


def test_default_engine_follows_entry_distance():
    # most entries of this power-law factor lie >= 8192 columns below the
    # diagonal (far fraction 0.74) with no common reach: the default is the
    # row engine; stencils (fraction 0), chains, and a 27-point grid whose
    # 10000-column planes put 9 of 13 entries beyond 8192 (but which has a
    # period, the plane + a row + 1) keep the lane engine
    for f, engine in ((mx.power_law_lower(150_000, seed=2), ROW),
                      (mx.laplacian3d_lower(40, 40, 20), S.ENGINE_LANE),
                      (mx.bidiagonal(20_000), S.ENGINE_LANE),
                      (mx.stencil27_lower(100, 100, 5), S.ENGINE_LANE)):
        plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
        assert plan.get_config(1)["engine"] == engine, f.name
        b = f.spmv(mx.x_uniform(f.n, seed=4))
        x, _ = plan_solve(f, b, plan=plan)
        assert_close(x, oracle_x(f, b), 1e-12)


# ------------------------------------------------ lanes over right-hand sides
# (lanes per row, lanes over right-hand sides, right-hand sides per lane):
# one pass over the factor for a tile of RL * RPL right-hand sides, whole rows
# of x read and written at once; rhs counts that are not a multiple of the
# tile leave lanes idle in the last tile.
RHS_LAYOUTS = [(8, 8, 1), (16, 16, 1), (32, 32, 1), (64, 64, 1), (64, 64, 2), (64, 64, 4)]


def _rl_id(t):
    return "G{}-RL{}-RPL{}".format(*t)


@pytest.mark.parametrize("layout", RHS_LAYOUTS, ids=_rl_id)
@pytest.mark.parametrize("rhs", [3, 8, 40, 130])
def test_row_engine_rhs_lanes(layout, rhs):
    G, RL, RPL = layout
    f = mx.laplacian3d_lower(14, 12, 10)
    X = mx.x_normal(f.n, rhs, seed=rhs)
    B = f.spmv(X)
    cfg = dict(engine=ROW, lanes_per_column=G, rhs_lanes=RL, rhs_per_lane=RPL, block=256)
    x, plan = plan_solve(f, B, rhs=rhs, cfg=cfg)
    assert_close(x, oracle_x(f, B, rhs=rhs), 1e-12)
    x2, _ = plan_solve(f, B, rhs=rhs, plan=plan)
    assert np.array_equal(x, x2)  # fixed summation order


@pytest.mark.parametrize("layout", [(16, 16, 1), (64, 64, 2)], ids=_rl_id)
def test_row_engine_rhs_lanes_powerlaw_backward_fp32(layout):
    G, RL, RPL = layout
    cfg = dict(engine=ROW, lanes_per_column=G, rhs_lanes=RL, rhs_per_lane=RPL, block=1024)
    f = mx.power_law_lower(12_000, window=1500, seed=13)
    X = mx.x_normal(f.n, 33, seed=14)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=33, cfg=cfg)
    assert_close(x, oracle_x(f, B, rhs=33), 1e-12)
    u = mx.upper_from_lower(f)
    Bu = u.spmv(X)
    x, _ = plan_solve(u, Bu, rhs=33, subst=1, cfg=cfg)
    assert_close(x, oracle_x(u, Bu, rhs=33, subst=1), 1e-12)
    x, _ = plan_solve(f, B, rhs=33, cfg=cfg, dtype=np.float32)
    assert_close(x, oracle_x(f, B, rhs=33, dtype=np.float32), 2e-5)


def test_row_engine_rhs_lanes_refinement():
    f = mx.laplacian3d_lower(24, 20, 18)
    X = mx.x_normal(f.n, 64, seed=15)
    B = f.spmv(X)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, flags=S.PLAN_VALIDATE | S.PLAN_MIXED)
    plan.set_config(64, engine=ROW, lanes_per_column=64, rhs_lanes=64, rhs_per_lane=1,
                    block=512)
    db = S.DeviceArray.from_host(B)
    dx = S.DeviceArray(B.shape, np.float64)
    plan.solve_ir(db, dx, rhs=64, iters=2)
    assert plan.check() == 0
    assert_close(dx.to_host(), oracle_x(f, B, rhs=64), 1e-10)


# ------------------------------------------------ ticket: any grid size works
@pytest.mark.parametrize("cfg", CFGS + [dict(engine=ROW, lanes_per_column=64, rhs_lanes=64,
                                             rhs_per_lane=2, block=256)],
                         ids=lambda c: "G{lanes_per_column}-B{block}".format(**c)
                         + ("-RL" if c.get("rhs_lanes", 1) > 1 else ""))
def test_row_engine_grid_beyond_resident_capacity(cfg):
    # four times as many workgroups as can be resident: items are claimed in
    # solve order through the ticket, so workgroups that start late only take
    # later items; a short spin cap turns any wait on a missing producer into
    # a bounded error instead of a hang
    cus = ctypes.c_int(0)
    S.check(S.lib().sptrsv_device_cu_count(ctypes.byref(cus)), "cu_count")
    per_cu = 2048 // cfg["block"] * 4  # at most 2048 threads per CU
    rhs = 4 if cfg.get("rhs_lanes", 1) == 1 else 128
    f = mx.power_law_lower(30_000, window=3000, seed=16)
    X = mx.x_normal(f.n, rhs, seed=17)
    B = f.spmv(X)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    plan.set_spin_cap(2_000_000)  # 2 s
    x, _ = plan_solve(f, B, rhs=rhs, cfg=dict(cfg, grid=cus.value * per_cu), plan=plan)
    assert_close(x, oracle_x(f, B, rhs=rhs), 1e-12)
