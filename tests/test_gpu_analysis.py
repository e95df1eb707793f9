"""Device level analysis, the drop-in's timing, the tuner's cache key and the
graph scratch rule.

Levels: sptrsv_plan_levels (the device pull sweep over the CSR) must equal
the oracle's findlevel restatement (findlevel.h:71-147) exactly -- the depth
and the min / avg / max level widths -- on factors whose depth has a closed
form (2D grid nx + ny - 1, 3D grid nx + ny + nz - 2, chain n), on random
power-law factors, and on upper factors solved backward.
Drop-in: sptrsv_syncfree_hip's *gflops is 2 nnz / (mean solve time); with the
per-configuration analysis outside the timed loop it must agree within 10%
with 2 nnz / (the plan's mean kernel time) measured with events here.
"""
import os

import numpy as np
import pytest

import sptrsv_amd as S
import sptrsv_oracle as orc
from sptrsv_amd import matrices as mx
from gpu_util import assert_close, oracle_x, plan_solve, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


def _chains_with_links(n, brk, extra, reach, seed):
    """Chains with random breaks (a fraction `brk` of the subdiagonal left
    out) and, for a fraction `extra` of the rows, one more input from up to
    `reach` columns back: runs of pure links broken wherever a column or a
    row has another entry, so the level sweep mixes run units and single
    rows, and the levels restart at the breaks."""
    rng = np.random.default_rng(seed)
    sub = np.flatnonzero(rng.random(n - 1) >= brk)
    rows_x = np.flatnonzero(rng.random(n) < extra)
    rows_x = rows_x[rows_x >= 2]
    src = rows_x - rng.integers(2, reach + 1, size=len(rows_x))
    keep = src >= 0
    rows_x, src = rows_x[keep], src[keep]
    cols = np.concatenate([np.arange(n), sub, src])
    rows = np.concatenate([np.arange(n), sub + 1, rows_x])
    vals = np.concatenate([np.full(n, 2.0), np.full(len(sub), -1.0),
                           np.full(len(src), -0.5)])
    cp, ri, vv = mx._from_columns(n, cols, rows, vals)
    return mx.Factor(cp, ri, vv, name=f"chains{n}_{brk}_{extra}_{reach}")


def _factors():
    return [
        ("lap2d_37x23", mx.laplacian2d_lower(37, 23), 37 + 23 - 1),
        ("lap3d_20x17x9", mx.laplacian3d_lower(20, 17, 9), 20 + 17 + 9 - 2),
        ("bidiag_5000", mx.bidiagonal(5000), 5000),
        ("powerlaw_40000", mx.power_law_lower(40_000, window=4000, seed=3), None),
        ("stencil27_12x10x8", mx.stencil27_lower(12, 10, 8), None),
        ("batch_20", mx.laplacian2d_batch(20, lo=8, hi=40, seed=4), None),
        # chain runs swept as one unit each (runs of 40-50 and of 3-10)
        ("chains_long_runs", _chains_with_links(30_000, 0.01, 0.02, 300, seed=1), None),
        ("chains_short_runs", _chains_with_links(30_000, 0.05, 0.2, 40, seed=2), None),
    ]


@pytest.mark.parametrize("idx", range(8))
@pytest.mark.parametrize("subst", [0, 1])
def test_device_levels_match_findlevel(idx, subst):
    name, f, closed = _factors()[idx]
    if subst:
        f = mx.upper_from_lower(f)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, substitution=subst)
    assert plan.info()["levels"] == 0
    got = plan.levels()
    want = orc.findlevel(f.colPtr, f.rowIdx)
    assert got == tuple(want), name
    if closed is not None:
        assert got[0] == closed
    assert plan.info()["levels"] == got[0]
    # the host entry point agrees too
    assert S.levels(f.colPtr, f.rowIdx) == got


@pytest.mark.parametrize("subst", [0, 1])
def test_device_levels_full_size_chain(subst):
    # the 10^6 bidiagonal chain is one run: its levels take one unit's sweep
    # (0.8 s as 10^6 single-row units), and the row engine, which claims
    # rows in level order, still solves it
    f = mx.bidiagonal(1_000_000)
    if subst:
        f = mx.upper_from_lower(f)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, substitution=subst)
    import time
    t0 = time.time()
    got = plan.levels()
    assert time.time() - t0 < 0.5
    assert got == (1_000_000, 1, 1, 1)
    g = _chains_with_links(200_000, 0.01, 0.02, 300, seed=3)
    plan2 = S.Plan.from_host(g.colPtr, g.rowIdx, g.val)
    assert plan2.levels() == tuple(orc.findlevel(g.colPtr, g.rowIdx))
    b = g.spmv(mx.x_uniform(g.n, seed=2))
    x, _ = plan_solve(g, b, plan=plan2, cfg=dict(engine=S.ENGINE_ROW, lanes_per_column=4,
                                                 rhs_lanes=1, rhs_per_lane=1, block=256,
                                                 window=1))
    assert_close(x, oracle_x(g, b), 1e-12)


def test_device_levels_full_size_powerlaw():
    f = mx.power_law_lower(1_000_000, seed=0)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    assert plan.levels() == tuple(orc.findlevel(f.colPtr, f.rowIdx))


def test_prepare_records_analysis_time():
    f = mx.laplacian3d_lower(40, 40, 40)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    assert plan.info()["prepare_ms"] == 0.0
    plan.prepare(1)
    t = plan.info()["prepare_ms"]
    assert t > 0.0
    b = f.spmv(mx.x_uniform(f.n, seed=1))
    db = S.DeviceArray.from_host(b)
    dx = S.DeviceArray(b.shape, np.float64)
    plan.solve(db, dx)
    assert plan.check() == 0
    assert plan.info()["prepare_ms"] == t  # nothing left to analyse
    assert_close(dx.to_host(), oracle_x(f, b), 1e-12)


def test_dropin_times_executor_launches_only():
    f = mx.laplacian3d_lower(96, 96, 96)
    b = f.spmv(mx.x_uniform(f.n, seed=2))
    x = np.zeros(f.n)
    S.set_bench_repeat(40)
    try:
        st, gflops = S.sptrsv_syncfree(f.colPtr, f.rowIdx, f.val, f.n, f.n, f.nnz, 0, 1,
                                       S.OPT_WARP_AUTO, x, b, None)
    finally:
        S.set_bench_repeat(100)
    assert st == 0
    assert_close(x, oracle_x(f, b), 1e-12)
    # the same solve through a plan, default configuration, events per launch
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    db = S.DeviceArray.from_host(b)
    dx = S.DeviceArray(b.shape, np.float64)
    plan.prepare(1)
    plan.solve(db, dx)
    ms = []
    for _ in range(40):
        e0, e1 = S.Event(), S.Event()
        e0.record()
        plan.solve(db, dx)
        e1.record()
        ms.append(e0.elapsed_ms(e1))
    assert plan.check() == 0
    expect = 2.0 * f.nnz / (1e6 * float(np.mean(ms)))
    assert abs(gflops - expect) <= 0.1 * expect, (gflops, expect)


def test_tune_cache_key_separates_depths(tmp_path):
    # equal n and nnz: a chain (depth n) and a factor whose second half reads
    # only the first half (depth 2)
    n = 6000
    h = n // 2
    chain = mx.bidiagonal(n)
    rows = np.arange(h, n)
    c = np.concatenate([np.arange(n), rows - h, (rows - h + 1) % h])[: 2 * n - 1]
    r = np.concatenate([np.arange(n), rows, rows])[: 2 * n - 1]
    v = np.concatenate([np.full(n, 4.0), np.full(n, -1.0)])[: 2 * n - 1]
    shallow = mx.Factor(*mx._from_columns(n, c, r, v))
    assert shallow.nnz == chain.nnz
    path = tmp_path / "tune.cache"
    os.environ["SPTRSV_TUNE_CACHE"] = str(path)
    try:
        depths = []
        for f in (chain, shallow):
            plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
            plan.autotune(1)
            depths.append(plan.levels()[0])
        keys = [line.split("|")[0] for line in open(path).read().splitlines()]
        assert len(keys) == 2 and keys[0] != keys[1], keys
        assert depths == [n, 2]
    finally:
        del os.environ["SPTRSV_TUNE_CACHE"]


def test_graph_keeps_scratch_in_place():
    # several default windows (5 planes of 1600 columns each), so the solve
    # uses cross-window slots, 8 times as many with 8 right-hand sides
    f = mx.laplacian3d_lower(40, 40, 16)
    B1 = f.spmv(mx.x_uniform(f.n, seed=3))
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    db = S.DeviceArray.from_host(B1)
    dx = S.DeviceArray(B1.shape, np.float64)
    g = plan.graph(db, dx)
    g.launch()
    assert plan.check() == 0
    ref = dx.to_host()
    assert_close(ref, oracle_x(f, B1), 1e-12)
    # 8 right-hand sides need more slot scratch than the graph's solve: the
    # plan refuses to reallocate while the graph exists
    X8 = mx.x_normal(f.n, 8, seed=4)
    B8 = f.spmv(X8)
    db8 = S.DeviceArray.from_host(B8)
    dx8 = S.DeviceArray(B8.shape, np.float64)
    with pytest.raises(S.SptrsvError) as e:
        plan.solve(db8, dx8, rhs=8)
    assert e.value.status == -11  # SPTRSV_ERR_UNSUPPORTED
    with pytest.raises(S.SptrsvError):
        plan.destroy()  # the graph still uses the plan
    for _ in range(3):
        dx.zero()
        g.launch()
        assert plan.check() == 0
        # LDS arrivals are summed in arrival order: replays agree to rounding
        assert_close(dx.to_host(), ref, 1e-12)
    g.destroy()
    plan.solve(db8, dx8, rhs=8)
    assert plan.check() == 0
    assert_close(dx8.to_host(), oracle_x(f, B8, rhs=8), 1e-12)


def test_device_levels_straddling_claims_full_size_batch():
    # grid lines of 700 and random widths straddle the sweep's claims of 16
    # rows; the sweep must still pipeline the lines (4 s when it did not)
    import time
    for f, closed in [(mx.laplacian2d_lower(700, 40), 700 + 40 - 1),
                      (mx.laplacian2d_batch(512, seed=0), None)]:
        plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, sys_ptr=f.sys_ptr)
        t = time.perf_counter()
        got = plan.levels()
        dt = time.perf_counter() - t
        assert got == tuple(orc.findlevel(f.colPtr, f.rowIdx))
        if closed is not None:
            assert got[0] == closed
        assert dt < 2.0, f"level sweep took {dt:.2f} s"
