"""GPU parity of every building block and executor form against the oracle.

Tolerances: integer/index work bit-exact; FP64 solves <= 1e-12 relative (L1
and Linf) against the serial oracle, bit-exact on the integer known-answer
factors; FP32 solves <= 2e-5; FP32 + FP64 refinement <= 1e-10.

On the x = 1 Laplacian known answers the group engine is bit-exact; the lane
engine solves chains inside a 64-column batch by pointer jumping, which
reassociates the chain's affine terms (x_r = r_r/d_r + c_r x_p), so there the
bound is KNOWN_ANSWER_TOL (a few units in the last place), still far inside
the 1e-12 parity bar.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import sptrsv_amd as S
import sptrsv_oracle as orc
from sptrsv_amd import matrices as mx
from gpu_util import assert_close, oracle_x, plan_solve, rel_errs, require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


# ------------------------------------------------------------ building blocks
@pytest.mark.parametrize("length", [1, 2, 2047, 2048, 2049, 100_003, 5_000_001])
def test_exclusive_scan(length):
    a = np.random.default_rng(length).integers(0, 5, size=length).astype(np.int32)
    np.testing.assert_array_equal(S.exclusive_scan(a), orc.exclusive_scan(a))


def _csr(m, n, density, seed):
    A = sp.random(m, n, density=density, random_state=seed, format="csr")
    A.sort_indices()
    return A


@pytest.mark.parametrize("m,n,density", [(1000, 800, 0.01), (20_000, 20_000, 0.0005),
                                         (5, 7, 0.6), (64, 1, 0.9)])
def test_csr2csc_matches_oracle(m, n, density):
    A = _csr(m, n, density, seed=m)
    p, i, v = S.csr2csc(m, n, A.indptr, A.indices, A.data)
    rp, ri, rv = orc.transpose(m, n, A.indptr, A.indices, A.data)
    np.testing.assert_array_equal(p, rp)
    np.testing.assert_array_equal(i, ri)
    np.testing.assert_array_equal(v, rv)


def test_csr2csc_long_columns_all_tiers():
    # column 0: 20000 rows (chunked merge tier), column 1: 3000 rows (LDS
    # tier), the rest short (register tier); rows listed out of order and one
    # duplicate, so only the stable transpose gives the expected order
    m = 20_000
    rows, cols = [], []
    for r in range(m):
        cols_r = [0]
        if r % 7 == 0 or r < 3000:
            cols_r.append(1)
        if r % 5 == 0:
            cols_r += [2 + (r % 13), 2 + (r % 11)]
        rows.append(cols_r[::-1])
    rowPtr = np.r_[0, np.cumsum([len(c) for c in rows])].astype(np.int32)
    colIdx = np.concatenate([np.array(c, np.int32) for c in rows])
    val = np.random.default_rng(0).standard_normal(len(colIdx))
    n = 16
    p, i, v = S.csr2csc(m, n, rowPtr, colIdx, val)
    rp, ri, rv = orc.transpose(m, n, rowPtr, colIdx, val)
    np.testing.assert_array_equal(p, rp)
    np.testing.assert_array_equal(i, ri)
    np.testing.assert_array_equal(v, rv)
    # float values and pattern-only transposes
    p2, i2, v2 = S.csr2csc(m, n, rowPtr, colIdx, val.astype(np.float32))
    np.testing.assert_array_equal(v2, rv.astype(np.float32))
    p3, i3, _ = S.csr2csc(m, n, rowPtr, colIdx)
    np.testing.assert_array_equal(i3, ri)


def test_analyse_matches_oracle():
    f = mx.power_law_lower(50_000, seed=2)
    np.testing.assert_array_equal(S.analyse(f.rowIdx, f.n), orc.analyser(f.rowIdx, f.n))


def test_validate_error_codes():
    f = mx.laplacian2d_lower(8, 8)
    assert S.validate(f.colPtr, f.rowIdx, f.val) == 0
    assert S.validate(f.colPtr, f.rowIdx, f.val, substitution=1) == -5  # not upper
    v = f.val.copy()
    v[f.colPtr[3]] = 0.0
    assert S.validate(f.colPtr, f.rowIdx, v) == -4
    r = f.rowIdx.copy()
    r[f.colPtr[5]] = 6  # diagonal replaced by a later row
    assert S.validate(f.colPtr, r, f.val) in (-3, -6)
    r = f.rowIdx.copy()
    r[f.colPtr[2] + 1], r[f.colPtr[2] + 2] = r[f.colPtr[2] + 2], r[f.colPtr[2] + 1]
    assert S.validate(f.colPtr, r, f.val) == -6
    r = f.rowIdx.copy()
    r[-1] = 64
    assert S.validate(f.colPtr, r, f.val) == -7
    u = mx.upper_from_lower(f)
    assert S.validate(u.colPtr, u.rowIdx, u.val, substitution=1) == 0
    assert S.validate(u.colPtr, u.rowIdx, u.val, substitution=0) == -5


# ------------------------------------------------------------ executor, rhs=1
RHS1_CONFIGS = [None] + [
    dict(lanes_per_column=g, rhs_lanes=1, rhs_per_lane=1, block=blk, window=(blk // g) * mult)
    for g in (1, 2, 4, 8, 16, 32, 64) for blk, mult in ((256, 1), (256, 4), (1024, 1))
] + [dict(lanes_per_column=4, rhs_lanes=1, rhs_per_lane=1, block=256, window=256,
          partition=S.NNZ_PARTITION)] + [
    dict(engine=S.ENGINE_LANE, block=blk, window=w, partition=part)
    for blk, w, part in ((128, 64, 0), (256, 1000, 0), (1024, 4096, 0), (1024, 8192, 0),
                         (512, 2048, S.NNZ_PARTITION))]


def _cfg_id(c):
    if c is None:
        return "default"
    if c.get("engine") == S.ENGINE_LANE:
        return "lane-B{block}-W{window}".format(**c) + ("-nnz" if c.get("partition") else "")
    return "G{lanes_per_column}-B{block}-W{window}".format(**c) + (
        "-nnz" if c.get("partition") else "")


KNOWN_ANSWER_TOL = 1e-14


def _is_lane(cfg):
    return cfg is None or cfg.get("engine") == S.ENGINE_LANE


@pytest.mark.parametrize("cfg", RHS1_CONFIGS, ids=_cfg_id)
def test_laplacian_known_answer(cfg):
    f = mx.laplacian2d_lower(96, 80)
    b = f.spmv(np.ones(f.n))
    x, _ = plan_solve(f, b, cfg=cfg)
    if _is_lane(cfg):  # the default configuration is the lane engine
        assert np.abs(x - 1.0).max() <= KNOWN_ANSWER_TOL
    else:
        assert np.array_equal(x, np.ones(f.n))


@pytest.mark.parametrize("cfg", RHS1_CONFIGS, ids=_cfg_id)
@pytest.mark.parametrize("backward", [0, 1])
def test_integer_unit_lower_bit_exact(cfg, backward):
    f = mx.integer_unit_lower(20_000, density=5, seed=3, backward=bool(backward))
    x_ref = np.random.default_rng(1).integers(1, 11, size=f.n).astype(np.float64)
    b = f.spmv(x_ref)
    x, _ = plan_solve(f, b, subst=backward, cfg=cfg)
    assert np.array_equal(x, x_ref)


@pytest.mark.parametrize("cfg", RHS1_CONFIGS, ids=_cfg_id)
def test_power_law_matches_oracle(cfg):
    f = mx.power_law_lower(30_000, window=5000, seed=4)
    b = f.spmv(mx.x_uniform(f.n, seed=0))
    x, _ = plan_solve(f, b, cfg=cfg)
    assert_close(x, oracle_x(f, b), 1e-12)


def test_backward_on_transposed_laplacian():
    f = mx.upper_from_lower(mx.laplacian3d_lower(20, 17, 9))
    b = f.spmv(mx.x_uniform(f.n, seed=3))
    for cfg in (None, dict(lanes_per_column=1, block=256, window=1024),
                dict(engine=S.ENGINE_LANE, block=512, window=700)):
        x, _ = plan_solve(f, b, subst=1, cfg=cfg)
        assert_close(x, oracle_x(f, b, subst=1), 1e-12)


def test_repeated_solves_restore_state():
    # the executor returns its counters and partial sums to zero itself
    f = mx.power_law_lower(20_000, window=3000, seed=9)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    ref = None
    for k in range(6):
        b = f.spmv(mx.x_uniform(f.n, seed=k))
        cfg = [dict(lanes_per_column=1, block=256), dict(engine=S.ENGINE_LANE, block=256),
               dict(lanes_per_column=64, block=256)][k % 3]
        x, _ = plan_solve(f, b, cfg=cfg, plan=plan)
        assert_close(x, oracle_x(f, b), 1e-12)


@pytest.mark.parametrize("cfg", [None, dict(engine=S.ENGINE_LANE, block=1024, window=4096)],
                         ids=["default", "lane"])
def test_fp32_plan(cfg):
    f = mx.power_law_lower(20_000, window=2000, seed=5)
    b = f.spmv(mx.x_uniform(f.n, seed=1))
    x, _ = plan_solve(f, b, dtype=np.float32, cfg=cfg)
    x_o = oracle_x(f, b, dtype=np.float32)
    assert_close(x, x_o, 2e-5)


def test_lane_engine_long_columns_overflow_registers():
    # columns with many more off-diagonal entries than the register prefetch
    f = mx.power_law_lower(20_000, window=400, seed=12)
    assert np.diff(f.colPtr).max() > 12
    b = f.spmv(mx.x_uniform(f.n, seed=2))
    for w in (64, 2048):
        x, _ = plan_solve(f, b, cfg=dict(engine=S.ENGINE_LANE, block=256, window=w))
        assert_close(x, oracle_x(f, b), 1e-12)


# ------------------------------------------------------------ multi-RHS
RHS_CONFIGS = [None,
               dict(lanes_per_column=8, rhs_lanes=8, rhs_per_lane=1, block=256, window=32),
               dict(lanes_per_column=64, rhs_lanes=64, rhs_per_lane=2, block=256, window=16),
               dict(lanes_per_column=16, rhs_lanes=16, rhs_per_lane=4, block=256, window=64),
               dict(lanes_per_column=64, rhs_lanes=1, rhs_per_lane=4, block=256, window=4),
               dict(lanes_per_column=1, rhs_lanes=1, rhs_per_lane=1, block=256, window=256),
               dict(lanes_per_column=4, rhs_lanes=1, rhs_per_lane=4, block=256, window=64),
               dict(engine=S.ENGINE_LANE, block=256, window=512),
               dict(engine=S.ENGINE_LANE, rhs_per_lane=2, block=512, window=1024),
               dict(engine=S.ENGINE_LANE, rhs_per_lane=4, block=512, window=2048),
               dict(engine=S.ENGINE_LANE, rhs_per_lane=8, block=256, window=300)]


def _rhs_id(c):
    if c is None:
        return "default"
    if c.get("engine") == S.ENGINE_LANE:
        return "lane-R{r}-B{block}-W{window}".format(r=c.get("rhs_per_lane", 1), **c)
    return "G{lanes_per_column}-RL{rhs_lanes}-r{rhs_per_lane}".format(**c)


@pytest.mark.parametrize("rhs", [2, 8, 33, 130])
@pytest.mark.parametrize("cfg", RHS_CONFIGS, ids=_rhs_id)
def test_multi_rhs_matches_oracle(rhs, cfg):
    f = mx.laplacian3d_lower(16, 12, 10)
    X = mx.x_normal(f.n, rhs, seed=1)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=rhs, cfg=cfg)
    assert_close(x, oracle_x(f, B, rhs=rhs), 1e-12)


@pytest.mark.parametrize("R", [2, 4, 8])
def test_multi_rhs_lane_forest_backward(R):
    # power-law batches are not chains: pointer jumping with R values per lane
    f = mx.upper_from_lower(mx.power_law_lower(6000, window=700, seed=4))
    X = mx.x_normal(f.n, 7, seed=5)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=7, subst=1,
                      cfg=dict(engine=S.ENGINE_LANE, rhs_per_lane=R, block=512, window=1000))
    assert_close(x, oracle_x(f, B, rhs=7, subst=1), 1e-12)


def test_multi_rhs_integer_bit_exact_backward():
    f = mx.integer_unit_lower(5000, density=3, seed=8, backward=True)
    X = np.random.default_rng(2).integers(1, 11, size=(f.n, 5)).astype(np.float64)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=5, subst=1)
    assert np.array_equal(x, X)


# ------------------------------------------------------------ drop-in ABI
@pytest.mark.parametrize("opt", [S.OPT_WARP_NNZ, S.OPT_WARP_RHS, S.OPT_WARP_AUTO])
@pytest.mark.parametrize("rhs", [1, 4])
@pytest.mark.parametrize("backward", [0, 1])
def test_dropin_matches_oracle(opt, rhs, backward):
    S.set_bench_repeat(3)
    base = mx.power_law_lower(8000, window=1000, seed=6)
    f = mx.upper_from_lower(base) if backward else base
    X = mx.x_normal(f.n, rhs, seed=2) if rhs > 1 else mx.x_uniform(f.n, seed=2)
    B = f.spmv(X)
    x = np.zeros_like(B)
    st, gflops = S.sptrsv_syncfree(f.colPtr, f.rowIdx, f.val, f.n, f.n, f.nnz, backward, rhs,
                                   opt, x, B, X)
    assert st == 0 and gflops > 0
    assert_close(x, oracle_x(f, B, rhs=rhs, subst=backward), 1e-12)


def test_dropin_f32_and_csr():
    S.set_bench_repeat(2)
    f = mx.laplacian2d_lower(64, 64)
    X = mx.x_uniform(f.n, seed=3)
    B = f.spmv(X)
    x32 = np.zeros(f.n, np.float32)
    st, _ = S.sptrsv_syncfree_f32(f.colPtr, f.rowIdx, f.val.astype(np.float32), f.n, f.n,
                                  f.nnz, 0, 1, 3, x32, B.astype(np.float32))
    assert st == 0
    assert_close(x32, oracle_x(f, B, dtype=np.float32), 2e-5)
    rowPtr, colIdx, v = mx.csc_to_csr(f)
    x = np.zeros(f.n)
    st, _ = S.sptrsv_syncfree_csr(rowPtr, colIdx, v, f.n, f.n, f.nnz, 0, 1, 3, x, B)
    assert st == 0
    assert_close(x, oracle_x(f, B), 1e-12)


def test_dropin_rejects_malformed_input():
    S.set_bench_repeat(1)
    f = mx.laplacian2d_lower(8, 8)
    B = f.spmv(np.ones(f.n))
    x = np.zeros(f.n)
    v = f.val.copy()
    v[f.colPtr[10]] = 0.0
    assert S.sptrsv_syncfree(f.colPtr, f.rowIdx, v, 64, 64, f.nnz, 0, 1, 3, x, B)[0] == -4
    # a lower factor handed to the backward solver
    assert S.sptrsv_syncfree(f.colPtr, f.rowIdx, f.val, 64, 64, f.nnz, 1, 1, 3, x, B)[0] == -5
    assert S.sptrsv_syncfree(f.colPtr, f.rowIdx, f.val, 64, 64, f.nnz, 0, 0, 3, x, B)[0] == -2


# ------------------------------------------------------------ refinement
@pytest.mark.parametrize("rhs", [1, 8])
def test_fp32_refinement_reaches_fp64_accuracy(rhs):
    f = mx.power_law_lower(40_000, window=4000, seed=7)
    X = mx.x_normal(f.n, rhs, seed=1) if rhs > 1 else mx.x_uniform(f.n, seed=0)
    B = f.spmv(X)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, flags=S.PLAN_VALIDATE | S.PLAN_MIXED)
    if rhs == 1:
        plan.set_config(1, engine=S.ENGINE_LANE, block=512, window=2048)
    db = S.DeviceArray.from_host(B)
    dx = S.DeviceArray(B.shape, np.float64)
    x64 = oracle_x(f, B, rhs=rhs)
    errs = []
    for iters in (0, 1, 2, 3):
        plan.solve_ir(db, dx, rhs=rhs, iters=iters)
        assert plan.check() == 0
        errs.append(rel_errs(dx.to_host(), x64)[0])
    assert errs[0] > 1e-9
    assert errs[2] <= 1e-10 and errs[3] <= 1e-10, errs


# ------------------------------------------------------------ batches, graphs
def test_batch_of_uneven_systems():
    f = mx.laplacian2d_batch(24, lo=8, hi=64, seed=3)
    X = mx.x_uniform(f.n, seed=5)
    B = f.spmv(X)
    x, plan = plan_solve(f, B)
    assert plan.info()["nsys"] == 24
    assert_close(x, oracle_x(f, B), 1e-12)


def test_batch_rejects_coupled_systems():
    f = mx.laplacian2d_lower(10, 10)
    with pytest.raises(S.SptrsvError):
        S.Plan.from_host(f.colPtr, f.rowIdx, f.val, sys_ptr=[0, 50, 100])


@pytest.mark.parametrize("mode", [0, 1])
def test_graph_replays_match_eager(mode):
    f = mx.laplacian2d_batch(12, lo=16, hi=96, seed=1)
    B = f.spmv(mx.x_uniform(f.n, seed=2))
    flags = S.PLAN_VALIDATE | (S.PLAN_MIXED if mode else 0)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, flags=flags, sys_ptr=f.sys_ptr)
    db = S.DeviceArray.from_host(B)
    dx_e = S.DeviceArray(B.shape, np.float64)
    if mode:
        plan.solve_ir(db, dx_e, iters=2)
    else:
        plan.solve(db, dx_e)
    assert plan.check() == 0
    eager = dx_e.to_host()
    dx = S.DeviceArray(B.shape, np.float64)
    g = plan.graph(db, dx, mode=mode, iters=2)
    for rep in range(4):
        dx.zero()
        g.launch()
        assert plan.check() == 0
        assert_close(dx.to_host(), eager, 1e-12)
        # eager work between replays
        plan.solve(db, dx_e) if not mode else plan.solve_ir(db, dx_e, iters=2)
        assert plan.check() == 0
    tol = 1e-12 if mode == 0 else 1e-10
    assert_close(eager, oracle_x(f, B), tol)


# ------------------------------------------------------------ failure handling
def test_spin_cap_timeout_then_recovery():
    f = mx.bidiagonal(200_000)
    B = f.spmv(mx.x_uniform(f.n, seed=1))
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    db = S.DeviceArray.from_host(B)
    dx = S.DeviceArray(B.shape, np.float64)
    for cfg in (dict(lanes_per_column=1, block=256, window=256),
                dict(engine=S.ENGINE_LANE, block=256, window=256)):
        plan.set_config(1, **cfg)
        plan.set_spin_cap(1)  # 1 us: the chain's later columns wait longer
        plan.solve(db, dx)
        assert plan.check() == -9  # SPTRSV_ERR_TIMEOUT; the plan state is reset
        plan.set_spin_cap(0)
        plan.solve(db, dx)
        assert plan.check() == 0
        assert_close(dx.to_host(), oracle_x(f, B), 1e-12)


def test_autotune_picks_a_working_config():
    f = mx.power_law_lower(20_000, window=2000, seed=1)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    cfg, ms = plan.autotune(rhs=1)
    assert ms > 0
    assert plan.get_config(1) == cfg
    B = f.spmv(mx.x_uniform(f.n, seed=3))
    x, _ = plan_solve(f, B, plan=plan)
    assert_close(x, oracle_x(f, B), 1e-12)


# ------------------------------------------------------------ reference inputs
@pytest.mark.parametrize("backward", [0, 1])
def test_matrix_market_factor(tmp_path, backward):
    import scipy.io
    from sptrsv_amd import mtx
    rng = np.random.default_rng(7)
    A = sp.random(3000, 3000, density=0.002, random_state=rng, format="coo")
    A = sp.tril(A + A.T + sp.eye(3000)).tocoo()
    path = str(tmp_path / "a.mtx")
    scipy.io.mmwrite(path, A, symmetry="symmetric")
    f = mtx.triangle_from_mtx(path, substitution=backward, seed=1)
    X = rng.integers(1, 4, size=f.n).astype(np.float64)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, subst=backward)
    assert_close(x, oracle_x(f, B, subst=backward), 1e-12)


# The FP32 solves of refinement read b in FP64 themselves (lane engine, row
# engine with lanes over right-hand sides) and the last one writes x = x_old
# + d beside d (row engine with lanes over right-hand sides, one-tile lane
# solves); the other layouts run the conversion and update passes.  Every
# layout, with 0, 1 (the update in place: x_old is x itself) and 2 steps,
# against the FP64 oracle.
IR_LAYOUTS = [
    ("lane-one-tile", 1, dict(engine=S.ENGINE_LANE, block=512)),
    ("lane-three-tiles", 3, dict(engine=S.ENGINE_LANE, block=512, rhs_per_lane=1)),
    ("row-rhs-lanes", 16, dict(engine=S.ENGINE_ROW, lanes_per_column=8, rhs_lanes=8,
                              rhs_per_lane=2, block=1024)),
    ("row-entry-lanes", 1, dict(engine=S.ENGINE_ROW, lanes_per_column=64, block=1024)),
    ("group", 1, dict(engine=S.ENGINE_GROUP, lanes_per_column=64, block=256, window=256)),
]


@pytest.mark.parametrize("name,rhs,cfg", IR_LAYOUTS, ids=[t[0] for t in IR_LAYOUTS])
def test_refinement_fused_and_pass_layouts(name, rhs, cfg):
    f = mx.laplacian3d_lower(24, 20, 16)
    X = mx.x_normal(f.n, rhs, seed=31) if rhs > 1 else mx.x_uniform(f.n, seed=31)
    B = f.spmv(X)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, flags=S.PLAN_VALIDATE | S.PLAN_MIXED)
    plan.set_config(rhs, **cfg)
    db = S.DeviceArray.from_host(B)
    dx = S.DeviceArray(B.shape, np.float64)
    x64 = oracle_x(f, B, rhs=rhs)
    errs = []
    for iters in (0, 1, 2):
        dx.zero()
        plan.solve_ir(db, dx, rhs=rhs, iters=iters)
        assert plan.check() == 0
        errs.append(rel_errs(dx.to_host(), x64)[0])
    assert 1e-12 < errs[0] <= 1e-5, errs   # an FP32 solve, rounded b
    assert errs[1] <= 1e-9 and errs[2] <= 1e-10, errs
