"""Edge cases of the executors on the GPU, against the oracle.

Sizes at the lane engine's batch (64 columns) and window boundaries, a single
column, a diagonal factor (no off-diagonal entries), an empty system, tiles
wider than the right-hand-side count (rhs_per_lane > rhs), FP32 plans and
batches of systems with several right-hand sides per lane, and hipGraph
replays of such a solve.  Tolerances as in test_gpu_parity.py: FP64 1e-12,
FP32 2e-5 relative to the serial oracle.
"""
import numpy as np
import pytest

import sptrsv_amd as S
from sptrsv_amd import matrices as mx
from gpu_util import assert_close, oracle_x, plan_solve, require_gpu

pytestmark = pytest.mark.gpu

LANE = S.ENGINE_LANE
CFGS = [None, dict(lanes_per_column=4, block=256, window=64),
        dict(engine=LANE, block=256, window=128), dict(engine=LANE, block=1024, window=8192)]


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


def _id(c):
    if c is None:
        return "default"
    return ("lane" if c.get("engine") == LANE else "group") + "-B{block}-W{window}".format(**c)


@pytest.mark.parametrize("cfg", CFGS, ids=_id)
@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 8191, 8192, 8193, 16_385])
def test_sizes_at_batch_and_window_boundaries(n, cfg):
    f = mx.bidiagonal(n) if n < 100 else mx.power_law_lower(n, window=min(n, 500), seed=n)
    b = f.spmv(mx.x_uniform(f.n, seed=n))
    x, _ = plan_solve(f, b, cfg=cfg)
    assert_close(x, oracle_x(f, b), 1e-12)


@pytest.mark.parametrize("cfg", CFGS, ids=_id)
@pytest.mark.parametrize("backward", [0, 1])
def test_diagonal_factor(cfg, backward):
    n = 3000
    d = np.random.default_rng(3).uniform(1.0, 2.0, n)
    f = mx.Factor(np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d)
    b = np.random.default_rng(4).uniform(-1, 1, n)
    x, _ = plan_solve(f, b, subst=backward, cfg=cfg)
    assert_close(x, b / d, 1e-15)


def test_empty_system():
    f = mx.Factor(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    db = S.DeviceArray((1,), np.float64)
    dx = S.DeviceArray((1,), np.float64)
    plan.solve(db, dx)
    assert plan.check() == 0


@pytest.mark.parametrize("R", [2, 4, 8])
@pytest.mark.parametrize("rhs", [1, 3, 5])
def test_lane_tile_wider_than_rhs(R, rhs):
    f = mx.laplacian3d_lower(12, 10, 9)
    X = mx.x_normal(f.n, rhs, seed=R) if rhs > 1 else mx.x_uniform(f.n, seed=R)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=rhs, cfg=dict(engine=LANE, rhs_per_lane=R, block=512, window=512))
    assert_close(x, oracle_x(f, B, rhs=rhs), 1e-12)


@pytest.mark.parametrize("R", [1, 4])
def test_fp32_plan_multi_rhs_lane(R):
    f = mx.laplacian2d_lower(70, 50)
    X = mx.x_normal(f.n, 6, seed=9)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=6, dtype=np.float32,
                      cfg=dict(engine=LANE, rhs_per_lane=R, block=256, window=1024))
    assert_close(x, oracle_x(f, B, rhs=6, dtype=np.float32), 2e-5)


@pytest.mark.parametrize("R", [2, 8])
def test_batch_multi_rhs_lane(R):
    f = mx.laplacian2d_batch(24, lo=8, hi=40, seed=5)
    X = mx.x_normal(f.n, 5, seed=6)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=5, cfg=dict(engine=LANE, rhs_per_lane=R, block=256, window=900))
    assert_close(x, oracle_x(f, B, rhs=5), 1e-12)


def test_graph_replays_multi_rhs_lane():
    f = mx.laplacian3d_lower(20, 16, 12)
    X = mx.x_normal(f.n, 4, seed=11)
    B = f.spmv(X)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    plan.set_config(4, engine=LANE, rhs_per_lane=4, block=512, window=2048)
    db = S.DeviceArray.from_host(B)
    dx = S.DeviceArray(B.shape, np.float64)
    plan.solve(db, dx, 4)
    assert plan.check() == 0
    ref = dx.to_host()
    assert_close(ref, oracle_x(f, B, rhs=4), 1e-12)
    stream = S.Stream()
    g = plan.graph(db, dx, rhs=4, stream=stream)
    for _ in range(3):
        dx.zero()
        g.launch(stream)
        stream.synchronize()
        assert plan.check(stream) == 0
        assert_close(dx.to_host(), ref, 1e-12)


# ------------------------------------------------------------ wide batches
# A run of pure links (each column's only off-diagonal entry feeds the next
# column, which has no other input) longer than 64 columns is one wide batch
# of up to 512 columns, 8 per lane, solved by one wavefront scan.
def _chain_with_breaks(n, every, seed):
    """Bidiagonal chain plus an extra entry every `every` columns, which
    breaks the run of pure links (the extra entry's column has two
    off-diagonal entries and its target row two inputs)."""
    rng = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.int64)
    extra_c = np.arange(0, n - 40, every, dtype=np.int64)
    c = np.concatenate([i, i[:-1], extra_c])
    r = np.concatenate([i, i[:-1] + 1, extra_c + 37])
    v = np.concatenate([rng.uniform(2.0, 3.0, n), rng.uniform(-1.0, 1.0, n - 1),
                        rng.uniform(-0.5, 0.5, len(extra_c))])
    p, ri, vv = mx._from_columns(n, c, r, v)
    return mx.Factor(p, ri, vv, name=f"chain{n}_e{every}")


@pytest.mark.parametrize("window", [256, 1000, 8192])
@pytest.mark.parametrize("n", [65, 100, 511, 512, 513, 1500, 20_000])
def test_wide_batches_bidiagonal(n, window):
    f = mx.bidiagonal(n)
    b = f.spmv(mx.x_uniform(n, seed=n))
    x, _ = plan_solve(f, b, cfg=dict(engine=LANE, block=512, window=window))
    assert_close(x, oracle_x(f, b), 1e-12)


def test_wide_batches_known_answer():
    f = mx.bidiagonal(100_000)  # diag 2, sub -1: x = 1 stays dyadic
    b = f.spmv(np.ones(f.n))
    x, _ = plan_solve(f, b, cfg=dict(engine=LANE, block=1024, window=8192))
    assert_close(x, np.ones(f.n), 1e-14)


@pytest.mark.parametrize("every", [50, 300, 2000])
@pytest.mark.parametrize("backward", [0, 1])
def test_wide_batches_broken_chains(every, backward):
    f = _chain_with_breaks(30_000, every, seed=every)
    if backward:
        f = mx.upper_from_lower(f)
    b = f.spmv(mx.x_uniform(f.n, seed=2))
    for cfg in (dict(engine=LANE, block=1024, window=8192), dict(engine=LANE, block=256, window=700)):
        x, _ = plan_solve(f, b, subst=backward, cfg=cfg)
        assert_close(x, oracle_x(f, b, subst=backward), 1e-12)


@pytest.mark.parametrize("R,rhs", [(1, 3), (2, 3), (4, 4), (8, 5)])
def test_wide_batches_multi_rhs(R, rhs):
    f = _chain_with_breaks(12_000, 900, seed=R)
    X = mx.x_normal(f.n, rhs, seed=R)
    B = f.spmv(X)
    window = 1024 if R == 8 else 2048  # LDS: (window + 64) * (8 R + 4) bytes
    x, _ = plan_solve(f, B, rhs=rhs, cfg=dict(engine=LANE, rhs_per_lane=R, block=512, window=window))
    assert_close(x, oracle_x(f, B, rhs=rhs), 1e-12)


def test_wide_batches_fp32_and_large_links():
    # every 7th link has |L(c+1, c)| > |L(c+1, c+1)|, which makes it impure:
    # the runs between those links are shorter than a wide batch needs only
    # where they are; FP32 plan on a pure chain
    n = 5000
    i = np.arange(n, dtype=np.int64)
    sub = np.where(i[:-1] % 7 == 6, -2.5, -1.0)
    p, ri, vv = mx._from_columns(n, np.concatenate([i, i[:-1]]), np.concatenate([i, i[:-1] + 1]),
                                 np.concatenate([np.full(n, 2.0), sub]))
    f = mx.Factor(p, ri, vv, name="chain_big_links")
    b = f.spmv(mx.x_uniform(n, seed=1))
    x, _ = plan_solve(f, b, cfg=dict(engine=LANE, block=512, window=8192))
    assert_close(x, oracle_x(f, b), 1e-12)
    f = mx.bidiagonal(n)
    b = f.spmv(mx.x_uniform(n, seed=2))
    x, _ = plan_solve(f, b, dtype=np.float32, cfg=dict(engine=LANE, block=512, window=8192))
    assert_close(x, oracle_x(f, b, dtype=np.float32), 2e-5)


# ------------------------------------------------------ batch deal and order
# Grid rows of 2..11 batches: the level-ordered deal of a window's batches to
# its wavefronts (anti-diagonals) instead of index round-robin.
@pytest.mark.parametrize("nx", [130, 200, 700, 769])
@pytest.mark.parametrize("backward", [0, 1])
def test_level_ordered_deal_on_wide_grids(nx, backward):
    f = mx.laplacian2d_lower(nx, 40)
    if backward:
        f = mx.upper_from_lower(f)
    b = f.spmv(mx.x_uniform(f.n, seed=nx))
    for cfg in (dict(engine=LANE, block=1024, window=8192), dict(engine=LANE, block=256, window=1024)):
        x, _ = plan_solve(f, b, subst=backward, cfg=cfg)
        assert_close(x, oracle_x(f, b, subst=backward), 1e-12)


@pytest.mark.parametrize("backward", [0, 1])
def test_batch_claim_order_largest_first(backward):
    # systems of very different sizes in no order: windows are claimed
    # largest system first, each system's windows in order
    f = mx.laplacian2d_batch(40, lo=8, hi=200, seed=9)
    if backward:
        sys_ptr = f.sys_ptr  # U = L^T keeps the system boundaries
        f = mx.upper_from_lower(f)
        f.sys_ptr = sys_ptr
    X = mx.x_uniform(f.n, seed=10)
    b = f.spmv(X)
    for cfg in (None, dict(engine=LANE, block=512, window=2048)):
        x, _ = plan_solve(f, b, subst=backward, cfg=cfg)
        assert_close(x, oracle_x(f, b, subst=backward), 1e-12)


@pytest.mark.parametrize("backward", [0, 1])
def test_batch_default_windows_16384(backward):
    # batches of systems default to 16384-column windows (byte counters make
    # room); systems wider than that (150^2 = 22500) are cut into pieces
    f = mx.laplacian2d_batch(6, lo=120, hi=180, seed=3)
    assert (np.diff(f.sys_ptr) > 16384).any() and (np.diff(f.sys_ptr) < 16384).any()
    if backward:
        sys_ptr = f.sys_ptr
        f = mx.upper_from_lower(f)
        f.sys_ptr = sys_ptr
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, substitution=backward,
                            sys_ptr=f.sys_ptr)
    assert plan.get_config(1)["window"] == 16384
    b = f.spmv(mx.x_uniform(f.n, seed=5))
    x, _ = plan_solve(f, b, subst=backward, plan=plan)
    assert_close(x, oracle_x(f, b, subst=backward), 1e-12)


def test_default_windows_follow_the_grid_period():
    # the default lane windows hold whole planes (or rows) of a grid: planes
    # up to 8192 columns are packed whole into 8192-column windows; a wider
    # plane gets a window of exactly one plane when it fits in LDS (FP64
    # planes of up to 16384 columns, with byte-wide arrival counters), else
    # it is split evenly (130^2 = 16900 -> 2 x 8450)
    for dims, window in (((60, 50, 30), 8192), ((100, 100, 12), 10000),
                         ((128, 128, 8), 16384), ((127, 127, 9), 16129),
                         ((130, 130, 3), 8450)):
        f = mx.laplacian3d_lower(*dims)
        plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
        assert plan.get_config(1)["window"] == window, dims
        b = f.spmv(mx.x_uniform(f.n, seed=3))
        x, _ = plan_solve(f, b, plan=plan)
        assert_close(x, oracle_x(f, b), 1e-12)


def _lap3d_with_long_row(nx, ny, nz, extra):
    """The 7-point Laplacian with `extra` more entries in its last row: half
    from the first plane (entries from earlier windows), half from the last
    plane (entries inside the row's own window)."""
    f = mx.laplacian3d_lower(nx, ny, nz)
    n, plane = f.n, nx * ny
    cols = np.repeat(np.arange(n), np.diff(f.colPtr))
    half = extra // 2
    xc = np.r_[np.arange(half), n - plane + np.arange(extra - half)]
    cols = np.r_[cols, xc]
    rows = np.r_[f.rowIdx, np.full(extra, n - 1)]
    vals = np.r_[f.val, np.full(extra, -1.0 / extra)]
    cp, ri, vv = mx._from_columns(n, cols, rows, vals)
    return mx.Factor(cp, ri, vv, name=f"lap3d_long{extra}")


@pytest.mark.parametrize("extra,window", [(250, 16384), (300, 8192)])
def test_byte_counters_follow_the_longest_row(extra, window):
    # byte-wide LDS arrival counters hold counts up to 255: a plane of
    # 16384 columns is one window while every row has at most 255 entries
    # (here the last row has 254 and its count climbs to 253), and is split
    # when one row has more
    f = _lap3d_with_long_row(128, 128, 3, extra)
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val)
    assert plan.get_config(1)["window"] == window
    b = f.spmv(mx.x_uniform(f.n, seed=4))
    x, _ = plan_solve(f, b, plan=plan)
    assert_close(x, oracle_x(f, b), 1e-12)


@pytest.mark.parametrize("backward", [0, 1])
@pytest.mark.parametrize("dims,cfg", [
    ((24, 20, 30), None),
    ((24, 20, 30), dict(engine=LANE, block=1024, window=2048)),
    ((70, 9, 20), dict(engine=LANE, block=512, window=4096)),
    ((64, 6, 80), dict(engine=LANE, block=1024, window=2048)),
])
def test_stencil27_windows_starting_mid_row(dims, cfg, backward):
    # the column reach of a 27-point factor is a plane + a row + 1, so
    # period-aligned windows start at every column of a grid row in turn;
    # batches must realign with the rows after a short first fragment (also
    # for rows wider than a wavefront: 70 = 64 + 6)
    f = mx.stencil27_lower(*dims)
    if backward:
        f = mx.upper_from_lower(f)
    b = f.spmv(mx.x_uniform(f.n, seed=5))
    x, _ = plan_solve(f, b, subst=backward, cfg=cfg)
    assert_close(x, oracle_x(f, b, subst=backward), 1e-12)


@pytest.mark.parametrize("rhs", [3, 8])
def test_stencil27_multi_rhs(rhs):
    f = mx.stencil27_lower(30, 20, 16)
    X = mx.x_normal(f.n, rhs, seed=6)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=rhs)
    assert_close(x, oracle_x(f, B, rhs=rhs), 1e-12)
