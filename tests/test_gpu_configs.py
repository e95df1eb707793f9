"""Parity at the benchmark workloads' full sizes.

2D 5-point Laplacian 256^2, x = 1: known answer within KNOWN_ANSWER_TOL
    (the default lane engine reassociates in-batch chains; see
    test_gpu_parity.py).
3D 7-point Laplacian 128^3: FP64 <= 1e-12 vs the serial oracle; x = 1 known
    answer within KNOWN_ANSWER_TOL; FP32 + FP64 refinement <= 1e-10.
Power-law factor, n = 1e6: FP64 <= 1e-12, refinement <= 1e-10.
The 128^3 factor with 8 / 32 / 128 right-hand sides: FP64 <= 1e-12,
    refinement <= 1e-10.
512 uneven 2D factors in one launch (and a hipGraph of it), and the 1e6
    bidiagonal chain: FP64 <= 1e-12.
"""
import numpy as np
import pytest

import sptrsv_amd as S
from sptrsv_amd import matrices as mx
from gpu_util import assert_close, oracle_x, plan_solve, require_gpu

KNOWN_ANSWER_TOL = 1e-14

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    require_gpu()


@pytest.fixture(scope="module")
def lap3d():
    return mx.laplacian3d_lower(128, 128, 128)


def _ir(f, B, rhs, iters=2):
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, flags=S.PLAN_VALIDATE | S.PLAN_MIXED)
    db = S.DeviceArray.from_host(B)
    dx = S.DeviceArray(B.shape, np.float64)
    plan.solve_ir(db, dx, rhs=rhs, iters=iters)
    assert plan.check() == 0
    return dx.to_host()


def test_lap2d_256_known_answer():
    f = mx.laplacian2d_lower(256, 256)
    b = f.spmv(np.ones(f.n))
    x, _ = plan_solve(f, b)
    assert np.abs(x - 1.0).max() <= KNOWN_ANSWER_TOL


def test_lap3d_128_fp64_and_known_answer(lap3d):
    f = lap3d
    b = f.spmv(mx.x_uniform(f.n, seed=0))
    x, plan = plan_solve(f, b)
    assert_close(x, oracle_x(f, b), 1e-12)
    b1 = f.spmv(np.ones(f.n))
    x1, _ = plan_solve(f, b1, plan=plan)
    assert np.abs(x1 - 1.0).max() <= KNOWN_ANSWER_TOL


def test_lap3d_128_refinement(lap3d):
    f = lap3d
    b = f.spmv(mx.x_uniform(f.n, seed=0))
    assert_close(_ir(f, b, 1), oracle_x(f, b), 1e-10)


def test_power_law_1e6():
    f = mx.power_law_lower(1_000_000, seed=0)
    assert f.nnz == 6_133_256
    b = f.spmv(mx.x_uniform(f.n, seed=1))
    x, _ = plan_solve(f, b)
    x_o = oracle_x(f, b)
    assert_close(x, x_o, 1e-12)
    assert_close(_ir(f, b, 1), x_o, 1e-10)


@pytest.mark.parametrize("rhs", [8, 32, 128])
def test_lap3d_128_multi_rhs(lap3d, rhs):
    f = lap3d
    X = mx.x_normal(f.n, rhs, seed=1)
    B = f.spmv(X)
    x, _ = plan_solve(f, B, rhs=rhs)
    x_o = oracle_x(f, B, rhs=rhs)
    assert_close(x, x_o, 1e-12)
    assert_close(_ir(f, B, rhs), x_o, 1e-10)


def test_512_uneven_lap2d_batch_and_graph():
    f = mx.laplacian2d_batch(512, seed=0)
    B = f.spmv(mx.x_uniform(f.n, seed=0))
    plan = S.Plan.from_host(f.colPtr, f.rowIdx, f.val, sys_ptr=f.sys_ptr)
    db = S.DeviceArray.from_host(B)
    dx = S.DeviceArray(B.shape, np.float64)
    g = plan.graph(db, dx)
    x_o = oracle_x(f, B)
    for _ in range(3):
        dx.zero()
        g.launch()
        assert plan.check() == 0
        assert_close(dx.to_host(), x_o, 1e-12)


def test_bidiagonal_chain_1e6():
    f = mx.bidiagonal(1_000_000)
    b = f.spmv(mx.x_uniform(f.n, seed=0))
    x, _ = plan_solve(f, b)
    assert_close(x, oracle_x(f, b), 1e-12)
