"""Pins the CPU oracle (CPU only, no GPU).

Each oracle function is checked against what the mathematics or an independent
library fixes (closed forms, scipy, brute force) and, where the reference's own
host routines build (oracle/_ref), against them bit for bit.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import sptrsv_oracle as orc
from sptrsv_amd import matrices as mx


def random_lower(n, density, seed, dominant=True):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=seed, format="coo")
    keep = A.row > A.col
    rows, cols = A.row[keep], A.col[keep]
    vals = rng.uniform(-1, 1, size=len(rows))
    absum = np.bincount(rows, weights=np.abs(vals), minlength=n)
    diag = (1.0 + absum) if dominant else rng.uniform(0.5, 2.0, n)
    L = sp.csc_matrix((np.r_[vals, diag], (np.r_[rows, np.arange(n)], np.r_[cols, np.arange(n)])),
                      shape=(n, n))
    L.sort_indices()
    return L


# ---------------------------------------------------------------- scan -----
@pytest.mark.parametrize("length", [0, 1, 2, 7, 1000, 4097])
def test_exclusive_scan_matches_cumsum(length):
    a = np.random.default_rng(length).integers(0, 50, size=length).astype(np.int32)
    got = orc.exclusive_scan(a)
    exp = np.r_[0, np.cumsum(a)[:-1]] if length else a
    np.testing.assert_array_equal(got, exp)
    r = orc.ref_exclusive_scan(a)
    if r is not None and length > 1:  # the reference leaves length-1 input unchanged
        np.testing.assert_array_equal(got, r)


# ------------------------------------------------------------ transpose ----
@pytest.mark.parametrize("m,n,density", [(50, 40, 0.1), (300, 300, 0.02), (1, 5, 1.0)])
def test_transpose_matches_scipy_and_reference(m, n, density):
    A = sp.random(m, n, density=density, random_state=7, format="csr")
    A.sort_indices()
    p, i, v = orc.transpose(m, n, A.indptr, A.indices, A.data)
    C = A.tocsc()
    C.sort_indices()
    np.testing.assert_array_equal(p, C.indptr)
    np.testing.assert_array_equal(i, C.indices)
    np.testing.assert_array_equal(v, C.data)
    r = orc.ref_transpose(m, n, A.indptr, A.indices, A.data)
    if r is not None:
        for a, b in zip((p, i, v), r):
            np.testing.assert_array_equal(a, b)


def test_transpose_is_stable_with_duplicates_and_unsorted_rows():
    # brute force: walk rows in order and append to per-column lists
    rowPtr = np.array([0, 3, 5, 8], np.int32)
    colIdx = np.array([2, 0, 2, 1, 1, 0, 2, 0], np.int32)
    val = np.arange(8, dtype=np.float64) + 1
    cols = [[] for _ in range(3)]
    for r in range(3):
        for j in range(rowPtr[r], rowPtr[r + 1]):
            cols[colIdx[j]].append((r, val[j]))
    p, i, v = orc.transpose(3, 3, rowPtr, colIdx, val)
    np.testing.assert_array_equal(p, np.r_[0, np.cumsum([len(c) for c in cols])])
    np.testing.assert_array_equal(i, [r for c in cols for r, _ in c])
    np.testing.assert_array_equal(v, [x for c in cols for _, x in c])


# ------------------------------------------------------------- analyser ----
def test_analyser_is_row_lengths():
    L = random_lower(500, 0.02, 3)
    got = orc.analyser(L.indices, 500)
    np.testing.assert_array_equal(got, np.diff(L.tocsr().indptr))
    r = orc.ref_analyser(L.indices, 500)
    if r is not None:
        np.testing.assert_array_equal(got, r)


# ------------------------------------------------------------- executor ----
def test_laplacian_x_one_known_answer():
    # b = L * 1, so every partial sum is a small integer and x is exactly 1
    for f in (mx.laplacian2d_lower(256, 256), mx.laplacian3d_lower(12, 9, 7)):
        b = f.spmv(np.ones(f.n))
        x = orc.solve(f.colPtr, f.rowIdx, f.val, b)
        assert np.array_equal(x, np.ones(f.n))


@pytest.mark.parametrize("backward", [False, True])
@pytest.mark.parametrize("rhs", [1, 3])
def test_integer_unit_lower_known_answer(backward, rhs):
    f = mx.integer_unit_lower(3000, density=4, seed=11, backward=backward)
    x_ref = np.random.default_rng(5).integers(1, 11, size=(f.n, rhs)).astype(np.float64)
    b = f.spmv(x_ref)
    x = orc.solve(f.colPtr, f.rowIdx, f.val, b.reshape(-1), substitution=int(backward), rhs=rhs)
    assert np.array_equal(x.reshape(f.n, rhs), x_ref)


@pytest.mark.parametrize("seed", [0, 1])
def test_solve_matches_scipy(seed):
    L = random_lower(2000, 0.003, seed)
    b = np.random.default_rng(seed).uniform(-1, 1, 2000)
    x = orc.solve(L.indptr, L.indices, L.data, b)
    exp = spla.spsolve_triangular(L.tocsr(), b, lower=True)
    assert np.max(np.abs(x - exp)) <= 1e-13 * np.max(np.abs(exp))
    # backward on U = L^T
    U = L.T.tocsc()
    U.sort_indices()
    xb = orc.solve(U.indptr, U.indices, U.data, b, substitution=1)
    expb = spla.spsolve_triangular(U.tocsr(), b, lower=False)
    assert np.max(np.abs(xb - expb)) <= 1e-13 * np.max(np.abs(expb))


def test_solve_matches_dense_solve_multi_rhs():
    n, rhs = 200, 4
    L = random_lower(n, 0.2, 9, dominant=False)
    B = np.random.default_rng(2).standard_normal((n, rhs))
    X = orc.solve(L.indptr, L.indices, L.data, B.reshape(-1), rhs=rhs).reshape(n, rhs)
    exp = np.linalg.solve(L.toarray(), B)
    assert np.max(np.abs(X - exp)) <= 1e-10 * np.max(np.abs(exp))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("backward", [0, 1])
def test_solve_bit_exact_with_reference_build(dtype, backward):
    if orc.ref(dtype) is None:
        pytest.skip("reference checkout not built into oracle/_ref")
    L = random_lower(3000, 0.002, 4)
    if backward:
        L = L.T.tocsc()
        L.sort_indices()
    rhs = 2
    b = np.random.default_rng(1).uniform(-1, 1, 3000 * rhs).astype(dtype)
    val = L.data.astype(dtype)
    x = orc.solve(L.indptr, L.indices, val, b, substitution=backward, rhs=rhs)
    r = orc.ref_solve(L.indptr, L.indices, val, b, substitution=backward, rhs=rhs)
    assert np.array_equal(x, r)


def test_solve_reports_unsatisfiable_wait():
    # an in-degree that never drains: the reference would spin forever
    f = mx.bidiagonal(10)
    indeg = orc.analyser(f.rowIdx, f.n)
    indeg[5] += 1
    with pytest.raises(RuntimeError):
        orc.solve(f.colPtr, f.rowIdx, f.val, np.ones(10), in_degree=indeg)


# ------------------------------------------------------------------ spmv ---
def test_spmv_matches_scipy():
    L = random_lower(700, 0.01, 2)
    x = np.random.default_rng(0).standard_normal((700, 3))
    b = orc.spmv(L.indptr, L.indices, L.data, x.reshape(-1), rhs=3).reshape(700, 3)
    np.testing.assert_allclose(b, L @ x, rtol=1e-13, atol=1e-13)
    f = mx.Factor(L.indptr, L.indices, L.data)
    np.testing.assert_allclose(f.spmv(x), L @ x, rtol=1e-13, atol=1e-13)


def test_stencil27_generator_matches_brute_force():
    nx, ny, nz = 5, 4, 3
    f = mx.stencil27_lower(nx, ny, nz)
    n = nx * ny * nz
    g = np.stack([np.arange(n) % nx, (np.arange(n) // nx) % ny, np.arange(n) // (nx * ny)], 1)
    near = np.abs(g[:, None, :] - g[None, :, :]).max(axis=2) <= 1
    M = np.where(near, -1.0, 0.0)
    np.fill_diagonal(M, 26.0)
    A = sp.csc_matrix((f.val, f.rowIdx, f.colPtr), shape=(n, n)).toarray()
    assert np.array_equal(A, np.tril(M))


def test_chain_far_generator_matches_brute_force():
    n, far = 40, 7
    f = mx.chain_far(n, far=far)
    A = sp.csc_matrix((f.val, f.rowIdx, f.colPtr), shape=(n, n)).toarray()
    M = 2.0 * np.eye(n) - np.eye(n, k=-1)
    M[np.arange(far, n), np.arange(0, n - far)] = -0.5
    assert np.array_equal(A, M)
    # every row is a chain step (depth n) and the chain is diagonally dominant
    assert orc.findlevel(f.colPtr, f.rowIdx)[0] == n
    x = mx.x_uniform(n, seed=2)
    np.testing.assert_allclose(orc.solve(f.colPtr, f.rowIdx, f.val, f.spmv(x)), x,
                               rtol=0, atol=1e-14)


def test_rel_err_formula():
    x = np.array([1.0, 2.0, -3.0])
    xr = np.array([1.0, 2.5, -2.0])
    assert orc.rel_err_l1(x, xr) == pytest.approx(1.5 / 5.5)
    assert orc.rel_err_l1(x, np.zeros(3)) == pytest.approx(6.0)


# --------------------------------------------------------------- levels ----
@pytest.mark.parametrize("f,expect", [
    (mx.laplacian2d_lower(256, 256), 511),
    (mx.laplacian2d_lower(17, 5), 21),
    (mx.laplacian3d_lower(9, 7, 5), 19),
    (mx.bidiagonal(1000), 1000),
    # 27-point: level(x, y, z) = x + 2y + 4z (every unit step back lowers it by 1)
    (mx.stencil27_lower(5, 4, 3), 4 + 6 + 8 + 1),
    (mx.stencil27_lower(9, 7, 6), 8 + 12 + 20 + 1),
])
def test_findlevel_closed_forms(f, expect):
    nl, pmin, pavg, pmax = orc.findlevel(f.colPtr, f.rowIdx)
    assert nl == expect
    assert pavg == f.n // nl
    r = orc.ref_findlevel(f.colPtr, f.rowIdx, f.val)
    if r is not None:
        assert r == (nl, pmin, pavg, pmax)


# ---------------------------------------------------------- partitioners ---
def test_equal_col_partition():
    c, d = orc.equal_col_partition(4, 10)
    assert c.tolist() == [3, 3, 3, 1] and d.tolist() == [0, 3, 6, 9]
    c, d = orc.equal_col_partition(3, 8)
    assert c.tolist() == [3, 3, 2] and d.tolist() == [0, 3, 6]


def _best_max_part(lens, nparts):
    best = None
    n = len(lens)
    for k in range(1, min(nparts, n) + 1):
        for cuts in itertools.combinations(range(1, n), k - 1):
            b = (0,) + cuts + (n,)
            m = max(sum(lens[b[i]:b[i + 1]]) for i in range(k))
            best = m if best is None else min(best, m)
    return best


@pytest.mark.parametrize("seed", range(12))
def test_equal_nnz_partition_is_optimal_on_tiny_inputs(seed):
    rng = np.random.default_rng(seed)
    ncols = int(rng.integers(3, 9))
    nparts = int(rng.integers(2, 4))
    lens = rng.integers(1, 7, size=ncols)
    ptr = np.r_[0, np.cumsum(lens)].astype(np.int32)
    c, d = orc.equal_nnz_partition(nparts, ptr)
    used = [(int(dd), int(cc)) for cc, dd in zip(c, d) if cc > 0]
    # contiguous cover of all columns, in order
    pos = 0
    for dd, cc in used:
        assert dd == pos
        pos += cc
    assert pos == ncols
    got = max(int(lens[dd:dd + cc].sum()) for dd, cc in used)
    assert got == _best_max_part(list(lens), nparts)


# ------------------------------------------------------------ refinement ---
def test_refinement_converges_to_fp64_solution():
    f = mx.laplacian2d_lower(64, 48)
    x_true = mx.x_uniform(f.n, seed=0)
    b = f.spmv(x_true)
    x64 = orc.solve(f.colPtr, f.rowIdx, f.val, b)
    x0 = orc.refine_f32(f.colPtr, f.rowIdx, f.val, b, iters=0)
    x2 = orc.refine_f32(f.colPtr, f.rowIdx, f.val, b, iters=2)
    assert orc.rel_err_l1(x0, x64) > 1e-9      # one FP32 solve is not enough
    assert orc.rel_err_l1(x2, x64) <= 1e-12    # two refinement steps are
