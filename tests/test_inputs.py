"""Reference input formats (Matrix Market and CSC text), host side.

The Matrix Market reader is checked against scipy.io.mmread, an independent
implementation, on general / symmetric / pattern / integer files; the triangle
extraction is checked structurally (strict triangle of the mirrored pattern
plus a unit diagonal, values in 1..10, rows ascending with the diagonal first
for forward and last for backward, as main.cu:152-246 leaves them).
"""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from sptrsv_amd import matrices as mx
from sptrsv_amd import mtx

import sptrsv_oracle as orc


def _write(tmp_path, name, A, **kw):
    path = str(tmp_path / name)
    scipy.io.mmwrite(path, A, **kw)
    return path if path.endswith(".mtx") else path + ".mtx"


def _dense(m, n, rows, cols, vals):
    D = np.zeros((m, n))
    np.add.at(D, (rows, cols), vals)
    return D


def _random(n, density, seed, symmetric=False, integer=False):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng, format="coo")
    if integer:
        A.data = rng.integers(-9, 10, size=A.nnz).astype(np.float64)
    if symmetric:
        A = sp.tril(A + A.T).tocoo()  # mmwrite(symmetric) stores one triangle
    return A


@pytest.mark.parametrize("kind", ["general", "symmetric", "pattern", "integer"])
def test_read_mtx_matches_scipy(tmp_path, kind):
    if kind == "general":
        A = _random(60, 0.08, 1)
        path = _write(tmp_path, "g.mtx", A)
    elif kind == "symmetric":
        A = _random(50, 0.1, 2, symmetric=True)
        path = _write(tmp_path, "s.mtx", A, symmetry="symmetric")
    elif kind == "pattern":
        A = _random(40, 0.1, 3, symmetric=True)
        path = _write(tmp_path, "p.mtx", A, field="pattern", symmetry="symmetric")
    else:
        A = _random(45, 0.1, 4, integer=True)
        path = _write(tmp_path, "i.mtx", A, field="integer")
    m, n, r, c, v = mtx.read_mtx(path)
    ref = scipy.io.mmread(path).toarray()
    assert (m, n) == ref.shape
    np.testing.assert_array_equal(_dense(m, n, r, c, v), ref)


@pytest.mark.parametrize("substitution", [mtx.FORWARD, mtx.BACKWARD])
def test_triangle_from_mtx_structure(tmp_path, substitution):
    A = _random(80, 0.06, 5, symmetric=True)
    path = _write(tmp_path, "a.mtx", A, symmetry="symmetric")
    f = mtx.triangle_from_mtx(path, substitution=substitution, seed=3)
    full = scipy.io.mmread(path).toarray()
    L = sp.csc_matrix((f.val, f.rowIdx, f.colPtr), shape=(f.n, f.n)).toarray()
    strict = np.tril(full != 0, -1) if substitution == mtx.FORWARD else np.triu(full != 0, 1)
    np.testing.assert_array_equal(L != 0, strict | np.eye(f.n, dtype=bool))
    np.testing.assert_array_equal(np.diag(L), np.ones(f.n))
    off = L[strict]
    assert off.min() >= 1 and off.max() <= 10 and np.all(off == np.round(off))
    for j in range(f.n):
        rows = f.rowIdx[f.colPtr[j]:f.colPtr[j + 1]]
        assert np.all(np.diff(rows) > 0)
        assert rows[0 if substitution == mtx.FORWARD else -1] == j
    # the oracle solves it exactly (unit diagonal, integers, integer x)
    x = np.random.default_rng(1).integers(1, 4, size=f.n).astype(np.float64)
    np.testing.assert_array_equal(
        orc.solve(f.colPtr, f.rowIdx, f.val, f.spmv(x), substitution=substitution), x)


def test_triangle_from_mtx_file_values(tmp_path):
    A = _random(30, 0.2, 6)
    path = _write(tmp_path, "v.mtx", A)
    f = mtx.triangle_from_mtx(path, values="file")
    L = sp.csc_matrix((f.val, f.rowIdx, f.colPtr), shape=(f.n, f.n)).toarray()
    full = scipy.io.mmread(path).toarray()
    np.testing.assert_array_equal(np.tril(L, -1), np.tril(full, -1))


def test_mtx_rejects_bad_input(tmp_path):
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(ValueError):
        mtx.read_mtx(str(bad))
    rect = _write(tmp_path, "r.mtx", sp.random(4, 6, density=0.5, random_state=0))
    with pytest.raises(ValueError):
        mtx.triangle_from_mtx(rect)


def test_csc_text_roundtrip(tmp_path):
    f = mx.laplacian2d_lower(9, 7)
    path = str(tmp_path / "l.csc")
    mtx.write_csc_text(f, path)
    g = mtx.read_csc_text(path, seed=2)
    np.testing.assert_array_equal(g.colPtr, f.colPtr)
    np.testing.assert_array_equal(g.rowIdx, f.rowIdx)
    assert g.val.min() >= 1 and g.val.max() <= 10
    (tmp_path / "short.csc").write_text("3 3 4\n1 2 3 5\n1 2\n")
    with pytest.raises(ValueError):
        mtx.read_csc_text(str(tmp_path / "short.csc"))


def test_triangle_from_mtx_file_diagonal(tmp_path):
    f0 = mx.laplacian2d_lower(12, 9)
    A = sp.csc_matrix((f0.val, f0.rowIdx, f0.colPtr), shape=(f0.n, f0.n))
    path = _write(tmp_path, "lap.mtx", sp.tril(A).tocoo(), symmetry="symmetric")
    f = mtx.triangle_from_mtx(path, values="file", diag="file")
    np.testing.assert_array_equal(f.colPtr, f0.colPtr)
    np.testing.assert_array_equal(f.rowIdx, f0.rowIdx)
    np.testing.assert_array_equal(f.val, f0.val)
    nodiag = _write(tmp_path, "nd.mtx", sp.random(20, 20, density=0.2, random_state=1))
    with pytest.raises(ValueError):
        mtx.triangle_from_mtx(nodiag, diag="file")
