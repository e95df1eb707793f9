"""Reference input formats for the triangular factor, host side.

* Matrix Market (``-mtx`` in the reference): the file is read as the
  reference's ``mmio_info``/``mmio_data`` read it (SpTRSV_cuda/mmio_highlevel.h
  :8-296): coordinate entries, 1-based, ``real`` / ``integer`` / ``pattern``
  (value 1.0) / ``complex`` (real part), and ``symmetric`` / ``hermitian``
  files mirrored across the diagonal.  The triangle is then extracted as in
  SpTRSV_cuda/main.cu:152-187: the strictly lower (forward) or strictly upper
  (backward) entries of each row, with values drawn uniformly from the
  integers 1..10 (the reference's ``rand() % 10 + 1``; it ignores the file's
  values), plus a unit diagonal, then transposed to CSC with each column's
  rows ascending (main.cu:197-218).
* CSC text (``-csc``, main.cu:247-286): ``m n nnz``, then n+1 1-based column
  pointers, then nnz 1-based row indices; values again drawn from 1..10.

The reference seeds ``rand()`` with the time of day (main.cu:140), so its
values are not reproducible; here they come from ``numpy.random.default_rng``
with an explicit seed.  ``values="file"`` keeps the file's off-diagonal
values instead, and ``diag="file"`` the file's diagonal instead of 1.  The
reference's recipe (random values, unit diagonal) makes the solution grow
like 10^(DAG depth) and overflow on large matrices; with the file's values a
diagonally dominant matrix gives a well-conditioned factor.
"""
import numpy as np

from .matrices import Factor, _from_columns

FORWARD, BACKWARD = 0, 1


def _banner(line):
    parts = line.strip().split()
    if len(parts) < 5 or parts[0].lower() != "%%matrixmarket":
        raise ValueError("not a Matrix Market file: " + line.strip()[:60])
    obj, fmt, field, sym = (p.lower() for p in parts[1:5])
    if obj != "matrix" or fmt != "coordinate":
        raise ValueError(f"only coordinate matrices are supported, got {obj} {fmt}")
    if field not in ("real", "double", "integer", "pattern", "complex"):
        raise ValueError(f"unsupported field {field}")
    if sym not in ("general", "symmetric", "hermitian", "skew-symmetric"):
        raise ValueError(f"unsupported symmetry {sym}")
    return field, sym


def read_mtx(path):
    """(m, n, rows, cols, vals): 0-based coordinates, symmetric entries
    mirrored (mmio_highlevel.h:8-121 counts, :85-105 mirrors)."""
    with open(path, "r") as fh:
        field, sym = _banner(fh.readline())
        line = fh.readline()
        while line.startswith("%") or not line.strip():
            line = fh.readline()
        m, n, nnz = (int(t) for t in line.split()[:3])
        ncol = {"pattern": 2, "complex": 4}.get(field, 3)
        data = np.loadtxt(fh, dtype=np.float64, ndmin=2, comments="%",
                          max_rows=nnz) if nnz else np.zeros((0, ncol))
    if data.shape[0] != nnz or (nnz and data.shape[1] < ncol):
        raise ValueError(f"{path}: expected {nnz} entries with {ncol} columns")
    rows = data[:, 0].astype(np.int64) - 1
    cols = data[:, 1].astype(np.int64) - 1
    vals = np.ones(nnz) if field == "pattern" else data[:, 2].astype(np.float64)
    if nnz and (rows.min() < 0 or cols.min() < 0 or rows.max() >= m or cols.max() >= n):
        raise ValueError(f"{path}: index out of range")
    if sym != "general":
        off = rows != cols
        mv = -vals[off] if sym == "skew-symmetric" else vals[off]
        rows, cols, vals = (np.concatenate([rows, cols[off]]),
                            np.concatenate([cols, rows[off]]),
                            np.concatenate([vals, mv]))
    return m, n, rows, cols, vals


def triangle_from_mtx(path, substitution=FORWARD, values="random", seed=0,
                      diag="unit"):
    """Triangular factor of the matrix in a Matrix Market file, as a CSC
    Factor (main.cu:152-187): strictly lower entries for forward
    substitution, strictly upper for backward, plus the diagonal (1, or the
    file's with diag="file")."""
    m, n, rows, cols, vals = read_mtx(path)
    if m != n:
        raise ValueError(f"{path}: matrix is {m} x {n}, not square")
    keep = cols < rows if substitution == FORWARD else cols > rows
    r, c, v = rows[keep], cols[keep], vals[keep]
    if values == "random":
        v = np.random.default_rng(seed).integers(1, 11, size=len(r)).astype(np.float64)
    elif values != "file":
        raise ValueError("values must be 'random' or 'file'")
    i = np.arange(n, dtype=np.int64)
    if diag == "unit":
        d = np.ones(n)
    elif diag == "file":
        on = rows == cols
        d = np.zeros(n)
        np.add.at(d, rows[on], vals[on])
        if np.any(d == 0):
            raise ValueError(f"{path}: {int(np.sum(d == 0))} rows have no (or a zero) diagonal")
    else:
        raise ValueError("diag must be 'unit' or 'file'")
    p, ri, vv = _from_columns(n, np.concatenate([c, i]), np.concatenate([r, i]),
                              np.concatenate([v, d]))
    name = path.rsplit("/", 1)[-1] + ("_L" if substitution == FORWARD else "_U")
    return Factor(p, ri, vv, name=name)


def read_csc_text(path, seed=0):
    """CSC text factor (main.cu:247-286): structure from the file, values
    drawn from the integers 1..10, each column's rows sorted ascending."""
    with open(path, "r") as fh:
        tok = np.array(fh.read().split(), dtype=np.int64)
    if tok.size < 3:
        raise ValueError(f"{path}: missing header")
    m, n, nnz = (int(t) for t in tok[:3])
    if m != n:
        raise ValueError(f"{path}: matrix is {m} x {n}, not square")
    if tok.size < 3 + n + 1 + nnz:
        raise ValueError(f"{path}: expected {n + 1} pointers and {nnz} row indices")
    colPtr = tok[3:3 + n + 1] - 1
    rowIdx = tok[3 + n + 1:3 + n + 1 + nnz] - 1
    if colPtr[0] != 0 or colPtr[-1] != nnz or np.any(np.diff(colPtr) < 0):
        raise ValueError(f"{path}: bad column pointers")
    if nnz and (rowIdx.min() < 0 or rowIdx.max() >= n):
        raise ValueError(f"{path}: row index out of range")
    col = np.repeat(np.arange(n, dtype=np.int64), np.diff(colPtr))
    vals = np.random.default_rng(seed).integers(1, 11, size=nnz).astype(np.float64)
    p, ri, vv = _from_columns(n, col, rowIdx, vals)
    return Factor(p, ri, vv, name=path.rsplit("/", 1)[-1])


def write_csc_text(f, path):
    """Write a factor's structure in the CSC text format (1-based)."""
    with open(path, "w") as fh:
        fh.write(f"{f.n} {f.n} {f.nnz}\n")
        fh.write(" ".join(str(int(v) + 1) for v in f.colPtr) + "\n")
        fh.write(" ".join(str(int(v) + 1) for v in f.rowIdx) + "\n")
