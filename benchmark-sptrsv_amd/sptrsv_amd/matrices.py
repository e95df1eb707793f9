"""Synthetic triangular factors of the benchmark workloads, built with numpy.

All factors come out in the solver's input format: CSC, 0-based int32 indices,
rows ascending within each column (so the diagonal is first for a lower
factor), values float64.  These replace the reference's Matrix Market loading
and random-value triangle extraction (SpTRSV_cuda/main.cu:142-246); the
right-hand side is b = L x as in main.cu:332-358.

Workloads (natural lexicographic ordering i = x + nx*(y + ny*z)):
  laplacian2d_lower(256, 256)         2D 5-point, diag 4, off -1 (n = 65,536)
  laplacian3d_lower(128, 128, 128)    3D 7-point, diag 6, off -1 (n = 2,097,152)
  power_law_lower(1_000_000)          Zipf(2) off-diagonals per row, capped at
                                      1000 and at the 50,000-column window
  laplacian2d_batch(512)              512 uneven 2D factors, sides log-uniform
                                      in [32, 256], diag 4 + U(0,1)
  bidiagonal(1_000_000)               diag 2, sub-diagonal -1
"""
import numpy as np


class Factor:
    """CSC triangular factor with optional system boundaries (batches)."""

    def __init__(self, colPtr, rowIdx, val, sys_ptr=None, name=""):
        self.colPtr = np.ascontiguousarray(colPtr, np.int32)
        self.rowIdx = np.ascontiguousarray(rowIdx, np.int32)
        self.val = np.ascontiguousarray(val, np.float64)
        self.n = len(self.colPtr) - 1
        self.nnz = len(self.rowIdx)
        self.sys_ptr = None if sys_ptr is None else np.ascontiguousarray(sys_ptr, np.int32)
        self.name = name

    def spmv(self, x):
        """b = L x with x row-major [n] or [n][rhs] (main.cu:347-358)."""
        x = np.asarray(x, np.float64)
        rhs = 1 if x.ndim == 1 else x.shape[1]
        cols = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.colPtr))
        contrib = self.val[:, None] * x.reshape(self.n, rhs)[cols]
        b = np.zeros((self.n, rhs))
        if self.nnz:
            _segment_add(b, self.rowIdx, contrib)
        return b.reshape(x.shape)

    def algorithmic_bytes(self, rhs=1, value_bytes=8):
        """CSC values + row indices + colPtr read once, b read once, x written once."""
        return (self.nnz * (value_bytes + 4) + (self.n + 1) * 4
                + 2 * self.n * rhs * value_bytes)


def _segment_add(b, rows, contrib):
    order = np.argsort(rows, kind="stable")
    r = rows[order]
    c = contrib[order]
    starts = np.flatnonzero(np.r_[True, r[1:] != r[:-1]])
    sums = np.add.reduceat(c, starts, axis=0)
    b[r[starts]] += sums


def _from_columns(n, cols_of_entries, rows, vals):
    """CSC from (col, row, val) triples; rows sorted ascending within a column."""
    order = np.lexsort((rows, cols_of_entries))
    counts = np.bincount(cols_of_entries, minlength=n)
    colPtr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=colPtr[1:])
    return colPtr.astype(np.int32), rows[order].astype(np.int32), vals[order]


def laplacian2d_lower(nx, ny, diag=4.0, off=-1.0, diag_values=None):
    """Lower triangle incl. diagonal of the 2D 5-point Laplacian."""
    n = nx * ny
    i = np.arange(n, dtype=np.int64)
    x = i % nx
    y = i // nx
    d = np.full(n, diag) if diag_values is None else np.asarray(diag_values, np.float64)
    cols = [i, i[x < nx - 1], i[y < ny - 1]]
    rows = [i, i[x < nx - 1] + 1, i[y < ny - 1] + nx]
    vals = [d, np.full((x < nx - 1).sum(), off), np.full((y < ny - 1).sum(), off)]
    c, r, v = (np.concatenate(a) for a in (cols, rows, vals))
    p, ri, vv = _from_columns(n, c, r, v)
    return Factor(p, ri, vv, name=f"lap2d_{nx}x{ny}")


def laplacian3d_lower(nx, ny, nz, diag=6.0, off=-1.0):
    """Lower triangle incl. diagonal of the 3D 7-point Laplacian."""
    n = nx * ny * nz
    i = np.arange(n, dtype=np.int64)
    x = i % nx
    y = (i // nx) % ny
    z = i // (nx * ny)
    mx, my, mz = x < nx - 1, y < ny - 1, z < nz - 1
    c = np.concatenate([i, i[mx], i[my], i[mz]])
    r = np.concatenate([i, i[mx] + 1, i[my] + nx, i[mz] + nx * ny])
    v = np.concatenate([np.full(n, diag), np.full(mx.sum(), off), np.full(my.sum(), off),
                        np.full(mz.sum(), off)])
    p, ri, vv = _from_columns(n, c, r, v)
    return Factor(p, ri, vv, name=f"lap3d_{nx}x{ny}x{nz}")


def stencil27_lower(nx, ny, nz, diag=26.0, off=-1.0):
    """Lower triangle incl. diagonal of the 3D 27-point stencil (every grid
    neighbour within one step along each axis): 13 entries below the
    diagonal in an interior column, reaching one plane, one row and one
    column ahead."""
    n = nx * ny * nz
    i = np.arange(n, dtype=np.int64)
    x = i % nx
    y = (i // nx) % ny
    z = i // (nx * ny)
    cols, rows, vals = [i], [i], [np.full(n, diag)]
    for dz in (0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                d = dx + nx * dy + nx * ny * dz
                if d <= 0:
                    continue
                ok = ((x + dx >= 0) & (x + dx < nx) & (y + dy >= 0) & (y + dy < ny)
                      & (z + dz < nz))
                cols.append(i[ok])
                rows.append(i[ok] + d)
                vals.append(np.full(int(ok.sum()), off))
    p, ri, vv = _from_columns(n, np.concatenate(cols), np.concatenate(rows),
                              np.concatenate(vals))
    return Factor(p, ri, vv, name=f"stencil27_{nx}x{ny}x{nz}")


def bidiagonal(n, diag=2.0, sub=-1.0):
    i = np.arange(n, dtype=np.int64)
    c = np.concatenate([i, i[:-1]])
    r = np.concatenate([i, i[:-1] + 1])
    v = np.concatenate([np.full(n, diag), np.full(n - 1, sub)])
    p, ri, vv = _from_columns(n, c, r, v)
    return Factor(p, ri, vv, name=f"bidiag_{n}")


def chain_far(n, far=9000, diag=2.0, sub=-1.0, farval=-0.5):
    """A bidiagonal chain whose rows also take an entry `far` columns back:
    L(i, i) = diag, L(i, i-1) = sub, L(i, i-far) = farval (i >= far).  The
    DAG depth is n (the chain), but no link is a pure first-order recurrence
    (row i has a second input and column i-1 a second entry), so a solver
    passes the chain batch by batch: its time over n/64 is the cost of one
    step from a 64-column batch to the next.  Diagonally dominant."""
    i = np.arange(n, dtype=np.int64)
    fi = i[far:]
    c = np.concatenate([i, i[:-1], fi - far])
    r = np.concatenate([i, i[:-1] + 1, fi])
    v = np.concatenate([np.full(n, diag), np.full(n - 1, sub), np.full(fi.size, farval)])
    p, ri, vv = _from_columns(n, c, r, v)
    return Factor(p, ri, vv, name=f"chain_far_{n}_{far}")


def power_law_lower(n, s=2.0, cap=1000, window=50000, seed=0):
    """Random power-law lower factor.

    Row i draws k_i = min(Zipf(s), cap, w_i) off-diagonal columns, w_i = i -
    max(0, i - window), uniformly without replacement from [i - w_i, i);
    values U(-1, 1); diagonal 1 + sum |row|.  Draw order with
    numpy.random.default_rng(seed): all k (zipf); the columns of rows with
    2 k_i <= w_i (vectorised integers, duplicates redrawn until none remain);
    the columns of the remaining rows, one rng.choice(w_i, k_i, replace=False)
    per row in ascending row order; then the values (uniform).  The x vector is
    drawn by the caller.
    """
    rng = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.int64)
    w = i - np.maximum(0, i - window)
    k = np.minimum(np.minimum(rng.zipf(s, size=n), cap), w).astype(np.int64)
    dense = 2 * k > w
    ks = np.where(dense, 0, k)
    rows = np.repeat(i, ks)
    lo = np.repeat(i - w, ks)
    span = np.repeat(w, ks)
    cols = lo + rng.integers(0, np.maximum(span, 1))
    # redraw duplicates within a row until every row's columns are distinct
    for _ in range(200):
        key = rows * (n + 1) + cols
        order = np.argsort(key, kind="stable")
        kk = key[order]
        dup = np.zeros(len(kk), bool)
        dup[1:] = kk[1:] == kk[:-1]
        if not dup.any():
            break
        idx = order[dup]
        cols[idx] = lo[idx] + rng.integers(0, span[idx])
    else:
        raise RuntimeError("power_law_lower: duplicate redraw did not converge")
    drows, dcols = [], []
    for r in np.flatnonzero(dense):
        pick = rng.choice(int(w[r]), size=int(k[r]), replace=False)
        drows.append(np.full(len(pick), r, np.int64))
        dcols.append(r - w[r] + pick.astype(np.int64))
    if drows:
        rows = np.concatenate([rows] + drows)
        cols = np.concatenate([cols] + dcols)
    vals = rng.uniform(-1.0, 1.0, size=len(rows))
    absum = np.bincount(rows, weights=np.abs(vals), minlength=n)
    diag = 1.0 + absum
    c = np.concatenate([i, cols])
    r = np.concatenate([i, rows])
    v = np.concatenate([diag, vals])
    p, ri, vv = _from_columns(n, c, r, v)
    return Factor(p, ri, vv, name=f"powerlaw_{n}_s{s}_seed{seed}")


def laplacian2d_batch(nsys=512, lo=32, hi=256, seed=0):
    """nsys independent 2D Laplacian lower factors as one block-diagonal factor.

    Sides nx, ny = floor(exp(U(ln lo, ln hi))) and diagonals 4 + U(0, 1), drawn
    from numpy.random.default_rng(seed) in that order (all sides, then all
    diagonal values, system by system).
    """
    rng = np.random.default_rng(seed)
    sides = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=(nsys, 2)))).astype(int)
    sizes = sides[:, 0] * sides[:, 1]
    sys_ptr = np.zeros(nsys + 1, np.int64)
    np.cumsum(sizes, out=sys_ptr[1:])
    diags = 4.0 + rng.uniform(0.0, 1.0, size=int(sys_ptr[-1]))
    ptrs, idxs, vals = [], [], []
    base_nnz = 0
    for s in range(nsys):
        nx, ny = int(sides[s, 0]), int(sides[s, 1])
        f = laplacian2d_lower(nx, ny, diag_values=diags[sys_ptr[s]:sys_ptr[s + 1]])
        ptrs.append(f.colPtr[:-1].astype(np.int64) + base_nnz)
        idxs.append(f.rowIdx.astype(np.int64) + sys_ptr[s])
        vals.append(f.val)
        base_nnz += f.nnz
    colPtr = np.concatenate(ptrs + [np.array([base_nnz])]).astype(np.int32)
    fac = Factor(colPtr, np.concatenate(idxs), np.concatenate(vals), sys_ptr=sys_ptr,
                 name=f"lap2d_batch_{nsys}_seed{seed}")
    fac.sides = sides
    return fac


def integer_unit_lower(n, density=4, seed=0, backward=False):
    """Known-answer factor in the reference's own style (main.cu:163-186):
    random off-diagonal pattern with integer values 1..10 and a unit diagonal.
    With integer x in 1..10 every partial sum is an exact integer well below
    2**53 as long as n*density*100 stays small, so any summation order gives
    the same FP64 result."""
    rng = np.random.default_rng(seed)
    i = np.arange(n, dtype=np.int64)
    k = np.minimum(rng.integers(0, density + 1, size=n), i if not backward else n - 1 - i)
    rows = np.repeat(i, k)
    if backward:
        cols = rows + 1 + rng.integers(0, np.maximum(np.repeat(n - 1 - i, k), 1))
    else:
        cols = rng.integers(0, np.maximum(np.repeat(i, k), 1))
    key = np.unique(rows * n + cols)
    rows, cols = key // n, key % n
    vals = rng.integers(1, 11, size=len(rows)).astype(np.float64)
    c = np.concatenate([i, cols])
    r = np.concatenate([i, rows])
    v = np.concatenate([np.ones(n), vals])
    p, ri, vv = _from_columns(n, c, r, v)
    return Factor(p, ri, vv, name=f"intunit_{n}_{'bwd' if backward else 'fwd'}")


def upper_from_lower(f):
    """U = L^T as a CSC factor (rows ascending, diagonal last): backward solves."""
    rows_of = f.rowIdx.astype(np.int64)
    cols_of = np.repeat(np.arange(f.n, dtype=np.int64), np.diff(f.colPtr))
    # entry (r, c) of L becomes (c, r) of U: U's column r holds row c
    p, ri, vv = _from_columns(f.n, rows_of, cols_of, f.val)
    return Factor(p, ri, vv, name=f.name + "_T")


def csc_to_csr(f):
    """Host CSR of a CSC factor (rows' columns ascending)."""
    cols_of = np.repeat(np.arange(f.n, dtype=np.int64), np.diff(f.colPtr))
    rows_of = f.rowIdx.astype(np.int64)
    order = np.lexsort((cols_of, rows_of))
    counts = np.bincount(rows_of, minlength=f.n)
    rowPtr = np.zeros(f.n + 1, np.int64)
    np.cumsum(counts, out=rowPtr[1:])
    return rowPtr.astype(np.int32), cols_of[order].astype(np.int32), f.val[order]


def x_uniform(n, rhs=1, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, size=(n, rhs) if rhs > 1 else n)
    return x


def x_normal(n, rhs, seed=1):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, rhs))
