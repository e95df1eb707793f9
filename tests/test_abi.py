"""The C-ABI library loads and exports what include/sptrsv_hip.h declares;
host-only entry points agree with the oracle (CPU only)."""
import os
import re

import numpy as np
import pytest

import sptrsv_amd as S
import sptrsv_oracle as orc
from sptrsv_amd import matrices as mx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sptrsv_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sptrsv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = S.lib()
    decl = declared_functions()
    assert len(decl) >= 40
    missing = [f for f in decl if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(S.EXPORTS) == decl


def test_constants_keep_reference_values():
    # SpTRSV_cuda/common.h:25-33
    assert (S.SUBSTITUTION_FORWARD, S.SUBSTITUTION_BACKWARD) == (0, 1)
    assert (S.OPT_WARP_NNZ, S.OPT_WARP_RHS, S.OPT_WARP_AUTO) == (1, 2, 3)
    assert (S.COL_PARTITION, S.NNZ_PARTITION) == (0, 1)
    hdr = open(os.path.join(ROOT, "include", "sptrsv_hip.h")).read()
    for name, val in [("SPTRSV_SUBSTITUTION_BACKWARD", 1), ("SPTRSV_OPT_WARP_AUTO", 3),
                      ("SPTRSV_ERR_NOT_SQUARE", "(-1)")]:
        assert re.search(rf"#define {name} {re.escape(str(val))}", hdr)


def test_status_strings_and_version():
    lib = S.lib()
    assert lib.sptrsv_status_string(0) == b"ok"
    assert b"square" in lib.sptrsv_status_string(-1)
    assert b"gfx950" in lib.sptrsv_version()


def test_partitioners_match_oracle():
    for nparts, cols in [(4, 10), (3, 8), (7, 100), (1, 5)]:
        a = S.equal_col_partition(nparts, cols)
        b = orc.equal_col_partition(nparts, cols)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    rng = np.random.default_rng(0)
    for trial in range(20):
        ncols = int(rng.integers(5, 400))
        ptr = np.r_[0, np.cumsum(rng.integers(1, 30, size=ncols))].astype(np.int32)
        nparts = int(rng.integers(2, 9))
        a = S.equal_nnz_partition(nparts, ptr)
        b = orc.equal_nnz_partition(nparts, ptr)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("f", [mx.laplacian2d_lower(40, 30), mx.laplacian3d_lower(6, 5, 4),
                               mx.power_law_lower(3000, seed=3), mx.bidiagonal(500)])
def test_levels_match_oracle(f):
    assert S.levels(f.colPtr, f.rowIdx) == orc.findlevel(f.colPtr, f.rowIdx)


def test_no_device_is_an_error_not_a_fallback():
    if S.device_count() > 0:
        pytest.skip("a GPU is present")
    f = mx.bidiagonal(16)
    x = np.zeros(16)
    st, _ = S.sptrsv_syncfree(f.colPtr, f.rowIdx, f.val, 16, 16, f.nnz, 0, 1, 3, x,
                              f.spmv(np.ones(16)))
    assert st == -10  # SPTRSV_ERR_NO_DEVICE
    assert not x.any()


def test_not_square_keeps_reference_code():
    f = mx.bidiagonal(8)
    x = np.zeros(8)
    st, _ = S.sptrsv_syncfree(f.colPtr, f.rowIdx, f.val, 8, 9, f.nnz, 0, 1, 3, x, np.ones(8))
    assert st == -1  # sptrsv_syncfree_cuda.h:423-427


def test_generators_match_config_sizes():
    f1 = mx.laplacian2d_lower(256, 256)
    assert (f1.n, f1.nnz) == (65536, 196096)
    assert f1.algorithmic_bytes() == 3663876
    f5b = mx.bidiagonal(1_000_000)
    assert (f5b.n, f5b.nnz) == (1_000_000, 1_999_999)
    assert f5b.algorithmic_bytes() == 43999992


def test_batch_generator_is_block_diagonal():
    f = mx.laplacian2d_batch(8, seed=0)
    assert f.sys_ptr[0] == 0 and f.sys_ptr[-1] == f.n
    cols = np.repeat(np.arange(f.n), np.diff(f.colPtr))
    sys_of = np.searchsorted(f.sys_ptr, np.arange(f.n), side="right") - 1
    assert np.array_equal(sys_of[cols], sys_of[f.rowIdx])
    assert np.all((f.sides >= 32) & (f.sides <= 256))
