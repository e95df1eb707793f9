"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py
with scipy's spsolve_triangular): the oracle on the CPU and the library on
the GPU must both reproduce them."""
import glob
import os

import numpy as np
import pytest

import sptrsv_oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    z = np.load(path)  # allow_pickle stays False
    return {k: z[k] for k in z.files}


def rel_inf(x, ref):
    return float(np.abs(x - ref).max() / np.abs(ref).max())


def test_fixtures_present():
    assert len(CASES) >= 5


@pytest.mark.parametrize("path", CASES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_golden(path):
    g = load(path)
    rhs, sub = int(g["rhs"]), int(g["substitution"])
    x = orc.solve(g["colPtr"], g["rowIdx"], g["val"], g["b"].reshape(-1), substitution=sub,
                  rhs=rhs).reshape(g["x"].shape)
    assert rel_inf(x, g["x"]) <= 1e-13
    if "intunit" in path:
        assert np.array_equal(x, g["x_true"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", CASES, ids=lambda p: os.path.basename(p)[:-4])
def test_library_reproduces_golden(path):
    from gpu_util import plan_solve, require_gpu
    from sptrsv_amd import matrices as mx
    require_gpu()
    g = load(path)
    rhs, sub = int(g["rhs"]), int(g["substitution"])
    f = mx.Factor(g["colPtr"], g["rowIdx"], g["val"])
    import sptrsv_amd as S
    for cfg in (None, dict(lanes_per_column=1, block=256, window=64),
                dict(lanes_per_column=64, block=1024, window=16),
                dict(engine=S.ENGINE_LANE, block=256, window=100)):
        x, _ = plan_solve(f, g["b"], rhs=rhs, subst=sub, cfg=cfg)
        assert rel_inf(x, g["x"]) <= 1e-12
        if "intunit" in path:
            assert np.array_equal(x, g["x_true"])
