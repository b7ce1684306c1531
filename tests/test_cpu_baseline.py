"""The C++/OpenMP CPU baseline (oracle/cpu_baseline.cpp) against the numpy
oracle's CG_solver on the same models: the same iteration count and the same
solution to rounding, with the three coarse storing states present (4 levels
give states 2, 2, 1) and with the level-3 quirk firing."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def cpu():
    from oracle import cpu_baseline
    cpu_baseline.build()
    return cpu_baseline


def _bc(oracle, m, nz, seed=3):
    rng = np.random.default_rng(seed)
    L1 = m.levels[0]
    nyn, nxn = L1.resY + 1, L1.resX + 1
    iz = (L1.nodMapBack - 1) // (nyn * nxn)
    fixed = np.nonzero(iz == 0)[0] + 1
    top = np.nonzero(iz == nz)[0] + 1
    loads = np.concatenate([top[:, None], rng.standard_normal((top.size, 3))], axis=1)
    oracle.apply_bc(m, loads, fixed)
    oracle.factor_coarsest(m)


@pytest.mark.parametrize("L", [3, 4])
def test_cpu_baseline_matches_oracle(oracle, cpu, L):
    rng = np.random.default_rng(7)
    vox = np.ones((16, 16, 24), np.int32)
    blocks = rng.random((8, 8, 12)) < 0.08
    inner = np.kron(blocks, np.ones((2, 2, 2), bool)).astype(bool)
    inner[:2] = inner[-2:] = False
    inner[:, :2] = inner[:, -2:] = False
    inner[:, :, :2] = inner[:, :, -2:] = False
    vox[inner] = 0
    m = oracle.build(vox, np.array([[0, 0, 0], [16, 16, 24.0]]), num_levels=L, omega=0.5)
    _bc(oracle, m, 24)
    ref = oracle.cg_solver(m, m.F, 1e-9, 500)
    x, its, conv, rel, sec = cpu.Prepared(m).solve(m.F, 1e-9, 500)
    assert conv and its == ref.its
    assert np.linalg.norm(x - ref.x) / np.linalg.norm(ref.x) < 1e-10
    assert abs(rel - ref.relres) <= 1e-6 * ref.relres
