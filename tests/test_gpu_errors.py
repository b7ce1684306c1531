"""Failure paths of the C ABI (status codes of include/gmgs.h).

The reference's own checks: too few voxels for the hierarchy
(DiscretizeDesignDomain.m:30-32), fewer than two levels (SetupSolver.m:19-21),
no loads / no constraints (ApplyBoundaryCondition.m:11-12,22); and the ones
this library adds: node ids out of range, non-finite inputs, a singular
coarsest matrix (the reference ignores the Cholesky flag, SetupSolver.m:259),
and call order.  A failed gmgs_set_bc leaves the context without BCs, so a
following solve is refused instead of running on half-rebuilt state.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def brick(pkg, shape=(8, 8, 16), L=3, **kw):
    vox = np.ones(shape, np.int32)
    ny, nx, nz = shape
    s = pkg.Solver(vox, np.array([[0, 0, 0], [nx, ny, nz]], float), num_levels=L, tol=1e-8,
                   max_it=500, **kw)
    nn = s.num_nodes(1)
    # bottom face nodes: full ids with iz = 0 are compact ids 1..(ny+1)(nx+1)
    fixed = np.arange(1, (ny + 1) * (nx + 1) + 1, dtype=np.int32)
    top = np.arange(nn - (ny + 1) * (nx + 1) + 1, nn + 1)
    loads = np.concatenate([top[:, None], np.tile([0.0, 0.0, -1.0], (top.size, 1))], axis=1)
    return s, fixed, loads


def status_of(excinfo):
    return excinfo.value.code


def test_create_errors(pkg):
    log("[errors] create")
    with pytest.raises(pkg.GMGSError) as e:        # 2x2x2 voxels < coarsestResolutionControl_
        pkg.Solver(np.ones((2, 2, 2), np.int32))
    assert status_of(e) == 2
    with pytest.raises(pkg.GMGSError) as e:        # one level asked for
        pkg.Solver(np.ones((4, 4, 4), np.int32), num_levels=1)
    assert status_of(e) == 3
    with pytest.raises(pkg.GMGSError) as e:        # no solid voxel
        pkg.Solver(np.zeros((8, 8, 8), np.int32), num_levels=2)
    assert status_of(e) == 2
    with pytest.raises(pkg.GMGSError) as e:
        pkg.Solver(np.ones((4, 4, 4), np.int32), num_levels=2, poisson_ratio=0.5)
    assert status_of(e) == 1
    with pytest.raises(pkg.GMGSError) as e:
        pkg.Solver(np.ones((4, 4, 4), np.int32), num_levels=2, bbox=None,
                   precision=1, flags=pkg.FLAG_CHEBYSHEV)
    assert status_of(e) == 1
    with pytest.raises(pkg.GMGSError) as e:
        pkg.Solver(np.ones((4, 4, 4), np.int32), np.array([[0, 0, 0], [np.nan, 1, 1]]), num_levels=2)
    assert status_of(e) == 7


def test_set_bc_errors_and_state(pkg):
    log("[errors] set_bc")
    s, fixed, loads = brick(pkg)
    with pytest.raises(pkg.GMGSError) as e:        # solve before any BCs
        s.solve()
    assert status_of(e) == 10
    nn = s.num_nodes(1)
    bad = loads.copy()
    bad[0, 0] = nn + 1
    with pytest.raises(pkg.GMGSError) as e:
        s.set_bc(fixed, bad)
    assert status_of(e) == 6
    bad = loads.copy()
    bad[0, 0] = 1.5                                # not an integer node id
    with pytest.raises(pkg.GMGSError) as e:
        s.set_bc(fixed, bad)
    assert status_of(e) == 6
    fx = fixed.copy()
    fx[3] = 0
    with pytest.raises(pkg.GMGSError) as e:
        s.set_bc(fx, loads)
    assert status_of(e) == 6
    bad = loads.copy()
    bad[2, 3] = np.nan
    with pytest.raises(pkg.GMGSError) as e:
        s.set_bc(fixed, bad)
    assert status_of(e) == 7
    with pytest.raises(pkg.GMGSError) as e:
        s.set_bc(fixed, np.zeros((0, 4)))
    assert status_of(e) == 4
    with pytest.raises(pkg.GMGSError) as e:
        s.set_bc(np.zeros(0, np.int32), loads)
    assert status_of(e) == 5
    # a valid set_bc, a solve, then a failing set_bc: the solution and the
    # BCs are gone, solve / stress / adjoint are refused
    s.set_bc(fixed, loads)
    U, info, _ = s.solve()
    assert info.converged
    with pytest.raises(pkg.GMGSError) as e:
        s.set_bc(fixed, np.zeros((0, 4)))
    assert status_of(e) == 4
    for call in (lambda: s.solve(), lambda: s.stress(None), lambda: s.adjoint(np.zeros(U.size))):
        with pytest.raises(pkg.GMGSError) as e:
            call()
        assert status_of(e) == 10
    # the context recovers with valid BCs and solves as before
    s.set_bc(fixed, loads)
    U2, info2, _ = s.solve()
    assert np.array_equal(U, U2) and info2.its == info.its
    s.close()


def test_singular_coarsest_matrix(pkg):
    """Two solid blocks separated by a gap wider than a coarsest element,
    only one of them fixed: the other floats, its rigid-body modes make the
    coarsest K_ff singular, and gmgs_set_bc reports it; the context is then
    without BCs."""
    log("[errors] singular coarsest matrix")
    vox = np.zeros((8, 8, 24), np.int32)
    vox[:, :, 0:8] = 1
    vox[:, :, 16:24] = 1
    s = pkg.Solver(vox, np.array([[0, 0, 0], [8, 8, 24.0]]), num_levels=3, tol=1e-8, max_it=200)
    nxy = 81
    fixed = np.arange(1, nxy + 1, dtype=np.int32)            # z = 0 face of the lower block
    nn = s.num_nodes(1)
    top = np.arange(nn - nxy + 1, nn + 1)                    # z = 24 face of the upper block
    loads = np.concatenate([top[:, None], np.tile([0.0, 0.0, -1.0], (top.size, 1))], axis=1)
    with pytest.raises(pkg.GMGSError) as e:
        s.set_bc(fixed, loads)
    assert status_of(e) == 8
    with pytest.raises(pkg.GMGSError) as e:
        s.solve()
    assert status_of(e) == 10
    # fixing a face of each block makes it solvable
    # compact ids run over the lower block's 9 node planes, then the upper
    # block's: its z = 16 face is ids nn/2 + 1 .. nn/2 + 81
    assert nn == 2 * 9 * nxy
    fixed2 = np.concatenate([fixed, np.arange(nn // 2 + 1, nn // 2 + nxy + 1)]).astype(np.int32)
    s.set_bc(fixed2, loads)
    U, info, _ = s.solve()
    assert info.converged and info.true_relres < 1e-7
    s.close()


def test_argument_and_adjoint_errors(pkg):
    log("[errors] arguments")
    s, fixed, loads = brick(pkg)
    s.set_bc(fixed, loads)
    with pytest.raises(pkg.GMGSError) as e:
        s.atx(np.zeros(3 * s.num_nodes(1)), level=9)
    assert status_of(e) == 1
    with pytest.raises(pkg.GMGSError) as e:
        s.set_modulus(np.ones(s.num_elements(1) + 1))
    assert status_of(e) == 1
    m = np.ones(s.num_elements(1))
    m[5] = -1.0
    with pytest.raises(pkg.GMGSError) as e:
        s.set_modulus(m)
    assert status_of(e) == 7
    U, info, _ = s.solve()
    g = np.zeros(U.size)
    g[7] = np.inf
    with pytest.raises(pkg.GMGSError) as e:         # rejected before b is touched
        s.adjoint(g, U)
    assert status_of(e) == 7
    U2, info2, _ = s.solve()                         # still solves K u = F
    assert np.array_equal(U, U2)
    # an adjoint with g = f leaves the load vector as it was
    lam, _, _ = s.adjoint(s.query("F"), U)
    assert np.linalg.norm(lam - U) / np.linalg.norm(U) < 1e-6
    U3, _, _ = s.solve()
    assert np.array_equal(U, U3)
    with pytest.raises(pkg.GMGSError) as e:
        s.query("SMOOTHER", s.num_levels)             # the coarsest level is not smoothed
    assert status_of(e) == 1
    s.close()
