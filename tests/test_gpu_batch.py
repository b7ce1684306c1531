"""Batched solve of uneven problems (gmgs_batch_solve) against one-by-one solves.

Each problem's kernels run in the same order in both modes, so the batched
displacements must be bitwise identical to the individual ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solvers(pkg, probs, cycle):
    solvers = []
    for p in probs:
        if cycle == "reference":
            kw = dict(flags=pkg.FLAG_NO_INDEX_TABLES)
        else:   # the symmetric Chebyshev cycle of the benchmark, V(1,1) on level 1
            kw = dict(flags=pkg.FLAG_NO_INDEX_TABLES | pkg.FLAG_SYM_VCYCLE | pkg.FLAG_NO_JACOBI_STEP
                      | pkg.FLAG_CHEBYSHEV | pkg.FLAG_EXACT_GALERKIN,
                      vcycle_sweeps=1, vcycle_sweeps_level2=2, vcycle_sweeps_coarse=4)
        s = pkg.Solver(p.vox, p.bbox, num_levels=p.num_levels, tol=1e-6, max_it=3000,
                       precision=pkg.GMGS_PREC_MIXED, **kw)
        s.set_bc(p.fixed, p.loads)
        solvers.append(s)
    return solvers


def _check(out, ref, probs):
    for (U, info), (Ur, ir), p in zip(out, ref, probs):
        print(f"  {p.name}: dofs {U.size} L {p.num_levels} its {info.its}/{ir.its} "
              f"relres {info.relres:.2e}", flush=True)
        assert info.its == ir.its and info.converged == ir.converged
        assert np.array_equal(U, Ur)
        assert info.true_relres < 1e-5


@pytest.mark.parametrize("cycle", ["reference", "symmetric"])
def test_batch_matches_individual(pkg, cycle):
    """Batched against one-by-one solves, twice (the second batch replays the
    kept graphs, including the one of the problems left once half converged),
    and again after one problem's BCs were set anew (its graphs are dropped)."""
    probs = pkg.workloads.batch_c4(seed=0, count=6)
    solvers = _solvers(pkg, probs, cycle)
    ref = [s.solve(raise_on_nonconvergence=False)[:2] for s in solvers]
    _check(pkg.batch_solve(solvers), ref, probs)
    _check(pkg.batch_solve(solvers), ref, probs)
    p = probs[2]
    loads = p.loads.copy()
    loads[:, 1:] *= 2.0                  # rows [node fx fy fz]
    solvers[2].set_bc(p.fixed, loads)
    ref[2] = solvers[2].solve(raise_on_nonconvergence=False)[:2]
    _check(pkg.batch_solve(solvers), ref, probs)
    for s in solvers:
        s.close()


def test_batch_rejects_a_context_twice(pkg):
    p = pkg.workloads.batch_c4(seed=0, count=1)[0]
    s = _solvers(pkg, [p], "reference")[0]
    with pytest.raises(pkg.GMGSError):
        pkg.batch_solve([s, s])
    s.close()
