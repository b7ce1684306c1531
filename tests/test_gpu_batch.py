"""Batched solve of uneven problems (gmgs_batch_solve) against one-by-one solves.

Each problem's kernels run in the same order in both modes, so the batched
displacements must be bitwise identical to the individual ones."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_matches_individual(pkg):
    probs = pkg.workloads.batch_c4(seed=0, count=6)
    solvers = []
    for p in probs:
        s = pkg.Solver(p.vox, p.bbox, num_levels=p.num_levels, tol=1e-6, max_it=3000,
                       precision=pkg.GMGS_PREC_MIXED, flags=pkg.FLAG_NO_INDEX_TABLES)
        s.set_bc(p.fixed, p.loads)
        solvers.append(s)
    ref = [s.solve(raise_on_nonconvergence=False)[:2] for s in solvers]
    out = pkg.batch_solve(solvers)
    for (U, info), (Ur, ir), p in zip(out, ref, probs):
        print(f"  {p.name}: dofs {U.size} L {p.num_levels} its {info.its}/{ir.its} "
              f"relres {info.relres:.2e}", flush=True)
        assert info.its == ir.its and info.converged == ir.converged
        assert np.array_equal(U, Ur)
        assert info.true_relres < 1e-5
    for s in solvers:
        s.close()
