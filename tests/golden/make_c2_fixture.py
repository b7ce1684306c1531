"""Writes tests/golden/c2_fp64_oracle.npz: the oracle's FP64 solve of the
169x147x256 ellipsoid-union problem (seed 0, 5 levels, 7,290,603 DOFs) with
the reference algorithm (kernel/CG_solver.m with the kernel/Vcycle.m cycle),
reduced to what a GPU test can compare without re-running it:

* ``its6``, ``relres_hist6``: iterations and iterationHist_ relres column at
  rtol 1e-6 (CG_solver.m:35-49), ``unorm6`` and ``u6`` = U on ``dof_idx``;
* ``its10``, ``unorm10``, ``u10``: the same at rtol 1e-10 (the reference to
  which a mixed-precision solve is compared at tight tolerance);
* ``stress10`` = GetCartesianStress of the rtol-1e-10 solution on the
  compact nodes ``node_idx`` (kernel/GetCartesianStress.m);
* ``dofs``, ``levels``, ``normB``.

The sample indices come from numpy's default_rng(2024).  Run from the
repository root (about 15 minutes on 8 cores, ~20 GB of memory):

    python tests/golden/make_c2_fixture.py
"""
from __future__ import annotations

import importlib.util
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(REPO))


def load_workloads():
    spec = importlib.util.spec_from_file_location("gmgs_workloads", REPO / "gmgs-3d_amd" / "workloads.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gmgs_workloads"] = mod          # dataclasses look the module up
    spec.loader.exec_module(mod)
    return mod


def main():
    from oracle import gmgs_oracle as oracle
    wl = load_workloads()
    t0 = time.time()
    p = wl.ellipsoids_c2(0)
    m = oracle.build(p.vox, p.bbox, num_levels=p.num_levels, omega=p.omega)
    oracle.apply_bc(m, p.loads, p.fixed)
    oracle.factor_coarsest(m)
    n = m.levels[0].numDOFs
    print(f"set-up {time.time() - t0:.0f} s: {n} DOFs, {m.numLevels} levels", flush=True)
    r6 = oracle.cg_solver(m, m.F, 1e-6, 20000)
    print(f"rtol 1e-6: {r6.its} its, {time.time() - t0:.0f} s", flush=True)
    r10 = oracle.cg_solver(m, m.F, 1e-10, 20000)
    print(f"rtol 1e-10: {r10.its} its, {time.time() - t0:.0f} s", flush=True)
    S = oracle.cartesian_stress(m, r10.x)
    rng = np.random.default_rng(2024)
    dof_idx = np.sort(rng.choice(n, size=10000, replace=False))
    node_idx = np.sort(rng.choice(m.levels[0].numNodes, size=2000, replace=False))
    out = REPO / "tests" / "golden" / "c2_fp64_oracle.npz"
    np.savez_compressed(out, dofs=n, levels=m.numLevels, normB=np.linalg.norm(m.F),
                        its6=r6.its, relres_hist6=r6.hist[:, 2], unorm6=np.linalg.norm(r6.x),
                        u6=r6.x[dof_idx], its10=r10.its, unorm10=np.linalg.norm(r10.x),
                        u10=r10.x[dof_idx], dof_idx=dof_idx, node_idx=node_idx,
                        stress10=S[node_idx], relres6=r6.relres, relres10=r10.relres)
    print(f"wrote {out} ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
