"""The reference's own demo models end to end: kitten and parts surface
meshes voxelised at finestResolutionControl_ 64 (CreateVoxilizedModel.m),
their shipped boundary-condition files (Demo.m:16-19,41-44; copies under
tests/golden/models), GetDeformation at the reference's default tolerance
through the GMGS3D mirror, against the oracle: iterations within 1,
iterationHist_ relres, U within 1e-6, stress within 1e-6."""
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODELS = Path(__file__).resolve().parent / "golden" / "models"


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["kitten", "parts"])
def test_demo_model_64(pkg, oracle, name):
    log(f"[models] {name} 64")
    g = pkg.GMGS3D()
    g.CreateVoxilizedModel(MODELS / f"{name}.ply.gz", 64)
    vox, bbox = g.voxelizedVolume_.copy(), g.boundingBox_.copy()   # before padding stretches it
    g.fixingCond_ = np.loadtxt(MODELS / f"demo_BC_{name}_64_fixedNodes.dat").astype(np.int32).ravel()
    g.loadingCond_ = np.loadtxt(MODELS / f"demo_BC_{name}_64_loadingVec.dat").reshape(-1, 4)
    U = g.GetDeformation()
    S = g.GetCartesianStress()
    m = oracle.build(vox, bbox, omega=g.weightFactorJacobi_)
    assert np.allclose(g.boundingBox_, m.boundingBox, rtol=1e-14)   # the padded box of both
    res = oracle.get_deformation(m, g.loadingCond_, g.fixingCond_, g.tol_, g.maxIT_)
    e = np.linalg.norm(U - res.x) / np.linalg.norm(res.x)
    its = g.iterationHist_.shape[0] + 1
    log(f"  {U.size} DOFs, {g.solver.num_levels} levels: its gpu {its} oracle {res.its}, U rel {e:.2e}")
    assert U.size == m.levels[0].numDOFs
    assert abs(its - res.its) <= 1 and e < 1e-6
    k = min(g.iterationHist_.shape[0], res.hist.shape[0])
    assert np.allclose(g.iterationHist_[:k, 2], res.hist[:k, 2], rtol=1e-6)
    Sr = oracle.cartesian_stress(m, res.x)
    assert np.abs(S - Sr).max() <= 1e-6 * np.abs(Sr).max()
    g.solver.close()
