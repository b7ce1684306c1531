"""CPU oracle: an FP64 numpy restatement of GMGS-3D's solve path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker / the timed CPU baseline.

It follows the reference MATLAB line by line (paths relative to the
reference's ``src/``), including its quirks:

* B1  a damped-Jacobi step on (x, r) at the top of every CG iteration
      (kernel/CG_solver.m:18-19);
* B2  V-cycle post-smoothing with the stale pre-smoothed residual
      (kernel/Vcycle.m:32);
* B3  the level-3 "unique element" test that inspects only present level-2
      children (kernel/SetupSolver.m:96-101);
* B4  duplicate load rows: the last row wins (common/ApplyBoundaryCondition.m:9);
* B12 child order: ``elementUpwardMap`` lists the 8 children column-major
      (kernel/BuildingMeshHierarchy.m:45-46) while the Galerkin scatter uses
      SubEleNodMat's counter-clockwise child order (kernel/SubEleNodMat.m:4-11,
      kernel/SetupSolver.m:28-31), so child columns 3<->4 and 7<->8 trade
      places in every computed coarse Ks (SetupSolver.m:56,80,118,145,168,189).

Index arrays are kept 1-based exactly as MATLAB holds them, so they can be
compared bit for bit with what the device builds.

Pins (see tests/test_oracle.py): Ke0 closed forms and rigid-body null space,
the literal transfer tables (tests/golden/transfer_tables.json), brute-force
index construction on tiny masks, AtX against an assembled sparse K
(DirectlyAssembleStifnessMatrix.m:7-23), Galerkin nesting and the B12
permutation identity, R = P^T, the CG solution against scipy spsolve, the
stress patch test.  Parity unpinned (no reference run exists): MATLAB's
summation order inside sparse()/mtimes and its Cholesky permutation.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import List, Optional

import numpy as np

GAUSS = 0.577350269189626          # literal from common/GaussianIntegral.m:11-16

# reordering of [x1..x8 y1..y8 z1..z8] -> interleaved (kernel/SetupSolver.m:14)
REORDER = np.array([1, 9, 17, 2, 10, 18, 3, 11, 19, 4, 12, 20, 5, 13, 21, 6, 14, 22,
                    7, 15, 23, 8, 16, 24]) - 1


# ------------------------------------------------------------------ helpers
def mfind(a: np.ndarray) -> np.ndarray:
    """MATLAB find() on an N-d array: 1-based column-major linear indices."""
    return np.flatnonzero(np.asarray(a).ravel(order="F")) + 1


def matlab_colon(a: float, d: float, b: float) -> np.ndarray:
    """a:d:b evaluated from both ends, as MATLAB's colon does."""
    n = int(math.floor((b - a) / d + 1e-10))
    k = np.arange(n + 1, dtype=np.float64)
    out = a + k * d
    half = k >= n / 2.0
    out[half] = b - (n - k[half]) * d
    return out


# ------------------------------------------------------------------ element primitives
def gauss_points():
    """common/GaussianIntegral.m:11-17."""
    g = GAUSS
    s = np.array([-g, g, g, -g, -g, g, g, -g])
    t = np.array([-g, -g, g, g, -g, -g, g, g])
    p = np.array([-g, -g, -g, -g, g, g, g, g])
    w = np.ones(8)
    return s, t, p, w


def shape_function(s, t, p):
    """common/ShapeFunction.m:15-23."""
    s, t, p = (np.asarray(v, dtype=np.float64).ravel() for v in (s, t, p))
    N = np.zeros((s.size, 8))
    N[:, 0] = 0.125 * (1 - s) * (1 - t) * (1 - p)
    N[:, 1] = 0.125 * (1 + s) * (1 - t) * (1 - p)
    N[:, 2] = 0.125 * (1 + s) * (1 + t) * (1 - p)
    N[:, 3] = 0.125 * (1 - s) * (1 + t) * (1 - p)
    N[:, 4] = 0.125 * (1 - s) * (1 - t) * (1 + p)
    N[:, 5] = 0.125 * (1 + s) * (1 - t) * (1 + p)
    N[:, 6] = 0.125 * (1 + s) * (1 + t) * (1 + p)
    N[:, 7] = 0.125 * (1 - s) * (1 + t) * (1 + p)
    return N


def de_shape_function(s, t, p):
    """common/DeShapeFunction.m:3-22: rows 3g..3g+2 = d/ds, d/dt, d/dp."""
    sgn_s = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=np.float64)
    sgn_t = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=np.float64)
    sgn_p = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=np.float64)
    n = len(s)
    d = np.zeros((3 * n, 8))
    for g in range(n):
        d[3 * g + 0] = 0.125 * sgn_s * (1 + sgn_t * t[g]) * (1 + sgn_p * p[g])
        d[3 * g + 1] = 0.125 * sgn_t * (1 + sgn_s * s[g]) * (1 + sgn_p * p[g])
        d[3 * g + 2] = 0.125 * sgn_p * (1 + sgn_s * s[g]) * (1 + sgn_t * t[g])
    return d


def hooke(E: float, nu: float) -> np.ndarray:
    """HookeLaw, common/ElementElasticityMatrix.m:11-24."""
    S = np.zeros((6, 6))
    c1 = (1 + nu) * (1 - 2 * nu)
    c2 = 2 * (1 + nu)
    S[0, 0] = E * (1 - nu) / c1
    S[0, 1] = E * nu / c1
    S[0, 2] = S[0, 1]
    S[1, 0] = S[0, 1]
    S[1, 1] = S[0, 0]
    S[1, 2] = S[1, 0]
    S[2, 0] = S[0, 2]
    S[2, 1] = S[1, 2]
    S[2, 2] = S[1, 1]
    S[3, 3] = E / c2
    S[4, 4] = S[3, 3]
    S[5, 5] = S[4, 4]
    return S


def element_elasticity(E, nu):
    """common/ElementElasticityMatrix.m:2-8: 48x48 block diagonal."""
    return np.kron(np.eye(8), hooke(E, nu))


def element_strain_matrix(dshape, invJ):
    """common/ElementStrainMatrix.m:3-94: 48x24, DOF order [x1 y1 z1 ...]."""
    der = invJ @ dshape
    B = np.zeros((48, 24))
    for g in range(8):
        dx, dy, dz = der[3 * g], der[3 * g + 1], der[3 * g + 2]
        r = 6 * g
        for n in range(8):
            c = 3 * n
            B[r + 0, c + 0] = dx[n]
            B[r + 1, c + 1] = dy[n]
            B[r + 2, c + 2] = dz[n]
            B[r + 3, c + 1] = dz[n]
            B[r + 3, c + 2] = dy[n]
            B[r + 4, c + 0] = dz[n]
            B[r + 4, c + 2] = dx[n]
            B[r + 5, c + 0] = dy[n]
            B[r + 5, c + 1] = dx[n]
    return B


def probe_jacobians(probe_coords):
    """Per-Gauss-point inverse Jacobians and determinants of the probe element
    (common/SetupStiffnessMatrix.m:11-18; kernel/GetCartesianStress.m:24-30)."""
    s, t, p, w = gauss_points()
    dshape = de_shape_function(s, t, p)
    invJ = np.zeros((24, 24))
    detJ = np.zeros(8)
    for g in range(8):
        J = dshape[3 * g:3 * g + 3] @ probe_coords
        detJ[g] = np.linalg.det(J)
        invJ[3 * g:3 * g + 3, 3 * g:3 * g + 3] = np.linalg.inv(J)
    return dshape, invJ, detJ, w


def element_stiffness(E, nu, probe_coords):
    """Ke0 = B'(D.*wgt)B, common/SetupStiffnessMatrix.m:4-22."""
    dshape, invJ, detJ, w = probe_jacobians(probe_coords)
    B = element_strain_matrix(dshape, invJ)
    D = element_elasticity(E, nu)
    wgt = np.repeat(w * detJ, 6)
    return B.T @ (D * wgt[None, :]) @ B


def outer_interpolation_matrix():
    """OTP = inv(sFM), kernel/GetCartesianStress.m:44-73."""
    s, t, p, _ = gauss_points()
    N = shape_function(s, t, p)
    sFM = np.kron(N, np.eye(6))
    return np.linalg.inv(sFM)


def transfer_tables():
    """Span-2 inNODE (27x8), inDOF (81x24) and SubEleNodMat (8x8, 1-based),
    Operator4MultiGridRestrictionAndInterpolation.m:5-113, SubEleNodMat.m:4-11,
    built from their definition: fine node j = 1 + a + 3b + 9c (a = row
    offset, b = x, c = z); coarse local corners in hex8 order."""
    corners = [(2, 0, 0), (2, 2, 0), (0, 2, 0), (0, 0, 0),
               (2, 0, 2), (2, 2, 2), (0, 2, 2), (0, 0, 2)]
    inNODE = np.zeros((27, 8))
    for c in range(3):
        for b in range(3):
            for a in range(3):
                j = a + 3 * b + 9 * c
                for k, (ka, kb, kc) in enumerate(corners):
                    inNODE[j, k] = ((1 - abs(a - ka) / 2.0) * (1 - abs(b - kb) / 2.0)
                                    * (1 - abs(c - kc) / 2.0))
    inDOF = np.kron(inNODE, np.eye(3))
    # children in counter-clockwise order (cy,cx,cz), as SubEleNodMat.m:4-11
    child = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
             (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    local = [(1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 0),
             (1, 0, 1), (1, 1, 1), (0, 1, 1), (0, 0, 1)]
    sub = np.zeros((8, 8), dtype=np.int64)
    for ci, (cy, cx, cz) in enumerate(child):
        for li, (ly, lx, lz) in enumerate(local):
            sub[ci, li] = 1 + (cy + ly) + 3 * (cx + lx) + 9 * (cz + lz)
    return inNODE, inDOF, sub


HEX_LOCAL = np.array([[1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 0],
                      [1, 0, 1], [1, 1, 1], [0, 1, 1], [0, 0, 1]])  # (row, x, z)


# ------------------------------------------------------------------ mesh structures
@dataclasses.dataclass
class Level:
    resX: int
    resY: int
    resZ: int
    eleSize: np.ndarray
    numElements: int = 0
    numNodes: int = 0
    numDOFs: int = 0
    eleMapBack: np.ndarray = None       # 1-based full ids (int32)
    eleMapForward: np.ndarray = None    # full grid -> 1-based compact (0 void)
    nodMapBack: np.ndarray = None       # level 1: full ids; >=2: finest-grid ids
    nodMapForward: np.ndarray = None
    eNodMat: np.ndarray = None          # (Ne, 8) compact 1-based
    eDofMat: np.ndarray = None          # (Ne, 24) 1-based interleaved
    nodesOnBoundary: np.ndarray = None
    elementsOnBoundary: np.ndarray = None
    numNod2Eles: np.ndarray = None
    elementUpwardMap: np.ndarray = None  # (Ne, 8) finer compact ids, 0 = void
    transferMat: np.ndarray = None       # (27, Ne) 1-based intermediate ids
    transferMatCoeffi: np.ndarray = None
    solidNodeMapCoarser2Finer: np.ndarray = None
    intermediateNumNodes: int = 0
    storingState: int = 0
    Ke: np.ndarray = None
    Ks: np.ndarray = None               # state 0: (24,24); 1: (Ne,24,24); 2: (1+nu,24,24)
    uniqueElements: np.ndarray = None   # 1-based
    universalElements: np.ndarray = None
    diagK: np.ndarray = None
    fixedDOFs: np.ndarray = None
    freeDOFs: np.ndarray = None


@dataclasses.dataclass
class Model:
    levels: List[Level]
    numLevels: int
    boundingBox: np.ndarray
    nodeCoords: np.ndarray
    eNodMatTemp: np.ndarray
    E: float = 1.0
    nu: float = 0.3
    omega: float = 0.4
    Ke0: np.ndarray = None
    cholL: np.ndarray = None
    coarsest_K: np.ndarray = None
    F: np.ndarray = None
    exact_galerkin: bool = False        # True: no B3/B12 (for comparison only)
    modulus: Optional[np.ndarray] = None


def _enodmat_full(nx, ny, nz):
    """eNodMatTemp: full-grid connectivity of every cell, 1-based
    (DiscretizeDesignDomain.m:76-84 / BuildingMeshHierarchy.m:101-110)."""
    ny1, nx1 = ny + 1, nx + 1
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    base = (iy + 1) + ny1 * ix + ny1 * nx1 * iz + 1
    base = base.ravel(order="F").astype(np.int64)
    S = ny1 * nx1
    tmp = np.array([0, ny + 1, ny, -1, S, S + ny + 1, S + ny, S - 1], dtype=np.int64)
    return base[:, None] + tmp[None, :]


def discretize(vox: np.ndarray, bbox: np.ndarray, control: float = 1.0e4,
               num_levels: Optional[int] = None) -> Model:
    """common/DiscretizeDesignDomain.m:29-120.  ``num_levels`` overrides the
    level rule (the north-star interface takes the level count directly)."""
    vox = np.asarray(vox, dtype=np.int32)
    nely, nelx, nelz = vox.shape
    if num_levels is None:
        numVoxels = nelx * nely * nelz
        if numVoxels < control:
            raise ValueError("There is no Sufficient Resolution to Build Mesh Hierarchy!")
        L = 0
        while numVoxels >= control:
            L += 1
            numVoxels = nelx / 2 ** L * nely / 2 ** L * nelz / 2 ** L
    else:
        L = num_levels - 1
    ax = int(math.ceil(nelx / 2 ** L)) * 2 ** L
    ay = int(math.ceil(nely / 2 ** L)) * 2 ** L
    az = int(math.ceil(nelz / 2 ** L)) * 2 ** L
    numLevels = L + 1
    v = np.zeros((ay, ax, az), dtype=np.int32)
    v[:nely, :nelx, :nelz] = vox
    bb = np.array(bbox, dtype=np.float64).reshape(2, 3).copy()
    bb[1] = bb[0] + (bb[1] - bb[0]) * np.array([ax, ay, az]) / np.array([nelx, nely, nelz])
    nx, ny, nz = ax, ay, az
    lev = Level(nx, ny, nz, (bb[1] - bb[0]) / np.array([nx, ny, nz]))
    lev.eleMapBack = mfind(v == 1).astype(np.int64)
    lev.numElements = lev.eleMapBack.size
    lev.eleMapForward = np.zeros(nx * ny * nz, dtype=np.int64)
    lev.eleMapForward[lev.eleMapBack - 1] = np.arange(1, lev.numElements + 1)
    eTemp = _enodmat_full(nx, ny, nz)
    eNod = eTemp[lev.eleMapBack - 1]
    lev.nodMapBack = np.unique(eNod)
    lev.numNodes = lev.nodMapBack.size
    lev.numDOFs = 3 * lev.numNodes
    lev.nodMapForward = np.zeros((nx + 1) * (ny + 1) * (nz + 1), dtype=np.int64)
    lev.nodMapForward[lev.nodMapBack - 1] = np.arange(1, lev.numNodes + 1)
    lev.eNodMat = lev.nodMapForward[eNod - 1]
    xs = matlab_colon(bb[0, 0], (bb[1, 0] - bb[0, 0]) / nx, bb[1, 0])
    ys = matlab_colon(bb[1, 1], (bb[0, 1] - bb[1, 1]) / ny, bb[0, 1])
    zs = matlab_colon(bb[0, 2], (bb[1, 2] - bb[0, 2]) / nz, bb[1, 2])
    Y, X, Z = np.meshgrid(ys, xs, zs, indexing="ij")
    coords = np.stack([X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")], axis=1)
    cnt = np.zeros(lev.numNodes, dtype=np.int64)
    for k in range(8):
        np.add.at(cnt, lev.eNodMat[:, k] - 1, 1)
    lev.numNod2Eles = cnt
    lev.nodesOnBoundary = mfind(cnt < 8)
    onb = np.zeros(lev.numNodes, dtype=np.int64)
    onb[lev.nodesOnBoundary - 1] = 1
    lev.elementsOnBoundary = mfind(onb[lev.eNodMat - 1].sum(axis=1) > 0)
    return Model([lev], numLevels, bb, coords, eTemp)


def apply_bc(model: Model, loads: np.ndarray, fixed: np.ndarray):
    """common/ApplyBoundaryCondition.m:7-23 (B4: last duplicate row wins)."""
    lev = model.levels[0]
    F = np.zeros((lev.numNodes, 3))
    loads = np.asarray(loads, dtype=np.float64).reshape(-1, 4)
    ids = loads[:, 0].astype(np.int64)
    for r in range(loads.shape[0]):            # sequential: last write wins
        F[ids[r] - 1] = loads[r, 1:4]
    model.F = F.reshape(-1)                    # reshape(F_', numDOFs, 1)
    f = 3 * np.asarray(fixed, dtype=np.int64).ravel()
    fixedDOFs = np.stack([f - 2, f - 1, f], axis=1).reshape(-1)
    lev.fixedDOFs = fixedDOFs
    lev.freeDOFs = np.setdiff1d(np.arange(1, lev.numDOFs + 1), fixedDOFs)
    return model.F


def build_hierarchy(model: Model):
    """kernel/BuildingMeshHierarchy.m:7-143."""
    L1 = model.levels[0]
    nodeVolume = np.arange(1, (L1.resX + 1) * (L1.resY + 1) * (L1.resZ + 1) + 1,
                           dtype=np.int64).reshape(L1.resY + 1, L1.resX + 1, L1.resZ + 1,
                                                   order="F")
    eTemp = model.eNodMatTemp
    for ii in range(2, model.numLevels + 1):
        fine = model.levels[ii - 2]
        nx, ny, nz = fine.resX // 2, fine.resY // 2, fine.resZ // 2
        lev = Level(nx, ny, nz, fine.eleSize * 2)
        fwd = fine.eleMapForward.reshape(2 * ny, 2 * nx, 2 * nz, order="F")
        tmpl = np.arange(1, 8 * nx * ny * nz + 1, dtype=np.int64).reshape(
            2 * ny, 2 * nx, 2 * nz, order="F")

        def children(a):
            b = a.reshape(ny, 2, nx, 2, nz, 2).transpose(4, 2, 0, 5, 3, 1)
            return b.reshape(nz * nx * ny, 8)

        up = children(fwd)                      # elementUpwardMap (all cells)
        upTemp = children(tmpl)
        fy, fx = 2 * ny + 1, 2 * nx + 1
        gy, gx, gz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
        gy, gx, gz = (v.ravel(order="F") for v in (gy, gx, gz))
        tmt = np.zeros((27, nx * ny * nz), dtype=np.int64)
        for c in range(3):
            for b in range(3):
                for a in range(3):
                    tmt[a + 3 * b + 9 * c] = ((2 * gy + a) + fy * (2 * gx + b)
                                              + fy * fx * (2 * gz + c) + 1)
        unempty = mfind(up.sum(axis=1) > 0)
        upTemp = upTemp[unempty - 1]
        nodesIncVoid = np.unique(eTemp[upTemp.reshape(-1, order="F") - 1])
        lev.intermediateNumNodes = nodesIncVoid.size
        tmt = tmt[:, unempty - 1]
        temp = np.zeros(fy * fx * (2 * nz + 1), dtype=np.int64)
        temp[nodesIncVoid - 1] = np.arange(1, lev.intermediateNumNodes + 1)
        lev.transferMat = temp[tmt - 1]
        coef = np.zeros(lev.intermediateNumNodes)
        for kk in range(27):
            coef[lev.transferMat[kk] - 1] += 1     # distinct within a row
        lev.transferMatCoeffi = coef
        nodesLast = np.unique(eTemp[fine.eleMapBack - 1])
        _, ia, _ = np.intersect1d(nodesIncVoid, nodesLast, assume_unique=True,
                                  return_indices=True)
        lev.solidNodeMapCoarser2Finer = ia + 1
        lev.eleMapForward = np.zeros(nx * ny * nz, dtype=np.int64)
        lev.eleMapBack = unempty
        lev.numElements = unempty.size
        lev.eleMapForward[unempty - 1] = np.arange(1, lev.numElements + 1)
        lev.elementUpwardMap = up[unempty - 1]
        eTemp = _enodmat_full(nx, ny, nz)
        eNod = eTemp[unempty - 1]
        nmb = np.unique(eNod)
        lev.numNodes = nmb.size
        lev.numDOFs = 3 * lev.numNodes
        lev.nodMapForward = np.zeros((nx + 1) * (ny + 1) * (nz + 1), dtype=np.int64)
        lev.nodMapForward[nmb - 1] = np.arange(1, lev.numNodes + 1)
        lev.eNodMat = lev.nodMapForward[eNod - 1]
        s = 2 ** (ii - 1)
        tmp = nodeVolume[::s, ::s, ::s].ravel(order="F")
        lev.nodMapBack = tmp[nmb - 1]
        cnt = np.zeros(lev.numNodes, dtype=np.int64)
        for k in range(8):
            np.add.at(cnt, lev.eNodMat[:, k] - 1, 1)
        lev.numNod2Eles = cnt
        lev.nodesOnBoundary = mfind(cnt < 8)
        onb = np.zeros(lev.numNodes, dtype=np.int64)
        onb[lev.nodesOnBoundary - 1] = 1
        lev.elementsOnBoundary = mfind(onb[lev.eNodMat - 1].sum(axis=1) > 0)
        model.levels.append(lev)
    model.eNodMatTemp = None


def edofmat(eNodMat):
    """kernel/SetupSolver.m:13-15, 199-200."""
    e = np.concatenate([3 * eNodMat - 2, 3 * eNodMat - 1, 3 * eNodMat], axis=1)
    return e[:, REORDER]


# Galerkin machinery ---------------------------------------------------------
_IN_NODE, _IN_DOF, _SUB = transfer_tables()
# DOFs (0-based, among the 81) of each SubEleNodMat child, interleaved
_CHILD_DOFS = np.stack([np.stack([3 * (_SUB[c] - 1) + k for k in range(3)], axis=1).reshape(-1)
                        for c in range(8)])
# column-major child order (elementUpwardMap) -> SubEleNodMat child index
COLMAJOR_TO_SUB = np.array([0, 1, 3, 2, 4, 5, 7, 6])


def galerkin(child_K: np.ndarray, present: np.ndarray, exact: bool = False):
    """Coarse element matrices Ks = P81' * sparse(sum of child K) * P81
    (kernel/SetupSolver.m:28-31,57-58).  ``child_K`` (n, 8, 24, 24) and
    ``present`` (n, 8) are in elementUpwardMap column order.  The reference
    places column j at SubEleNodMat child j (quirk B12); ``exact=True`` puts
    it at its true geometric position instead."""
    n = child_K.shape[0]
    K81 = np.zeros((n, 81, 81))
    for j in range(8):
        slot = COLMAJOR_TO_SUB[j] if exact else j
        d = _CHILD_DOFS[slot]
        K81[:, d[:, None], d[None, :]] += child_K[:, j] * present[:, j, None, None]
    return np.einsum("ai,nab,bj->nij", _IN_DOF, K81, _IN_DOF, optimize=True)


def setup_solver(model: Model, E: float = 1.0, nu: float = 0.3):
    """kernel/SetupSolver.m:8-254 (the BC-dependent coarsest factorisation is
    in ``factor_coarsest``)."""
    model.E, model.nu = E, nu
    L1 = model.levels[0]
    probe = L1.nodMapBack[L1.eNodMat[0] - 1]
    Ke0 = element_stiffness(E, nu, model.nodeCoords[probe - 1])
    model.Ke0 = Ke0
    L1.storingState = 0
    L1.Ke = Ke0
    L1.Ks = Ke0
    L1.eDofMat = edofmat(L1.eNodMat)
    if model.numLevels < 2:
        raise ValueError("There is no sufficient resolution to invoke GMGS!")
    Lc = model.numLevels
    state_prev = None                       # eleStiffnessMatStateOnFinerLevel (1-based)
    exact = model.exact_galerkin
    for ii in range(2, Lc + 1):
        lev, fine = model.levels[ii - 1], model.levels[ii - 2]
        lev.Ke = Ke0 * (lev.eleSize[0] / L1.eleSize[0])
        up = lev.elementUpwardMap
        present = (up != 0).astype(np.float64)
        ne = lev.numElements

        def child_mats(rows):
            """Ks of the finer children of coarse elements ``rows`` (0-based)."""
            u = up[rows]
            out = np.zeros((rows.size, 8, 24, 24))
            if fine.storingState == 0:
                out[:] = fine.Ke
            elif fine.storingState == 2:
                idx = np.where(u > 0, state_prev[np.maximum(u, 1) - 1], 1)
                out[:] = fine.Ks[idx - 1]
            else:
                idx = np.maximum(u, 1)
                out[:] = fine.Ks[idx - 1]
            return out

        if ii == 2 and Lc > 2 or ii == 3 and Lc > 3:
            lev.storingState = 2
            if ii == 2:                            # SetupSolver.m:43-44
                uniq = np.unique(np.nonzero(up == 0)[0] + 1)
            else:
                prev_up = fine.elementUpwardMap
                flags = np.zeros(ne, dtype=bool)
                for jj in range(ne):               # SetupSolver.m:96-101 (B3)
                    kids = up[jj][up[jj] != 0]
                    if exact:
                        flags[jj] = (up[jj] == 0).any() or (prev_up[kids - 1] == 0).any()
                    else:
                        flags[jj] = (prev_up[kids - 1] == 0).any()
                uniq = np.nonzero(flags)[0] + 1
            univ = np.setdiff1d(np.arange(1, ne + 1), uniq)
            Ks = np.zeros((1 + uniq.size, 24, 24))
            Ks[0] = lev.Ke
            for a in range(0, uniq.size, 2048):
                rows = uniq[a:a + 2048] - 1
                Ks[1 + a:1 + a + rows.size] = galerkin(child_mats(rows), present[rows], exact)
            st = np.ones(ne, dtype=np.int64)
            st[uniq - 1] = np.arange(2, 2 + uniq.size)
            state_prev = st
            lev.Ks, lev.uniqueElements, lev.universalElements = Ks, uniq, univ
        else:
            lev.storingState = 1
            rows = np.arange(ne)
            Ks = np.zeros((ne, 24, 24))
            chunk = 2048
            for a in range(0, ne, chunk):
                r = rows[a:a + chunk]
                Ks[r] = galerkin(child_mats(r), present[r], exact)
            lev.Ks = Ks
    for ii in range(2, Lc + 1):
        model.levels[ii - 1].eDofMat = edofmat(model.levels[ii - 1].eNodMat)
    for ii in range(1, Lc):                    # SetupSolver.m:204-244
        lev = model.levels[ii - 1]
        d = np.zeros(lev.numDOFs)
        ed = lev.eDofMat - 1
        if lev.storingState == 0:
            dk = np.diag(lev.Ks)
            for j in range(24):
                np.add.at(d, ed[:, j], dk[j])
        elif lev.storingState == 1:
            dk = np.einsum("eii->ei", lev.Ks)
            for j in range(24):
                np.add.at(d, ed[:, j], dk[:, j])
        else:
            st = np.ones(lev.numElements, dtype=np.int64)
            if lev.uniqueElements.size:
                st[lev.uniqueElements - 1] = np.arange(2, 2 + lev.uniqueElements.size)
            dk = np.einsum("eii->ei", lev.Ks)[st - 1]
            for j in range(24):
                np.add.at(d, ed[:, j], dk[:, j])
        lev.diagK = d
    return model


def coarse_fixed_dofs(model: Model):
    """kernel/SetupSolver.m:246-254: coarse DOF fixed iff the restricted 0/1
    indicator is non-zero."""
    L1 = model.levels[0]
    ind = np.zeros(L1.numDOFs)
    ind[L1.fixedDOFs - 1] = 1
    for ii in range(2, model.numLevels + 1):
        c = restrict(model, ind, ii)
        lev = model.levels[ii - 1]
        lev.freeDOFs = mfind(c == 0)
        lev.fixedDOFs = np.setdiff1d(np.arange(1, lev.numDOFs + 1), lev.freeDOFs)
        ind = c


def assemble_coarsest(model: Model) -> np.ndarray:
    """kernel/DirectlyAssembleStifnessMatrix.m:24-33 (dense here)."""
    lev = model.levels[-1]
    K = np.zeros((lev.numDOFs, lev.numDOFs))
    ed = lev.eDofMat - 1
    for e in range(lev.numElements):
        K[np.ix_(ed[e], ed[e])] += lev.Ks[e]
    return K


def factor_coarsest(model: Model):
    """kernel/SetupSolver.m:257-259 (dense Cholesky; the reference's
    fill-reducing permutation does not change the solution)."""
    coarse_fixed_dofs(model)
    lev = model.levels[-1]
    K = assemble_coarsest(model)
    fr = lev.freeDOFs - 1
    Kff = K[np.ix_(fr, fr)]
    model.coarsest_K = Kff
    model.cholL = np.linalg.cholesky(Kff)


def assemble_level1(model: Model, modulus=None):
    """Global level-1 K (DirectlyAssembleStifnessMatrix.m:7-23, 'just for
    test'), scipy sparse."""
    import scipy.sparse as sp
    lev = model.levels[0]
    ed = lev.eDofMat - 1
    ne = lev.numElements
    w = np.ones(ne) if modulus is None else np.asarray(modulus)
    rows = np.repeat(ed, 24, axis=1).reshape(-1)
    cols = np.tile(ed, (1, 24)).reshape(-1)
    vals = (w[:, None] * model.Ke0.reshape(1, -1)).reshape(-1)
    return sp.csr_matrix((vals, (rows, cols)), shape=(lev.numDOFs, lev.numDOFs))


# ------------------------------------------------------------------ operators
def atx(model: Model, x: np.ndarray, ii: int = 1) -> np.ndarray:
    """kernel/AtX.m:1-58 (column-wise scatter passes are conflict-free)."""
    lev = model.levels[ii - 1]
    v = np.zeros(lev.numDOFs)
    ed = lev.eDofMat - 1
    if lev.storingState == 0:
        if model.modulus is not None and ii == 1:
            ve = (x[ed] @ lev.Ks) * model.modulus[:, None]
        else:
            ve = x[ed] @ lev.Ks
        for k in range(24):
            v[ed[:, k]] += ve[:, k]
    elif lev.storingState == 1:
        ve = np.einsum("ei,eij->ej", x[ed], lev.Ks)
        for k in range(24):
            v[ed[:, k]] += ve[:, k]
    else:
        if lev.uniqueElements.size:
            eu = ed[lev.uniqueElements - 1]
            ve = np.einsum("ei,eij->ej", x[eu], lev.Ks[1:])
            for k in range(24):
                v[eu[:, k]] += ve[:, k]
        ev = ed[lev.universalElements - 1]
        ve = x[ev] @ lev.Ks[0]
        for k in range(24):
            v[ev[:, k]] += ve[:, k]
    v[lev.fixedDOFs - 1] = 0.0
    return v


def jacobi(model: Model, x, r, ii=1, with_residual=True):
    """kernel/JacobiSmoother.m:5-9."""
    lev = model.levels[ii - 1]
    dx = model.omega * r / lev.diagK
    xn = x + dx
    if with_residual:
        return xn, r - atx(model, dx, ii)
    return xn


def restrict(model: Model, rf: np.ndarray, ii: int) -> np.ndarray:
    """kernel/RestrictResidual.m:3-17."""
    lev, fine = model.levels[ii - 1], model.levels[ii - 2]
    rF = rf.reshape(fine.numNodes, 3)
    r1 = np.zeros((lev.intermediateNumNodes, 3))
    r1[lev.solidNodeMapCoarser2Finer - 1] = rF
    r1 = r1 / lev.transferMatCoeffi[:, None]
    rc = np.zeros((lev.numNodes, 3))
    for j in range(3):
        t = r1[:, j][lev.transferMat - 1]          # 27 x Ne
        t = t.T @ _IN_NODE                          # Ne x 8
        for k in range(8):
            rc[lev.eNodMat[:, k] - 1, j] += t[:, k]
    return rc.reshape(-1)


def interpolate(model: Model, xc: np.ndarray, ii: int) -> np.ndarray:
    """kernel/InterpolationDeviation.m:3-17."""
    lev, fine = model.levels[ii - 1], model.levels[ii - 2]
    xC = xc.reshape(lev.numNodes, 3)
    xf = np.zeros((lev.intermediateNumNodes, 3))
    for j in range(3):
        t = xC[:, j][lev.eNodMat - 1]               # Ne x 8
        t1 = _IN_NODE @ t.T                         # 27 x Ne
        for k in range(27):
            xf[lev.transferMat[k] - 1, j] += t1[k]
    xf = xf / lev.transferMatCoeffi[:, None]
    xf = xf[lev.solidNodeMapCoarser2Finer - 1]
    return xf.reshape(-1)


def coarsest_solve(model: Model, r: np.ndarray) -> np.ndarray:
    """kernel/Vcycle.m:21-23."""
    from scipy.linalg import solve_triangular
    lev = model.levels[-1]
    x = np.zeros(lev.numDOFs)
    fr = lev.freeDOFs - 1
    y = solve_triangular(model.cholL, r[fr], lower=True)
    x[fr] = solve_triangular(model.cholL.T, y, lower=False)
    return x


def vcycle(model: Model, r: np.ndarray) -> np.ndarray:
    """kernel/Vcycle.m:1-36."""
    Lc = model.numLevels
    xs = [0.0] * Lc                           # varVcycle(:).x = 0 (Vcycle.m:7-8)
    rs = [None] * Lc
    rs[0] = r
    for ii in range(2, Lc + 1):
        xs[ii - 2], rs[ii - 2] = jacobi(model, xs[ii - 2], rs[ii - 2], ii - 1)
        rs[ii - 1] = restrict(model, rs[ii - 2], ii)
    xs[Lc - 1] = coarsest_solve(model, rs[Lc - 1])
    for ii in range(Lc, 1, -1):
        xs[ii - 2] = xs[ii - 2] + interpolate(model, xs[ii - 1], ii)
        xs[ii - 2] = jacobi(model, xs[ii - 2], rs[ii - 2], ii - 1, with_residual=False)
    d = xs[0].copy()
    d[model.levels[0].fixedDOFs - 1] = 0.0
    return d


def vcycle_sym(model: Model, r: np.ndarray, nu: int = 2, weights=None) -> np.ndarray:
    """Symmetric V(nu, nu) cycle on the free DOFs (no reference counterpart).

    The levels, transfers and coarsest solve of kernel/Vcycle.m:1-36, with
    three changes: each level works on its free DOFs only (smoothing
    increments, smoothed residuals and the prolonged correction are zero on the
    level's fixed DOFs), the post-smoothing starts from the residual
    recomputed after the coarse correction (Vcycle.m:27-33 reuses the
    pre-smoothed one), and each level smooths nu times on the way down and up.
    Of the free-DOF restriction only the zero smoothing weight on fixed DOFs
    changes the result: a coarse DOF is fixed whenever a fine fixed DOF lies in
    its support (SetupSolver.m:247-254), so fixed-DOF residual values only ever
    reach coarse fixed DOFs and P x_c vanishes on fine fixed DOFs; the other
    two masks are kept as plain statements of the intent.
    nu is an int or a dict of per-level counts; weights[ii], if given, lists
    the sweep weights w_s of level ii and sets its count (default: the Jacobi
    weight of JacobiSmoother.m for every sweep).  In the form the device
    evaluates it: R_s = R_{s-1} - A (w_{s-1} R_{s-1}/d), e = P x_c,
    Q_0 = R_nu - A e, Q_s = Q_{s-1} - A (w_{s-1} Q_{s-1}/d),
    x = sum_{s<nu} w_s (R_s + Q_s)/d + e.
    """
    Lc = model.numLevels
    rin = [None] * Lc
    Rs = [None] * Lc
    xs = [None] * Lc
    rin[0] = r

    def fixed(ii):
        return model.levels[ii - 1].fixedDOFs - 1

    def dinv(ii):
        w = 1.0 / model.levels[ii - 1].diagK
        w[fixed(ii)] = 0.0
        return w

    def wts(ii):
        if weights is not None:
            return list(weights[ii])
        return [model.omega] * (nu[ii] if isinstance(nu, dict) else nu)

    def smooth(rr, ii, w):
        out = rr - atx(model, w * dinv(ii) * rr, ii)
        out[fixed(ii)] = 0.0
        return out

    for ii in range(1, Lc):
        w = wts(ii)
        R = [rin[ii - 1]]
        for s in range(len(w)):
            R.append(smooth(R[-1], ii, w[s]))
        Rs[ii - 1] = R
        rin[ii] = restrict(model, R[-1], ii + 1)
    xs[Lc - 1] = coarsest_solve(model, rin[Lc - 1])
    for ii in range(Lc - 1, 0, -1):
        w = wts(ii)
        R = Rs[ii - 1]
        e = interpolate(model, xs[ii], ii + 1)
        e[fixed(ii)] = 0.0
        Q = [R[-1] - atx(model, e, ii)]
        for s in range(len(w) - 1):
            Q.append(smooth(Q[-1], ii, w[s]))
        acc = sum(w[s] * (R[s] + Q[s]) for s in range(len(w)))
        xs[ii - 1] = dinv(ii) * acc + e
    return xs[0]


def chebyshev_weights(lmax: float, nu: int, lower: float = 0.05, safety: float = 1.1) -> list:
    """Sweep weights of a degree-nu Chebyshev smoother (no reference counterpart).

    nu Jacobi sweeps x += w_i r/d have the error propagator prod_i (I - w_i D^-1 A),
    a polynomial in D^-1 A; with w_i the reciprocal roots of the Chebyshev
    polynomial T_nu shifted to [lower b, b], b = safety lmax, it is
    T_nu((theta - l)/delta) / T_nu(theta/delta) (theta, delta: centre and
    half-width of the interval), the polynomial of least maximum on it.
    """
    b = safety * lmax
    a = lower * b
    theta, delta = 0.5 * (b + a), 0.5 * (b - a)
    return [1.0 / (theta + delta * np.cos((2 * i - 1) * np.pi / (2 * nu))) for i in range(1, nu + 1)]


def estimate_lmax(model: Model, ii: int, steps: int = 20, r0=None) -> float:
    """Top Ritz value of D^-1 A on the free DOFs of level ii from 'steps' Jacobi-
    preconditioned CG iterations (the CG coefficients give the Lanczos matrix:
    T_kk = 1/alpha_k + beta_{k-1}/alpha_{k-1}, T_k,k+1 = sqrt(beta_k)/alpha_k).
    The start vector r0 defaults to a seeded uniform one on the free DOFs."""
    lv = model.levels[ii - 1]
    f = lv.fixedDOFs - 1
    dinv = 1.0 / lv.diagK
    dinv[f] = 0.0
    r = (np.random.default_rng(0).uniform(-1, 1, lv.numDOFs) if r0 is None else np.array(r0, float))
    r[f] = 0.0
    z = dinv * r
    rz = r @ z
    p = z.copy()
    al, be = [], []
    for _ in range(steps):
        q = atx(model, p, ii)
        q[f] = 0.0
        pq = p @ q
        if not pq > 0:
            break
        a = rz / pq
        r = r - a * q
        z = dinv * r
        rzn = r @ z
        al.append(a)
        be.append(rzn / rz)
        if not rzn > 1e-24 * rz:
            break
        p = z + (rzn / rz) * p
        rz = rzn
    n = len(al)
    T = np.zeros((n, n))
    for i in range(n):
        T[i, i] = 1.0 / al[i] + (be[i - 1] / al[i - 1] if i else 0.0)
        if i + 1 < n:
            T[i, i + 1] = T[i + 1, i] = np.sqrt(be[i]) / al[i]
    return float(np.linalg.eigvalsh(T)[-1])


@dataclasses.dataclass
class CGResult:
    x: np.ndarray
    its: int
    hist: np.ndarray          # rows [its, elapsed_s, relres] of non-final iterations
    relres: float
    converged: bool


def cg_solver(model: Model, b: np.ndarray, tol: float, max_it: int,
              flexible: bool = False, jacobi_step: bool = True, cycle=None) -> CGResult:
    """kernel/CG_solver.m:1-56.  ``flexible=True`` swaps in the Polak-Ribiere
    beta, ``jacobi_step=False`` drops the Jacobi step of CG_solver.m:18-19 and
    ``cycle`` replaces the V-cycle (e.g. vcycle_sym); none of these has a
    reference counterpart."""
    if cycle is None:
        cycle = vcycle
    n = b.size
    normB = np.linalg.norm(b)
    its = 0
    x = np.zeros(n)
    r1 = b - atx(model, x)
    z1 = np.zeros(n)
    p1 = np.zeros(n)
    r0 = None
    hist = []
    t0 = time.perf_counter()
    resnorm = np.inf
    converged = False
    while its <= max_it:
        if jacobi_step:
            x, r1 = jacobi(model, x, r1, 1)
        z2 = cycle(model, r1)
        its += 1
        if its == 1:
            p2 = z2
        else:
            if flexible:
                beta = (r1 @ z2 - r1 @ z1) / (r0 @ z1)
            else:
                beta = (r1 @ z2) / (r0 @ z1)
            p2 = z2 + beta * p1
        q = atx(model, p2)
        alpha = (r1 @ z2) / (p2 @ q)
        x = x + alpha * p2
        r2 = r1 - alpha * q
        resnorm = np.linalg.norm(r2) / normB
        if resnorm < tol:
            converged = True
            break
        z1, p1, r0, r1 = z2, p2, r1, r2
        hist.append([its, time.perf_counter() - t0, resnorm])
    return CGResult(x, its, np.array(hist).reshape(-1, 3), float(resnorm), converged)


def cartesian_stress(model: Model, U: np.ndarray) -> np.ndarray:
    """kernel/GetCartesianStress.m:1-42: nodal [sxx syy szz syz szx sxy]."""
    L1 = model.levels[0]
    OTP = outer_interpolation_matrix()
    probe = L1.nodMapBack[L1.eNodMat[0] - 1]
    dshape, invJ, _, _ = probe_jacobians(model.nodeCoords[probe - 1])
    B = element_strain_matrix(dshape, invJ)
    D = element_elasticity(model.E, model.nu)
    M = OTP @ (D @ B)                        # 48 x 24
    ue = U[L1.eDofMat - 1]                   # Ne x 24
    mid = ue @ M.T                           # Ne x 48
    if model.modulus is not None:
        mid = mid * model.modulus[:, None]
    S = np.zeros((L1.numNodes, 6))
    for k in range(8):
        S[L1.eNodMat[:, k] - 1] += mid[:, 6 * k:6 * k + 6]
    return S / L1.numNod2Eles[:, None]


# ------------------------------------------------------------------ SIMP / adjoint
def setup_modulus(model: Model, modulus: np.ndarray):
    """Per-element modulus scale on level 1 (no reference counterpart): level-1
    element matrices become E_e*Ke0 and every coarse level is a per-element
    Galerkin product (storing state 1) of the scaled children."""
    model.modulus = np.asarray(modulus, dtype=np.float64)
    L1 = model.levels[0]
    Lc = model.numLevels
    prevK = model.Ke0[None] * model.modulus[:, None, None]
    for ii in range(2, Lc + 1):
        lev = model.levels[ii - 1]
        up = lev.elementUpwardMap
        present = (up != 0).astype(np.float64)
        child = prevK[np.maximum(up, 1) - 1]
        lev.storingState = 1
        lev.Ks = galerkin(child, present, model.exact_galerkin)
        lev.uniqueElements = None
        prevK = lev.Ks
    for ii in range(1, Lc):
        lev = model.levels[ii - 1]
        d = np.zeros(lev.numDOFs)
        ed = lev.eDofMat - 1
        if ii == 1:
            dk = model.modulus[:, None] * np.diag(model.Ke0)[None, :]
        else:
            dk = np.einsum("eii->ei", lev.Ks)
        for j in range(24):
            np.add.at(d, ed[:, j], dk[:, j])
        lev.diagK = d


def element_energy_gradient(model: Model, lam: np.ndarray, u: np.ndarray) -> np.ndarray:
    """dJ/dE_e = -lambda_e' Ke0 u_e (no reference counterpart)."""
    ed = model.levels[0].eDofMat - 1
    return -np.einsum("ei,ij,ej->e", lam[ed], model.Ke0, u[ed])


# ------------------------------------------------------------------ driver
def build(vox, bbox, control=1.0e4, num_levels=None, E=1.0, nu=0.3, omega=0.4,
          exact_galerkin=False) -> Model:
    m = discretize(vox, bbox, control, num_levels)
    m.omega = omega
    m.exact_galerkin = exact_galerkin
    build_hierarchy(m)
    setup_solver(m, E, nu)
    return m


def get_deformation(model: Model, loads, fixed, tol, max_it=20000, flexible=False):
    """kernel/GetDeformation.m:1-21 (BC-dependent setup re-run every call)."""
    b = apply_bc(model, loads, fixed)
    factor_coarsest(model)
    return cg_solver(model, b, tol, max_it, flexible)
