"""Synthetic voxel problems for the GMGS solver (inputs only, no solving).

Every problem is returned in the reference's own input convention
(src/common/CreateVoxilizedModel.m:27-36, src/GlobalVariables.m:29-30):

* ``vox``      int32 array of shape (ny, nx, nz); MATLAB column-major linear
               order (y fastest) is ``vox.ravel(order="F")``; value 1 = solid.
* ``bbox``     (2, 3) float64 ``[[xmin ymin zmin], [xmax ymax zmax]]``.
* ``fixed``    compact 1-based node ids (fixingCond_).
* ``loads``    (m, 4) float64 rows ``[node fx fy fz]`` (loadingCond_).

Node ids refer to the padded finest mesh that DiscretizeDesignDomain builds
(src/common/DiscretizeDesignDomain.m:29-93): compact id = 1-based rank of the
full-grid node id among nodes touched by a solid element.  The few lines of
that numbering repeated here are needed to place boundary conditions; the
solver itself rebuilds the mesh from ``vox``.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Problem:
    name: str
    vox: np.ndarray            # int32 (ny, nx, nz)
    bbox: np.ndarray           # (2, 3)
    fixed: np.ndarray          # int32 compact 1-based node ids
    loads: np.ndarray          # (m, 4) float64
    num_levels: int            # 0 = reference rule from coarsest control
    coarsest_control: float = 1.0e4
    tol: float = 1.0e-6
    omega: float = 0.4
    E: float = 1.0
    nu: float = 0.3
    max_it: int = 20000
    meta: dict = dataclasses.field(default_factory=dict)


# --------------------------------------------------------------------------
# level count and padding (DiscretizeDesignDomain.m:29-50)
# --------------------------------------------------------------------------
def level_rule(nx: int, ny: int, nz: int, control: float) -> int:
    """Level count from coarsestResolutionControl_ (DiscretizeDesignDomain.m:33-41)."""
    num = nx * ny * nz
    if num < control:
        raise ValueError("insufficient resolution for the mesh hierarchy")
    L = 0
    while num >= control:
        L += 1
        num = nx / 2 ** L * ny / 2 ** L * nz / 2 ** L
    return L + 1


def padded_dims(nx: int, ny: int, nz: int, num_levels: int):
    s = 2 ** (num_levels - 1)
    return (int(math.ceil(nx / s)) * s, int(math.ceil(ny / s)) * s,
            int(math.ceil(nz / s)) * s)


def pad_vox(vox: np.ndarray, num_levels: int) -> np.ndarray:
    ny, nx, nz = vox.shape
    ax, ay, az = padded_dims(nx, ny, nz, num_levels)
    out = np.zeros((ay, ax, az), dtype=np.int32)
    out[:ny, :nx, :nz] = vox
    return out


def solid_node_table(vox_padded: np.ndarray):
    """Boolean node grid (ny+1, nx+1, nz+1): node touched by a solid element,
    plus the per-node count of solid incident elements."""
    s = (vox_padded == 1).astype(np.int32)
    ny, nx, nz = s.shape
    cnt = np.zeros((ny + 1, nx + 1, nz + 1), dtype=np.int32)
    for dy in (0, 1):
        for dx in (0, 1):
            for dz in (0, 1):
                cnt[dy:dy + ny, dx:dx + nx, dz:dz + nz] += s
    return cnt > 0, cnt


def compact_ids(solid_nodes: np.ndarray) -> np.ndarray:
    """Full-grid node grid -> compact 1-based id (0 for unused nodes)."""
    flat = solid_nodes.ravel(order="F")
    ids = np.zeros(flat.size, dtype=np.int64)
    ids[flat] = np.arange(1, int(flat.sum()) + 1)
    return ids.reshape(solid_nodes.shape, order="F")


# --------------------------------------------------------------------------
# ellipsoid-union solids
# --------------------------------------------------------------------------
def ellipsoid_union(nx, ny, nz, n_ell, rng, semi_lo=0.15, semi_hi=0.40):
    """Union of ``n_ell`` axis-aligned ellipsoids, voxel centres at i+1/2,
    reduced to its largest 6-connected component.  Grid axes are
    (y-row, x, z) as in the voxel array."""
    from scipy import ndimage

    ext = np.array([nx, ny, nz], dtype=np.float64)
    centres = rng.uniform(0.0, 1.0, size=(n_ell, 3)) * ext
    semis = rng.uniform(semi_lo, semi_hi, size=(n_ell, 3)) * ext
    yc = np.arange(ny, dtype=np.float64) + 0.5
    xc = np.arange(nx, dtype=np.float64) + 0.5
    zc = np.arange(nz, dtype=np.float64) + 0.5
    mask = np.zeros((ny, nx, nz), dtype=bool)
    for k in range(n_ell):
        cx, cy, cz = centres[k]
        ax, ay, az = semis[k]
        ty = ((yc - cy) / ay) ** 2
        tx = ((xc - cx) / ax) ** 2
        tz = ((zc - cz) / az) ** 2
        mask |= (ty[:, None, None] + tx[None, :, None] + tz[None, None, :]) <= 1.0
    lab, n = ndimage.label(mask)
    if n == 0:
        raise RuntimeError("empty ellipsoid union")
    counts = np.bincount(lab.ravel())
    counts[0] = 0
    keep = int(np.argmax(counts))
    return (lab == keep).astype(np.int32)


def slab_fixity_and_surface_loads(vox, num_levels, n_loads, rng, slab_frac=0.02):
    """Fix every solid node with z <= zmin(solid) + slab_frac*nz; load
    ``n_loads`` distinct boundary nodes (fewer than 8 solid incident elements,
    DiscretizeDesignDomain.m:103-108) outside the slab with unit forces in
    normalised N(0, I3) directions."""
    ny, nx, nz = vox.shape
    vp = pad_vox(vox, num_levels)
    solid, cnt = solid_node_table(vp)
    ids = compact_ids(solid)
    zs = np.nonzero((vp == 1).any(axis=(0, 1)))[0]
    zmin = int(zs.min())
    zlim = zmin + slab_frac * nz
    iz = np.arange(solid.shape[2])
    slab = solid & (iz[None, None, :] <= zlim)
    fixed = np.sort(ids[slab]).astype(np.int32)
    boundary = solid & (cnt < 8) & ~slab
    cand = np.sort(ids[boundary])
    n_loads = min(n_loads, cand.size)
    nodes = rng.choice(cand, size=n_loads, replace=False)
    d = rng.standard_normal((n_loads, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    loads = np.concatenate([nodes[:, None].astype(np.float64), d], axis=1)
    return fixed, loads


# --------------------------------------------------------------------------
# named problems
# --------------------------------------------------------------------------
def box_c1() -> Problem:
    """Solid 40x40x64 box, z=0 face fixed, total -1 in z on the z=max face,
    4 levels (coarsestResolutionControl_=1000), omega 0.6, rtol 1e-3."""
    nx, ny, nz = 40, 40, 64
    vox = np.ones((ny, nx, nz), dtype=np.int32)
    bbox = np.array([[0.0, 0.0, 0.0], [nx, ny, nz]])
    nface = (nx + 1) * (ny + 1)
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    fixed = np.arange(1, nface + 1, dtype=np.int32)
    top = np.arange(nn - nface + 1, nn + 1)
    loads = np.zeros((nface, 4))
    loads[:, 0] = top
    loads[:, 3] = -1.0 / nface
    return Problem("box40x40x64", vox, bbox, fixed, loads, num_levels=0,
                   coarsest_control=1000.0, tol=1.0e-3, omega=0.6)


def ellipsoid_problem(nx, ny, nz, num_levels, seed=0, n_ell=12, n_loads=1000,
                      tol=1e-6, name=None) -> Problem:
    rng = np.random.default_rng(seed)
    vox = ellipsoid_union(nx, ny, nz, n_ell, rng)
    fixed, loads = slab_fixity_and_surface_loads(vox, num_levels, n_loads, rng)
    bbox = np.array([[0.0, 0.0, 0.0], [nx, ny, nz]], dtype=np.float64)
    return Problem(name or f"ellipsoids{nx}x{ny}x{nz}s{seed}", vox, bbox, fixed,
                   loads, num_levels=num_levels, tol=tol, omega=0.4)


def ellipsoids_c2(seed=0) -> Problem:
    return ellipsoid_problem(169, 147, 256, 5, seed=seed)


def ellipsoids_c3(seed=0) -> Problem:
    return ellipsoid_problem(337, 294, 512, 6, seed=seed)


def batch_c4(seed=0, count=64, coarsest_control=1.0e3):
    """64 uneven problems; problem i uses default_rng([seed, i]).  The level
    count follows the level rule with coarsestResolutionControl_ = 1000 (as
    the 40x40x64 box): the solver's coarsest solve applies an explicit dense
    inverse, whose bytes grow with the square of the coarsest level's DOFs
    (with 10^4 the 64 problems' inverses held 23.6 GB)."""
    out = []
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        nx, ny, nz = (int(v) for v in rng.integers(24, 97, size=3))
        n_ell = int(rng.integers(4, 9))
        n_loads = int(rng.integers(50, 201))
        vox = ellipsoid_union(nx, ny, nz, n_ell, rng)
        L = level_rule(nx, ny, nz, coarsest_control)
        fixed, loads = slab_fixity_and_surface_loads(vox, L, n_loads, rng)
        bbox = np.array([[0.0, 0.0, 0.0], [nx, ny, nz]], dtype=np.float64)
        out.append(Problem(f"batch{seed}_{i}", vox, bbox, fixed, loads,
                           num_levels=L, tol=1e-6, omega=0.4))
    return out


def cantilever_c5(seed=0, nx=128, ny=64, nz=64, num_levels=5) -> Problem:
    """Full 128x64x64 cantilever, x=0 face fixed, total -1 in z spread over
    the nodes of the x=max, z=0 edge; per-element SIMP densities in meta."""
    rng = np.random.default_rng(seed)
    vox = np.ones((ny, nx, nz), dtype=np.int32)
    bbox = np.array([[0.0, 0.0, 0.0], [nx, ny, nz]])
    nyn, nxn = ny + 1, nx + 1
    iy, ix, iz = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), np.arange(nz + 1),
                             indexing="ij")
    full = iy + nyn * ix + nyn * nxn * iz + 1          # box: compact == full
    fixed = np.sort(full[:, 0, :].ravel()).astype(np.int32)
    edge = np.sort(full[:, nx, 0].ravel())
    loads = np.zeros((edge.size, 4))
    loads[:, 0] = edge
    loads[:, 3] = -1.0 / edge.size
    rho = rng.uniform(0.01, 1.0, size=nx * ny * nz)      # MATLAB element order
    p = Problem("cantilever128x64x64", vox, bbox, fixed, loads,
                num_levels=num_levels, tol=1e-6, omega=0.4)
    p.meta["rho"] = rho
    p.meta["g"] = rng.standard_normal(3 * full.size)
    return p


def small_ellipsoids(nx=24, ny=20, nz=16, num_levels=3, seed=1, n_ell=5, n_loads=40,
                     tol=1e-8) -> Problem:
    """A few-thousand-DOF irregular solid for quick parity tests."""
    return ellipsoid_problem(nx, ny, nz, num_levels, seed=seed, n_ell=n_ell,
                             n_loads=n_loads, tol=tol,
                             name=f"small{nx}x{ny}x{nz}s{seed}")
