"""Surface mesh -> voxel model: the reference's model input path.

* ``read_ply`` / ``read_obj``: triangle meshes (common/ReadTriangularPolygonMesh_plyFormat.m,
  common/ReadTriangularPolygonMesh_objFormat.m; the OBJ reader maps (x, y, z) to
  (x, -z, y) as the reference does).
* ``resolution``: the per-axis voxel counts from the finest-resolution control
  (common/AdaptData4Voxelize.m:5-9): round(maxRes * extent / max extent).
* ``voxelize``: ray-casting voxelisation after A. H. Aitkenhead's published
  method (the reference vendors it as common/Voxelize.m): rays through the voxel
  centres along x, y and z; along each ray the crossings with the facets are
  paired and the voxel centres strictly between a pair are inside; a ray that
  meets the mesh an odd number of times, or passes through a vertex whose
  facets do not all face one way along the ray, takes its column from the
  majority (at least 4) of its 8 neighbouring rays; a voxel is solid when at
  least two of the three ray directions say so.
* ``create_voxelized_model``: common/CreateVoxilizedModel.m:11-26 -- the grid
  above, flipped in y (voxel row 0 is the max-y side), and the mesh's
  bounding box.

The predicates and formulas that decide borderline rays (a ray through an
edge or a vertex) are the published method's: which side of each edge the
ray lies on, by strict comparisons against the edge line evaluated at the
ray; the crossing depth from the facet's plane equation; crossings rounded to
1e-12 and de-duplicated.  Only with the same decisions do the voxel models
(and the node and DOF counts of the reference's published examples) come out
the same.
"""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np


# ------------------------------------------------------------------ readers
def _text(path) -> str:
    p = Path(path)
    if p.suffix == ".gz":
        import gzip
        return gzip.decompress(p.read_bytes()).decode()
    return p.read_text()


def _suffix(path) -> str:
    p = Path(path)
    return (p.with_suffix("") if p.suffix == ".gz" else p).suffix.lower()


def read_ply(path):
    """ASCII PLY with 'element vertex', 'element face' and triangles
    '3 i j k' (0-based); returns (nodes (n, 3) float64, faces (m, 3) 1-based).
    A '.gz' file is read through gzip."""
    lines = _text(path).splitlines()
    nv = nf = None
    k = 0
    while k < len(lines):
        tok = lines[k].split()
        k += 1
        if len(tok) >= 3 and tok[0] == "element" and tok[1] == "vertex":
            nv = int(tok[2])
        elif len(tok) >= 3 and tok[0] == "element" and tok[1] == "face":
            nf = int(tok[2])
        elif tok and tok[0] == "end_header":
            break
    if nv is None or nf is None:
        raise ValueError(f"{path}: not an ASCII PLY triangle mesh")
    nodes = np.array([lines[k + i].split()[:3] for i in range(nv)], dtype=np.float64)
    faces = np.array([lines[k + nv + i].split()[1:4] for i in range(nf)], dtype=np.int64) + 1
    return nodes, faces


def read_obj(path):
    """OBJ 'v x y z' / 'f i j k' lines; coordinates mapped to (x, -z, y)."""
    v, f = [], []
    for line in _text(path).splitlines():
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "v":
            v.append([float(t) for t in tok[1:4]])
        elif tok[0] == "f":
            f.append([int(t.split("/")[0]) for t in tok[1:4]])
    nodes = np.array(v, dtype=np.float64)[:, [0, 2, 1]]
    nodes[:, 1] = -nodes[:, 1]
    return nodes, np.array(f, dtype=np.int64)


def read_mesh(path):
    suffix = _suffix(path)
    if suffix == ".ply":
        return read_ply(path)
    if suffix == ".obj":
        return read_obj(path)
    raise ValueError("Wrong Input!")


def _round_half_away(x):
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def resolution(nodes, max_res):
    """(nx, ny, nz) = round(maxRes * extent / max extent) per axis."""
    ext = nodes.max(axis=0) - nodes.min(axis=0)
    r = _round_half_away(max_res * (ext / ext.max()))
    return tuple(int(v) for v in r)


# ------------------------------------------------------------------ ray casting
def _colon(a, d, b):
    """MATLAB a:d:b: n = floor((b - a)/d) steps, elements computed from both
    ends of a .. a + n d."""
    n = int(math.floor((b - a) / d + 1e-10))
    k = np.arange(n + 1, dtype=np.float64)
    last = a + n * d
    out = a + k * d
    hi = k >= n / 2.0
    out[hi] = last - (n - k[hi]) * d
    return out


def _centres(lo, hi, count):
    w = (hi - lo) / (count + 0.5)
    return _colon(lo + w / 2, w, hi - w / 2)


def _nearest(g, v):
    d = np.abs(g - v)
    return int(np.flatnonzero(d == d.min())[0])


def _side_ok(vi_u, vi_v, ja_u, ja_v, jb_u, jb_v, ru, rv):
    """Is the ray (ru, rv) strictly on the same side of the edge a-b as the
    opposite vertex i?  The edge line is evaluated at u = u_i and u = ru."""
    with np.errstate(divide="ignore", invalid="ignore"):
        pv = ja_v - ((ja_v - jb_v) * (ja_u - vi_u) / (ja_u - jb_u))
        pr = ja_v - ((ja_v - jb_v) * (ja_u - ru) / (ja_u - jb_u))
    return ((pv > vi_v) & (pr > rv)) | ((pv < vi_v) & (pr < rv))


def _cast(gu, gv, gw, M):
    """Voxels inside along rays parallel to w through (gu[i], gv[j]); M is
    (F, 3, 3) [facet, (u, v, w), vertex].  Returns bool (nu, nv, nw)."""
    nu, nv, nw = len(gu), len(gv), len(gw)
    out = np.zeros((nu, nv, nw), dtype=bool)
    umin, umax = M[:, 0, :].min(), M[:, 0, :].max()
    vmin, vmax = M[:, 1, :].min(), M[:, 1, :].max()
    wmin, wmax = M[:, 2, :].min(), M[:, 2, :].max()
    iu0, iu1 = sorted((_nearest(gu, umin), _nearest(gu, umax)))
    iv0, iv1 = sorted((_nearest(gv, vmin), _nearest(gv, vmax)))
    fmin, fmax = M.min(axis=2), M.max(axis=2)
    U, V, W = M[:, 0, :], M[:, 1, :], M[:, 2, :]
    # plane Au + Bv + Cw + D = 0 of every facet
    A = V[:, 0] * (W[:, 1] - W[:, 2]) + V[:, 1] * (W[:, 2] - W[:, 0]) + V[:, 2] * (W[:, 0] - W[:, 1])
    B = W[:, 0] * (U[:, 1] - U[:, 2]) + W[:, 1] * (U[:, 2] - U[:, 0]) + W[:, 2] * (U[:, 0] - U[:, 1])
    Cc = U[:, 0] * (V[:, 1] - V[:, 2]) + U[:, 1] * (V[:, 2] - V[:, 0]) + U[:, 2] * (V[:, 0] - V[:, 1])
    D = (-U[:, 0] * (V[:, 1] * W[:, 2] - V[:, 2] * W[:, 1]) - U[:, 1] * (V[:, 2] * W[:, 0] - V[:, 0] * W[:, 2])
         - U[:, 2] * (V[:, 0] * W[:, 1] - V[:, 1] * W[:, 0]))
    Cc = np.where(np.abs(Cc) < 1e-14, 0.0, Cc)
    # unit normals' w component (for rays through a vertex)
    nrm = np.cross(M[:, :, 1] - M[:, :, 0], M[:, :, 2] - M[:, :, 0])
    with np.errstate(divide="ignore", invalid="ignore"):
        nw_comp = nrm[:, 2] / np.linalg.norm(nrm, axis=1)
    corrections = []
    for j in range(iv0, iv1 + 1):
        rv = gv[j]
        rows = np.flatnonzero((fmin[:, 1] <= rv) & (fmax[:, 1] >= rv))
        if rows.size == 0:
            continue
        for i in range(iu0, iu1 + 1):
            ru = gu[i]
            cand = rows[(fmin[rows, 0] <= ru) & (fmax[rows, 0] >= ru)]
            if cand.size == 0:
                continue
            crossed = []
            # rays through a vertex: counted once per group of facets sharing
            # vertices when their normals all point one way along the ray
            on_vertex = ((U[cand] == ru) & (V[cand] == rv)).any(axis=1)
            vlist = cand[on_vertex]
            skip = False
            if vlist.size:
                keys = [{tuple(M[f, :, k]) for k in range(3)} for f in vlist]
                checked = np.zeros(vlist.size, dtype=bool)
                while not checked.all():
                    a = int(np.flatnonzero(~checked)[0])
                    checked[a] = True
                    adj = np.array([bool(keys[a] & keys[b]) for b in range(vlist.size)])
                    checked |= adj
                    nz = nw_comp[vlist[adj]]
                    if nz.max() < 0 or nz.min() > 0:
                        crossed.append(vlist[a])
                    else:
                        skip = True
                        corrections.append((i, j))
                        break
            if skip:
                continue
            u1, u2, u3 = U[cand, 0], U[cand, 1], U[cand, 2]
            v1, v2, v3 = V[cand, 0], V[cand, 1], V[cand, 2]
            ok = (_side_ok(u1, v1, u2, v2, u3, v3, ru, rv)
                  & _side_ok(u2, v2, u3, v3, u1, v1, ru, rv)
                  & _side_ok(u3, v3, u1, v1, u2, v2, ru, rv))
            crossed = np.concatenate([np.array(crossed, dtype=np.int64), cand[ok]])
            with np.errstate(divide="ignore", invalid="ignore"):
                z = (-D[crossed] - A[crossed] * ru - B[crossed] * rv) / Cc[crossed]
            z = z[(z >= wmin - 1e-12) & (z <= wmax + 1e-12)]
            z = np.unique(_round_half_away(z * 1e12) / 1e12)
            if z.size % 2 == 0:
                for p in range(z.size // 2):
                    out[i, j, (gw > z[2 * p]) & (gw < z[2 * p + 1])] = True
            elif z.size:
                corrections.append((i, j))
    if corrections:
        c = np.array(corrections)
        pad = c[:, 0].min() == 0 or c[:, 0].max() == nu - 1 or c[:, 1].min() == 0 or c[:, 1].max() == nv - 1
        g = np.pad(out, ((1, 1), (1, 1), (0, 0))) if pad else out
        off = 1 if pad else 0
        for i, j in corrections:           # in order: later rays see earlier corrections
            a, b = i + off, j + off
            s = (g[a - 1, b - 1].astype(np.int32) + g[a - 1, b] + g[a - 1, b + 1] + g[a, b - 1] + g[a, b + 1]
                 + g[a + 1, b - 1] + g[a + 1, b] + g[a + 1, b + 1])
            g[a, b, s >= 4] = True
        out = g[1:-1, 1:-1] if pad else g
    return out


def voxelize(nx, ny, nz, tri):
    """Solid voxels of the closed triangle mesh ``tri`` (F, 3, 3) [facet,
    (x, y, z), vertex] on an nx x ny x nz grid spanning its bounding box;
    returns bool (ny, nx, nz) with y ascending."""
    lo = tri.min(axis=(0, 2))
    hi = tri.max(axis=(0, 2))
    gx, gy, gz = (_centres(lo[a], hi[a], n) for a, n in enumerate((nx, ny, nz)))
    votes = np.zeros((len(gx), len(gy), len(gz)), dtype=np.int32)
    votes += np.transpose(_cast(gy, gz, gx, tri[:, [1, 2, 0], :]), (2, 0, 1))   # rays along x
    votes += np.transpose(_cast(gz, gx, gy, tri[:, [2, 0, 1], :]), (1, 2, 0))   # rays along y
    votes += _cast(gx, gy, gz, tri)                                           # rays along z
    return np.transpose(votes >= 1.5, (1, 0, 2))


def create_voxelized_model(path, max_res):
    """(vox int32 (ny, nx, nz), bbox (2, 3)) of CreateVoxilizedModel(file, maxRes)."""
    nodes, faces = read_mesh(path)
    nx, ny, nz = resolution(nodes, max_res)
    tri = np.stack([nodes[faces[:, k] - 1] for k in range(3)], axis=2)
    vox = voxelize(nx, ny, nz, tri)[::-1].astype(np.int32)      # flip(., 1)
    bbox = np.stack([nodes.min(axis=0), nodes.max(axis=0)])
    return np.ascontiguousarray(vox), bbox
