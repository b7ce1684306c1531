"""On-disk formats of the reference's results and models.

* ``export_dataset_3dtsv``: the ``.carti`` Cartesian-grid stress dataset of
  common/ExportDataset_3DTSV.m:11-41 (resolution, bounds, 0-based element
  ids, loads and fixed nodes as 0-based finest-grid node ids, the nodal
  stress field in ``%.6e``).
* ``wrap_voxel_fea_model``: the wrapped voxel model of
  common/WrapVoxelFEAmodel.m:7-25 (1-based element ids, fixed positions with
  their x/y/z fixity flags, loads with 1-based finest-grid node ids).

The text is the reference's byte for byte, including what MATLAB's
``fprintf`` does with a format that has more conversions than arguments:
output stops right before the first conversion without data, so
``fprintf(fid, '%s %s ', 'Node Forces:')`` writes ``"Node Forces: "`` (the
header lines of both files end in a space).  ``%d`` of an integer-valued
double prints the integer.
"""
from __future__ import annotations

import io
from pathlib import Path

import numpy as np


def _rows(buf: io.StringIO, fmt: str, a: np.ndarray):
    """MATLAB fprintf(fid, fmt, A') for a matrix A with one row per line."""
    a = np.asarray(a)
    if a.size:
        np.savetxt(buf, a.reshape(a.shape[0], -1), fmt=fmt.rstrip("\n").split(" "), delimiter=" ",
                   newline="\n")


def _ints(a) -> np.ndarray:
    return np.asarray(a, dtype=np.int64).reshape(-1)


def dataset_3dtsv_text(res, bbox, ele_map_back, nod_map_back, fixing_cond, loading_cond, stress) -> str:
    """Text of ExportDataset_3DTSV.m:14-40.  ``res`` = (resX, resY, resZ) of
    the finest level, ``bbox`` = (2, 3) boundingBox_, ``ele_map_back`` and
    ``nod_map_back`` the finest level's 1-based maps, ``fixing_cond`` compact
    1-based node ids, ``loading_cond`` (m, 4) rows [node fx fy fz], ``stress``
    (numNodes, 6) cartesianStressField_."""
    nmb = _ints(nod_map_back)
    loads = np.asarray(loading_cond, dtype=np.float64).reshape(-1, 4)
    fixed = _ints(fixing_cond)
    S = np.asarray(stress, dtype=np.float64).reshape(-1, 6)
    b = np.asarray(bbox, dtype=np.float64).reshape(2, 3)
    out = io.StringIO()
    out.write("DATASET CARTESIAN_GRID \n")                         # '%s %s' with one argument
    out.write("Resolution: %d %d %d\n" % tuple(int(v) for v in res))
    out.write("LowerBound: %.6f %.6f %.6f\n" % tuple(b[0]))
    out.write("UpperBound: %.6f %.6f %.6f\n" % tuple(b[1]))
    ele = _ints(ele_map_back)
    out.write("ELEMENTS %d int \n" % ele.size)
    _rows(out, "%d", (ele - 1)[:, None])
    out.write("Number of Stress Fields: %d\n" % 1)
    out.write("Node Forces: %d\n" % loads.shape[0])
    if loads.shape[0]:
        ids = nmb[loads[:, 0].astype(np.int64) - 1] - 1
        out.write("".join("%d %.6f %.6f %.6f\n" % (i, fx, fy, fz)
                          for i, (fx, fy, fz) in zip(ids, loads[:, 1:4])))
    out.write("Fixed Nodes: %d\n" % fixed.size)
    if fixed.size:
        _rows(out, "%d", (nmb[fixed - 1] - 1)[:, None])
    out.write("Cartesian Stress: %d\n" % nmb.size)
    _rows(out, "%.6e %.6e %.6e %.6e %.6e %.6e", S)
    return out.getvalue()


def wrapped_model_text(res, ele_map_back, nod_map_back, fixing_cond, loading_cond) -> str:
    """Text of WrapVoxelFEAmodel.m:9-24."""
    nmb = _ints(nod_map_back)
    loads = np.asarray(loading_cond, dtype=np.float64).reshape(-1, 4)
    fixed = _ints(fixing_cond)
    ele = _ints(ele_map_back)
    out = io.StringIO()
    out.write("domain type: 3D \n")                                # '%s %s %s' with one argument
    out.write("resolution: %d %d %d\n" % tuple(int(v) for v in res))
    out.write("valid elements: %d\n" % ele.size)
    _rows(out, "%d", ele[:, None])
    out.write("fixed position: %d\n" % fixed.size)
    if fixed.size:
        _rows(out, "%d %d %d %d", np.stack([nmb[fixed - 1], np.ones(fixed.size, np.int64),
                                           np.ones(fixed.size, np.int64), np.ones(fixed.size, np.int64)], 1))
    out.write("loading condition: %d\n" % loads.shape[0])
    if loads.shape[0]:
        ids = nmb[loads[:, 0].astype(np.int64) - 1]
        out.write("".join("%d %.6f %.6f %.6f\n" % (i, fx, fy, fz)
                          for i, (fx, fy, fz) in zip(ids, loads[:, 1:4])))
    return out.getvalue()


def export_dataset_3dtsv(path, *args) -> Path:
    """Write dataset_3dtsv_text(*args) to ``path`` (the reference writes
    ../data/dataset_TSV.carti)."""
    p = Path(path)
    p.write_text(dataset_3dtsv_text(*args))
    return p


def wrap_voxel_fea_model(path, *args) -> Path:
    """Write wrapped_model_text(*args) to ``path`` (the reference writes
    ../data/wrappedVoxelFEAmodel.txt)."""
    p = Path(path)
    p.write_text(wrapped_model_text(*args))
    return p
