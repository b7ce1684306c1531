"""ctypes binding of libgmgs.so (include/gmgs.h) and a host-side mirror of
the reference's MATLAB entry points.

The low-level class :class:`Solver` owns one C context.  The class
:class:`GMGS3D` keeps the reference's global state under the reference's
names (src/GlobalVariables.m:3-38) and exposes its entry points:
``CreateVoxilizedModel`` (common/CreateVoxilizedModel.m, cuboid branch),
``DiscretizeDesignDomain`` (common/DiscretizeDesignDomain.m),
``GetDeformation`` (kernel/GetDeformation.m) and ``GetCartesianStress``
(kernel/GetCartesianStress.m), plus the operator handles ``AtX``, ``Vcycle``,
``JacobiSmoother``, ``RestrictResidual`` and ``InterpolationDeviation``.
"""
from __future__ import annotations

import ctypes as C
import os
import warnings
from pathlib import Path
from typing import Optional

import numpy as np

# the in-tree build; GMGS_LIBRARY names another build of the same library
_LIB_PATH = Path(os.environ.get("GMGS_LIBRARY") or Path(__file__).resolve().parent / "libgmgs.so")

GMGS_PREC_FP64 = 0
GMGS_PREC_MIXED = 1
FLAG_EXACT_GALERKIN = 1
FLAG_NO_INDEX_TABLES = 2
FLAG_FR_BETA = 4
FLAG_PR_BETA = 8
FLAG_SIMPLE_KERNELS = 16
FLAG_AUTOTUNE = 32
FLAG_NO_GRAPHS = 64
FLAG_SYM_VCYCLE = 128
FLAG_NO_JACOBI_STEP = 256
FLAG_CHEBYSHEV = 512
FLAG_ASSEMBLED_COARSE = 1024
FLAG_SPLIT_COARSE = 2048
FLAG_UNFUSED_PROLONG = 4096
FLAG_UNFUSED_RESIDUAL = 8192

Q = dict(NUM_LEVELS=0, RES=1, NUM_ELEMENTS=2, NUM_NODES=3, ELE_MAP_BACK=4, ELE_MAP_FORWARD=5,
         NOD_MAP_BACK=6, NOD_MAP_FORWARD=7, ENODMAT=8, NUM_NOD2ELES=9, NODES_ON_BOUNDARY=10,
         ELEMENTS_ON_BOUNDARY=11, ELEMENT_UPWARD_MAP=12, TRANSFER_MAT=13, TRANSFER_MAT_COEFFI=14,
         SOLID_NODE_MAP_C2F=15, INTERMEDIATE_NUM_NODES=16, STORING_STATE=17, UNIQUE_ELEMENTS=18,
         KS=19, DIAGK=20, FIXED_DOFS=21, FREE_DOFS=22, KE0=23, BBOX=24, F=25, NUM_TILES=26,
         NUM_FREE_COARSEST=27, TILE_CONFIG=28, SMOOTHER=29, MAX_IT=30)
_Q_DOUBLE = {Q["TRANSFER_MAT_COEFFI"], Q["KS"], Q["DIAGK"], Q["KE0"], Q["BBOX"], Q["F"], Q["SMOOTHER"]}
_Q_INT64 = {Q["NUM_LEVELS"], Q["RES"], Q["NUM_ELEMENTS"], Q["NUM_NODES"], Q["INTERMEDIATE_NUM_NODES"],
            Q["STORING_STATE"], Q["NUM_TILES"], Q["NUM_FREE_COARSEST"], Q["TILE_CONFIG"], Q["MAX_IT"]}

STATUS = {0: "ok", 1: "invalid argument", 2: "insufficient resolution", 3: "fewer than two levels",
          4: "no loads", 5: "no constraints", 6: "node id out of range", 7: "non-finite input",
          8: "coarsest matrix not positive definite", 9: "maximum iterations exceeded",
          10: "invalid call order", 11: "device error", 12: "PCG breakdown"}


class GMGSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code} {STATUS.get(code, '?')}] {msg}")
        self.code = code


class NotConverged(GMGSError):
    pass


class Params(C.Structure):
    _fields_ = [("youngs_modulus", C.c_double), ("poisson_ratio", C.c_double), ("tol", C.c_double),
                ("jacobi_weight", C.c_double), ("coarsest_resolution_control", C.c_double),
                ("max_it", C.c_int32), ("num_levels", C.c_int32), ("precision", C.c_int32),
                ("flags", C.c_int32), ("device", C.c_int32), ("vcycle_sweeps", C.c_int32),
                ("vcycle_sweeps_coarse", C.c_int32), ("vcycle_sweeps_level2", C.c_int32),
                ("reserved", C.c_int32 * 4)]


class SolveInfo(C.Structure):
    _fields_ = [("its", C.c_int32), ("converged", C.c_int32), ("relres", C.c_double),
                ("true_relres", C.c_double), ("solve_seconds", C.c_double), ("hist_len", C.c_int64)]


_lib = None


def library_path() -> Path:
    return _LIB_PATH


def load_library() -> C.CDLL:
    """Load libgmgs.so; raises if it has not been built (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise RuntimeError(f"{_LIB_PATH} is missing: build it with gmgs-3d_amd/build.py")
    lib = C.CDLL(str(_LIB_PATH))
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    lib.gmgs_default_params.argtypes = [C.POINTER(Params)]
    lib.gmgs_default_params.restype = None
    lib.gmgs_status_string.argtypes = [C.c_int]
    lib.gmgs_status_string.restype = C.c_char_p
    lib.gmgs_create.argtypes = [C.POINTER(Params), ip, C.c_int64, C.c_int64, C.c_int64, dp,
                                C.POINTER(P)]
    lib.gmgs_set_bc.argtypes = [P, ip, C.c_int64, dp, C.c_int64]
    lib.gmgs_set_modulus.argtypes = [P, dp, C.c_int64]
    lib.gmgs_solve.argtypes = [P, dp, C.POINTER(SolveInfo), dp, C.c_int64]
    lib.gmgs_stress.argtypes = [P, dp, dp]
    lib.gmgs_adjoint.argtypes = [P, dp, dp, dp, dp, C.POINTER(SolveInfo)]
    lib.gmgs_atx.argtypes = [P, C.c_int, dp, dp]
    lib.gmgs_vcycle.argtypes = [P, dp, dp]
    lib.gmgs_smooth.argtypes = [P, C.c_int, dp, dp, dp]
    lib.gmgs_restrict.argtypes = [P, C.c_int, dp, dp]
    lib.gmgs_interpolate.argtypes = [P, C.c_int, dp, dp]
    lib.gmgs_query_size.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_int64)]
    lib.gmgs_query.argtypes = [P, C.c_int, C.c_int, P, C.c_int64]
    lib.gmgs_batch_solve.argtypes = [C.POINTER(P), C.c_int, C.POINTER(dp), C.POINTER(SolveInfo)]
    lib.gmgs_batch_solve.restype = C.c_int
    lib.gmgs_copy_probe.argtypes = [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    lib.gmgs_copy_probe.restype = C.c_int
    lib.gmgs_kernel_timing.argtypes = [P, C.c_int]
    lib.gmgs_kernel_timing.restype = C.c_int
    lib.gmgs_kernel_timing_read.argtypes = [P, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_int64)]
    lib.gmgs_kernel_timing_read.restype = C.c_int
    lib.gmgs_time_operator.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_double)]
    lib.gmgs_time_operator.restype = C.c_int
    lib.gmgs_atx_device_noalloc.argtypes = [P, C.c_int, C.c_void_p, C.c_void_p]
    lib.gmgs_atx_device_noalloc.restype = C.c_int
    lib.gmgs_last_error.argtypes = [P]
    lib.gmgs_last_error.restype = C.c_char_p
    lib.gmgs_destroy.argtypes = [P]
    lib.gmgs_destroy.restype = None
    for name in ("gmgs_create", "gmgs_set_bc", "gmgs_set_modulus", "gmgs_solve", "gmgs_stress",
                 "gmgs_adjoint", "gmgs_atx", "gmgs_vcycle", "gmgs_smooth", "gmgs_restrict",
                 "gmgs_interpolate", "gmgs_query_size", "gmgs_query"):
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def exported_symbols():
    return ["gmgs_default_params", "gmgs_status_string", "gmgs_create", "gmgs_set_bc",
            "gmgs_set_modulus", "gmgs_solve", "gmgs_stress", "gmgs_adjoint", "gmgs_atx",
            "gmgs_vcycle", "gmgs_smooth", "gmgs_restrict", "gmgs_interpolate", "gmgs_query_size",
            "gmgs_query", "gmgs_last_error", "gmgs_destroy", "gmgs_kernel_timing",
            "gmgs_kernel_timing_read", "gmgs_batch_solve", "gmgs_copy_probe",
            "gmgs_time_operator", "gmgs_atx_device_noalloc"]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def default_params(**kw) -> Params:
    lib = load_library()
    p = Params()
    lib.gmgs_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


class Solver:
    """One solver context on one GPU (thin wrapper of the C ABI)."""

    def __init__(self, vox: np.ndarray, bbox=None, **params):
        self._lib = load_library()
        self._ctx = C.c_void_p()
        self.params = default_params(**params)
        v = np.asfortranarray(np.asarray(vox, dtype=np.int32))
        if v.ndim != 3:
            raise ValueError("vox must be (ny, nx, nz)")
        ny, nx, nz = v.shape
        flat = np.ascontiguousarray(v.ravel(order="F"))
        bb = None
        if bbox is not None:
            b = np.asarray(bbox, dtype=np.float64).reshape(2, 3)
            bb = np.ascontiguousarray(b.reshape(-1))
        st = self._lib.gmgs_create(C.byref(self.params), _iptr(flat), ny, nx, nz,
                                   _dptr(bb) if bb is not None else None, C.byref(self._ctx))
        if st != 0:
            msg = self._lib.gmgs_last_error(None).decode()
            raise GMGSError(st, msg)

    def _check(self, st):
        if st == 0:
            return
        msg = self._lib.gmgs_last_error(self._ctx).decode()
        if st == 9:
            raise NotConverged(st, msg)
        raise GMGSError(st, msg)

    def close(self):
        if self._ctx:
            self._lib.gmgs_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- queries
    def query(self, name: str, level: int = 1):
        key = Q[name]
        n = C.c_int64()
        self._check(self._lib.gmgs_query_size(self._ctx, key, level, C.byref(n)))
        if key in _Q_DOUBLE:
            out = np.zeros(max(n.value, 1), dtype=np.float64)
        elif key in _Q_INT64:
            out = np.zeros(max(n.value, 1), dtype=np.int64)
        else:
            out = np.zeros(max(n.value, 1), dtype=np.int32)
        self._check(self._lib.gmgs_query(self._ctx, key, level, out.ctypes.data_as(C.c_void_p),
                                         out.size))
        out = out[:n.value]
        if key in _Q_INT64 and n.value == 1:
            return int(out[0])
        return out

    @property
    def num_levels(self) -> int:
        return self.query("NUM_LEVELS")

    def smoother(self, level: int = 1):
        """(lambda_max estimate of D^-1 A on the free DOFs or 0, [w_1 .. w_nu]): the
        sweep weights of the symmetric cycle on a smoothed level, after set_bc."""
        q = self.query("SMOOTHER", level)
        return float(q[0]), [float(w) for w in q[1:] if w != 0.0]

    def num_nodes(self, level=1) -> int:
        return self.query("NUM_NODES", level)

    def num_elements(self, level=1) -> int:
        return self.query("NUM_ELEMENTS", level)

    def level_arrays(self, level: int) -> dict:
        """meshHierarchy_(level) index fields, MATLAB layouts."""
        ne = self.num_elements(level)
        d = dict(numElements=ne, numNodes=self.num_nodes(level),
                 eleMapBack=self.query("ELE_MAP_BACK", level),
                 eleMapForward=self.query("ELE_MAP_FORWARD", level),
                 nodMapBack=self.query("NOD_MAP_BACK", level),
                 nodMapForward=self.query("NOD_MAP_FORWARD", level),
                 eNodMat=self.query("ENODMAT", level).reshape(ne, 8, order="F"),
                 numNod2Eles=self.query("NUM_NOD2ELES", level),
                 nodesOnBoundary=self.query("NODES_ON_BOUNDARY", level),
                 elementsOnBoundary=self.query("ELEMENTS_ON_BOUNDARY", level))
        if level >= 2:
            d["elementUpwardMap"] = self.query("ELEMENT_UPWARD_MAP", level).reshape(ne, 8, order="F")
            d["transferMat"] = self.query("TRANSFER_MAT", level).reshape(27, ne, order="F")
            d["transferMatCoeffi"] = self.query("TRANSFER_MAT_COEFFI", level)
            d["solidNodeMapCoarser2Finer"] = self.query("SOLID_NODE_MAP_C2F", level)
            d["intermediateNumNodes"] = self.query("INTERMEDIATE_NUM_NODES", level)
        return d

    # ---- set-up and solve
    def set_bc(self, fixed, loads):
        fx = np.ascontiguousarray(np.asarray(fixed, dtype=np.int32).ravel())
        ld = np.asarray(loads, dtype=np.float64).reshape(-1, 4)
        ldc = np.ascontiguousarray(ld.ravel(order="F"))
        self._check(self._lib.gmgs_set_bc(self._ctx, _iptr(fx), fx.size, _dptr(ldc), ld.shape[0]))

    def set_modulus(self, modulus):
        if modulus is None:
            self._check(self._lib.gmgs_set_modulus(self._ctx, None, 0))
            return
        m = np.ascontiguousarray(np.asarray(modulus, dtype=np.float64).ravel())
        self._check(self._lib.gmgs_set_modulus(self._ctx, _dptr(m), m.size))

    def solve(self, raise_on_nonconvergence=True, copy=True):
        """GetDeformation: returns (U, info, iteration history).  With
        ``copy=False`` the solution stays on the device (U is None); a later
        ``stress()`` or ``adjoint()`` uses it there."""
        n = 3 * self.num_nodes(1)
        U = np.zeros(n) if copy else None
        info = SolveInfo()
        cap = int(self.params.max_it) + 2
        hist = np.zeros(3 * cap)
        st = self._lib.gmgs_solve(self._ctx, _dptr(U) if copy else None, C.byref(info),
                                  _dptr(hist), cap)
        if st == 9 and not raise_on_nonconvergence:
            warnings.warn("Exceed the maximum iterate numbers")
        else:
            self._check(st)
        h = hist[:3 * info.hist_len].reshape(-1, 3)
        return U, info, h

    def stress(self, U=None):
        nn = self.num_nodes(1)
        S = np.zeros(6 * nn)
        u = None if U is None else np.ascontiguousarray(np.asarray(U, dtype=np.float64))
        self._check(self._lib.gmgs_stress(self._ctx, _dptr(u) if u is not None else None, _dptr(S)))
        return S.reshape(nn, 6, order="F")

    def adjoint(self, g, U=None):
        n = 3 * self.num_nodes(1)
        gg = np.ascontiguousarray(np.asarray(g, dtype=np.float64))
        lam = np.zeros(n)
        dJ = np.zeros(self.num_elements(1))
        info = SolveInfo()
        u = None if U is None else np.ascontiguousarray(np.asarray(U, dtype=np.float64))
        self._check(self._lib.gmgs_adjoint(self._ctx, _dptr(gg), _dptr(u) if u is not None else None,
                                           _dptr(lam), _dptr(dJ), C.byref(info)))
        return lam, dJ, info

    # ---- kernel timing
    KERNELS = dict(matvec64=0, jacobi64=1, vsmooth1=2, restrict1=3, prolong1=4, iteration=5, matvec64_p=6)

    def kernel_timing(self, enable=True):
        self._check(self._lib.gmgs_kernel_timing(self._ctx, 1 if enable else 0))

    def kernel_times(self):
        out = {}
        for name, k in self.KERNELS.items():
            ms = C.c_double()
            n = C.c_int64()
            self._check(self._lib.gmgs_kernel_timing_read(self._ctx, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def time_operator(self, kernel="matvec64", reps=20):
        """Mean duration (ms) of ``reps`` back-to-back launches of a level-1
        FP64 operator ("matvec64", "jacobi64", or "matvec64_p": the matvec
        with the direction update fused, where the solver uses it) on
        device-resident vectors."""
        ms = C.c_double()
        self._check(self._lib.gmgs_time_operator(self._ctx, self.KERNELS[kernel], int(reps),
                                                 C.byref(ms)))
        return ms.value

    # ---- operator handles (CG_solver(@AtX, @Vcycle, ...))
    def atx(self, x, level=1):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        y = np.zeros_like(x)
        self._check(self._lib.gmgs_atx(self._ctx, level, _dptr(x), _dptr(y)))
        return y

    def vcycle(self, r):
        r = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
        z = np.zeros_like(r)
        self._check(self._lib.gmgs_vcycle(self._ctx, _dptr(r), _dptr(z)))
        return z

    def smooth(self, r, level=1):
        r = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
        dx = np.zeros_like(r)
        ro = np.zeros_like(r)
        self._check(self._lib.gmgs_smooth(self._ctx, level, _dptr(r), _dptr(dx), _dptr(ro)))
        return dx, ro

    def restrict(self, r_fine, level):
        r = np.ascontiguousarray(np.asarray(r_fine, dtype=np.float64))
        out = np.zeros(3 * self.num_nodes(level))
        self._check(self._lib.gmgs_restrict(self._ctx, level, _dptr(r), _dptr(out)))
        return out

    def interpolate(self, x_coarse, level):
        x = np.ascontiguousarray(np.asarray(x_coarse, dtype=np.float64))
        out = np.zeros(3 * self.num_nodes(level - 1))
        self._check(self._lib.gmgs_interpolate(self._ctx, level, _dptr(x), _dptr(out)))
        return out


def copy_probe(nbytes: int, reps: int = 10, device: int = 0) -> float:
    """Mean duration (ms) of a plain device copy of nbytes."""
    lib = load_library()
    ms = C.c_double()
    st = lib.gmgs_copy_probe(device, nbytes, reps, C.byref(ms))
    if st != 0:
        raise GMGSError(st, lib.gmgs_last_error(None).decode())
    return ms.value


def batch_solve(solvers, copy=True):
    """Solve every solver's problem together (gmgs_batch_solve); returns
    [(U, SolveInfo)] in order (U is None with ``copy=False``: the solutions
    stay on the device)."""
    lib = load_library()
    n = len(solvers)
    ctxs = (C.c_void_p * n)(*[s._ctx.value for s in solvers])
    Us = [np.zeros(3 * s.num_nodes(1)) if copy else None for s in solvers]
    ptrs = (C.POINTER(C.c_double) * n)(*[_dptr(u) if copy else None for u in Us])
    infos = (SolveInfo * n)()
    st = lib.gmgs_batch_solve(ctxs, n, ptrs, infos)
    if st not in (0, 9):
        raise GMGSError(st, lib.gmgs_last_error(solvers[0]._ctx).decode())
    return [(Us[i], infos[i]) for i in range(n)]


class GMGS3D:
    """The reference's global workspace and entry points (src/GlobalVariables.m)."""

    def __init__(self):
        self.tol_ = 1.0e-3
        self.maxIT_ = 20000
        self.weightFactorJacobi_ = 0.4
        self.moduls_ = 1.0
        self.poissonRatio_ = 0.3
        self.coarsestResolutionControl_ = 10000.0
        self.numLevels_ = 0            # 0: the level rule; >0 overrides it
        self.precision = GMGS_PREC_FP64
        self.flags = 0
        self.voxelizedVolume_ = None
        self.boundingBox_ = None
        self.nelx_ = self.nely_ = self.nelz_ = 0
        self.loadingCond_ = np.zeros((0, 4))
        self.fixingCond_ = np.zeros(0, dtype=np.int32)
        self.U_ = None
        self.cartesianStressField_ = None
        self.iterationHist_ = np.zeros((0, 3))
        self.solver: Optional[Solver] = None

    def CreateVoxilizedModel(self, *args):
        """common/CreateVoxilizedModel.m: (triFaceModelFile, finestResolutionControl)
        voxelises a .ply / .obj surface mesh (:11-26), (nx, ny, nz) builds a
        solid cuboid (:27-32)."""
        if len(args) == 2:
            from .voxelize import create_voxelized_model
            self.voxelizedVolume_, self.boundingBox_ = create_voxelized_model(args[0], args[1])
            self.nely_, self.nelx_, self.nelz_ = self.voxelizedVolume_.shape
        elif len(args) == 3:
            nx, ny, nz = (int(v) for v in args)
            self.nelx_, self.nely_, self.nelz_ = nx, ny, nz
            self.voxelizedVolume_ = np.ones((ny, nx, nz), dtype=np.int32)
            self.boundingBox_ = np.array([[0.0, 0.0, 0.0], [nx, ny, nz]])
        else:
            raise ValueError("Wrong input for voxelization!")
        if self.solver is not None:
            self.solver.close()
            self.solver = None
        return self.voxelizedVolume_

    def DiscretizeDesignDomain(self):
        """Mesh, hierarchy and BC-independent solver set-up on the device."""
        if self.solver is not None:
            self.solver.close()
        self.solver = Solver(self.voxelizedVolume_, self.boundingBox_,
                             youngs_modulus=self.moduls_, poisson_ratio=self.poissonRatio_,
                             tol=self.tol_, jacobi_weight=self.weightFactorJacobi_,
                             coarsest_resolution_control=self.coarsestResolutionControl_,
                             max_it=self.maxIT_, num_levels=self.numLevels_,
                             precision=self.precision, flags=self.flags)
        self.boundingBox_ = self.solver.query("BBOX").reshape(2, 3)
        return self.solver

    def GetDeformation(self):
        """kernel/GetDeformation.m: BCs are re-applied on every call."""
        if self.solver is None:
            self.DiscretizeDesignDomain()
        self.solver.set_bc(self.fixingCond_, self.loadingCond_)
        U, info, hist = self.solver.solve(raise_on_nonconvergence=False)
        self.U_ = U
        self.iterationHist_ = hist
        return U

    def GetCartesianStress(self):
        if self.U_ is None:
            warnings.warn("No Deformation Field Available, Please Compute Deformation First "
                          "by running FUNCTION GetDeformation!")
            return None
        self.cartesianStressField_ = self.solver.stress(self.U_)
        return self.cartesianStressField_

    # on-disk formats (common/ExportDataset_3DTSV.m, common/WrapVoxelFEAmodel.m)
    def _finest_maps(self):
        res = self.solver.query("RES", 1)
        return (res, self.solver.query("ELE_MAP_BACK", 1), self.solver.query("NOD_MAP_BACK", 1))

    def ExportDataset_3DTSV(self, fileName="dataset_TSV.carti"):
        """The .carti stress dataset of the last GetCartesianStress."""
        from .export import export_dataset_3dtsv
        if self.cartesianStressField_ is None:
            warnings.warn("No Cartesian Stress Available!")
            return None
        res, emb, nmb = self._finest_maps()
        return export_dataset_3dtsv(fileName, res, self.boundingBox_, emb, nmb, self.fixingCond_,
                                    self.loadingCond_, self.cartesianStressField_)

    def WrapVoxelFEAmodel(self, fileName="wrappedVoxelFEAmodel.txt"):
        from .export import wrap_voxel_fea_model
        if self.solver is None:
            self.DiscretizeDesignDomain()
        res, emb, nmb = self._finest_maps()
        return wrap_voxel_fea_model(fileName, res, emb, nmb, self.fixingCond_, self.loadingCond_)

    # operator handles
    def AtX(self, x, level=1):
        return self.solver.atx(x, level)

    def Vcycle(self, r):
        return self.solver.vcycle(r)

    def JacobiSmoother(self, x, r, level=1):
        dx, r2 = self.solver.smooth(r, level)
        return x + dx, r2

    def RestrictResidual(self, r, ii):
        return self.solver.restrict(r, ii)

    def InterpolationDeviation(self, x, ii):
        return self.solver.interpolate(x, ii)
