"""ctypes driver of oracle/cpu_baseline.cpp (the reference algorithm's solve in
C++/OpenMP) on a model built by the oracle's set-up.

TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg and tests/ use it; the
product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_baseline.cpp"
LIB = HERE / "libgmgs_cpu.so"


class CpuLevel(C.Structure):
    _fields_ = [("ne", C.c_int64), ("nn", C.c_int64), ("edof", C.c_void_p), ("state", C.c_int32),
                ("Ks", C.c_void_p), ("kidx", C.c_void_p), ("diag", C.c_void_p), ("fixed", C.c_void_p),
                ("nfixed", C.c_int64), ("enod", C.c_void_p), ("ninter", C.c_int64), ("tmat", C.c_void_p),
                ("tcoef", C.c_void_p), ("c2f", C.c_void_p)]


def build() -> Path:
    """gcc -O3 -fopenmp for any x86-64 host with AVX2/FMA (x86-64-v3)."""
    cmd = [os.environ.get("CXX", "g++"), "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared",
           "-std=c++17", str(SRC), "-o", str(LIB) + ".tmp"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


_lib = None


def library():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing: run oracle/cpu_baseline.py build() first")
        lib = C.CDLL(str(LIB))
        lib.cpu_cg_solve.argtypes = [C.POINTER(CpuLevel), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_int64, C.c_double, C.c_void_p, C.c_double, C.c_int, C.c_void_p,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.cpu_cg_solve.restype = C.c_int
        lib.cpu_threads.restype = C.c_int
        _lib = lib
    return _lib


def threads() -> int:
    return library().cpu_threads()


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class Prepared:
    """The model's arrays in the layout cpu_baseline.cpp reads (kept alive here)."""

    def __init__(self, model):
        self.keep = []
        L = model.numLevels
        self.levels = (CpuLevel * L)()
        for li, lev in enumerate(model.levels):
            c = self.levels[li]
            c.ne, c.nn = lev.numElements, lev.numNodes
            c.edof = self._p(_i32(lev.eDofMat - 1))
            c.enod = self._p(_i32(lev.eNodMat - 1))
            st = 0 if li == 0 else lev.storingState
            c.state = st
            c.Ks = self._p(_f64(lev.Ks))
            if st == 2:
                k = np.zeros(lev.numElements, np.int32)
                k[lev.uniqueElements - 1] = np.arange(1, lev.uniqueElements.size + 1)
                c.kidx = self._p(k)
            c.diag = self._p(_f64(lev.diagK)) if li < L - 1 else None
            fx = _i32(lev.fixedDOFs - 1)
            c.fixed, c.nfixed = self._p(fx), fx.size
            if li > 0:
                c.ninter = lev.intermediateNumNodes
                c.tmat = self._p(_i32(lev.transferMat - 1))
                c.tcoef = self._p(_f64(lev.transferMatCoeffi))
                c.c2f = self._p(_i32(lev.solidNodeMapCoarser2Finer - 1))
        from oracle import gmgs_oracle
        self.in_node = _f64(gmgs_oracle._IN_NODE)
        self.cholL = _f64(model.cholL)
        self.cfree = _i32(model.levels[-1].freeDOFs - 1)
        self.L = L
        self.omega = model.omega
        self.n = model.levels[0].numDOFs

    def _p(self, a):
        self.keep.append(a)
        return a.ctypes.data

    def solve(self, b, tol, max_it):
        """(x, its, converged, relres, seconds) of CG_solver.m on b."""
        lib = library()
        bb = _f64(b)
        x = np.zeros(self.n)
        rel, sec = C.c_double(), C.c_double()
        its = lib.cpu_cg_solve(self.levels, self.L, self.in_node.ctypes.data, self.cholL.ctypes.data,
                               self.cfree.ctypes.data, self.cfree.size, self.omega, bb.ctypes.data, tol,
                               max_it, x.ctypes.data, C.byref(rel), C.byref(sec))
        return x, abs(its), its > 0, rel.value, sec.value


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"
