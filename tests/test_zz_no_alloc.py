"""gmgs_atx_device_noalloc (no device allocation, file-scope scratch array)
against gmgs_atx: the same kernels on the same operator and input, so the
results must be bitwise equal.  Two calls in a row, the second on a smaller
level (the scratch keeps the first call's contents: the input's void and
halo entries must be cleared again), then a level that does not fit the
scratch array is refused.

Device buffers come from the HIP runtime libgmgs.so is linked against (its
own symbol lookup), not from torch."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DeviceArray:
    """A hipMalloc'd double array owned by the test (the entry point itself
    allocates nothing)."""

    def __init__(self, lib, n):
        self.lib, self.n, self.p = lib, n, C.c_void_p()
        assert lib.hipMalloc(C.byref(self.p), C.c_size_t(8 * n)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.lib.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), 1) == 0

    def download(self):
        out = np.empty(self.n)
        assert self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.lib.hipFree(self.p)


def _ellipsoid_solver(pkg, nx, ny, nz, L, seed):
    p = pkg.workloads.ellipsoid_problem(nx, ny, nz, L, seed=seed, n_ell=5, n_loads=60)
    s = pkg.Solver(p.vox, p.bbox, num_levels=L, tol=1e-8, max_it=200)
    s.set_bc(p.fixed, p.loads)
    return s


def test_noalloc_atx_matches_atx(pkg):
    lib = pkg.load_library()
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipFree.argtypes = [C.c_void_p]
    s = _ellipsoid_solver(pkg, 48, 40, 64, 4, seed=2)
    rng = np.random.default_rng(17)
    # level 1, then level 2 (smaller) of the same context, then level 1 of a
    # smaller context
    small = _ellipsoid_solver(pkg, 24, 20, 32, 3, seed=5)
    for solver, level in ((s, 1), (s, 2), (small, 1)):
        n = 3 * solver.num_nodes(level)
        x = rng.standard_normal(n)
        ref = solver.atx(x, level)
        dx, dy = DeviceArray(lib, n), DeviceArray(lib, n)
        dx.upload(x)
        dy.upload(np.full(n, np.nan))
        log(f"[noalloc] level {level}: {n} DOFs")
        st = lib.gmgs_atx_device_noalloc(solver._ctx, level, dx.p, dy.p)
        assert st == 0, lib.gmgs_last_error(solver._ctx)
        y = dy.download()
        dx.free()
        dy.free()
        assert np.array_equal(y, ref)
    small.close()
    # a level whose two padded vectors exceed the scratch array is refused
    big = pkg.Solver(np.ones((160, 160, 160), np.int32), num_levels=5, tol=1e-6, max_it=10)
    nn = big.num_nodes(1)
    fixed = np.arange(1, 161 * 161 + 1, dtype=np.int32)
    big.set_bc(fixed, np.array([[nn, 0.0, 0.0, -1.0]]))
    n = 3 * nn
    dx, dy = DeviceArray(lib, n), DeviceArray(lib, n)
    st = lib.gmgs_atx_device_noalloc(big._ctx, 1, dx.p, dy.p)
    assert st == 1 and b"scratch" in lib.gmgs_last_error(big._ctx)
    # level 2 of the same context fits
    n2 = 3 * big.num_nodes(2)
    x2 = rng.standard_normal(n2)
    d2x, d2y = DeviceArray(lib, n2), DeviceArray(lib, n2)
    d2x.upload(x2)
    assert lib.gmgs_atx_device_noalloc(big._ctx, 2, d2x.p, d2y.p) == 0
    assert np.array_equal(d2y.download(), big.atx(x2, 2))
    for d in (dx, dy, d2x, d2y):
        d.free()
    big.close()
    s.close()
