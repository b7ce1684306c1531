"""C-ABI checks that need no GPU: the library loads, exports every function
include/gmgs.h declares, and rejects invalid parameters before touching a device."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (REPO / "include" / "gmgs.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(gmgs_\w+)\s*\(", text, re.M)))


def test_library_exports_header(pkg):
    lib = pkg.load_library()
    names = declared_functions()
    assert len(names) >= 17
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(pkg.exported_symbols())


def test_default_params_mirror_globals(pkg):
    p = pkg.default_params()
    # src/GlobalVariables.m:3-9,23
    assert (p.tol, p.max_it, p.jacobi_weight) == (1e-3, 20000, 0.4)
    assert (p.youngs_modulus, p.poisson_ratio) == (1.0, 0.3)
    assert p.coarsest_resolution_control == 1e4


@pytest.mark.parametrize("bad", [dict(poisson_ratio=0.5), dict(tol=0.0), dict(precision=7),
                                 dict(jacobi_weight=-1.0), dict(max_it=-2)])
def test_invalid_params_rejected_without_device(pkg, bad):
    with pytest.raises(pkg.GMGSError) as e:
        pkg.Solver(np.ones((4, 4, 4), np.int32), **bad)
    assert e.value.code == 1


def test_status_strings(pkg):
    lib = pkg.load_library()
    assert lib.gmgs_status_string(0) == b"ok"
    assert lib.gmgs_status_string(8).startswith(b"coarsest")
