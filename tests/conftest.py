import importlib.util
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def load_package():
    """Import the ``gmgs-3d_amd`` directory as the ``gmgs3d_amd`` package."""
    if "gmgs3d_amd" in sys.modules:
        return sys.modules["gmgs3d_amd"]
    root = REPO / "gmgs-3d_amd"
    spec = importlib.util.spec_from_file_location(
        "gmgs3d_amd", root / "__init__.py", submodule_search_locations=[str(root)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gmgs3d_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    from oracle import gmgs_oracle
    return gmgs_oracle
