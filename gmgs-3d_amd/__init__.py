"""MI355X-native geometric-multigrid PCG for voxel hex8 elasticity."""
from . import workloads  # noqa: F401
