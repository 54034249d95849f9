"""h2g: MI355X-native hierarchical (H2) matvec with H2Factory's map interface.

The product path is ``libh2g.so`` (HIP kernels for gfx950 behind the C ABI in
``include/h2g.h``); this package is its host side: the method builder (``method``,
``tree``), the map API (``maps``), the serving loop (``serve``) and the autotuner
(``autotune``).
"""

from .maps import (ArgumentError, ConcurrentWorkspaceError, DimensionMismatch, GpuH2Core, H2Workspace, Matrix,
                   adjoint, farmul_, mul_, nearmul_, similarworkspace, transpose)
from .method import MapArrays, build_map

__all__ = [
    "ArgumentError", "ConcurrentWorkspaceError", "DimensionMismatch", "GpuH2Core", "H2Workspace", "Matrix",
    "adjoint", "farmul_", "mul_", "nearmul_", "similarworkspace", "transpose", "MapArrays", "build_map",
]
