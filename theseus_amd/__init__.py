"""MI355X-native DGSEM right-hand-side engine (gfx950 HIP kernels behind a C-ABI).

The compute lives in libtheseus_rhs.so (built by ``theseus_amd.build``); this
package holds the host-side mirror of the reference's operator interface.
"""
from .config import RunTime, numerical_flux_name  # noqa: F401
from .mesh import BoxMesh, box_mesh, gll_nodes  # noqa: F401
from .operator import EulerOperator, MakeRHSOperator, NSOperator, RHSOperator, StabilityEstimate  # noqa: F401

__all__ = [
    "RunTime", "numerical_flux_name", "BoxMesh", "box_mesh", "gll_nodes",
    "RHSOperator", "EulerOperator", "NSOperator", "MakeRHSOperator", "StabilityEstimate",
]
