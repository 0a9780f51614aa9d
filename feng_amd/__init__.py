"""feng_amd: MI355X-native assemble-and-solve path of the feNG finite-element
library (gfx950 HIP kernels behind a C ABI, see include/fe_hip.h)."""
from .api import (  # noqa: F401
    Fn, LinearSystemHIP, Mesh, Numbering, SolveResult, essential_values,
)
