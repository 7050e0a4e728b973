"""MI355X (gfx950) linear-algebra core of the Sovereign LP solver's
interior-point path: HIP kernels behind a C ABI (include/sovereign_hip.h),
with a Python mirror of the reference's interfaces in `api`."""
from . import problems  # noqa: F401

__all__ = ["problems", "api"]
