"""MI355X-native SFWI per-shot elastic modeling and gradient path.

The compute path is libsfwi_amd.so (HIP for gfx950) behind the C ABI in
include/sfwi_abi.h; this package holds its ctypes binding (`_lib`), the
host-side mirror of the reference interface (`elastic`) and the seeded
synthetic workloads (`workloads`).
"""
from . import workloads  # noqa: F401

__all__ = ["workloads", "elastic", "_lib"]
