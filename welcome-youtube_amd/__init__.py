"""MI355X (gfx950) 2D DPD step: HIP kernels behind a C ABI (include/oercel_mi.h),
driven from Python through ctypes."""
from . import abi, params  # noqa: F401
