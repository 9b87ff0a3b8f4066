"""MI355X-native IB3D hot path: HIP kernels behind a C ABI (include/ib3d.h),
with host-side case I/O and a ctypes binding that mirrors the reference's
subroutines."""
from . import cases  # noqa: F401

__all__ = ["cases", "native"]
