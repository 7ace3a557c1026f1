"""fft_method_test_amd -- MI355X-native FFT permeability hot path (Bignonnet MINRES + Brinkman).

The compute lives in libfftperm.so (HIP, gfx950, C ABI in include/fftperm.h); this
package is the Python host layer over it: ctypes bindings (_lib), a mirror of the
reference's permeability classes (permeability), the Brinkman API (brinkman),
seeded geometry generators (geometry) and voxel/velocity file I/O (voxelio).
"""
from . import _lib
from ._lib import PRECISION_FP64, PRECISION_MIXED, SUPPORT_FULLSOLID, SUPPORT_INTERFACE, device_info

__all__ = ["_lib", "PRECISION_FP64", "PRECISION_MIXED", "SUPPORT_FULLSOLID", "SUPPORT_INTERFACE", "device_info"]
