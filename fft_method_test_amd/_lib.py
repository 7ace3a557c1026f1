"""ctypes loader for libfftperm.so (the C ABI declared in include/fftperm.h).

The library is built in-tree (python -m fft_method_test_amd.build).  There is no
CPU fallback: if the library is missing or the GPU is unavailable, calls raise.
Error codes map back to the exception types the reference throws:
FFTPERM_EINVAL -> ValueError (std::invalid_argument), FFTPERM_ERANGE -> IndexError
(std::out_of_range), everything else -> RuntimeError (std::runtime_error).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfftperm.so")

OK, EINVAL, ERANGE, ERUNTIME, EDEVICE, EUNSUPPORTED = range(6)
SUPPORT_FULLSOLID, SUPPORT_INTERFACE = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1
PASS_XFWD, PASS_XINV, PASS_YPASS, PASS_ZGREEN, PASS_PLANE = 0, 1, 2, 3, 4
MIXED_CG_RELIABLE, MIXED_MINRES_REFINED, MIXED_CR_RELIABLE, MIXED_CR_FP32_VECTORS = 0, 1, 2, 3
# tile variants the specialised passes offer at a length (codes as in fftperm_profile_stats)
ZGREEN_CODES = (11, 12, 21, 22, 41, 42, 81, 82, 161, 162, 44, 84, 164, 324, 45, 85, 165, 86, 87, 88, 166, 167, 168, 326, 327)

EXPORTED_SYMBOLS = (
    "fftperm_last_error", "fftperm_device_info", "fftperm_create", "fftperm_destroy",
    "fftperm_set_medium", "fftperm_set_medium_dev", "fftperm_support_noalloc", "fftperm_solve_batch",
    "fftperm_porosity",
    "fftperm_get_support",
    "fftperm_green_tensor", "fftperm_fft_forward", "fftperm_fft_inverse", "fftperm_spectral_green",
    "fftperm_green_apply", "fftperm_background_force", "fftperm_rhs", "fftperm_operator_apply",
    "fftperm_operator_apply_fp32", "fftperm_operator_apply_fp32v", "fftperm_check_edges",
    "fftperm_probe_copy",
    "fftperm_minres", "fftperm_set_refinement", "fftperm_set_mixed_solver", "fftperm_pin_variant", "fftperm_recover",
    "fftperm_solve_direction", "fftperm_solve_tensor", "fftperm_residual_norm", "fftperm_symmetry_gap",
    "fftperm_profile", "fftperm_profile_read", "fftperm_set_graphs", "fftperm_unet_create", "fftperm_unet_destroy",
    "fftperm_unet_num_params", "fftperm_unet_get_params", "fftperm_unet_set_params", "fftperm_unet_loss_grad",
    "fftperm_unet_predict", "fftperm_solve_tensor_warm", "fftperm_brinkman_solve", "fftperm_set_tensor_batch",
    "fftperm_time_apply",
)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_uint64), ("ny", C.c_uint64), ("nz", C.c_uint64),
                ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_uint64), ("final_rel_residual", C.c_double), ("converged", C.c_int),
                ("history_len", C.c_uint64), ("refinement_sweeps", C.c_uint64)]


class ProfileStats(C.Structure):
    _fields_ = [("zgreen_ms", C.c_double * 2), ("zgreen_launches", C.c_uint64 * 2),
                ("apply_ms", C.c_double * 2), ("apply_launches", C.c_uint64 * 2),
                ("N", C.c_uint64), ("Nh", C.c_uint64), ("nb", C.c_uint64), ("zgreen_tile", C.c_int * 2),
                ("xfwd_tile", C.c_int * 2), ("xinv_tile", C.c_int * 2), ("ypass_tile", C.c_int * 2),
                ("zgreen_rhs", C.c_uint64 * 2), ("apply_rhs", C.c_uint64 * 2)]


class BrinkmanInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double), ("converged", C.c_int)]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p

_lib = None


def load():
    """Load libfftperm.so and declare the signatures (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m fft_method_test_amd.build`")
    # One HIP runtime per process: torch ships its own libamdhip64 and binds to whichever copy of
    # that soname is loaded first.  Importing torch before this library makes both use torch's
    # copy; loading this library first left torch with no usable device.
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    L.fftperm_last_error.restype = C.c_char_p
    L.fftperm_device_info.argtypes = [C.c_char_p, C.c_size_t]
    L.fftperm_create.argtypes = [C.POINTER(Grid), C.c_double, C.c_int, C.c_int, C.POINTER(_vp)]
    L.fftperm_destroy.argtypes = [_vp]
    L.fftperm_set_medium.argtypes = [_vp, _u8p, C.c_int, C.POINTER(C.c_uint64)]
    L.fftperm_set_medium_dev.argtypes = [_vp, _vp, C.c_int, C.POINTER(C.c_uint64)]
    L.fftperm_support_noalloc.argtypes = [C.c_uint64] * 3 + [_vp, C.c_int, _vp, _vp, C.c_uint64, _vp, _vp,
                                                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.fftperm_solve_batch.argtypes = [_vp, C.c_uint32, C.POINTER(C.c_void_p), C.c_int, _dp, C.c_double, C.c_uint64,
                                      _dp, _vp]
    L.fftperm_porosity.argtypes = [_vp, C.POINTER(C.c_double)]
    L.fftperm_get_support.argtypes = [_vp, _vp, _vp]
    L.fftperm_green_tensor.argtypes = [_vp, _dp]
    L.fftperm_fft_forward.argtypes = [_vp, _dp, _dp]
    L.fftperm_fft_inverse.argtypes = [_vp, _dp, _dp]
    L.fftperm_spectral_green.argtypes = [_vp, _dp, _dp, _dp]
    L.fftperm_green_apply.argtypes = [_vp, _dp, _dp]
    L.fftperm_background_force.argtypes = [_vp, _dp, _dp]
    L.fftperm_rhs.argtypes = [_vp, _dp, _dp]
    L.fftperm_operator_apply.argtypes = [_vp, _dp, _dp]
    L.fftperm_operator_apply_fp32.argtypes = [_vp, _dp, _dp]
    L.fftperm_operator_apply_fp32v.argtypes = [_vp, _dp, _dp]
    L.fftperm_check_edges.argtypes = [C.c_int, C.c_int, C.c_int]
    L.fftperm_probe_copy.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.fftperm_minres.argtypes = [_vp, _dp, _vp, C.c_double, C.c_uint64, _dp, _vp, C.POINTER(SolveInfo)]
    L.fftperm_set_refinement.argtypes = [_vp, C.c_double, C.c_double, C.c_uint32]
    L.fftperm_set_mixed_solver.argtypes = [_vp, C.c_int, C.c_double]
    L.fftperm_pin_variant.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.fftperm_set_graphs.argtypes = [_vp, C.c_int]
    L.fftperm_set_tensor_batch.argtypes = [_vp, C.c_int]
    L.fftperm_time_apply.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.fftperm_recover.argtypes = [_vp, _dp, _dp, _vp, _dp]
    L.fftperm_solve_direction.argtypes = [_vp, _dp, C.c_double, C.c_uint64, _vp, _dp, _vp, C.POINTER(SolveInfo)]
    L.fftperm_solve_tensor.argtypes = [_vp, C.c_double, C.c_uint64, _dp, _vp]
    L.fftperm_residual_norm.argtypes = [_vp, _dp, _dp, C.POINTER(C.c_double)]
    L.fftperm_symmetry_gap.argtypes = [_vp, _dp, _dp, C.POINTER(C.c_double)]
    L.fftperm_profile.argtypes = [_vp, C.c_int]
    L.fftperm_profile_read.argtypes = [_vp, C.POINTER(ProfileStats)]
    L.fftperm_brinkman_solve.argtypes = [C.c_uint64] * 3 + [_dp, _dp, C.c_double, C.c_double, C.c_int, _dp, C.c_int,
                                                             _dp, _vp, C.POINTER(BrinkmanInfo)]
    _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.fftperm_unet_create.argtypes = [C.c_uint64, C.POINTER(_vp)]
    L.fftperm_unet_destroy.argtypes = [_vp]
    L.fftperm_unet_num_params.argtypes = [_vp, C.POINTER(C.c_uint64)]
    L.fftperm_unet_get_params.argtypes = [_vp, _fp]
    L.fftperm_unet_set_params.argtypes = [_vp, _fp]
    L.fftperm_unet_loss_grad.argtypes = [_vp, _vp, _vp, C.c_uint32, C.c_float, C.POINTER(C.c_double), _vp]
    L.fftperm_unet_predict.argtypes = [_vp, _vp, C.c_uint32, _vp]
    L.fftperm_solve_tensor_warm.argtypes = [_vp, _vp, C.c_double, C.c_uint64, _dp, _vp, _vp]
    for name in EXPORTED_SYMBOLS:
        getattr(L, name).restype = C.c_char_p if name == "fftperm_last_error" else C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    """Raise the exception type the reference would throw for a failed call."""
    if rc == OK:
        return
    msg = load().fftperm_last_error().decode(errors="replace")
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ERANGE:
        raise IndexError(msg)
    raise RuntimeError(f"fftperm error {rc}: {msg}")


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f64(a, n=None):
    out = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if n is not None and out.size != n:
        raise ValueError(f"expected {n} values, got {out.size}")
    return out


def device_info() -> str:
    buf = C.create_string_buffer(256)
    check(load().fftperm_device_info(buf, 256))
    return buf.value.decode()
