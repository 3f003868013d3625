"""ctypes binding of libfftm_mi.so (the C ABI in include/fftm_mi.h).

Loading never falls back to anything: if the shared library is missing the
import of the transform API raises, naming the build command.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfftm_mi.so")

SUCCESS = 0
INVALID_PLAN = 1
ALLOC_FAILED = 2
INVALID_TYPE = 3
INVALID_VALUE = 4
INTERNAL_ERROR = 5
EXEC_FAILED = 6
SETUP_FAILED = 7
INVALID_SIZE = 8
UNALIGNED_DATA = 9
INVALID_DEVICE = 11
NO_WORKSPACE = 13
NOT_IMPLEMENTED = 14
NOT_SUPPORTED = 16

R2C, C2R, C2CF, C2CB = 1, 2, 3, 4
F32, F64 = 0, 1
F32_H16 = 2  # FP32 transforms, FP16-stored intermediate spectra
F32_H16X2 = 3  # ... stored as hi + lo FP16 pairs (8 bytes per bin)
FLAG_OVERWRITE_INPUT = 1
FLAG_NO_PLANE_FUSION = 2

LAYOUT_3D = 0
LAYOUT_4D_PENCIL = 1
LAYOUT_4D_SLAB = 2
LAYOUT_4D_XZWY = 3

_STATUS_NAMES = {
    SUCCESS: "success", INVALID_PLAN: "invalid plan", ALLOC_FAILED: "allocation failed",
    INVALID_TYPE: "invalid type", INVALID_VALUE: "invalid value", INTERNAL_ERROR: "internal error",
    EXEC_FAILED: "execution failed", SETUP_FAILED: "setup failed", INVALID_SIZE: "invalid size",
    UNALIGNED_DATA: "unaligned data", INVALID_DEVICE: "invalid device", NO_WORKSPACE: "missing workspace",
    NOT_IMPLEMENTED: "not implemented", NOT_SUPPORTED: "not supported",
}


class FftmMiError(RuntimeError):
    """A non-zero status from the C ABI (hipfftResult numbering)."""

    def __init__(self, status: int, where: str, message: str):
        self.status = status
        super().__init__(f"{where}: {_STATUS_NAMES.get(status, 'unknown')} ({status}): {message}")


OP_POISSON, OP_GAUSSIAN, OP_SHELL, OP_DERIVATIVE, OP_POISSON_GRADIENT, OP_BANDPASS = 1, 2, 3, 4, 5, 6


class SpectralOp(ctypes.Structure):
    """fftm_mi_spectral_op."""
    _fields_ = [("kind", ctypes.c_int32), ("nshells", ctypes.c_int32), ("scale", ctypes.c_double),
                ("sigma", ctypes.c_double), ("coeffs", ctypes.c_void_p), ("axis", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("klo", ctypes.c_double), ("khi", ctypes.c_double)]


class PlanDesc(ctypes.Structure):
    """fftm_mi_plan_desc (mirrors fftm::wrap::fft_plan_descriptor)."""
    _fields_ = [
        ("dir", ctypes.c_int32), ("precision", ctypes.c_int32), ("rank", ctypes.c_int32),
        ("default_layout", ctypes.c_int32), ("n", ctypes.c_int64 * 3), ("inembed", ctypes.c_int64 * 3),
        ("istride", ctypes.c_int64), ("idist", ctypes.c_int64), ("onembed", ctypes.c_int64 * 3),
        ("ostride", ctypes.c_int64), ("odist", ctypes.c_int64), ("batch", ctypes.c_int64),
        ("flags", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


# every symbol include/fftm_mi.h declares, with (restype, argtypes)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_szp = ctypes.POINTER(ctypes.c_size_t)

SIGNATURES = {
    "fftm_mi_plan_validate": (ctypes.c_int, [ctypes.POINTER(PlanDesc), _szp]),
    "fftm_mi_plan_create": (ctypes.c_int, [ctypes.POINTER(PlanDesc), ctypes.POINTER(_vp)]),
    "fftm_mi_plan_get_work_size": (ctypes.c_int, [_vp, _szp]),
    "fftm_mi_plan_set_work_area": (ctypes.c_int, [_vp, _vp]),
    "fftm_mi_plan_set_stream": (ctypes.c_int, [_vp, _vp]),
    "fftm_mi_plan_exec": (ctypes.c_int, [_vp, _i32, _vp, _vp, _i32]),
    "fftm_mi_plan_get_descriptor": (ctypes.c_int, [_vp, ctypes.POINTER(PlanDesc)]),
    "fftm_mi_plan_num_passes": (ctypes.c_int, [_vp, _i32p]),
    "fftm_mi_plan_destroy": (ctypes.c_int, [_vp]),
    "fftm_mi_transpose_4d": (ctypes.c_int, [_i32, _i64p, _i32p, _vp, _vp, _vp, _i32]),
    "fftm_mi_ffts_validate": (ctypes.c_int, [_i32, _i64p, _i32, _i32, _szp]),
    "fftm_mi_ffts_create": (ctypes.c_int, [_i32, _i64p, _i32, _i32, ctypes.POINTER(_vp)]),
    "fftm_mi_ffts_get_work_size": (ctypes.c_int, [_vp, _szp]),
    "fftm_mi_ffts_set_work_area": (ctypes.c_int, [_vp, _vp]),
    "fftm_mi_ffts_set_stream": (ctypes.c_int, [_vp, _vp]),
    "fftm_mi_ffts_forward": (ctypes.c_int, [_vp, _vp, _vp, _i32]),
    "fftm_mi_ffts_backward": (ctypes.c_int, [_vp, _vp, _vp, _i32]),
    "fftm_mi_ffts_num_passes": (ctypes.c_int, [_vp, _i32p, _i32p]),
    "fftm_mi_ffts_run_pass": (ctypes.c_int, [_vp, _i32, _i32, _vp, _vp, _i32]),
    "fftm_mi_ffts_pass_bytes": (ctypes.c_int, [_vp, _i32, _i32, _i64p]),
    "fftm_mi_ffts_set_pass_variant": (ctypes.c_int, [_vp, _i32, _i32, _i32]),
    "fftm_mi_ffts_get_pass_variant": (ctypes.c_int, [_vp, _i32, _i32, _i32p]),
    "fftm_mi_ffts_destroy": (ctypes.c_int, [_vp]),
    "fftm_mi_ffts_r2c_adjoint": (ctypes.c_int, [_vp, _vp, _vp, _i32]),
    "fftm_mi_ffts_c2r_adjoint": (ctypes.c_int, [_vp, _vp, _vp, _i32]),
    "fftm_mi_ffts_spectral_multiply": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i32, _i32]),
    "fftm_mi_ffts_shell_gradient": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32]),
    "fftm_mi_ffts_shell_gradient_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i32]),
    "fftm_mi_ffts_spectral_apply": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32]),
    "fftm_mi_ffts_filter": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32]),
    "fftm_mi_ffts_autotune": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, ctypes.c_char_p]),
    "fftm_mi_ffts_load_tuning": (ctypes.c_int, [_vp, ctypes.c_char_p, _i32p]),
    "fftm_mi_hardware_signature": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "fftm_mi_fnv1a64": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p]),
    "fftm_mi_fill_random_real": (ctypes.c_int, [_i32, _i32, _i64p, _i64p, _i64p, ctypes.c_uint64, _vp, _vp, _i32]),
    "fftm_mi_last_error": (ctypes.c_char_p, []),
    "fftm_mi_version": (ctypes.c_int, []),
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m fftm_amd._build` "
                              "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().fftm_mi_last_error().decode(errors="replace")


def check(status: int, where: str) -> None:
    if status != SUCCESS:
        raise FftmMiError(status, where, last_error())


def i64_array(values):
    return (ctypes.c_int64 * len(values))(*[int(v) for v in values])


def i32_array(values):
    return (ctypes.c_int32 * len(values))(*[int(v) for v in values])
