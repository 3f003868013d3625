"""Shared helpers for the parity tests (inputs, tolerances, error reporting)."""

from __future__ import annotations

import math

import numpy as np

import oracle

# Tolerances, stated once.
#  * round trip: RMS(C2R(R2C(u))/N - u) <= 10 eps(T), the reference benchmark's default epsilon
#    (source/tests/detail/fft_benchmark_common.h:28-32).
#  * forward/backward vs the FP64 oracle, relative L2 (||a - e|| / ||e||) and relative max
#    (max|a - e| / max|e|):
#      FP64: <= 1e-12 (reference cross-frontend threshold 1e-11, test_fftm_3D_compare.cu:53,
#            tightened);  FP32: <= 2e-6 (rel-L2) and 1e-5 (rel-max).
EPS = {np.float32: float(np.finfo(np.float32).eps), np.float64: float(np.finfo(np.float64).eps)}
TOL_REL_L2 = {np.float32: 2e-6, np.float64: 1e-12}
TOL_REL_MAX = {np.float32: 1e-5, np.float64: 1e-12}


def roundtrip_tol(dtype) -> float:
    return 10.0 * EPS[np.dtype(dtype).type]


def torch_dtypes(np_dtype):
    import torch
    if np.dtype(np_dtype) == np.float32:
        return torch.float32, torch.complex64
    return torch.float64, torch.complex128


def to_gpu(a: np.ndarray):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def from_gpu(t) -> np.ndarray:
    return t.detach().cpu().numpy()


def check_close(actual, expected, dtype, what, report=None):
    """Assert the relative errors against the FP64 oracle; returns (rms_rel, max_rel)."""
    rms, mx = oracle.rel_errors(actual, expected)
    t = np.dtype(dtype).type
    if report is not None:
        report.append((what, rms, mx))
    print(f"[parity] {what}: rel_l2={rms:.3e} rel_max={mx:.3e}")
    assert rms <= TOL_REL_L2[t], f"{what}: rel-L2 {rms:.3e} > {TOL_REL_L2[t]:.1e}"
    assert mx <= TOL_REL_MAX[t], f"{what}: rel-max {mx:.3e} > {TOL_REL_MAX[t]:.1e}"
    return rms, mx


def rms_diff(a, b) -> float:
    d = np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)
    return float(math.sqrt(np.mean(d * d)))
