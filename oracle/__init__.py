"""CPU parity oracle for the ``ffts`` R2C/C2R path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
package ``fftm_amd`` never imports it.

Two layers:

* ``liboracle.so`` (``fft_oracle.c``): FP64 arithmetic -- splitmix64 inputs,
  one hipFFT advanced-layout plan (``plan_exec``), dense ``rfftn``/``irfftn``
  and a brute-force long-double DFT.
* this module: the ``ffts`` stage pipelines (which plan runs on which stage
  array, and the index permutations between stages), restated with numpy
  from ``source/ffts.hpp`` and ``source/detail/direct_transpose_4d.h``.

Arrays here are *memory-order* numpy arrays (C order, slowest axis first).  An
SCFD arranger ``custom_index_fast_arranger<a,b,c,d>`` lists logical indices
from fastest to slowest, so its memory order is the reversed list.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

R2C, C2R, C2CF, C2CB = 1, 2, 3, 4   # source/fft_direction.h:9-15 (same values)

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH) or (
            os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "fft_oracle.c"))):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True)
    return _LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.fo_splitmix64.restype = ctypes.c_uint64
        L.fo_splitmix64.argtypes = [ctypes.c_uint64]
        L.fo_unit_random_f64.restype = ctypes.c_double
        L.fo_unit_random_f64.argtypes = [ctypes.c_uint64]
        L.fo_unit_random_f32.restype = ctypes.c_float
        L.fo_unit_random_f32.argtypes = [ctypes.c_uint64]
        L.fo_fill_random_real_3d.restype = None
        L.fo_fill_random_real_3d.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_int64] * 6 + [ctypes.c_uint64]
        L.fo_fill_random_real_4d.restype = None
        L.fo_fill_random_real_4d.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_int64] * 8 + [ctypes.c_uint64]
        L.fo_fftw_input_value.restype = ctypes.c_double
        L.fo_fftw_input_value.argtypes = [ctypes.c_uint64]
        L.fo_fft1d_c2c.restype = None
        L.fo_fft1d_c2c.argtypes = [ctypes.c_int64, _f64p, _f64p, ctypes.c_int]
        L.fo_plan_exec.restype = ctypes.c_int
        L.fo_plan_exec.argtypes = [ctypes.c_int, ctypes.c_int, _i64p, _i64p, ctypes.c_int64, ctypes.c_int64,
                                   _i64p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _f64p, _f64p]
        for name in ("fo_rfftn_any", "fo_irfftn_any"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, _i64p, _f64p, _f64p]
        L.fo_brute_r2c.restype = None
        L.fo_brute_r2c.argtypes = [ctypes.c_int, _i64p, _f64p, _f64p]
        L.fo_max_threads.restype = ctypes.c_int
        L.fo_set_threads.restype = None
        L.fo_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _i64(seq):
    arr = (ctypes.c_int64 * len(seq))(*[int(v) for v in seq])
    return arr


def _dptr(a: np.ndarray):
    assert a.dtype in (np.float64, np.complex128) and a.flags.c_contiguous
    return a.ctypes.data_as(_f64p)


# --------------------------------------------------------------------------
# Inputs (fft_benchmark_common.h:201-257; fftw_mpi_benchmark.cpp:215-219)
# --------------------------------------------------------------------------

def splitmix64(x: int) -> int:
    return int(lib().fo_splitmix64(ctypes.c_uint64(x & 0xFFFFFFFFFFFFFFFF)))


def unit_random(state: int, dtype=np.float64):
    if np.dtype(dtype) == np.float32:
        return np.float32(lib().fo_unit_random_f32(state & 0xFFFFFFFFFFFFFFFF))
    return float(lib().fo_unit_random_f64(state & 0xFFFFFFFFFFFFFFFF))


def fill_random_real_3d(nx, ny, nz, seed, dtype=np.float64):
    """Logical (x,y,z) field of fill_random_real_3d_functor, returned in memory
    order [x][y][z] (arranger 210: z fastest) -- the ffts 3D real layout."""
    dt = np.dtype(dtype)
    out = np.empty((nx, ny, nz), dtype=dt)
    lib().fo_fill_random_real_3d(out.ctypes.data, 0 if dt == np.float32 else 1,
                                 nx, ny, nz, ny * nz, nz, 1, seed & 0xFFFFFFFFFFFFFFFF)
    return out


def fill_random_real_4d_logical(nx, ny, nz, nw, seed, dtype=np.float64):
    """Logical array u[x,y,z,w] of fill_random_real_4d_functor (any arrangement
    is a transpose of this)."""
    dt = np.dtype(dtype)
    out = np.empty((nx, ny, nz, nw), dtype=dt)
    lib().fo_fill_random_real_4d(out.ctypes.data, 0 if dt == np.float32 else 1,
                                 nx, ny, nz, nw, ny * nz * nw, nz * nw, nw, 1, seed & 0xFFFFFFFFFFFFFFFF)
    return out


def fftw_input_value(global_index: int) -> float:
    return float(lib().fo_fftw_input_value(global_index))


# --------------------------------------------------------------------------
# One plan (hipfftMakePlanMany64 + hipfftExec* semantics)
# --------------------------------------------------------------------------

def _span(rank, ext, emb, stride, dist, batch):
    lin = 0
    for k in range(rank):
        lin = lin * emb[k] + (ext[k] - 1)
    return (batch - 1) * dist + lin * stride + 1


def plan_exec(ptype, n, inembed, istride, idist, onembed, ostride, odist, batch, data_in,
              out=None):
    """Run one advanced-layout plan on FP64 host memory.

    ``data_in``: flat float64 (real side) or complex128 (complex side) array
    addressed with (inembed, istride, idist).  Returns the flat output array
    sized to the output layout's span (or fills ``out``)."""
    rank = len(n)
    n = [int(v) for v in n]
    ext_real = list(n)
    ext_cplx = list(n)
    if ptype in (R2C, C2R):
        ext_cplx[-1] = n[-1] // 2 + 1
    ext_out = ext_real if ptype == C2R else ext_cplx
    if onembed is None:
        onembed_l = ext_out
        ostride, odist = 1, int(np.prod(ext_out))
    else:
        onembed_l = onembed
    span = _span(rank, ext_out, onembed_l, ostride, odist, batch)
    out_dtype = np.float64 if ptype == C2R else np.complex128
    if out is None:
        out = np.zeros(span, dtype=out_dtype)
    in_dtype = np.float64 if ptype == R2C else np.complex128
    src = np.ascontiguousarray(np.asarray(data_in).reshape(-1).astype(in_dtype, copy=False))
    rc = lib().fo_plan_exec(ptype, rank, _i64(n), None if inembed is None else _i64(inembed),
                            istride, idist, None if onembed is None else _i64(onembed),
                            ostride, odist, batch, _dptr(src), _dptr(out))
    if rc != 0:
        raise ValueError(f"fo_plan_exec failed rc={rc}")
    return out


def rfftn(x: np.ndarray) -> np.ndarray:
    """Dense row-major R2C of rank 1..4 (last axis halved), unnormalised, FP64."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    shape = x.shape
    oshape = shape[:-1] + (shape[-1] // 2 + 1,)
    out = np.empty(oshape, dtype=np.complex128)
    rc = lib().fo_rfftn_any(len(shape), _i64(shape), _dptr(x), _dptr(out))
    if rc != 0:
        raise ValueError(rc)
    return out


def irfftn(X: np.ndarray, shape) -> np.ndarray:
    """Dense row-major C2R (unnormalised: irfftn * N), rank 1..4."""
    X = np.ascontiguousarray(X, dtype=np.complex128)
    shape = tuple(int(v) for v in shape)
    assert X.shape == shape[:-1] + (shape[-1] // 2 + 1,)
    out = np.empty(shape, dtype=np.float64)
    rc = lib().fo_irfftn_any(len(shape), _i64(shape), _dptr(X), _dptr(out))
    if rc != 0:
        raise ValueError(rc)
    return out


def fft1d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().fo_fft1d_c2c(x.size, _dptr(x), _dptr(out), sign)
    return out


def brute_r2c(x: np.ndarray) -> np.ndarray:
    """Definition-by-sum long-double DFT (tiny shapes only)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    shape = x.shape
    out = np.empty(shape[:-1] + (shape[-1] // 2 + 1,), dtype=np.complex128)
    lib().fo_brute_r2c(len(shape), _i64(shape), _dptr(x), _dptr(out))
    return out


# --------------------------------------------------------------------------
# Arrangers (source/detail/array_arrangers.h:29-57, source/ffts.hpp:104-155)
# --------------------------------------------------------------------------

def to_memory(logical: np.ndarray, arranger) -> np.ndarray:
    """Logical array -> C-order memory array for custom_index_fast_arranger<arranger...>."""
    return np.ascontiguousarray(np.transpose(logical, tuple(reversed(arranger))))


def from_memory(mem: np.ndarray, arranger) -> np.ndarray:
    order = tuple(reversed(arranger))
    inv = np.argsort(order)
    return np.transpose(mem, tuple(inv))


def direct_transpose(logical: np.ndarray, dst_axes) -> np.ndarray:
    """direct_transpose_4d_functor (direct_transpose_4d.h:17-32):
    out(c[A0], c[A1], c[A2], c[A3]) = in(c)."""
    return np.transpose(logical, tuple(dst_axes))


# --------------------------------------------------------------------------
# ffts 3D (ffts.hpp:248-263, 356-368, 677-695)
# --------------------------------------------------------------------------

def ffts3d_forward(real_mem: np.ndarray) -> np.ndarray:
    """real [x][y][z] -> spectrum [x][y][kz] via plan "forward_3d" (ffts.hpp:680-686)."""
    nx, ny, nz = real_mem.shape
    nzh = nz // 2 + 1
    out = plan_exec(R2C, [nx, ny, nz], [nx, ny, nz], 1, nx * ny * nz, [nx, ny, nzh], 1,
                    nx * ny * nzh, 1, real_mem.astype(np.float64).reshape(-1))
    return out.reshape(nx, ny, nzh)


def ffts3d_backward(hat_mem: np.ndarray, nz: int) -> np.ndarray:
    """spectrum [x][y][kz] -> real [x][y][z] via plan "inverse_3d" (ffts.hpp:688-694), x N."""
    nx, ny, nzh = hat_mem.shape
    out = plan_exec(C2R, [nx, ny, nz], [nx, ny, nzh], 1, nx * ny * nzh, [nx, ny, nz], 1,
                    nx * ny * nz, 1, hat_mem.astype(np.complex128).reshape(-1))
    return out.reshape(nx, ny, nz)


# --------------------------------------------------------------------------
# ffts 4D (ffts.hpp:123-155, 697-776, 809-867)
# --------------------------------------------------------------------------

A0123 = (0, 1, 2, 3)
A0213 = (0, 2, 1, 3)
A3210 = (3, 2, 1, 0)
A1032 = (1, 0, 3, 2)
A2103 = (2, 1, 0, 3)


def ffts4d_real_layout(strategy: str):
    return A0123 if strategy == "pencil" else A3210


def _exec_mem(mem: np.ndarray, *plan_args, ptype):
    flat = plan_exec(ptype, *plan_args, mem.reshape(-1))
    return flat


def ffts4d_forward(real_mem: np.ndarray, strategy: str, nx: int, ny: int, nz: int, nw: int) -> np.ndarray:
    """Reference ``ffts::forward`` 4D, stage by stage, on memory arrays.

    pencil (ffts.hpp:809-825): real [w][z][y][x] -> spectrum [y][z][wh][x]
    slab   (ffts.hpp:827-837): real [x][y][z][w] -> spectrum [x][y][z][wh]
    """
    nwh = nw // 2 + 1
    r = real_mem.astype(np.float64)
    if strategy == "pencil":
        w_stride = nx * ny * nz
        z_stride = nx * ny * nwh
        y_stride = nx * nz * nwh
        # forward_w (ffts.hpp:709-711): stage0 xyzw, arranger 0123 -> memory [wh][z][y][x]
        st0 = plan_exec(R2C, [nw], [1], w_stride, 1, [1], w_stride, 1, nx * ny * nz, r.reshape(-1))
        st0 = st0.reshape(nwh, nz, ny, nx)
        # xyzw_to_xywz <0,1,3,2>: stage1 logical (x,y,w,z), arranger 0123
        lg = direct_transpose(from_memory(st0, A0123), (0, 1, 3, 2))
        st1 = to_memory(lg, A0123)                     # [z][wh][y][x]
        # forward_z (ffts.hpp:717-719), in place
        st1 = plan_exec(C2CF, [nz], [1], z_stride, 1, [1], z_stride, 1, nx * ny * nwh,
                        st1.reshape(-1)).reshape(st1.shape)
        # xywz_to_xzwy <0,3,2,1>: stage2 logical (x,z,w,y), arranger 0213
        lg = direct_transpose(from_memory(st1, A0123), (0, 3, 2, 1))
        st2 = to_memory(lg, A0213)                     # [y][z][wh][x]
        # forward_y (ffts.hpp:725-727), in place
        st2 = plan_exec(C2CF, [ny], [1], y_stride, 1, [1], y_stride, 1, nx * nz * nwh,
                        st2.reshape(-1)).reshape(st2.shape)
        # xzwy_to_yzwx <3,1,2,0>: out logical (y,z,w,x), arranger 3210
        lg = direct_transpose(from_memory(st2, A0213), (3, 1, 2, 0))
        out = to_memory(lg, A3210)                     # [y][z][wh][x]
        # forward_x (ffts.hpp:733-736), contiguous
        out = plan_exec(C2CF, [nx], [1], 1, nx, [1], 1, nx, ny * nz * nwh,
                        out.reshape(-1)).reshape(out.shape)
        return out
    if strategy == "slab":
        xy_stride = nz * nwh
        # forward_zw (ffts.hpp:751-756): stage0 xyzw arranger 3210 -> [x][y][z][wh]
        st0 = plan_exec(R2C, [nz, nw], [nz, nw], 1, nz * nw, [nz, nwh], 1, nz * nwh, nx * ny,
                        r.reshape(-1)).reshape(nx, ny, nz, nwh)
        # xyzw_to_zwxy <2,3,0,1>: stage1 logical (z,w,x,y), arranger 1032
        lg = direct_transpose(from_memory(st0, A3210), (2, 3, 0, 1))
        st1 = to_memory(lg, A1032)                     # [x][y][z][wh]
        # forward_xy (ffts.hpp:765-769), in place
        st1 = plan_exec(C2CF, [nx, ny], [nx, ny], xy_stride, 1, [nx, ny], xy_stride, 1, nz * nwh,
                        st1.reshape(-1)).reshape(st1.shape)
        # zwxy_to_yzwx <3,0,1,2>: out logical (y,z,w,x), arranger 2103
        lg = direct_transpose(from_memory(st1, A1032), (3, 0, 1, 2))
        return to_memory(lg, A2103)                    # [x][y][z][wh]
    raise ValueError(strategy)


def ffts4d_backward(hat_mem: np.ndarray, strategy: str, nx: int, ny: int, nz: int, nw: int) -> np.ndarray:
    """Reference ``ffts::backward`` 4D (ffts.hpp:839-867), unnormalised (x N)."""
    nwh = nw // 2 + 1
    h = hat_mem.astype(np.complex128)
    if strategy == "pencil":
        w_stride = nx * ny * nz
        z_stride = nx * ny * nwh
        y_stride = nx * nz * nwh
        h = plan_exec(C2CB, [nx], [1], 1, nx, [1], 1, nx, ny * nz * nwh, h.reshape(-1)).reshape(h.shape)
        # yzwx_to_xzwy <3,1,2,0>
        lg = direct_transpose(from_memory(h, A3210), (3, 1, 2, 0))
        st2 = to_memory(lg, A0213)
        st2 = plan_exec(C2CB, [ny], [1], y_stride, 1, [1], y_stride, 1, nx * nz * nwh,
                        st2.reshape(-1)).reshape(st2.shape)
        # xzwy_to_xywz <0,3,2,1>
        lg = direct_transpose(from_memory(st2, A0213), (0, 3, 2, 1))
        st1 = to_memory(lg, A0123)
        st1 = plan_exec(C2CB, [nz], [1], z_stride, 1, [1], z_stride, 1, nx * ny * nwh,
                        st1.reshape(-1)).reshape(st1.shape)
        # xywz_to_xyzw <0,1,3,2>
        lg = direct_transpose(from_memory(st1, A0123), (0, 1, 3, 2))
        st0 = to_memory(lg, A0123)                     # [wh][z][y][x]
        out = plan_exec(C2R, [nw], [1], w_stride, 1, [1], w_stride, 1, nx * ny * nz, st0.reshape(-1))
        return out.reshape(nw, nz, ny, nx)
    if strategy == "slab":
        xy_stride = nz * nwh
        # yzwx_to_zwxy <1,2,3,0>
        lg = direct_transpose(from_memory(h, A2103), (1, 2, 3, 0))
        st1 = to_memory(lg, A1032)
        st1 = plan_exec(C2CB, [nx, ny], [nx, ny], xy_stride, 1, [nx, ny], xy_stride, 1, nz * nwh,
                        st1.reshape(-1)).reshape(st1.shape)
        # zwxy_to_xyzw <2,3,0,1>
        lg = direct_transpose(from_memory(st1, A1032), (2, 3, 0, 1))
        st0 = to_memory(lg, A3210)
        out = plan_exec(C2R, [nz, nw], [nz, nwh], 1, nz * nwh, [nz, nw], 1, nz * nw, nx * ny,
                        st0.reshape(-1))
        return out.reshape(nx, ny, nz, nw)
    raise ValueError(strategy)


# One-rank fftm "native_xzwy" spectrum (fftm.hpp:2941-2961, 4018-4079, 4519-4560;
# mpi_transpose_4d_same_xw.h:437-497, 986-1060): slab zw-R2C, reorder
# [x][y][z][wh] -> [y][z][wh][x], then the 2D (y,x) C2C plan
# n={ny,nx}, inembed={ny, nx*nz*wh}, istride 1, idist nx, batch nz*wh.

def native_xzwy_forward(real_xyzw_mem: np.ndarray, nx, ny, nz, nw) -> np.ndarray:
    nwh = nw // 2 + 1
    st0 = plan_exec(R2C, [nz, nw], [nz, nw], 1, nz * nw, [nz, nwh], 1, nz * nwh, nx * ny,
                    real_xyzw_mem.astype(np.float64).reshape(-1)).reshape(nx, ny, nz, nwh)
    packed = np.ascontiguousarray(np.transpose(st0, (1, 2, 3, 0)))  # [y][z][wh][x]
    out = plan_exec(C2CF, [ny, nx], [ny, nx * nz * nwh], 1, nx, [ny, nx * nz * nwh], 1, nx,
                    nz * nwh, packed.reshape(-1))
    return out.reshape(ny, nz, nwh, nx)


def native_xzwy_backward(hat_yzwx_mem: np.ndarray, nx, ny, nz, nw) -> np.ndarray:
    nwh = nw // 2 + 1
    h = plan_exec(C2CB, [ny, nx], [ny, nx * nz * nwh], 1, nx, [ny, nx * nz * nwh], 1, nx,
                  nz * nwh, hat_yzwx_mem.astype(np.complex128).reshape(-1)).reshape(ny, nz, nwh, nx)
    st0 = np.ascontiguousarray(np.transpose(h, (3, 0, 1, 2)))      # [x][y][z][wh]
    out = plan_exec(C2R, [nz, nw], [nz, nwh], 1, nz * nwh, [nz, nw], 1, nz * nw, nx * ny,
                    st0.reshape(-1))
    return out.reshape(nx, ny, nz, nw)


# --------------------------------------------------------------------------
# Error metrics reported for every result
# --------------------------------------------------------------------------

def rel_errors(actual: np.ndarray, expected: np.ndarray):
    """(rms_rel, max_rel): ||a-e||_2/||e||_2 and max|a-e| / max|e|."""
    a = np.asarray(actual).astype(np.complex128 if np.iscomplexobj(actual) or np.iscomplexobj(expected)
                                  else np.float64)
    e = np.asarray(expected).astype(a.dtype)
    d = a - e
    ne = float(np.sqrt(np.sum(np.abs(e) ** 2)))
    nd = float(np.sqrt(np.sum(np.abs(d) ** 2)))
    me = float(np.max(np.abs(e))) if e.size else 0.0
    md = float(np.max(np.abs(d))) if d.size else 0.0
    rms = nd / ne if ne > 0 else nd
    mx = md / me if me > 0 else md
    return rms, mx


# --------------------------------------------------------------------------
# Spectral operators (restated for the adjoint / filter tests)
# --------------------------------------------------------------------------

def wavenumbers(n: int, halved: bool) -> np.ndarray:
    """k = i <= n/2 ? i : i - n on full axes (Nyquist positive), k = i on the halved axis
    (versioned_ffts_3d_test_driver.h:140-142)."""
    if halved:
        return np.arange(n // 2 + 1)
    i = np.arange(n)
    return np.where(i <= n // 2, i, i - n)


def shell_index_logical(dims, nshells: int) -> np.ndarray:
    """shell(k) = min(S-1, floor(S |k| / kmax)), kmax = sqrt(sum (n_a/2)^2), on the logical
    half spectrum (last axis halved); returned as an int array of the logical half shape.
    Evaluated exactly: the largest s with s^2 kmax^2 <= S^2 |k|^2 (integers throughout),
    so bins where S |k| / kmax is an integer land in the upper shell."""
    d = len(dims)
    ks = [wavenumbers(n, a == d - 1).astype(np.int64) for a, n in enumerate(dims)]
    grids = np.meshgrid(*ks, indexing="ij")
    k2 = np.zeros(grids[0].shape, dtype=np.int64)
    for g in grids:
        k2 += g * g
    kmax2 = max(1, sum((n // 2) ** 2 for n in dims))
    rhs = np.int64(nshells) * np.int64(nshells) * k2
    s = np.floor(np.sqrt(rhs.astype(np.float64) / kmax2)).astype(np.int64)
    for _ in range(2):
        s = np.where((s > 0) & (s * s * kmax2 > rhs), s - 1, s)
    for _ in range(2):
        s = np.where((s + 1) * (s + 1) * kmax2 <= rhs, s + 1, s)
    return np.minimum(s, nshells - 1)


def realaxis_weights_logical(dims) -> np.ndarray:
    """w_k = 1 on the DC and (even n) Nyquist planes of the real axis, 2 elsewhere."""
    n = dims[-1]
    w = np.full(n // 2 + 1, 2.0)
    w[0] = 1.0
    if n % 2 == 0:
        w[-1] = 1.0
    return np.broadcast_to(w, tuple(dims[:-1]) + (n // 2 + 1,))


def spectral_operator_logical(dims, kind: str, scale: float = 1.0, sigma: float = 0.0, coeffs=None,
                              axis: int = 0, band=(0.0, 0.0)) -> np.ndarray:
    """H on the logical half spectrum (last axis halved), FP64:
    poisson  -scale/|k|^2 with H(0) = 0 (solve_poisson_3d_functor,
             versioned_ffts_3d_test_driver.h:128-154),
    gaussian scale*exp(-|k|^2/(2 sigma^2)),
    shell    scale*coeffs[shell(k)],
    derivative        i*scale*k_axis (poisson_3d_derivative_spectra_functor,
                      versioned_ffts_3d_test_driver.h:156-184),
    poisson_gradient  -i*scale*k_axis/|k|^2, 0 at k = 0,
    bandpass          scale where band[0] <= |k| <= band[1].
    Pinned in tests by closed forms: the periodic Poisson problem and its gradient
    (versioned_ffts_3d_test_driver.h:187-209), derivatives of sine products and
    Gaussian- or band-filtered plane waves."""
    d = len(dims)
    ks = [wavenumbers(n, a == d - 1) for a, n in enumerate(dims)]
    grids = np.meshgrid(*ks, indexing="ij")
    k2 = np.zeros(grids[0].shape, dtype=np.float64)
    for g in grids:
        k2 += g.astype(np.float64) ** 2
    if kind == "poisson":
        with np.errstate(divide="ignore"):
            h = np.where(k2 > 0, -scale / np.where(k2 > 0, k2, 1.0), 0.0)
        return h
    if kind == "gaussian":
        return scale * np.exp(-k2 / (2.0 * sigma * sigma))
    if kind == "shell":
        c = np.asarray(coeffs, dtype=np.float64)
        return scale * c[shell_index_logical(dims, len(c))]
    if kind == "derivative":
        return 1j * scale * grids[axis].astype(np.float64)
    if kind == "poisson_gradient":
        ka = grids[axis].astype(np.float64)
        return np.where(k2 > 0, -1j * scale * ka / np.where(k2 > 0, k2, 1.0), 0.0)
    if kind == "bandpass":
        return np.where((k2 >= band[0] ** 2) & (k2 <= band[1] ** 2), scale, 0.0)
    raise ValueError(kind)


def filter_logical(x: np.ndarray, h: np.ndarray) -> np.ndarray:
    """backward(H * forward(x)) on a logical real array, unnormalised like the ffts
    backward (x N): N * irfftn(H * rfftn(x)) over all axes, last axis halved, in FP64."""
    axes = tuple(range(x.ndim))
    X = np.fft.rfftn(x.astype(np.float64), axes=axes)
    return np.fft.irfftn(X * h, s=x.shape, axes=axes) * float(np.prod(x.shape))
