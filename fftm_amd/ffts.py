"""``ffts``: the single-GPU 2D/3D/4D R2C/C2R frontend of fftm, on the MI355X backend.

Mirrors ``fftm::ffts<BaseFFT, Backend, Strategy4D>`` (source/ffts.hpp:159-1167):
``init(nx, ny[, nz[, nw]])`` once, then ``forward(real, spectrum)`` /
``backward(spectrum, real)``, unnormalised (C2R(R2C(u)) = N u).

Memory layouts (C-order shapes of the torch tensors; ffts.hpp:104-155):

* 2D   real [x][y]        spectrum [x][y/2+1]
* 3D   real [x][y][z]     spectrum [x][y][z/2+1]
* 4D   pencil-pencil: real [w][z][y][x]   spectrum [y][z][w/2+1][x]
       slab-slab:     real [x][y][z][w]   spectrum [x][y][z][w/2+1]
       slab + native_xzwy (the one-rank fftm spectral layout, fftm.hpp:2941-2961):
                      real [x][y][z][w]   spectrum [y][z][w/2+1][x]

Two executions of the same mathematics:

* ``execution="fused"`` (default): the whole transform is one C-ABI call that runs one
  HBM pass per transformed axis; the stage transposes of the reference are folded
  into the passes' address maps.
* ``execution="staged"``: the reference's own stage sequence -- the named plans of
  ``add_{2,3,4}d_plans_`` (ffts.hpp:659-776) executed through ``fft_wrap_many`` with
  the stage transposes of ``forward_4d_`` / ``backward_4d_`` (ffts.hpp:809-867) run by
  the LDS-tiled transpose kernel.  Kept for drop-in parity checks and to measure the
  cost the fusion removes.

Like the reference, the 4D staged backward uses its input as scratch
(ffts.hpp:377); the fused backward keeps its input unless
``ffts_init_options.destructive_backward`` is set.
"""

from __future__ import annotations

import ctypes
import enum
import sys
import time
from collections import OrderedDict
from dataclasses import dataclass
from typing import Optional, Sequence

from . import _lib
from .wrap import LogicError, device_ptr, direction, fft_wrap_many, stream_handle, transpose_4d


class transpose_backend(enum.Enum):
    direct = "direct"
    memcpy = "memcpy"


class transform_strategy_4d(enum.Enum):
    pencil_pencil = "pencil_pencil"
    slab_slab = "slab_slab"
    slab_native_xzwy = "slab_native_xzwy"


@dataclass(frozen=True)
class strategy_4d_pencil_pencil:
    backend: transpose_backend = transpose_backend.direct
    family = transform_strategy_4d.pencil_pencil

    def name(self):
        return "pencil-pencil-direct" if self.backend == transpose_backend.direct else "pencil-pencil-memcpy"


@dataclass(frozen=True)
class strategy_4d_slab_slab:
    backend: transpose_backend = transpose_backend.direct
    family = transform_strategy_4d.slab_slab

    def name(self):
        return "slab-slab-direct" if self.backend == transpose_backend.direct else "slab-slab-memcpy"


@dataclass(frozen=True)
class strategy_4d_slab_native_xzwy:
    """Slab real layout in, fftm's one-rank native xzwy spectrum out (fused only)."""
    family = transform_strategy_4d.slab_native_xzwy

    def name(self):
        return "slab-native-xzwy"


@dataclass
class ffts_init_options:
    """ffts_init_options (ffts.hpp:64-72) plus backend switches."""
    profiling_key: str = "ffts_prof"
    memory_profiling_key: str = "ffts_mem"
    print_profile_summary_on_destroy: bool = True
    print_profile_totals_on_destroy: bool = True
    print_memory_profile_on_destroy: bool = True
    print_memory_totals_on_destroy: bool = True
    # backend switches (not in the reference)
    execution: str = "fused"          # "fused" | "staged"
    sync: bool = True                 # device-synchronise after each call (reference safe-call behaviour)
    destructive_backward: bool = False
    # "fp32": spectra between passes in the transform precision; "fp16": FP32 transforms
    # whose intermediate half spectra are stored in FP16 with per-line power-of-two
    # exponents (3D, fused execution, lengths 128..2048); "fp16x2": the same with every
    # stored bin a hi + lo FP16 pair (FP32-level precision, FP32's bytes; lengths 128..1024)
    spectrum_storage: str = "fp32"


class _profiler:
    """Host-timer scope profiler (the role of scfd::utils::profiler, common/profiler.h:21-268)."""

    def __init__(self):
        self.enabled = False
        self.key = ""
        self.stats = OrderedDict()

    def scope(self, name):
        prof = self

        class _S:
            def __enter__(s):
                s.t0 = time.perf_counter()

            def __exit__(s, *a):
                if prof.enabled:
                    dt = time.perf_counter() - s.t0
                    cnt, tot = prof.stats.get(name, (0, 0.0))
                    prof.stats[name] = (cnt + 1, tot + dt)
        return _S()

    def log_print(self, out):
        for name, (cnt, tot) in self.stats.items():
            out.write(f"{self.key}: {name}: calls={cnt} total={tot * 1e3:.3f} ms avg={tot * 1e3 / max(cnt, 1):.3f} ms\n")

    def log_print_totals(self, out):
        tot = sum(t for _, t in self.stats.values())
        out.write(f"{self.key}: total={tot * 1e3:.3f} ms\n")


class _memory_profiler:
    """Bytes profiler (common/memory_profiler.h:14-110), device bucket only."""

    def __init__(self):
        self.enabled = False
        self.key = ""
        self.entries = OrderedDict()

    def set_bytes(self, name, nbytes):
        cur, peak = self.entries.get(name, (0, 0))
        self.entries[name] = (int(nbytes), max(peak, int(nbytes)))

    def total(self):
        return sum(c for c, _ in self.entries.values()), sum(p for _, p in self.entries.values())


class ffts:
    """Single-GPU R2C/C2R transforms (ffts.hpp:159-1167)."""

    def __init__(self, real="float64", strategy_4d=None):
        import torch
        self.real = {"float32": torch.float32, "float64": torch.float64}.get(real, real)
        if self.real not in (torch.float32, torch.float64):
            raise TypeError("ffts: real type must be float32 or float64")
        self.complex = torch.complex64 if self.real == torch.float32 else torch.complex128
        self.strategy_4d = strategy_4d if strategy_4d is not None else strategy_4d_pencil_pencil()
        self._init_done = False
        self._dim = 0
        self.nx = self.ny = self.nz = self.nw = 0
        self._profiler = _profiler()
        self._mem = _memory_profiler()
        self._options = ffts_init_options()
        self._handle = None
        self._work = None
        self._base = None
        self._stage = {}

    def strategy_name(self) -> str:
        return self.strategy_4d.name()

    # ------------------------------------------------------------------ init
    def init(self, nx, ny, nz=None, nw=None, options: Optional[ffts_init_options] = None):
        """init(nx, ny[, nz[, nw]]) (ffts.hpp:232-285)."""
        if self._init_done:
            raise LogicError("ffts::init can only be called once per instance.")
        self._configure(options or ffts_init_options())
        dims = [d for d in (nx, ny, nz, nw) if d is not None]
        dim = len(dims)
        for d in dims:
            if int(d) < 1:
                raise ValueError("ffts::init: sizes must be positive")
        if self._options.spectrum_storage in ("fp16", "fp16x2") and (dim != 3 or self._options.execution != "fused"):
            raise ValueError("ffts: FP16 spectrum storage is implemented for fused 3D transforms")
        with self._profiler.scope(f"ffts::init<{dim}>"):
            self.nx, self.ny = int(nx), int(ny)
            self.nz = int(nz) if nz is not None else 0
            self.nw = int(nw) if nw is not None else 0
            self._dim = dim
            if dim == 2:
                self._add_2d_plans()
            elif dim == 3:
                if self._options.execution == "fused":
                    self._create_fused(_lib.LAYOUT_3D, [self.nx, self.ny, self.nz])
                else:
                    self._add_3d_plans()
            elif dim == 4:
                fam = self.strategy_4d.family
                if self._options.execution == "fused" or fam == transform_strategy_4d.slab_native_xzwy:
                    layout = {transform_strategy_4d.pencil_pencil: _lib.LAYOUT_4D_PENCIL,
                              transform_strategy_4d.slab_slab: _lib.LAYOUT_4D_SLAB,
                              transform_strategy_4d.slab_native_xzwy: _lib.LAYOUT_4D_XZWY}[fam]
                    self._create_fused(layout, [self.nx, self.ny, self.nz, self.nw])
                else:
                    self._init_4d_staged()
            else:
                raise LogicError("ffts::init: dimension must be 2, 3 or 4")
        self._init_done = True
        self._update_memory_profile()

    def init_array(self, grid: Sequence[int], options=None):
        """init<Dim>(std::array) (ffts.hpp:226-230)."""
        self.init(*grid, options=options)

    def _configure(self, options):
        self._options = options
        self._profiler.enabled = bool(options.profiling_key)
        self._profiler.key = options.profiling_key
        self._mem.enabled = bool(options.memory_profiling_key)
        self._mem.key = options.memory_profiling_key

    # fused C-ABI handle ----------------------------------------------------
    def _create_fused(self, layout, n):
        import torch
        L = _lib.lib()
        prec = _lib.F32 if self.real == torch.float32 else _lib.F64
        if self._options.spectrum_storage in ("fp16", "fp16x2"):
            if self.real != torch.float32:
                raise TypeError("ffts: FP16 spectrum storage needs float32 transforms")
            prec = _lib.F32_H16 if self._options.spectrum_storage == "fp16" else _lib.F32_H16X2
        elif self._options.spectrum_storage != "fp32":
            raise ValueError("ffts: spectrum_storage is 'fp32', 'fp16' or 'fp16x2'")
        flags = _lib.FLAG_OVERWRITE_INPUT if self._options.destructive_backward else 0
        h = ctypes.c_void_p()
        _lib.check(L.fftm_mi_ffts_create(layout, _lib.i64_array(n), prec, flags, ctypes.byref(h)),
                   "fftm_mi_ffts_create")
        self._handle = h
        ws = ctypes.c_size_t()
        _lib.check(L.fftm_mi_ffts_get_work_size(h, ctypes.byref(ws)), "fftm_mi_ffts_get_work_size")
        if ws.value:
            self._work = torch.empty(int((ws.value + 255) // 256 * 256), dtype=torch.uint8, device="cuda")
            _lib.check(L.fftm_mi_ffts_set_work_area(h, device_ptr(self._work)), "fftm_mi_ffts_set_work_area")

    def set_stream(self, stream):
        """Bind all work of this instance to a stream (default: the null stream)."""
        if self._handle is not None:
            _lib.check(_lib.lib().fftm_mi_ffts_set_stream(self._handle, stream_handle(stream)),
                       "fftm_mi_ffts_set_stream")
        if self._base is not None:
            for name, _ in self._base.plan_descriptors():
                self._base.set_stream(name, stream)
        self._stream = stream

    def num_passes(self):
        """(forward, backward) HBM passes of the fused transforms."""
        f, b = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(_lib.lib().fftm_mi_ffts_num_passes(self._handle, ctypes.byref(f), ctypes.byref(b)),
                   "fftm_mi_ffts_num_passes")
        return int(f.value), int(b.value)

    # staged (reference plan sequence) ----------------------------------------
    def _new_base(self):
        self._base = fft_wrap_many(self.real)
        self._base.set_memory_profiler(self._mem if self._mem.enabled else None, "ffts/base_fft")
        if not self._options.sync:
            self._base.set_hot_exec_no_sync_enabled(True)

    def _activate_shared_work_area(self):
        import torch
        size = (self._base.activate_work_size() + 255) // 256 * 256
        self._work = torch.empty(max(size, 256), dtype=torch.uint8, device="cuda")
        self._base.set_external_work_area(self._work)

    def _add_2d_plans(self):
        """ffts.hpp:659-675."""
        self._new_base()
        nx, ny = self.nx, self.ny
        nyh = ny // 2 + 1
        self._base.add_plan_2D(direction.R2C, "forward_2d", nx, ny, nx, ny, 1, nx * ny, nx, nyh, 1, nx * nyh, 1)
        self._base.add_plan_2D(direction.C2R, "inverse_2d", nx, ny, nx, nyh, 1, nx * nyh, nx, ny, 1, nx * ny, 1)
        self._activate_shared_work_area()

    def _add_3d_plans(self):
        """ffts.hpp:677-695."""
        self._new_base()
        nx, ny, nz = self.nx, self.ny, self.nz
        nzh = nz // 2 + 1
        self._base.add_plan_3D(direction.R2C, "forward_3d", nx, ny, nz, nx, ny, nz, 1, nx * ny * nz, nx, ny, nzh,
                               1, nx * ny * nzh, 1)
        self._base.add_plan_3D(direction.C2R, "inverse_3d", nx, ny, nz, nx, ny, nzh, 1, nx * ny * nzh, nx, ny, nz,
                               1, nx * ny * nz, 1)
        self._activate_shared_work_area()

    def _init_4d_staged(self):
        """init_4d_storage_ + add_4d_plans_ (ffts.hpp:697-807, 1099-1119)."""
        import torch
        self._new_base()
        nx, ny, nz, nw = self.nx, self.ny, self.nz, self.nw
        nwh = nw // 2 + 1
        fam = self.strategy_4d.family
        if fam == transform_strategy_4d.pencil_pencil:
            # stage0 xyzw [wh][z][y][x] and stage2 xzwy [y][z][wh][x] share one buffer; stage1 xywz [z][wh][y][x]
            shared = torch.empty(nx * ny * nz * nwh, dtype=self.complex, device="cuda")
            self._stage = {
                "stage0": shared.view(nwh, nz, ny, nx),
                "stage2": shared.view(ny, nz, nwh, nx),
                "stage1": torch.empty((nz, nwh, ny, nx), dtype=self.complex, device="cuda"),
            }
            real_batch = nx * ny * nz
            z_batch = nx * ny * nwh
            y_batch = nx * nz * nwh
            x_batch = ny * nz * nwh
            w_stride, z_stride, y_stride = nx * ny * nz, nx * ny * nwh, nx * nz * nwh
            b = self._base
            b.add_plan_1D(direction.R2C, "forward_w", nw, 1, w_stride, 1, 1, w_stride, 1, real_batch)
            b.add_plan_1D(direction.C2R, "inverse_w", nw, 1, w_stride, 1, 1, w_stride, 1, real_batch)
            b.add_plan_1D(direction.C2CF, "forward_z", nz, 1, z_stride, 1, 1, z_stride, 1, z_batch)
            b.add_plan_1D(direction.C2CB, "inverse_z", nz, 1, z_stride, 1, 1, z_stride, 1, z_batch)
            b.add_plan_1D(direction.C2CF, "forward_y", ny, 1, y_stride, 1, 1, y_stride, 1, y_batch)
            b.add_plan_1D(direction.C2CB, "inverse_y", ny, 1, y_stride, 1, 1, y_stride, 1, y_batch)
            b.add_plan_1D(direction.C2CF, "forward_x", nx, 1, 1, nx, 1, 1, nx, x_batch)
            b.add_plan_1D(direction.C2CB, "inverse_x", nx, 1, 1, nx, 1, 1, nx, x_batch)
        else:
            # stage0 xyzw and stage1 zwxy both [x][y][z][wh] (stage1 owned, stage0 shared)
            self._stage = {
                "stage0": torch.empty((nx, ny, nz, nwh), dtype=self.complex, device="cuda"),
                "stage1": torch.empty((nx, ny, nz, nwh), dtype=self.complex, device="cuda"),
            }
            zw_batch, xy_batch, xy_stride = nx * ny, nz * nwh, nz * nwh
            b = self._base
            b.add_plan_2D(direction.R2C, "forward_zw", nz, nw, nz, nw, 1, nz * nw, nz, nwh, 1, nz * nwh, zw_batch)
            b.add_plan_2D(direction.C2R, "inverse_zw", nz, nw, nz, nwh, 1, nz * nwh, nz, nw, 1, nz * nw, zw_batch)
            b.add_plan_2D(direction.C2CF, "forward_xy", nx, ny, nx, ny, xy_stride, 1, nx, ny, xy_stride, 1, xy_batch)
            b.add_plan_2D(direction.C2CB, "inverse_xy", nx, ny, nx, ny, xy_stride, 1, nx, ny, xy_stride, 1, xy_batch)
        self._activate_shared_work_area()

    # ------------------------------------------------------------- arrays
    def real_shape(self, dim=None):
        dim = dim or self._dim
        if dim == 2:
            return (self.nx, self.ny)
        if dim == 3:
            return (self.nx, self.ny, self.nz)
        if self.strategy_4d.family == transform_strategy_4d.pencil_pencil:
            return (self.nw, self.nz, self.ny, self.nx)
        return (self.nx, self.ny, self.nz, self.nw)

    def complex_shape(self, dim=None):
        dim = dim or self._dim
        if dim == 2:
            return (self.nx, self.ny // 2 + 1)
        if dim == 3:
            return (self.nx, self.ny, self.nz // 2 + 1)
        nwh = self.nw // 2 + 1
        if self.strategy_4d.family == transform_strategy_4d.slab_slab:
            return (self.nx, self.ny, self.nz, nwh)
        return (self.ny, self.nz, nwh, self.nx)

    def real_array(self):
        import torch
        return torch.empty(self.real_shape(), dtype=self.real, device="cuda")

    def complex_array(self):
        import torch
        return torch.empty(self.complex_shape(), dtype=self.complex, device="cuda")

    # ------------------------------------------------------------ queries
    def is_initialized(self) -> bool:
        return self._init_done

    def dimension(self) -> int:
        self._ensure_initialized()
        return self._dim

    def _ensure_initialized(self):
        if not self._init_done:
            raise LogicError("ffts: call init() before using the transform.")

    def _ensure_dimension(self, expected):
        self._ensure_initialized()
        if self._dim != expected:
            raise LogicError(f"ffts: initialized for {self._dim}D but used as {expected}D.")

    def _check(self, t, shape, dtype, what):
        if tuple(t.shape) != tuple(shape) or t.dtype != dtype or not t.is_contiguous() or not t.is_cuda:
            raise ValueError(f"ffts: {what} must be a contiguous {dtype} GPU tensor of shape {tuple(shape)}, "
                             f"got {t.dtype} {tuple(t.shape)}")

    # ---------------------------------------------------------- transforms
    def forward(self, inp, out):
        """R2C, unnormalised (ffts.hpp:342-375)."""
        dim = self._dim
        self._ensure_dimension(dim if dim in (2, 3, 4) else -1)
        self._check(inp, self.real_shape(), self.real, "forward input")
        self._check(out, self.complex_shape(), self.complex, "forward output")
        with self._profiler.scope(f"ffts::forward_{dim}d"):
            if self._handle is not None:
                _lib.check(_lib.lib().fftm_mi_ffts_forward(self._handle, device_ptr(inp), device_ptr(out),
                                                           1 if self._options.sync else 0), "fftm_mi_ffts_forward")
            elif dim == 2:
                self._base.exec("forward_2d", inp, out)
            elif dim == 3:
                self._base.exec("forward_3d", inp, out)
            else:
                self._forward_4d_staged(inp, out)

    def backward(self, inp, out):
        """C2R, unnormalised (x N) (ffts.hpp:349-383)."""
        dim = self._dim
        self._ensure_dimension(dim if dim in (2, 3, 4) else -1)
        self._check(inp, self.complex_shape(), self.complex, "backward input")
        self._check(out, self.real_shape(), self.real, "backward output")
        with self._profiler.scope(f"ffts::backward_{dim}d"):
            if self._handle is not None:
                _lib.check(_lib.lib().fftm_mi_ffts_backward(self._handle, device_ptr(inp), device_ptr(out),
                                                            1 if self._options.sync else 0),
                           "fftm_mi_ffts_backward")
            elif dim == 2:
                self._base.exec("inverse_2d", inp, out)
            elif dim == 3:
                self._base.exec("inverse_3d", inp, out)
            else:
                self._backward_4d_staged(inp, out)

    # ------------------------------------------------------------- tuning
    def autotune(self, warmup: int = 3, iterations: int = 10, cache_path: Optional[str] = None):
        """Measure every tile-shape variant of every pass and keep the fastest valid one
        (reference autotune protocol; fused execution only).  Returns the chosen variants
        as {(backward, index): variant}."""
        import torch
        self._need_handle("autotune")
        u = torch.rand(self.real_shape(), dtype=self.real, device="cuda")
        X = self.complex_array()
        _lib.check(_lib.lib().fftm_mi_ffts_autotune(self._handle, device_ptr(u), device_ptr(X), int(warmup),
                                                    int(iterations),
                                                    None if cache_path is None else cache_path.encode()),
                   "fftm_mi_ffts_autotune")
        return self.pass_variants()

    def load_tuning(self, cache_path: str) -> int:
        self._need_handle("load_tuning")
        n = ctypes.c_int32()
        _lib.check(_lib.lib().fftm_mi_ffts_load_tuning(self._handle, cache_path.encode(), ctypes.byref(n)),
                   "fftm_mi_ffts_load_tuning")
        return int(n.value)

    def pass_variants(self):
        self._need_handle("pass_variants")
        nf, nb = self.num_passes()
        out = {}
        for bw, n in ((0, nf), (1, nb)):
            for i in range(n):
                v = ctypes.c_int32()
                _lib.check(_lib.lib().fftm_mi_ffts_get_pass_variant(self._handle, bw, i, ctypes.byref(v)),
                           "fftm_mi_ffts_get_pass_variant")
                out[(bw, i)] = int(v.value)
        return out

    def set_pass_variant(self, backward: int, index: int, variant: int):
        self._need_handle("set_pass_variant")
        _lib.check(_lib.lib().fftm_mi_ffts_set_pass_variant(self._handle, int(backward), int(index), int(variant)),
                   "fftm_mi_ffts_set_pass_variant")

    # -------------------------------------------- adjoints / spectral operators
    # Spectrum inner product <X, Y> = Re sum conj(X) Y.  With w_k = 1 on the DC and
    # Nyquist planes of the real axis and 2 elsewhere: R2C^T(G) = C2R(G / w) and
    # C2R^T(v) = w R2C(v) (fused into the real-axis pass).  Fused execution only.
    def _need_handle(self, what):
        if self._handle is None:
            raise LogicError(f"ffts::{what}: available with execution='fused' for 3D/4D")

    def r2c_adjoint(self, spectrum, real_out):
        """R2C^T: spectrum (complex layout) -> real layout."""
        self._need_handle("r2c_adjoint")
        self._check(spectrum, self.complex_shape(), self.complex, "r2c_adjoint input")
        self._check(real_out, self.real_shape(), self.real, "r2c_adjoint output")
        _lib.check(_lib.lib().fftm_mi_ffts_r2c_adjoint(self._handle, device_ptr(spectrum), device_ptr(real_out),
                                                       1 if self._options.sync else 0), "fftm_mi_ffts_r2c_adjoint")

    def c2r_adjoint(self, real_in, spectrum_out):
        """C2R^T: real layout -> spectrum (complex layout)."""
        self._need_handle("c2r_adjoint")
        self._check(real_in, self.real_shape(), self.real, "c2r_adjoint input")
        self._check(spectrum_out, self.complex_shape(), self.complex, "c2r_adjoint output")
        _lib.check(_lib.lib().fftm_mi_ffts_c2r_adjoint(self._handle, device_ptr(real_in), device_ptr(spectrum_out),
                                                       1 if self._options.sync else 0), "fftm_mi_ffts_c2r_adjoint")

    def spectral_multiply(self, inp, out, filter=None, shell_coeffs=None):
        """out = H * inp with a real per-bin filter or radial shell coefficients."""
        self._need_handle("spectral_multiply")
        self._check(inp, self.complex_shape(), self.complex, "spectral_multiply input")
        self._check(out, self.complex_shape(), self.complex, "spectral_multiply output")
        ns = 0
        if filter is not None:
            self._check(filter, self.complex_shape(), self.real, "spectral_multiply filter")
        if shell_coeffs is not None:
            if shell_coeffs.dtype != self.real or not shell_coeffs.is_cuda or shell_coeffs.dim() != 1:
                raise ValueError("spectral_multiply: shell_coeffs must be a 1D GPU tensor of the real type")
            ns = shell_coeffs.numel()
        _lib.check(_lib.lib().fftm_mi_ffts_spectral_multiply(
            self._handle, device_ptr(inp), device_ptr(out), device_ptr(filter) if filter is not None else None,
            device_ptr(shell_coeffs) if shell_coeffs is not None else None, ns, 1 if self._options.sync else 0),
            "fftm_mi_ffts_spectral_multiply")

    def _op(self, kind, scale, sigma, coeffs, axis=None, band=None):
        kinds = {"poisson": _lib.OP_POISSON, "gaussian": _lib.OP_GAUSSIAN, "shell": _lib.OP_SHELL,
                 "derivative": _lib.OP_DERIVATIVE, "poisson_gradient": _lib.OP_POISSON_GRADIENT,
                 "bandpass": _lib.OP_BANDPASS}
        if kind not in kinds:
            raise ValueError(f"spectral operator kind must be one of {sorted(kinds)}")
        op = _lib.SpectralOp()
        op.kind = kinds[kind]
        op.scale = float(scale)
        op.sigma = float(sigma or 0.0)
        op.nshells = 0
        op.coeffs = None
        op.axis = int(axis) if axis is not None else -1
        if band is not None:
            op.klo, op.khi = float(band[0]), float(band[1])
        if kind == "shell":
            if coeffs is None or coeffs.dtype != self.real or not coeffs.is_cuda or coeffs.dim() != 1:
                raise ValueError("shell operator: coeffs must be a 1D GPU tensor of the real type")
            op.nshells = coeffs.numel()
            op.coeffs = device_ptr(coeffs)
        return op

    def spectral_apply(self, inp, out, kind, scale=1.0, sigma=None, coeffs=None, axis=None, band=None):
        """out = H(k) * inp with an analytic operator: "poisson" (-scale/|k|^2, 0 at k = 0),
        "gaussian" (scale * exp(-|k|^2 / (2 sigma^2))), "shell" (scale * coeffs[shell(k)]),
        "derivative" (i scale k_axis), "poisson_gradient" (-i scale k_axis / |k|^2) or
        "bandpass" (scale for band[0] <= |k| <= band[1]).  axis is the logical axis
        (x, y, z[, w] = 0, 1, 2[, 3])."""
        self._need_handle("spectral_apply")
        self._check(inp, self.complex_shape(), self.complex, "spectral_apply input")
        self._check(out, self.complex_shape(), self.complex, "spectral_apply output")
        op = self._op(kind, scale, sigma, coeffs, axis, band)
        _lib.check(_lib.lib().fftm_mi_ffts_spectral_apply(self._handle, device_ptr(inp), device_ptr(out),
                                                          ctypes.byref(op), 1 if self._options.sync else 0),
                   "fftm_mi_ffts_spectral_apply")

    def filter(self, inp, out, kind, scale=1.0, sigma=None, coeffs=None, axis=None, band=None):
        """out = backward(H * forward(inp)) in one call (the last forward axis, the multiply and
        the first backward axis fused into one pass); scale folds in the 1/N normalisation.
        Operators as for spectral_apply."""
        self._need_handle("filter")
        self._check(inp, self.real_shape(), self.real, "filter input")
        self._check(out, self.real_shape(), self.real, "filter output")
        op = self._op(kind, scale, sigma, coeffs, axis, band)
        with self._profiler.scope("ffts::filter"):
            _lib.check(_lib.lib().fftm_mi_ffts_filter(self._handle, device_ptr(inp), device_ptr(out),
                                                      ctypes.byref(op), 1 if self._options.sync else 0),
                       "fftm_mi_ffts_filter")

    def shell_gradient(self, grad_spectrum, spectrum, nshells):
        """g[s] = sum_{k in shell s} Re(conj(G_k) X_k)."""
        import torch
        self._need_handle("shell_gradient")
        self._check(grad_spectrum, self.complex_shape(), self.complex, "shell_gradient G")
        self._check(spectrum, self.complex_shape(), self.complex, "shell_gradient X")
        g = torch.empty(nshells, dtype=self.real, device="cuda")
        _lib.check(_lib.lib().fftm_mi_ffts_shell_gradient(self._handle, device_ptr(grad_spectrum),
                                                          device_ptr(spectrum), device_ptr(g), int(nshells),
                                                          1 if self._options.sync else 0),
                   "fftm_mi_ffts_shell_gradient")
        return g

    # 4D stage sequences (ffts.hpp:809-867); transposes as dst = transpose(src, perm)
    # on memory arrays (the arranger algebra of direct_transpose_4d.h:42-120).
    def _tr(self, src, dst, perm):
        transpose_4d(src, dst, perm, stream=getattr(self, "_stream", None), sync=self._options.sync)

    def _forward_4d_staged(self, inp, out):
        b, st = self._base, self._stage
        if self.strategy_4d.family == transform_strategy_4d.pencil_pencil:
            b.exec("forward_w", inp, st["stage0"])                 # [wh][z][y][x]
            self._tr(st["stage0"], st["stage1"], (1, 0, 2, 3))       # xyzw -> xywz: [z][wh][y][x]
            b.exec("forward_z", st["stage1"], st["stage1"])
            self._tr(st["stage1"], st["stage2"], (2, 0, 1, 3))       # xywz -> xzwy: [y][z][wh][x]
            b.exec("forward_y", st["stage2"], st["stage2"])
            out.copy_(st["stage2"])                                  # xzwy -> yzwx: identity in memory
            b.exec("forward_x", out, out)
        else:
            b.exec("forward_zw", inp, st["stage0"])
            st["stage1"].copy_(st["stage0"])                         # xyzw -> zwxy: identity in memory
            b.exec("forward_xy", st["stage1"], st["stage1"])
            out.copy_(st["stage1"])                                  # zwxy -> yzwx: identity in memory

    def _backward_4d_staged(self, inp, out):
        b, st = self._base, self._stage
        if self.strategy_4d.family == transform_strategy_4d.pencil_pencil:
            b.exec("inverse_x", inp, inp)                          # destructive on the input (ffts.hpp:377)
            st["stage2"].copy_(inp)                                  # yzwx -> xzwy: identity in memory
            b.exec("inverse_y", st["stage2"], st["stage2"])
            self._tr(st["stage2"], st["stage1"], (1, 2, 0, 3))       # xzwy -> xywz: [z][wh][y][x]
            b.exec("inverse_z", st["stage1"], st["stage1"])
            self._tr(st["stage1"], st["stage0"], (1, 0, 2, 3))       # xywz -> xyzw: [wh][z][y][x]
            b.exec("inverse_w", st["stage0"], out)
        else:
            st["stage1"].copy_(inp)                                  # yzwx -> zwxy: identity in memory
            b.exec("inverse_xy", st["stage1"], st["stage1"])
            st["stage0"].copy_(st["stage1"])                         # zwxy -> xyzw: identity in memory
            b.exec("inverse_zw", st["stage0"], out)

    # ------------------------------------------------------- observability
    def _update_memory_profile(self):
        if not self._mem.enabled:
            return
        self._mem.set_bytes("ffts/shared_work_buffer", 0 if self._work is None else self._work.numel())
        s02 = self._stage.get("stage0")
        s1 = self._stage.get("stage1")
        csz = 8 if self.real.itemsize == 4 else 16
        self._mem.set_bytes("ffts/scratch_stage0_stage2_4d",
                            0 if s02 is None else s02.untyped_storage().nbytes())
        self._mem.set_bytes("ffts/scratch_stage1_4d", 0 if s1 is None else s1.numel() * csz)

    def print_profile(self, out=sys.stdout):
        if self._profiler.enabled:
            self._profiler.log_print(out)

    def print_profile_totals(self, out=sys.stdout):
        if self._profiler.enabled:
            self._profiler.log_print_totals(out)

    def print_memory_profile(self, out=sys.stdout):
        if self._mem.enabled:
            out.write("Memory profile:")
            for k, (c, p) in self._mem.entries.items():
                out.write(f"\n  {k}: current={c} B, peak={p} B")
            out.write("\n")

    def print_memory_totals(self, out=sys.stdout):
        if self._mem.enabled:
            c, p = self._mem.total()
            out.write(f"Memory profile totals:\n  total: current={c} B, peak={p} B\n")

    def memory_profile_buckets(self):
        return dict(self._mem.entries)

    def is_memory_profiling_enabled(self) -> bool:
        return self._mem.enabled

    def release(self):
        """Destroy plans before their work area (ffts.hpp:199-224)."""
        if self._handle is not None:
            _lib.lib().fftm_mi_ffts_destroy(self._handle)
            self._handle = None
        if self._base is not None:
            self._base.release()
            self._base = None
        self._work = None
        self._stage = {}

    def close(self):
        if self._init_done:
            o = self._options
            if o.print_profile_summary_on_destroy:
                self.print_profile()
            if o.print_profile_totals_on_destroy:
                self.print_profile_totals()
            if o.print_memory_profile_on_destroy:
                self.print_memory_profile()
            if o.print_memory_totals_on_destroy:
                self.print_memory_totals()
            self._init_done = False
        self.release()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
