"""Host mirror of fftm's backend-neutral ``external_wrap`` boundary over the C ABI.

* ``fft``            -- one plan: the surface of ``fftm::wrap::fft_base<T>`` /
                         ``hipfft::fft<T,D>`` (source/external_wrap/hipfft_wrap.h:472-496, 499-1240).
* ``fft_wrap_many``  -- the named-plan container ``fft_wrap_many<BaseFFTWrap,T>``,
                         subset used by ``ffts`` (source/external_wrap/fft_wrap_many.h:120-292,
                         375-383, 1606-1680, 1753-1802).
* ``runtime_api``    -- the members of ``hip_runtime_api`` that ffts uses
                         (hipfft_wrap.h:42-257): device_synchronize, stream creation, events.

Device memory is whatever the caller passes: a CUDA/HIP ``torch.Tensor`` or a raw
integer device address.  Error behaviour follows the reference: API misuse raises
``LogicError`` (std::logic_error), unknown plan names raise ``KeyError``
(std::map::at -> std::out_of_range), backend failures raise ``FftmMiError``
(std::runtime_error from FFTM_HIPFFT_SAFE_CALL).
"""

from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from . import _lib


class LogicError(RuntimeError):
    """API misuse (the reference throws std::logic_error)."""


class direction(enum.IntEnum):
    """fftm::direction (source/fft_direction.h:9-15)."""
    R2C = 1
    C2R = 2
    C2CF = 3
    C2CB = 4


def fft_direction_name(d: direction) -> str:
    return {direction.R2C: "R2C", direction.C2R: "C2R", direction.C2CF: "C2CF",
            direction.C2CB: "C2CB"}.get(direction(d), "unknown")


@dataclass
class fft_plan_descriptor:
    """fftm::wrap::fft_plan_descriptor (source/external_wrap/fft_plan_descriptor.h:15-32)."""
    dir: direction = direction.R2C
    backend_type: int = 0
    rank: int = 0
    default_layout: bool = False
    n: Tuple[int, int, int] = (0, 0, 0)
    inembed: Tuple[int, int, int] = (0, 0, 0)
    istride: int = 0
    idist: int = 0
    onembed: Tuple[int, int, int] = (0, 0, 0)
    ostride: int = 0
    odist: int = 0
    batch: int = 0
    work_size: int = 0
    work_area_token: int = 0
    stream_token: int = 0


def _precision_of(real) -> int:
    import torch
    if real in (torch.float32, "float32", "float", "f32"):
        return _lib.F32
    if real in (torch.float64, "float64", "double", "f64"):
        return _lib.F64
    raise TypeError(f"unsupported real type {real!r} (float32 / float64)")


def device_ptr(x) -> int:
    """Raw device address of a tensor (or pass an int through)."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if not x.is_cuda:
            raise ValueError("fftm_amd: tensor must live on the GPU")
        return int(x.data_ptr())
    raise TypeError(f"cannot take a device address of {type(x).__name__}")


def stream_handle(stream) -> int:
    if stream is None:
        return 0
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


class runtime_api:
    """Members of hip_runtime_api (hipfft_wrap.h:42-257) used around ffts."""

    @staticmethod
    def device_synchronize():
        import torch
        torch.cuda.synchronize()

    @staticmethod
    def create_nonblocking_stream():
        import torch
        return torch.cuda.Stream()

    @staticmethod
    def stream_synchronize(stream):
        stream.synchronize()

    @staticmethod
    def stream_ready(stream) -> bool:
        return bool(stream.query())


class fft:
    """One plan -- fft<T,D> (hipfft_wrap.h:499-1240) on the MI355X backend.

    ``fft(real, D, n, inembed, istride, idist, onembed, ostride, odist, batch)``
    for the advanced layout (constructors :520-580), or ``fft(real, D, n, batch=b)``
    for the default layout (:582-602).  ``n``/``inembed``/``onembed`` are sequences
    of length rank (1..3)."""

    def __init__(self, real, D, n: Sequence[int], inembed: Optional[Sequence[int]] = None,
                 istride: int = 1, idist: int = 0, onembed: Optional[Sequence[int]] = None,
                 ostride: int = 1, odist: int = 0, batch: int = 1, flags: int = 0):
        self._handle = None
        L = _lib.lib()
        D = direction(D)
        rank = len(n)
        if not 1 <= rank <= 3:
            raise LogicError("fft: supports only 1D, 2D, and 3D plans")
        desc = _lib.PlanDesc()
        desc.dir = int(D)
        desc.precision = _precision_of(real)
        desc.rank = rank
        desc.default_layout = 1 if inembed is None else 0
        for k in range(rank):
            desc.n[k] = int(n[k])
            if inembed is not None:
                desc.inembed[k] = int(inembed[k])
                desc.onembed[k] = int(onembed[k])
        desc.istride, desc.idist = int(istride), int(idist)
        desc.ostride, desc.odist = int(ostride), int(odist)
        desc.batch = int(batch)
        desc.flags = int(flags)
        h = ctypes.c_void_p()
        _lib.check(L.fftm_mi_plan_create(ctypes.byref(desc), ctypes.byref(h)), "fftm_mi_plan_create")
        self._handle = h
        self._desc = desc
        self._dir = D
        self._real = real
        self._work_token = 0
        self._stream_token = 0
        ws = ctypes.c_size_t()
        _lib.check(L.fftm_mi_plan_get_work_size(h, ctypes.byref(ws)), "fftm_mi_plan_get_work_size")
        self.work_size = int(ws.value)

    # -- fft_base<T> surface ------------------------------------------------
    def exec(self, inp, out):
        self._exec(self._dir, inp, out, sync=True)

    def exec_no_sync(self, inp, out):
        self._exec(self._dir, inp, out, sync=False)

    def exec_direction(self, exec_dir, inp, out):
        self._exec(direction(exec_dir), inp, out, sync=True)

    def exec_direction_no_sync(self, exec_dir, inp, out):
        self._exec(direction(exec_dir), inp, out, sync=False)

    def _exec(self, d, inp, out, sync):
        self._check_alive()
        if d != self._dir and not (self._dir in (direction.C2CF, direction.C2CB)
                                   and d in (direction.C2CF, direction.C2CB)):
            raise LogicError("fft::exec_typed_direction: incompatible execution direction")
        _lib.check(_lib.lib().fftm_mi_plan_exec(self._handle, int(d), device_ptr(inp), device_ptr(out),
                                                1 if sync else 0), "fftm_mi_plan_exec")

    def get_direction(self) -> direction:
        return self._dir

    def descriptor(self) -> fft_plan_descriptor:
        d = self._desc
        r = d.rank
        pad = lambda a: tuple(int(a[k]) if k < r else 0 for k in range(3))
        return fft_plan_descriptor(
            dir=self._dir, backend_type=int(self._dir), rank=r, default_layout=bool(d.default_layout),
            n=pad(d.n), inembed=pad(d.inembed), istride=int(d.istride), idist=int(d.idist),
            onembed=pad(d.onembed), ostride=int(d.ostride), odist=int(d.odist), batch=int(d.batch),
            work_size=self.work_size, work_area_token=self._work_token, stream_token=self._stream_token)

    def get_work_size(self) -> int:
        return self.work_size

    def set_work_area(self, work_area):
        self._check_alive()
        ptr = device_ptr(work_area)
        self._work_token = ptr
        self._work_ref = work_area
        _lib.check(_lib.lib().fftm_mi_plan_set_work_area(self._handle, ptr), "fftm_mi_plan_set_work_area")

    def set_stream(self, stream):
        self._check_alive()
        h = stream_handle(stream)
        self._stream_token = h
        _lib.check(_lib.lib().fftm_mi_plan_set_stream(self._handle, h), "fftm_mi_plan_set_stream")

    def recreate_with_stream_and_work_area(self, stream, work_area):
        # plans hold no state beyond the stream and the work area
        self.set_stream(stream)
        self.set_work_area(work_area)

    def num_passes(self) -> int:
        n = ctypes.c_int32()
        _lib.check(_lib.lib().fftm_mi_plan_num_passes(self._handle, ctypes.byref(n)), "fftm_mi_plan_num_passes")
        return int(n.value)

    def destroy(self):
        if self._handle is not None:
            _lib.lib().fftm_mi_plan_destroy(self._handle)
            self._handle = None

    def _check_alive(self):
        if self._handle is None:
            raise LogicError("fft: plan has been destroyed")

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class fft_wrap_many:
    """Named-plan container (fft_wrap_many.h:25-1927): the plans ffts uses, raw-pointer
    execution, plan sequences executed at element offsets (fft_wrap_many.h:435-780),
    plans bound to their own non-blocking streams (fft_wrap_many.h:396-433), C2C plan
    arrays (one plan per element offset, each on its own stream; :784-868, 1260-1530)
    and R2C/C2R plan bundles (lanes of 2D forward/inverse plan pairs with their own
    streams and work areas; :188-237, 310-378, 1815-1855).

    The plan-array handle and stream tokens (:1270-1296) return the plans' opaque handle
    and stream values.  The minimal-reference plan arrays and the opt0 local plan contexts
    (:1297-1365), diagnostics of fftm's MPI transposes, raise ``NotImplementedError``."""

    def __init__(self, real):
        self.real = real
        self._plans: Dict[str, fft] = {}
        self._work_area_size = 0
        self._activated = False
        self._external_activated = False
        self._hot_exec_no_sync = False
        self._work = None
        self._memory_profiler = None
        self._memory_profile_prefix = ""
        self._plan_arrays = []
        self._bundles = []

    # -- plan creation (fft_wrap_many.h:120-186) ----------------------------
    def add_plan(self, D, name: str, n, inembed, istride, idist, onembed, ostride, odist, batch):
        if self._activated:
            raise LogicError("fft_wrap_many::add_plan: cannot add new plans after activation.")
        if name in self._plans:
            return  # std::map::emplace keeps the existing entry
        self._plans[name] = fft(self.real, D, n, inembed, istride, idist, onembed, ostride, odist, batch)

    def add_plan_1D(self, D, name, n, inembed, istride, idist, onembed, ostride, odist, batch):
        self.add_plan(D, name, [n], [inembed], istride, idist, [onembed], ostride, odist, batch)

    def add_plan_1D_default(self, D, name, n, batch):
        if self._activated:
            raise LogicError("fft_wrap_many::add_plan_1D_default: cannot add new plans after activation.")
        if name not in self._plans:
            self._plans[name] = fft(self.real, D, [n], batch=batch)

    def add_plan_2D(self, D, name, n0, n1, inembed0, inembed1, istride, idist, onembed0, onembed1, ostride,
                    odist, batch):
        self.add_plan(D, name, [n0, n1], [inembed0, inembed1], istride, idist, [onembed0, onembed1], ostride,
                      odist, batch)

    def add_plan_3D(self, D, name, n0, n1, n2, inembed0, inembed1, inembed2, istride, idist, onembed0,
                    onembed1, onembed2, ostride, odist, batch):
        self.add_plan(D, name, [n0, n1, n2], [inembed0, inembed1, inembed2], istride, idist,
                      [onembed0, onembed1, onembed2], ostride, odist, batch)

    # -- work area (fft_wrap_many.h:239-292) ---------------------------------
    def activate(self):
        if self._activated:
            raise LogicError("fft_wrap_many::activate: cannot activate again.")
        if self._external_activated:
            raise LogicError("fft_wrap_many::activate: cannot use external and internal activation, the class "
                             "is externally activated.")
        self.activate_work_size()
        if self._work_area_size > 0:
            import torch
            self._work = torch.empty(self._work_area_size, dtype=torch.uint8, device="cuda")
            for p in self._plans.values():
                p.set_work_area(self._work)
            self._bind_bundle_work_areas(device_ptr(self._work))
        self._update_memory_profile()
        self._activated = True

    def activate_work_size(self) -> int:
        for p in self._plans.values():
            self._work_area_size = max(self._work_area_size, p.get_work_size())
        for b in self._bundles:
            self._work_area_size = max(self._work_area_size, self._bundle_work_size(b))
        return self._work_area_size

    def set_external_work_area(self, work_area):
        if self._activated:
            raise LogicError("fft_wrap_many::set_external_work_area: cannot set external work_area with activated "
                             "local work_area.")
        for p in self._plans.values():
            p.set_work_area(work_area)
        self._bind_bundle_work_areas(device_ptr(work_area))
        self._external_activated = True
        self._update_memory_profile()

    def get_work_size(self) -> int:
        return self._work_area_size

    def set_memory_profiler(self, profiler, prefix: str):
        self._memory_profiler = profiler
        self._memory_profile_prefix = prefix
        self._update_memory_profile()

    def set_hot_exec_no_sync_enabled(self, enabled: bool):
        self._hot_exec_no_sync = bool(enabled)

    # -- execution (fft_wrap_many.h:375-395) ---------------------------------
    def exec(self, name: str, inp, out):
        plan = self._plans[name]
        if self._hot_exec_no_sync:
            plan.exec_no_sync(inp, out)
        else:
            plan.exec(inp, out)

    def exec_named_no_sync(self, name: str, inp, out):
        self._plans[name].exec_no_sync(inp, out)

    def synchronize_device_execution(self):
        runtime_api.device_synchronize()

    # -- per-plan accessors (fft_wrap_many.h:1628-1680) ----------------------
    def work_size(self, name: str) -> int:
        return self._plans[name].get_work_size()

    def set_work_area(self, name: str, work_area):
        self._plans[name].set_work_area(work_area)

    def set_stream(self, name: str, stream):
        self._plans[name].set_stream(stream)

    def recreate_plan_with_stream_and_work_area(self, name: str, stream, work_area):
        self._plans[name].recreate_with_stream_and_work_area(stream, work_area)

    def plan_descriptor(self, name: str) -> fft_plan_descriptor:
        return self._plans[name].descriptor()

    def plan_descriptors(self) -> List[Tuple[str, fft_plan_descriptor]]:
        return [(k, self._plans[k].descriptor()) for k in sorted(self._plans)]

    def plan(self, name: str) -> fft:
        return self._plans[name]

    # -- lifecycle (fft_wrap_many.h:1753-1802) -------------------------------
    def release(self):
        for p in self._plans.values():
            p.destroy()
        self._plans.clear()
        for arr in getattr(self, "_plan_arrays", []):
            for p in arr["plans"]:
                p.destroy()
        for b in getattr(self, "_bundles", []):
            for p in b["forward"] + b["inverse"]:
                p.destroy()
        self._plan_arrays = []
        self._bundles = []
        self._work = None
        self._work_area_size = 0
        self._activated = False
        self._external_activated = False
        self._hot_exec_no_sync = False
        self._update_memory_profile()

    def _update_memory_profile(self):
        if self._memory_profiler is None or not self._memory_profile_prefix:
            return
        self._memory_profiler.set_bytes(self._memory_profile_prefix + "/work_area",
                                        0 if self._external_activated else self._work_area_size)

    # -- raw execution (fft_wrap_many.h:435-451) ------------------------------
    def exec_raw(self, name: str, in_ptr: int, out_ptr: int):
        p = self._plans[name]
        p._exec(p.get_direction(), int(in_ptr), int(out_ptr), sync=not self._hot_exec_no_sync)

    def exec_raw_direction(self, name: str, exec_dir, in_ptr: int, out_ptr: int):
        self._plans[name]._exec(direction(exec_dir), int(in_ptr), int(out_ptr), sync=not self._hot_exec_no_sync)

    # -- owned streams (fft_wrap_many.h:396-433) -------------------------------
    def bind_plan_to_owned_nonblocking_stream(self, name: str):
        if not hasattr(self, "_owned_streams"):
            self._owned_streams = {}
        if name not in self._owned_streams:
            self._owned_streams[name] = runtime_api.create_nonblocking_stream()
        self._plans[name].set_stream(self._owned_streams[name])

    def synchronize_owned_plan_stream(self, name: str):
        streams = getattr(self, "_owned_streams", {})
        if name not in streams:
            raise KeyError("fft_wrap_many: plan has no owned execution stream")
        runtime_api.stream_synchronize(streams[name])

    # -- plan sequences (fft_wrap_many.h:453-780) -------------------------------
    def make_plan_sequence(self, names) -> int:
        """A fixed list of named plans, executed in order; returns its id."""
        if not hasattr(self, "_sequences"):
            self._sequences = []
        self._sequences.append([self._plans[n] for n in names])
        return len(self._sequences) - 1

    def plan_sequence_size(self, sequence_id: int) -> int:
        return len(self._sequence(sequence_id))

    def _sequence(self, sequence_id):
        seqs = getattr(self, "_sequences", [])
        if sequence_id < 0 or sequence_id >= len(seqs):
            raise IndexError("fft_wrap_many: invalid plan sequence id")
        return seqs[sequence_id]

    def _run_sequence(self, sequence_id, offsets, exec_dir, inp, out, sync, what):
        seq = self._sequence(sequence_id)
        if len(seq) != len(offsets):
            raise LogicError(f"fft_wrap_many::{what}: plan/offset count mismatch.")
        # the offsets count elements of the input / output arrays (typed pointer arithmetic)
        ib, ob = inp.element_size(), out.element_size()
        ip, op = device_ptr(inp), device_ptr(out)
        for plan, off in zip(seq, offsets):
            d = plan.get_direction() if exec_dir is None else direction(exec_dir)
            plan._exec(d, ip + int(off) * ib, op + int(off) * ob, sync=sync)

    def exec_plan_sequence_offsets(self, sequence_id, offsets, inp, out):
        self._run_sequence(sequence_id, offsets, None, inp, out, True, "exec_plan_sequence_offsets")

    def exec_plan_sequence_offsets_direction(self, sequence_id, offsets, exec_dir, inp, out):
        self._run_sequence(sequence_id, offsets, exec_dir, inp, out, True, "exec_plan_sequence_offsets_direction")

    def exec_plan_sequence_offsets_direction_repeated(self, sequence_id, offsets, exec_dir, inp, out, repeats):
        for _ in range(int(repeats)):
            self.exec_plan_sequence_offsets_direction(sequence_id, offsets, exec_dir, inp, out)

    def exec_plan_sequence_offsets_no_sync(self, sequence_id, offsets, inp, out):
        self._run_sequence(sequence_id, offsets, None, inp, out, False, "exec_plan_sequence_offsets_no_sync")

    def exec_plan_sequence_offsets_direction_no_sync(self, sequence_id, offsets, exec_dir, inp, out):
        self._run_sequence(sequence_id, offsets, exec_dir, inp, out, False,
                           "exec_plan_sequence_offsets_direction_no_sync")

    def exec_plan_sequence_offsets_direction_no_sync_repeated(self, sequence_id, offsets, exec_dir, inp, out,
                                                              repeats):
        for _ in range(int(repeats)):
            self.exec_plan_sequence_offsets_direction_no_sync(sequence_id, offsets, exec_dir, inp, out)

    # the reference's "tight" and "opaque" forms are faster dispatches of the same
    # sequence (typed C2C plans, raw plan handles); here they are the same loop
    exec_plan_sequence_offsets_tight = exec_plan_sequence_offsets
    exec_plan_sequence_offsets_tight_direction = exec_plan_sequence_offsets_direction
    exec_plan_sequence_offsets_opaque_direction = exec_plan_sequence_offsets_direction

    # -- C2C plan arrays (fft_wrap_many.h:784-868, 1260-1530) -------------------
    # One C2C plan per element offset, each on its own non-blocking stream (or on the
    # caller's streams); the work areas are bound explicitly, plan i at base + i * stride.
    @staticmethod
    def invalid_c2c_plan_array_id() -> int:
        return (1 << 64) - 1

    def _new_c2c_plan_array(self, what, n, inembed, istride, idist, onembed, ostride, odist, batch, offsets,
                            streams=None):
        if self._activated:
            raise LogicError(f"fft_wrap_many::{what}: cannot add plan arrays after activation.")
        arr = {"plans": [], "streams": [], "external_streams": [], "offsets": [int(o) for o in offsets],
               "work_stride": 0, "descriptor": None}
        for i in range(len(arr["offsets"])):
            plan = fft(self.real, direction.C2CF, [n], [inembed], istride, idist, [onembed], ostride, odist, batch)
            if streams is None:
                st = runtime_api.create_nonblocking_stream()
                arr["streams"].append(st)
            else:
                st = streams[i]
                arr["external_streams"].append(st)
            plan.set_stream(st)
            arr["work_stride"] = max(arr["work_stride"], plan.get_work_size())
            if i == 0:
                arr["descriptor"] = plan.descriptor()
            arr["plans"].append(plan)
        self._plan_arrays.append(arr)
        return len(self._plan_arrays) - 1

    def make_c2c_plan_array_1D(self, n, inembed, istride, idist, onembed, ostride, odist, batch, offsets) -> int:
        return self._new_c2c_plan_array("make_c2c_plan_array_1D", n, inembed, istride, idist, onembed, ostride,
                                        odist, batch, offsets)

    def make_raw_c2c_plan_array_1D(self, n, inembed, istride, idist, onembed, ostride, odist, batch,
                                   offsets) -> int:
        # the reference's raw-handle arrays dispatch the same plans through hipFFT handles
        return self._new_c2c_plan_array("make_raw_c2c_plan_array_1D", n, inembed, istride, idist, onembed,
                                        ostride, odist, batch, offsets)

    def make_raw_c2c_plan_array_1D_with_streams(self, n, inembed, istride, idist, onembed, ostride, odist, batch,
                                                offsets, streams) -> int:
        if len(streams) < len(offsets):
            raise LogicError("fft_wrap_many::make_raw_c2c_plan_array_1D_with_streams: stream count mismatch")
        return self._new_c2c_plan_array("make_raw_c2c_plan_array_1D_with_streams", n, inembed, istride, idist,
                                        onembed, ostride, odist, batch, offsets, streams)

    def _plan_array(self, array_id):
        if array_id < 0 or array_id >= len(self._plan_arrays):
            raise IndexError("fft_wrap_many: invalid C2C plan-array id")
        return self._plan_arrays[array_id]

    def c2c_plan_array_size(self, array_id) -> int:
        return len(self._plan_array(array_id)["plans"])

    def c2c_plan_array_work_size(self, array_id) -> int:
        return self._plan_array(array_id)["work_stride"]

    def c2c_plan_array_descriptor(self, array_id) -> fft_plan_descriptor:
        return self._plan_array(array_id)["descriptor"]

    def bind_c2c_plan_array_work_areas(self, array_id, base_work_area, work_stride_bytes: int):
        arr = self._plan_array(array_id)
        stride = max(int(work_stride_bytes), arr["work_stride"])
        base = device_ptr(base_work_area)
        for i, plan in enumerate(arr["plans"]):
            plan.set_work_area(base + i * stride if base else 0)
        arr["bound_stride"] = stride
        arr["work_ref"] = base_work_area

    def bind_c2c_plan_array_streams(self, array_id, streams):
        arr = self._plan_array(array_id)
        if len(streams) < len(arr["plans"]):
            raise LogicError("fft_wrap_many::bind_c2c_plan_array_streams: stream count mismatch")
        arr["external_streams"] = list(streams[:len(arr["plans"])])
        for plan, st in zip(arr["plans"], arr["external_streams"]):
            plan.set_stream(st)

    def synchronize_c2c_plan_array_streams(self, array_id):
        arr = self._plan_array(array_id)
        for st in (arr["external_streams"] or arr["streams"]):
            runtime_api.stream_synchronize(st)

    def _run_plan_array(self, array_id, exec_dir, in_offsets, out_offsets, inp, out, sync):
        arr = self._plan_array(array_id)
        ib, ob = inp.element_size(), out.element_size()
        ip, op = device_ptr(inp), device_ptr(out)
        d = direction(exec_dir)
        for plan, io, oo in zip(arr["plans"], in_offsets, out_offsets):
            plan._exec(d, ip + int(io) * ib, op + int(oo) * ob, sync=sync)

    def exec_c2c_plan_array_direction(self, array_id, exec_dir, inp, out):
        offs = self._plan_array(array_id)["offsets"]
        self._run_plan_array(array_id, exec_dir, offs, offs, inp, out, True)

    def exec_c2c_plan_array_direction_no_sync(self, array_id, exec_dir, inp, out):
        offs = self._plan_array(array_id)["offsets"]
        self._run_plan_array(array_id, exec_dir, offs, offs, inp, out, False)

    def exec_c2c_plan_array_direction_dual_offsets_no_sync(self, array_id, exec_dir, in_offsets, out_offsets,
                                                           inp, out):
        n = len(self._plan_array(array_id)["plans"])
        if n != len(in_offsets) or n != len(out_offsets):
            raise LogicError("fft_wrap_many::exec_c2c_plan_array_direction_dual_offsets_no_sync: offset count "
                             "mismatch")
        self._run_plan_array(array_id, exec_dir, in_offsets, out_offsets, inp, out, False)

    # -- R2C/C2R plan bundles (fft_wrap_many.h:188-237, 310-378, 1815-1855) -----
    @staticmethod
    def invalid_r2c_c2r_plan_bundle_id() -> int:
        return (1 << 64) - 1

    def make_r2c_c2r_plan_bundle_2D(self, n0, n1, forward_inembed0, forward_inembed1, forward_istride,
                                    forward_idist, forward_onembed0, forward_onembed1, forward_ostride,
                                    forward_odist, forward_batch, inverse_inembed0, inverse_inembed1,
                                    inverse_istride, inverse_idist, inverse_onembed0, inverse_onembed1,
                                    inverse_ostride, inverse_odist, inverse_batch, lane_count) -> int:
        if self._activated or self._external_activated:
            raise LogicError("fft_wrap_many::make_r2c_c2r_plan_bundle_2D: cannot add plan bundles after activation.")
        if int(lane_count) <= 0:
            raise ValueError("fft_wrap_many::make_r2c_c2r_plan_bundle_2D: lane count must be positive")
        b = {"forward": [], "inverse": [], "streams": [], "work_stride": 0, "bound_stride": 0}
        for _ in range(int(lane_count)):
            st = runtime_api.create_nonblocking_stream()
            f = fft(self.real, direction.R2C, [n0, n1], [forward_inembed0, forward_inembed1], forward_istride,
                    forward_idist, [forward_onembed0, forward_onembed1], forward_ostride, forward_odist,
                    forward_batch)
            g = fft(self.real, direction.C2R, [n0, n1], [inverse_inembed0, inverse_inembed1], inverse_istride,
                    inverse_idist, [inverse_onembed0, inverse_onembed1], inverse_ostride, inverse_odist,
                    inverse_batch)
            f.set_stream(st)
            g.set_stream(st)
            b["work_stride"] = max(b["work_stride"], f.get_work_size(), g.get_work_size())
            b["forward"].append(f)
            b["inverse"].append(g)
            b["streams"].append(st)
        self._bundles.append(b)
        return len(self._bundles) - 1

    @staticmethod
    def _aligned_bundle_stride(nbytes: int) -> int:
        return 0 if nbytes == 0 else (nbytes + 255) // 256 * 256

    def _bundle_work_size(self, b) -> int:
        return self._aligned_bundle_stride(b["work_stride"]) * len(b["streams"])

    def _bind_bundle_work_areas(self, base: int):
        for b in self._bundles:
            stride = self._aligned_bundle_stride(b["work_stride"])
            for lane in range(len(b["streams"])):
                w = base + lane * stride if base else 0
                b["forward"][lane].set_work_area(w)
                b["inverse"][lane].set_work_area(w)
            b["bound_stride"] = stride

    def _bundle(self, bundle_id):
        if bundle_id < 0 or bundle_id >= len(self._bundles):
            raise IndexError("fft_wrap_many: invalid R2C/C2R plan-bundle id")
        return self._bundles[bundle_id]

    def r2c_c2r_plan_bundle_lane_count(self, bundle_id) -> int:
        return len(self._bundle(bundle_id)["streams"])

    def r2c_c2r_plan_bundle_work_size(self, bundle_id) -> int:
        return self._bundle_work_size(self._bundle(bundle_id))

    def _bundle_lane(self, bundle_id, lane, key, what):
        b = self._bundle(bundle_id)
        if lane < 0 or lane >= len(b[key]):
            raise IndexError(f"fft_wrap_many: R2C/C2R plan-bundle {what} lane is out of range")
        return b, b[key][lane]

    def exec_r2c_c2r_plan_bundle_forward_lane_no_sync(self, bundle_id, lane, in_offset, out_offset, inp, out):
        _, plan = self._bundle_lane(bundle_id, lane, "forward", "forward")
        plan._exec(direction.R2C, device_ptr(inp) + int(in_offset) * inp.element_size(),
                   device_ptr(out) + int(out_offset) * out.element_size(), sync=False)

    def exec_r2c_c2r_plan_bundle_inverse_lane_no_sync(self, bundle_id, lane, in_offset, out_offset, inp, out):
        _, plan = self._bundle_lane(bundle_id, lane, "inverse", "inverse")
        plan._exec(direction.C2R, device_ptr(inp) + int(in_offset) * inp.element_size(),
                   device_ptr(out) + int(out_offset) * out.element_size(), sync=False)

    def r2c_c2r_plan_bundle_lane_ready(self, bundle_id, lane) -> bool:
        _, st = self._bundle_lane(bundle_id, lane, "streams", "readiness")
        return runtime_api.stream_ready(st)

    def synchronize_r2c_c2r_plan_bundle_lane(self, bundle_id, lane):
        _, st = self._bundle_lane(bundle_id, lane, "streams", "synchronization")
        runtime_api.stream_synchronize(st)

    def synchronize_r2c_c2r_plan_bundle(self, bundle_id):
        for st in self._bundle(bundle_id)["streams"]:
            runtime_api.stream_synchronize(st)

    # plan-array tokens (fft_wrap_many.h:1270-1296): the first / last plan handle and
    # stream as integers, 0 for an empty array
    def _array_token(self, array_id, which, kind):
        arr = self._plan_array(array_id)
        if not arr["plans"]:
            return 0
        if kind == "handle":
            h = arr["plans"][which]._handle
            return int(getattr(h, "value", h) or 0)
        streams = arr["external_streams"] or arr["streams"]
        return stream_handle(streams[which])

    def c2c_plan_array_first_handle_token(self, array_id) -> int:
        return self._array_token(array_id, 0, "handle")

    def c2c_plan_array_last_handle_token(self, array_id) -> int:
        return self._array_token(array_id, -1, "handle")

    def c2c_plan_array_first_stream_token(self, array_id) -> int:
        return self._array_token(array_id, 0, "stream")

    def c2c_plan_array_last_stream_token(self, array_id) -> int:
        return self._array_token(array_id, -1, "stream")

    # -- fftm's MPI-side diagnostics ---------------------------------------------
    def _mpi_diagnostics(self, *a, **k):
        raise NotImplementedError("fft_wrap_many: the minimal-reference plan arrays and the opt0 local plan "
                                  "contexts are diagnostics of fftm's MPI transposes and are not mirrored")

    make_minimal_reference_c2c_plan_array_1D_diagnostic = _mpi_diagnostics
    make_reference_opt0_local_plan_context_1D = _mpi_diagnostics

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def fill_random_real(dst, logical_dims: Sequence[int], seed: int, logical_strides: Optional[Sequence[int]] = None,
                     start: Optional[Sequence[int]] = None, stream=None, sync: bool = False):
    """fill_random_real_3d_functor / fill_random_real_4d_functor of the benchmark harness
    (source/tests/detail/fft_benchmark_common.h:201-257) run on the device into `dst`
    (float32 or float64, 3 or 4 logical axes x, y, z[, w]).  `logical_strides` are the
    element strides of the logical axes in `dst` (default: C order of `logical_dims`, i.e.
    the 3D / slab real layouts); `start` is a slab's logical origin."""
    import torch
    nd = len(logical_dims)
    if nd not in (3, 4):
        raise ValueError("fill_random_real: 3 or 4 logical axes")
    if dst.dtype not in (torch.float32, torch.float64):
        raise TypeError("fill_random_real: dst must be float32 or float64")
    if logical_strides is None:
        logical_strides = [1] * nd
        for a in range(nd - 2, -1, -1):
            logical_strides[a] = logical_strides[a + 1] * int(logical_dims[a + 1])
    span = 1 + sum((int(n) - 1) * int(s) for n, s in zip(logical_dims, logical_strides))
    if span > dst.numel() or not dst.is_contiguous():
        raise ValueError("fill_random_real: the strides address elements outside the contiguous tensor")
    st = None if start is None else _lib.i64_array(start)
    L = _lib.lib()
    _lib.check(L.fftm_mi_fill_random_real(_lib.F32 if dst.dtype == torch.float32 else _lib.F64, nd,
                                          _lib.i64_array(logical_dims), _lib.i64_array(logical_strides), st,
                                          int(seed) & 0xFFFFFFFFFFFFFFFF, device_ptr(dst), stream_handle(stream),
                                          1 if sync else 0), "fftm_mi_fill_random_real")
    return dst


def transpose_4d(src, dst, perm: Sequence[int], stream=None, sync: bool = False):
    """dst = transpose(src, perm) for 4D memory arrays (torch tensors, contiguous)."""
    if src.dim() != 4:
        raise ValueError("transpose_4d: src must be a 4D tensor")
    exp = tuple(src.shape[p] for p in perm)
    if tuple(dst.shape) != exp:
        raise ValueError(f"transpose_4d: destination shape {tuple(dst.shape)} != expected {exp}")
    if not (src.is_contiguous() and dst.is_contiguous()) or src.dtype != dst.dtype:
        raise ValueError("transpose_4d: tensors must be contiguous and of one dtype")
    L = _lib.lib()
    _lib.check(L.fftm_mi_transpose_4d(src.element_size(), _lib.i64_array(src.shape), _lib.i32_array(perm),
                                      device_ptr(src), device_ptr(dst), stream_handle(stream), 1 if sync else 0),
               "fftm_mi_transpose_4d")
