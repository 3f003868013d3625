"""The named-plan container beyond ffts' own use: raw-pointer execution, plan sequences
executed at element offsets (fft_wrap_many.h:435-780), owned non-blocking streams
(fft_wrap_many.h:396-433), and the 2D ffts (ffts.hpp:232-246, 659-675)."""

import numpy as np
import pytest

import oracle
from helpers import check_close, from_gpu, to_gpu

pytestmark = pytest.mark.gpu


def _rng_cplx(shape, seed):
    g = np.random.default_rng(seed)
    return (g.uniform(-1, 1, size=shape) + 1j * g.uniform(-1, 1, size=shape)).astype(np.complex64)


def test_plan_sequence_offsets(gpu):
    import torch
    import fftm_amd
    D = fftm_amd.direction
    w = fftm_amd.fft_wrap_many(torch.float32)
    for i, (a, b) in enumerate([(16, 16), (8, 32), (16, 16)]):
        w.add_plan_2D(D.C2CF, f"p{i}", a, b, a, b, 1, a * b, a, b, 1, a * b, 1)
    w.activate()
    sid = w.make_plan_sequence(["p0", "p1", "p2"])
    assert w.plan_sequence_size(sid) == 3
    c = _rng_cplx(768, 3)
    x = to_gpu(c)
    y = torch.zeros_like(x)
    offs = [0, 256, 512]
    w.exec_plan_sequence_offsets(sid, offs, x, y)
    yh = from_gpu(y)
    for (a, b), o in zip([(16, 16), (8, 32), (16, 16)], offs):
        ref = np.fft.fft2(c[o:o + a * b].astype(np.complex128).reshape(a, b)).ravel()
        check_close(yh[o:o + a * b], ref, np.float32, f"sequence plan at {o}")
    # the same sequence executed backward through the direction form, repeated
    z = torch.zeros_like(x)
    w.exec_plan_sequence_offsets_direction_repeated(sid, offs, D.C2CB, y, z, 2)
    check_close(from_gpu(z) / 256.0, c, np.float32, "sequence round trip")
    with pytest.raises(fftm_amd.LogicError):
        w.exec_plan_sequence_offsets(sid, [0, 256], x, y)
    with pytest.raises(IndexError):
        w.plan_sequence_size(7)
    with pytest.raises(NotImplementedError):
        w.make_c2c_plan_array()
    w.release()


def test_exec_raw_and_owned_stream(gpu):
    import torch
    import fftm_amd
    D = fftm_amd.direction
    w = fftm_amd.fft_wrap_many(torch.float32)
    w.add_plan_1D(D.C2CF, "f", 64, 64, 1, 64, 64, 1, 64, 4)
    w.activate()
    w.bind_plan_to_owned_nonblocking_stream("f")
    c = _rng_cplx(256, 5)
    x = to_gpu(c)
    y = torch.zeros_like(x)
    w.set_hot_exec_no_sync_enabled(True)
    w.exec_raw("f", x.data_ptr(), y.data_ptr())
    w.synchronize_owned_plan_stream("f")
    check_close(from_gpu(y), np.fft.fft(c.astype(np.complex128).reshape(4, 64), axis=1).ravel(), np.float32,
                "exec_raw on an owned stream")
    w.exec_raw_direction("f", D.C2CB, y.data_ptr(), x.data_ptr())
    w.synchronize_owned_plan_stream("f")
    check_close(from_gpu(x) / 64.0, c, np.float32, "exec_raw_direction")
    w.release()


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("dims", [(64, 48), (12, 10), (256, 384)])
def test_ffts_2d(gpu, dt, dims):
    import torch
    import fftm_amd
    f = fftm_amd.ffts(dt)
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    npdt = np.float32 if dt == "float32" else np.float64
    x = np.random.default_rng(9).uniform(-1, 1, size=dims).astype(npdt)
    u = to_gpu(x)
    X = f.complex_array()
    f.forward(u, X)
    check_close(from_gpu(X), np.fft.rfft2(x.astype(np.float64)), npdt, f"ffts 2D forward {dims}")
    y = f.real_array()
    f.backward(X, y)
    check_close(from_gpu(y) / np.prod(dims), x.astype(np.float64), npdt, f"ffts 2D round trip {dims}")
    f.close()
