"""The named-plan container beyond ffts' own use: raw-pointer execution, plan sequences
executed at element offsets (fft_wrap_many.h:435-780), owned non-blocking streams
(fft_wrap_many.h:396-433), C2C plan arrays (:784-868, 1260-1530), R2C/C2R plan
bundles (:188-237, 310-378, 1815-1855), and the 2D ffts (ffts.hpp:232-246, 659-675)."""

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


def test_c2c_plan_array(gpu):
    import torch
    import fftm_amd
    D = fftm_amd.direction
    w = fftm_amd.fft_wrap_many(torch.float32)
    n, batch = 64, 4
    offs = [0, 512, 1024]  # elements of the arrays the plans run on
    aid = w.make_c2c_plan_array_1D(n, n, 1, n, n, 1, n, batch, offs)
    assert w.c2c_plan_array_size(aid) == 3
    assert w.c2c_plan_array_descriptor(aid).batch == batch
    w.activate()
    with pytest.raises(fftm_amd.LogicError):
        w.make_c2c_plan_array_1D(n, n, 1, n, n, 1, n, batch, offs)
    ws = w.c2c_plan_array_work_size(aid)
    work = torch.empty(3 * max(ws, 256), dtype=torch.uint8, device="cuda")
    w.bind_c2c_plan_array_work_areas(aid, work, ws)
    c = _rng_cplx(1536, 9)
    x = to_gpu(c)
    y = torch.zeros_like(x)
    w.exec_c2c_plan_array_direction(aid, D.C2CF, x, y)
    w.synchronize_c2c_plan_array_streams(aid)
    yh = from_gpu(y)
    for o in offs:
        ref = np.fft.fft(c[o:o + n * batch].astype(np.complex128).reshape(batch, n), axis=1).ravel()
        check_close(yh[o:o + n * batch], ref, np.float32, f"plan array at {o}")
    assert np.all(yh[n * batch:512] == 0), "elements between the plans' segments are left alone"
    # inverse on the caller's streams, each plan from its input offset to another output offset
    streams = [torch.cuda.Stream() for _ in range(3)]
    w.bind_c2c_plan_array_streams(aid, streams)
    outs = [256, 768, 1280]
    z = torch.zeros_like(x)
    torch.cuda.synchronize()
    w.exec_c2c_plan_array_direction_dual_offsets_no_sync(aid, D.C2CB, offs, outs, y, z)
    w.synchronize_c2c_plan_array_streams(aid)
    zh = from_gpu(z)
    for o, q in zip(offs, outs):
        check_close(zh[q:q + n * batch] / n, c[o:o + n * batch], np.float32, f"plan array inverse {o}->{q}")
    with pytest.raises(fftm_amd.LogicError):
        w.bind_c2c_plan_array_streams(aid, streams[:2])
    with pytest.raises(fftm_amd.LogicError):
        w.exec_c2c_plan_array_direction_dual_offsets_no_sync(aid, D.C2CB, offs, outs[:2], y, z)
    with pytest.raises(IndexError):
        w.c2c_plan_array_size(5)
    assert w.c2c_plan_array_first_handle_token(aid) != w.c2c_plan_array_last_handle_token(aid)
    assert w.c2c_plan_array_first_stream_token(aid) == streams[0].cuda_stream
    assert w.c2c_plan_array_last_stream_token(aid) == streams[2].cuda_stream
    with pytest.raises(NotImplementedError):
        w.make_minimal_reference_c2c_plan_array_1D_diagnostic(n, n, 1, n, n, 1, n, batch, offs)
    w.release()


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_r2c_c2r_plan_bundle(gpu, dt):
    import torch
    import fftm_amd
    real = torch.float32 if dt == "float32" else torch.float64
    npt = np.float32 if dt == "float32" else np.float64
    w = fftm_amd.fft_wrap_many(real)
    n0, n1, batch, lanes = 32, 48, 2, 3
    nh = n1 // 2 + 1
    args = (n0, n1, n0, n1, 1, n0 * n1, n0, nh, 1, n0 * nh, batch,   # forward: real -> half spectrum
            n0, nh, 1, n0 * nh, n0, n1, 1, n0 * n1, batch)           # inverse: half spectrum -> real
    with pytest.raises(ValueError):
        w.make_r2c_c2r_plan_bundle_2D(*args, 0)
    bid = w.make_r2c_c2r_plan_bundle_2D(*args, lanes)
    assert w.r2c_c2r_plan_bundle_lane_count(bid) == lanes
    ws = w.r2c_c2r_plan_bundle_work_size(bid)
    assert ws % 256 == 0  # lanes at a 256-byte aligned stride
    w.activate()
    assert w.get_work_size() >= ws
    with pytest.raises(fftm_amd.LogicError):
        w.make_r2c_c2r_plan_bundle_2D(*args, lanes)
    rl, cl = batch * n0 * n1, batch * n0 * nh  # elements per lane
    xh = np.random.default_rng(13).uniform(-1, 1, lanes * rl).astype(npt)
    x = to_gpu(xh)
    X = torch.zeros(lanes * cl, dtype=torch.complex64 if dt == "float32" else torch.complex128, device="cuda")
    y = torch.zeros_like(x)
    torch.cuda.synchronize()
    for lane in range(lanes):
        w.exec_r2c_c2r_plan_bundle_forward_lane_no_sync(bid, lane, lane * rl, lane * cl, x, X)
    w.synchronize_r2c_c2r_plan_bundle(bid)
    Xh = from_gpu(X)
    for lane in range(lanes):
        ref = np.fft.rfft2(xh[lane * rl:(lane + 1) * rl].astype(np.float64).reshape(batch, n0, n1)).ravel()
        check_close(Xh[lane * cl:(lane + 1) * cl], ref, npt, f"bundle lane {lane} forward")
    for lane in range(lanes):
        w.exec_r2c_c2r_plan_bundle_inverse_lane_no_sync(bid, lane, lane * cl, lane * rl, X, y)
        w.synchronize_r2c_c2r_plan_bundle_lane(bid, lane)
        assert w.r2c_c2r_plan_bundle_lane_ready(bid, lane)
    check_close(from_gpu(y) / (n0 * n1), xh, npt, "bundle round trip")
    with pytest.raises(IndexError):
        w.exec_r2c_c2r_plan_bundle_forward_lane_no_sync(bid, lanes, 0, 0, x, X)
    with pytest.raises(IndexError):
        w.r2c_c2r_plan_bundle_lane_count(4)
    w.release()
