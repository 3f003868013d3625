"""GPU parity of single plans (the fft<T,D> boundary) against the FP64 oracle.

Each case restates a plan descriptor that ffts (or its callers) builds
(source/ffts.hpp:659-776, source/tests/test_fft_3D.cu:46-51) and compares the
C-ABI execution with ``oracle.plan_exec`` on the same seeded input.
"""

import numpy as np
import pytest

import oracle
from helpers import check_close, from_gpu, to_gpu, torch_dtypes

pytestmark = pytest.mark.gpu

DT = [np.float32, np.float64]


def _plan(dt, D, n, inembed, istride, idist, onembed, ostride, odist, batch, flags=0):
    import fftm_amd
    real = torch_dtypes(dt)[0]
    return fftm_amd.fft(real, D, n, inembed, istride, idist, onembed, ostride, odist, batch, flags=flags)


def _run(plan, inp, out_shape, out_dtype, work=None):
    import torch
    out = torch.zeros(out_shape, dtype=out_dtype, device="cuda")
    if plan.get_work_size():
        work = torch.empty(plan.get_work_size(), dtype=torch.uint8, device="cuda")
        plan.set_work_area(work)
    plan.exec(inp, out)
    return out


def _rng_real(shape, seed):
    return np.random.default_rng(seed).uniform(-1, 1, size=shape)


def _rng_cplx(shape, seed):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, size=shape) + 1j * g.uniform(-1, 1, size=shape)


SIZES_3D = [(12, 10, 8), (8, 6, 10), (16, 16, 16), (7, 5, 9), (32, 64, 16), (128, 128, 128), (2, 3, 4), (1, 1, 8)]


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("shape", SIZES_3D)
def test_plan_3d_r2c_c2r(gpu, dt, shape):
    """ffts "forward_3d"/"inverse_3d" descriptors (ffts.hpp:680-694)."""
    import fftm_amd
    nx, ny, nz = shape
    nzh = nz // 2 + 1
    rt, ct = torch_dtypes(dt)
    print(f"case plan_3d dt={np.dtype(dt).name} shape={shape}", flush=True)
    x = oracle.fill_random_real_3d(nx, ny, nz, 0x1234ABCDEF987654, dt)
    fwd = _plan(dt, fftm_amd.direction.R2C, [nx, ny, nz], [nx, ny, nz], 1, nx * ny * nz, [nx, ny, nzh], 1,
                nx * ny * nzh, 1)
    X = _run(fwd, to_gpu(x), (nx, ny, nzh), ct)
    Xo = oracle.ffts3d_forward(x)
    check_close(from_gpu(X), Xo, dt, f"R2C {shape} {np.dtype(dt).name}")
    inv = _plan(dt, fftm_amd.direction.C2R, [nx, ny, nz], [nx, ny, nzh], 1, nx * ny * nzh, [nx, ny, nz], 1,
                nx * ny * nz, 1)
    Xin = X.clone()
    y = _run(inv, X, (nx, ny, nz), rt)
    assert np.array_equal(from_gpu(X), from_gpu(Xin)), "C2R must keep its input when a work area is bound"
    yo = oracle.ffts3d_backward(Xo, nz)
    check_close(from_gpu(y), yo, dt, f"C2R {shape} {np.dtype(dt).name}")


@pytest.mark.parametrize("dt", DT)
def test_plan_c2r_non_hermitian_input(gpu, dt):
    """C2R of an arbitrary (non-Hermitian) half spectrum follows irfftn semantics."""
    import fftm_amd
    nx, ny, nz = 6, 8, 10
    nzh = nz // 2 + 1
    X = _rng_cplx((nx, ny, nzh), 7)
    inv = _plan(dt, fftm_amd.direction.C2R, [nx, ny, nz], [nx, ny, nzh], 1, nx * ny * nzh, [nx, ny, nz], 1,
                nx * ny * nz, 1)
    ct = torch_dtypes(dt)[1]
    y = _run(inv, to_gpu(X.astype(np.complex64 if dt == np.float32 else np.complex128)), (nx, ny, nz),
             torch_dtypes(dt)[0])
    ref = np.fft.irfftn(X, s=(nx, ny, nz), axes=(0, 1, 2)) * nx * ny * nz
    check_close(from_gpu(y), ref, dt, "C2R non-Hermitian vs irfftn")
    assert ct is not None


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("dims", [(16, 8, 12, 16), (64, 64, 64, 64), (12, 10, 8, 10)])
def test_plan_4d_pencil_descriptors(gpu, dt, dims):
    """The eight 1D plans of the 4D pencil strategy (ffts.hpp:697-742)."""
    import fftm_amd
    import torch
    nx, ny, nz, nw = dims
    if nx * ny * nz * nw > 2 ** 22 and dt == np.float64:
        pytest.skip("oracle time")
    nwh = nw // 2 + 1
    D = fftm_amd.direction
    rt, ct = torch_dtypes(dt)
    w_stride, z_stride, y_stride = nx * ny * nz, nx * ny * nwh, nx * nz * nwh
    print(f"case 4d-pencil plans dt={np.dtype(dt).name} dims={dims}", flush=True)
    # forward_w: R2C along w, strided
    x = _rng_real(nw * nz * ny * nx, 11)
    p = _plan(dt, D.R2C, [nw], [1], w_stride, 1, [1], w_stride, 1, nx * ny * nz)
    out = _run(p, to_gpu(x.astype(dt)), (nwh * w_stride,), ct)
    ref = oracle.plan_exec(oracle.R2C, [nw], [1], w_stride, 1, [1], w_stride, 1, nx * ny * nz, x.astype(dt))
    check_close(from_gpu(out), ref, dt, "forward_w")
    # inverse_w
    p = _plan(dt, D.C2R, [nw], [1], w_stride, 1, [1], w_stride, 1, nx * ny * nz)
    y = _run(p, out, (nw * w_stride,), rt)
    ref2 = oracle.plan_exec(oracle.C2R, [nw], [1], w_stride, 1, [1], w_stride, 1, nx * ny * nz, from_gpu(out))
    check_close(from_gpu(y), ref2, dt, "inverse_w")
    # forward/inverse z, y (strided C2C) and x (contiguous C2C)
    c = _rng_cplx(nx * ny * nz * nwh, 12)
    cdt = np.complex64 if dt == np.float32 else np.complex128
    for name, n, stride, dist, batch in [("z", nz, z_stride, 1, nx * ny * nwh), ("y", ny, y_stride, 1, nx * nz * nwh),
                                         ("x", nx, 1, nx, ny * nz * nwh)]:
        for d, od in ((D.C2CF, oracle.C2CF), (D.C2CB, oracle.C2CB)):
            p = _plan(dt, d, [n], [1], stride, dist, [1], stride, dist, batch)
            t = to_gpu(c.astype(cdt))
            p.exec(t, t)  # in place, as ffts runs them
            ref = oracle.plan_exec(od, [n], [1], stride, dist, [1], stride, dist, batch, c.astype(cdt))
            check_close(from_gpu(t), ref, dt, f"{d.name}_{name}")
    torch.cuda.synchronize()


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("dims", [(16, 8, 12, 16), (32, 32, 32, 32), (12, 10, 8, 10)])
def test_plan_4d_slab_descriptors(gpu, dt, dims):
    """The four 2D plans of the 4D slab strategy (ffts.hpp:744-776)."""
    import fftm_amd
    nx, ny, nz, nw = dims
    nwh = nw // 2 + 1
    D = fftm_amd.direction
    rt, ct = torch_dtypes(dt)
    cdt = np.complex64 if dt == np.float32 else np.complex128
    print(f"case 4d-slab plans dt={np.dtype(dt).name} dims={dims}", flush=True)
    x = _rng_real(nx * ny * nz * nw, 21).astype(dt)
    p = _plan(dt, D.R2C, [nz, nw], [nz, nw], 1, nz * nw, [nz, nwh], 1, nz * nwh, nx * ny)
    X = _run(p, to_gpu(x), (nx * ny * nz * nwh,), ct)
    ref = oracle.plan_exec(oracle.R2C, [nz, nw], [nz, nw], 1, nz * nw, [nz, nwh], 1, nz * nwh, nx * ny, x)
    check_close(from_gpu(X), ref, dt, "forward_zw")
    p = _plan(dt, D.C2R, [nz, nw], [nz, nwh], 1, nz * nwh, [nz, nw], 1, nz * nw, nx * ny)
    y = _run(p, X, (nx * ny * nz * nw,), rt)
    ref = oracle.plan_exec(oracle.C2R, [nz, nw], [nz, nwh], 1, nz * nwh, [nz, nw], 1, nz * nw, nx * ny, from_gpu(X))
    check_close(from_gpu(y), ref, dt, "inverse_zw")
    s = nz * nwh
    c = _rng_cplx(nx * ny * nz * nwh, 22).astype(cdt)
    for d, od in ((D.C2CF, oracle.C2CF), (D.C2CB, oracle.C2CB)):
        p = _plan(dt, d, [nx, ny], [nx, ny], s, 1, [nx, ny], s, 1, s)
        t = to_gpu(c)
        p.exec(t, t)
        ref = oracle.plan_exec(od, [nx, ny], [nx, ny], s, 1, [nx, ny], s, 1, s, c)
        check_close(from_gpu(t), ref, dt, f"{d.name}_xy")


@pytest.mark.parametrize("dt", DT)
def test_plan_c2c_exec_direction(gpu, dt):
    """A C2CF plan executes C2CB through exec_direction (hipfft_wrap.h:686-712)."""
    import fftm_amd
    D = fftm_amd.direction
    cdt = np.complex64 if dt == np.float32 else np.complex128
    c = _rng_cplx((24, 40), 5).astype(cdt)
    p = _plan(dt, D.C2CF, [24, 40], [24, 40], 1, 24 * 40, [24, 40], 1, 24 * 40, 1)
    t = to_gpu(c)
    o = to_gpu(np.zeros_like(c))
    p.exec_direction(D.C2CB, t, o)
    check_close(from_gpu(o), np.fft.ifftn(c) * c.size, dt, "exec_direction C2CB")
    p.exec(t, o)
    check_close(from_gpu(o), np.fft.fftn(c), dt, "exec C2CF")
    with pytest.raises(fftm_amd.LogicError):
        p.exec_direction(D.R2C, t, o)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 3, 5, 6, 7, 10, 12, 14, 22,
                               26, 30, 36, 100, 1000, 1344, 2100, 48, 96, 160, 384, 480, 896, 1536, 1680, 1920,
                               2000])
def test_plan_1d_lengths(gpu, dt, n):
    """Every compiled power-of-two length, compiled 3/5/7-smooth lengths and any-radix
    lengths (2,3,5,7,11,13), both mappings, R2C/C2R pairing with an odd batch (phantom
    column)."""
    import fftm_amd
    D = fftm_amd.direction
    rt, ct = torch_dtypes(dt)
    cdt = np.complex64 if dt == np.float32 else np.complex128
    nh = n // 2 + 1
    batch = 7
    x = _rng_real((batch, n), n).astype(dt)
    # contiguous lines
    p = _plan(dt, D.R2C, [n], [n], 1, n, [nh], 1, nh, batch)
    X = _run(p, to_gpu(x), (batch, nh), ct)
    check_close(from_gpu(X), np.fft.rfft(x.astype(np.float64), axis=1), dt, f"R2C n={n} contiguous")
    p = _plan(dt, D.C2R, [n], [nh], 1, nh, [n], 1, n, batch)
    y = _run(p, X, (batch, n), rt)
    check_close(from_gpu(y), np.fft.irfft(from_gpu(X).astype(np.complex128), n=n, axis=1) * n, dt,
                f"C2R n={n} contiguous")
    # strided lines (column mapping): batch contiguous, stride = batch
    xt = np.ascontiguousarray(x.T)
    p = _plan(dt, D.R2C, [n], [n], batch, 1, [nh], batch, 1, batch)
    X2 = _run(p, to_gpu(xt), (nh, batch), ct)
    check_close(from_gpu(X2), np.fft.rfft(x.astype(np.float64), axis=1).T, dt, f"R2C n={n} strided")
    c = _rng_cplx((batch, n), n + 1).astype(cdt)
    for d, f in ((D.C2CF, np.fft.fft), (D.C2CB, lambda a, axis: np.fft.ifft(a, axis=axis) * n)):
        p = _plan(dt, d, [n], [n], 1, n, [n], 1, n, batch)
        o = to_gpu(np.zeros_like(c))
        p.exec(to_gpu(c), o)
        check_close(from_gpu(o), f(c.astype(np.complex128), axis=1), dt, f"{d.name} n={n} contiguous")
        ct_ = np.ascontiguousarray(c.T)
        p = _plan(dt, d, [n], [n], batch, 1, [n], batch, 1, batch)
        o = to_gpu(np.zeros_like(ct_))
        p.exec(to_gpu(ct_), o)
        check_close(from_gpu(o), f(c.astype(np.complex128), axis=1).T, dt, f"{d.name} n={n} strided")


def test_plan_errors_on_gpu(gpu):
    """Error paths of exec that need a live plan."""
    import torch
    import fftm_amd
    D = fftm_amd.direction
    p = fftm_amd.fft(torch.float32, D.C2R, [8, 8], [8, 5], 1, 40, [8, 8], 1, 64, 1)
    assert p.get_work_size() == 40 * 8
    a = torch.zeros(40, dtype=torch.complex64, device="cuda")
    b = torch.zeros(64, dtype=torch.float32, device="cuda")
    with pytest.raises(fftm_amd.FftmMiError) as e:
        p.exec(a, b)  # no work area bound
    assert e.value.status == 13
    q = fftm_amd.fft(torch.float32, D.C2R, [8, 8], [8, 5], 1, 40, [8, 8], 1, 64, 1, flags=1)
    assert q.get_work_size() == 0
    q.exec(a, b)


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("n", [32, 64, 512])
@pytest.mark.parametrize("layout", ["odd_dist", "offset_base"])
def test_plan_paired_access_fallbacks(gpu, dt, n, layout):
    """Contiguous lines move two consecutive elements per load / store when every line
    start is even and the base is 16-byte aligned; odd line pitches and an offset base
    take the element-wise path.  Both give the same transform."""
    import torch
    import fftm_amd
    D = fftm_amd.direction
    rt, ct = torch_dtypes(dt)
    cdt = np.complex64 if dt == np.float32 else np.complex128
    nh, batch = n // 2 + 1, 5
    print(f"case paired-access dt={np.dtype(dt).name} n={n} layout={layout}", flush=True)
    pitch = n + 1 if layout == "odd_dist" else n
    off = 1 if layout == "offset_base" else 0
    x = _rng_real((batch, pitch), n + 3).astype(dt)
    buf = to_gpu(np.concatenate([np.zeros(off, dt), x.ravel()]))
    p = _plan(dt, D.R2C, [n], [pitch], 1, pitch, [nh], 1, nh, batch)
    X = _run(p, buf[off:], (batch, nh), ct)
    check_close(from_gpu(X), np.fft.rfft(x[:, :n].astype(np.float64), axis=1), dt, f"R2C {layout} n={n}")
    # C2R into an odd-pitch / offset real output
    yb = torch.zeros(off + batch * pitch, dtype=rt, device="cuda")
    q = _plan(dt, D.C2R, [n], [nh], 1, nh, [pitch], 1, pitch, batch)
    work = torch.empty(max(q.get_work_size(), 1), dtype=torch.uint8, device="cuda")
    q.set_work_area(work)
    q.exec(X, yb[off:])
    y = from_gpu(yb[off:]).reshape(batch, pitch)[:, :n]
    check_close(y, np.fft.irfft(from_gpu(X).astype(np.complex128), n=n, axis=1) * n, dt, f"C2R {layout} n={n}")
    # C2C along contiguous lines with the same pitch / offset
    c = _rng_cplx((batch, pitch), n + 4).astype(cdt)
    cb = to_gpu(np.concatenate([np.zeros(off, cdt), c.ravel()]))
    ob = torch.zeros(off + batch * pitch, dtype=ct, device="cuda")
    r = _plan(dt, D.C2CF, [n], [pitch], 1, pitch, [pitch], 1, pitch, batch)
    r.exec(cb[off:], ob[off:])
    o = from_gpu(ob[off:]).reshape(batch, pitch)[:, :n]
    check_close(o, np.fft.fft(c[:, :n].astype(np.complex128), axis=1), dt, f"C2C {layout} n={n}")
