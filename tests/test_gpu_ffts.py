"""GPU parity of the ffts frontend (3D and the 4D strategies) against the FP64 oracle,
and the reference's own versioned checks (v1 determinism, v3 round trip, v4 periodic
Laplacian) -- source/tests/detail/versioned_ffts_{3d,4d}_test_driver.h.
"""

import math

import numpy as np
import pytest

import oracle
from helpers import check_close, from_gpu, roundtrip_tol, rms_diff, to_gpu, torch_dtypes

pytestmark = pytest.mark.gpu

DT = [np.float32, np.float64]
EXECUTIONS = ["fused", "staged"]


def _ffts(dt, dims, strategy=None, execution="fused", destructive=False):
    import fftm_amd
    f = fftm_amd.ffts("float32" if dt == np.float32 else "float64", strategy_4d=strategy)
    opts = fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="", execution=execution,
                                      destructive_backward=destructive)
    f.init(*dims, options=opts)
    return f


def _strategy(name):
    import fftm_amd
    return {"pencil": fftm_amd.strategy_4d_pencil_pencil(), "slab": fftm_amd.strategy_4d_slab_slab(),
            "xzwy": fftm_amd.strategy_4d_slab_native_xzwy()}[name]


@pytest.mark.parametrize("execution", EXECUTIONS)
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("dims", [(12, 10, 8), (16, 16, 16), (64, 32, 48), (128, 128, 128), (9, 15, 6),
                                  (96, 160, 384), (480, 36, 100)])
def test_ffts_3d_parity(gpu, dt, dims, execution):
    nx, ny, nz = dims
    print(f"case ffts3d {execution} dt={np.dtype(dt).name} dims={dims}", flush=True)
    f = _ffts(dt, dims, execution=execution)
    x = oracle.fill_random_real_3d(nx, ny, nz, 0x1234ABCDEF987654, dt)
    u = to_gpu(x)
    X = f.complex_array()
    f.forward(u, X)
    Xo = oracle.ffts3d_forward(x)
    check_close(from_gpu(X), Xo, dt, f"ffts3d forward {dims} {execution}")
    y = f.real_array()
    f.backward(X, y)
    check_close(from_gpu(y), oracle.ffts3d_backward(Xo, nz), dt, f"ffts3d backward {dims} {execution}")
    rt = rms_diff(from_gpu(y) / (nx * ny * nz), x)
    print(f"[parity] ffts3d roundtrip rms={rt:.3e}")
    assert rt <= roundtrip_tol(dt)
    f.close()


@pytest.mark.parametrize("execution", EXECUTIONS)
@pytest.mark.parametrize("strategy", ["pencil", "slab", "xzwy"])
@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("dims", [(12, 10, 8, 10), (16, 16, 16, 16), (8, 12, 16, 6), (32, 16, 8, 24), (48, 20, 24, 40)])
def test_ffts_4d_parity(gpu, dt, dims, strategy, execution):
    if strategy == "xzwy" and execution == "staged":
        pytest.skip("native_xzwy exists only as a fused transform")
    nx, ny, nz, nw = dims
    print(f"case ffts4d {strategy} {execution} dt={np.dtype(dt).name} dims={dims}", flush=True)
    f = _ffts(dt, dims, _strategy(strategy), execution)
    u_log = oracle.fill_random_real_4d_logical(nx, ny, nz, nw, 0x1234ABCDEF987654, dt)
    if strategy == "pencil":
        real_mem = oracle.to_memory(u_log, oracle.A0123)     # [w][z][y][x]
        Xo = oracle.ffts4d_forward(real_mem, "pencil", nx, ny, nz, nw)
    elif strategy == "slab":
        real_mem = oracle.to_memory(u_log, oracle.A3210)     # [x][y][z][w]
        Xo = oracle.ffts4d_forward(real_mem, "slab", nx, ny, nz, nw)
    else:
        real_mem = oracle.to_memory(u_log, oracle.A3210)
        Xo = oracle.native_xzwy_forward(real_mem, nx, ny, nz, nw)
    u = to_gpu(real_mem)
    X = f.complex_array()
    f.forward(u, X)
    check_close(from_gpu(X), Xo, dt, f"ffts4d {strategy} forward {dims} {execution}")
    Xkeep = X.clone()
    y = f.real_array()
    f.backward(X, y)
    if strategy == "pencil":
        yo = oracle.ffts4d_backward(Xo, "pencil", nx, ny, nz, nw)
    elif strategy == "slab":
        yo = oracle.ffts4d_backward(Xo, "slab", nx, ny, nz, nw)
    else:
        yo = oracle.native_xzwy_backward(Xo, nx, ny, nz, nw)
    check_close(from_gpu(y), yo, dt, f"ffts4d {strategy} backward {dims} {execution}")
    if execution == "fused":
        assert np.array_equal(from_gpu(X), from_gpu(Xkeep)), "fused backward keeps its input"
    rt = rms_diff(from_gpu(y) / (nx * ny * nz * nw), real_mem)
    assert rt <= roundtrip_tol(dt)
    f.close()


@pytest.mark.parametrize("dt", DT)
def test_ffts_4d_destructive_backward(gpu, dt):
    import fftm_amd
    dims = (16, 8, 8, 16)
    f = _ffts(dt, dims, fftm_amd.strategy_4d_pencil_pencil(), "fused", destructive=True)
    assert f._work is None
    x = oracle.to_memory(oracle.fill_random_real_4d_logical(*dims, 99, dt), oracle.A0123)
    X = f.complex_array()
    f.forward(to_gpu(x), X)
    y = f.real_array()
    f.backward(X, y)
    assert rms_diff(from_gpu(y) / np.prod(dims), x) <= roundtrip_tol(dt)


@pytest.mark.parametrize("dt", DT)
def test_ffts_v1_determinism(gpu, dt):
    """v1: two instances on the same analytic input agree (versioned_ffts_3d_test_driver.h:333-427)."""
    nx, ny, nz = 24, 20, 16
    h = [2 * math.pi / n for n in (nx, ny, nz)]
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    x, y, z = h[0] * i, h[1] * j, h[2] * k
    u = (np.sin(x) + 0.5 * np.cos(y) - 0.25 * np.sin(2 * z) + 0.125 * np.sin(x + y - z)).astype(dt)
    a, b = _ffts(dt, (nx, ny, nz)), _ffts(dt, (nx, ny, nz))
    Xa, Xb = a.complex_array(), b.complex_array()
    a.forward(to_gpu(u), Xa)
    b.forward(to_gpu(u), Xb)
    assert np.array_equal(from_gpu(Xa), from_gpu(Xb))


@pytest.mark.parametrize("dt", DT)
@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 24, 20)])
def test_ffts_v4_periodic_laplacian(gpu, dt, dims):
    """v4: spectral solve of Laplace(u) = -3 sin x sin y sin z, u and grad u vs exact
    (versioned_ffts_3d_test_driver.h:128-210, 560-713): max(L2, H1) <= 10 eps scaled."""
    import torch
    nx, ny, nz = dims
    L = 2 * math.pi
    hx, hy, hz = L / nx, L / ny, L / nz
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    x, y, z = hx * i, hy * j, hz * k
    exact = np.sin(x) * np.sin(y) * np.sin(z)
    rhs = -3.0 * exact
    f = _ffts(dt, dims)
    rt, ct = torch_dtypes(dt)
    R = f.complex_array()
    f.forward(to_gpu(rhs.astype(dt)), R)
    kx = torch.fft.fftfreq(nx, d=1.0 / nx).to(rt).cuda()
    ky = torch.fft.fftfreq(ny, d=1.0 / ny).to(rt).cuda()
    kz = torch.arange(nz // 2 + 1, dtype=rt).cuda()
    # k = i <= n/2 ? i : i - n  (Nyquist positive, the driver's convention)
    if nx % 2 == 0:
        kx[nx // 2] = nx // 2
    if ny % 2 == 0:
        ky[ny // 2] = ny // 2
    K2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    scale = torch.where(K2 == 0, torch.zeros_like(K2), -1.0 / torch.where(K2 == 0, torch.ones_like(K2), K2))
    S = (R * scale).to(ct)
    norm = 1.0 / (nx * ny * nz)
    sol = f.real_array()
    f.backward(S.clone(), sol)
    u = from_gpu(sol) * norm
    grads = []
    for kk in (kx[:, None, None], ky[None, :, None], kz[None, None, :]):
        D = (S * (1j * kk)).to(ct)
        g = f.real_array()
        f.backward(D, g)
        grads.append(from_gpu(g) * norm)
    ex_d = [np.cos(x) * np.sin(y) * np.sin(z), np.sin(x) * np.cos(y) * np.sin(z), np.sin(x) * np.sin(y) * np.cos(z)]
    cell = hx * hy * hz
    l2 = math.sqrt(np.sum((u - exact) ** 2) * cell)
    h1 = math.sqrt(sum(np.sum((g - e) ** 2) for g, e in zip(grads, ex_d)) * cell)
    print(f"[parity] v4 laplacian {dims} {np.dtype(dt).name}: L2={l2:.3e} H1={h1:.3e}")
    bound = 100 * float(np.finfo(dt).eps)
    assert max(l2, h1) <= bound


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_transpose_4d_integer_coded(gpu, dt):
    """The stage permutations of direct_transpose_4d (direct_transpose_4d.h:42-120) and the
    xzwy reorder, on integer-coded data: exact equality with numpy.transpose."""
    import itertools
    import fftm_amd
    import torch
    dims = (5, 12, 33, 40)
    src = torch.arange(int(np.prod(dims)), dtype=torch.float64 if dt == np.float64 else torch.float32).reshape(dims)
    cpl = torch.complex(src, -src)
    for t in (src, cpl):
        tg = t.cuda()
        for perm in itertools.permutations(range(4)):
            dst = torch.empty(tuple(dims[p] for p in perm), dtype=t.dtype, device="cuda")
            fftm_amd.transpose_4d(tg, dst, perm, sync=True)
            assert torch.equal(dst.cpu(), t.permute(perm).contiguous()), f"perm {perm} {t.dtype}"
