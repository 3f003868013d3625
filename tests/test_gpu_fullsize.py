"""The benchmark workloads at their full sizes (512^3 FP32 3D, 64^4 FP32 4D pencil), checked
through properties that need no CPU transform of the whole field:

* the closed-form sparse spectrum of the reference's analytic test field
  (versioned_ffts_3d_test_driver.h:62-67, versioned_ffts_4d_test_driver.h:64-70),
  compared at every bin;
* the round trip C2R(R2C(u)) = N u, to the reference's benchmark epsilon 10 eps(T)
  (fft_benchmark_common.h:28-32);
* Parseval with Hermitian weights on the halved axis, and linearity;
* the tuned variants giving bitwise the same spectrum as the defaults (within the FP32
  tolerance where the tuner picks another radix schedule)."""

import math

import pytest

pytestmark = pytest.mark.gpu

EPS32 = 2.0 ** -23


def _grid(n, dev):
    import torch
    return torch.arange(n, device=dev, dtype=torch.float64) * (2 * math.pi / n)


def _sample_3d(nx, ny, nz):
    """sin x + cos(y)/2 - sin(2z)/4 + sin(x + y - z)/8 on [x][y][z], FP32."""
    x = _grid(nx, "cuda")[:, None, None]
    y = _grid(ny, "cuda")[None, :, None]
    z = _grid(nz, "cuda")[None, None, :]
    return (x.sin() + 0.5 * y.cos() - 0.25 * (2 * z).sin() + 0.125 * (x + y - z).sin()).float()


def _max_dev(X, bins):
    """max |X - E| where E is zero except at `bins` ({index tuple: value})."""
    D = X.clone()
    for idx, v in bins.items():
        D[idx] -= v
    return float(D.abs().max())


def test_512_cubed_fp32_properties(gpu):
    import torch
    import fftm_amd
    n = 512
    N = n ** 3
    f = fftm_amd.ffts("float32")
    f.init(n, n, n, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    print("case fullsize 512^3 fp32", flush=True)

    # closed-form spectrum: six nonzero bins, every other one zero
    u = _sample_3d(n, n, n)
    X = f.complex_array()
    f.forward(u, X)
    bins = {(1, 0, 0): -0.5j * N, (n - 1, 0, 0): 0.5j * N, (0, 1, 0): 0.25 * N, (0, n - 1, 0): 0.25 * N,
            (0, 0, 2): 0.125j * N, (n - 1, n - 1, 1): 1j * N / 16}
    dev = _max_dev(X, bins)
    print(f"closed-form spectrum: max |X - E| / N = {dev / N:.3e}", flush=True)
    assert dev <= 1e-6 * N
    del u

    # round trip, Parseval and linearity on uniform(-1, 1) fields
    g = torch.Generator(device="cuda").manual_seed(2024)
    a = torch.rand((n, n, n), generator=g, device="cuda") * 2 - 1
    b = torch.rand((n, n, n), generator=g, device="cuda") * 2 - 1
    A, B = f.complex_array(), f.complex_array()
    f.forward(a, A)
    f.forward(b, B)
    y = f.real_array()
    f.backward(A, y)
    rms = float(torch.sqrt(torch.mean((y.double() / N - a.double()) ** 2)))
    print(f"round trip RMS {rms:.3e}", flush=True)
    assert rms <= 10 * EPS32
    w = torch.full((n // 2 + 1,), 2.0, device="cuda", dtype=torch.float64)
    w[0] = w[-1] = 1.0
    lhs = float(torch.sum(a.double() ** 2)) * N
    rhs = float(torch.sum((A.abs().double() ** 2) * w))
    print(f"Parseval rel {abs(lhs - rhs) / lhs:.3e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * lhs
    C = f.complex_array()
    f.forward(2.5 * a - 0.75 * b, C)
    lin = float((C - (2.5 * A - 0.75 * B)).abs().max() / C.abs().max())
    print(f"linearity rel-max {lin:.3e}", flush=True)
    assert lin <= 1e-5
    del b, B, C, y

    # tuned variants: bitwise the same spectrum in the default radix schedules (variant
    # numbers below 32); the alternative schedules within the FP32 parity tolerance
    chosen = f.autotune(warmup=1, iterations=3)
    A2 = f.complex_array()
    f.forward(a, A2)
    print(f"tuned variants {chosen}", flush=True)
    if all(v < 32 for v in chosen.values()):
        assert torch.equal(A2, A)
    else:
        assert float((A2 - A).abs().max() / A.abs().max()) <= 1e-5
    f.close()


def test_64_fourth_fp32_pencil_properties(gpu):
    import torch
    import fftm_amd
    n = 64
    N = n ** 4
    f = fftm_amd.ffts("float32", strategy_4d=fftm_amd.strategy_4d_pencil_pencil())
    f.init(n, n, n, n, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    print("case fullsize 64^4 fp32 pencil", flush=True)
    # real [w][z][y][x]; spectrum [y][z][w/2+1][x]
    x = _grid(n, "cuda")[None, None, None, :]
    y = _grid(n, "cuda")[None, None, :, None]
    z = _grid(n, "cuda")[None, :, None, None]
    w = _grid(n, "cuda")[:, None, None, None]
    u = (x.sin() + 0.5 * y.cos() - 0.25 * (2 * z).sin() + 0.125 * w.cos() + 0.0625 * (x + y - z + w).sin()).float()
    X = f.complex_array()
    f.forward(u, X)
    # logical bins (x, y, z, w) of versioned_ffts_4d_test_driver.h:64-70, stored at [y][z][w][x]
    logical = {(1, 0, 0, 0): -0.5j * N, (n - 1, 0, 0, 0): 0.5j * N, (0, 1, 0, 0): 0.25 * N,
               (0, n - 1, 0, 0): 0.25 * N, (0, 0, 2, 0): 0.125j * N, (0, 0, n - 2, 0): -0.125j * N,
               (0, 0, 0, 1): N / 16, (1, 1, n - 1, 1): -1j * N / 32}
    bins = {(ky, kz, kw, kx): v for (kx, ky, kz, kw), v in logical.items()}
    dev = _max_dev(X, bins)
    print(f"closed-form spectrum: max |X - E| / N = {dev / N:.3e}", flush=True)
    assert dev <= 1e-6 * N
    yb = f.real_array()
    f.backward(X, yb)
    rms = float(torch.sqrt(torch.mean((yb.double() / N - u.double()) ** 2)))
    print(f"round trip RMS {rms:.3e}", flush=True)
    assert rms <= 10 * EPS32
    f.close()
