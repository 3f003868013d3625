"""The 4D Gaussian low-pass workload at its full sizes and the 1024^3 mixed-precision
workload, checked through properties that need no CPU transform of the whole field.

4D (64^4 and 128^4 FP32; pencil, slab and the native xzwy layout; ffts.hpp:697-776):
16 plane waves cos(k.x + phi) through the fused filter (forward, x exp(-|k|^2 / (2 sigma^2))
with sigma_k = 8, backward, 1/N) give exp(-|k|^2 / 128) cos(k.x + phi) exactly, so the
output is compared with that closed form everywhere; the round trip of uniform(-1, 1) noise
is held to 10 eps (fft_benchmark_common.h:28-32).

3D 1024^3 (ffts.hpp:677-695): the closed-form sparse spectrum of the reference's analytic
field (versioned_ffts_3d_test_driver.h:62-67), the round trip and Parseval in FP32; the
FP16-stored spectra path on a uniform(-1,1) + k^-5/3 random-phase field against the FP32
path at the FP16-storage tolerance (two roundings of 2^-11 forward, four for the round trip).
"""

import math

import pytest

pytestmark = pytest.mark.gpu

EPS32 = 2.0 ** -23
SIGMA = 8.0
# plane waves through the filter: FP32 transforms of O(1) data, ~log2 N roundings of 2^-24 per
# element relative to the field's largest value
TOL_WAVES = 2e-5
# FP16 storage against FP32
TOL_H16_FWD = 1.5e-3
TOL_H16_RT = 3e-3


def _grid(n):
    import torch
    return torch.arange(n, device="cuda", dtype=torch.float64) * (2 * math.pi / n)


def _plane_waves(n, layout, seed):
    """(field, expected filter output) in the layout's real memory order: pencil
    [w][z][y][x], slab / xzwy [x][y][z][w]; k uniform in [-12, 12]^4."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    g = _grid(n)
    ax = [g[:, None, None, None], g[None, :, None, None], g[None, None, :, None], g[None, None, None, :]]
    # logical axis a (x, y, z, w) sits at memory axis mem[a]
    mem = [3, 2, 1, 0] if layout == "pencil" else [0, 1, 2, 3]
    waves = torch.zeros((n,) * 4, device="cuda", dtype=torch.float32)
    expect = torch.zeros_like(waves)
    for _ in range(16):
        k = rng.integers(-12, 13, size=4)
        ph = float(rng.uniform(0, 2 * math.pi))
        arg = (float(k[0]) * ax[mem[0]] + float(k[1]) * ax[mem[1]]) + (float(k[2]) * ax[mem[2]] +
                                                                       float(k[3]) * ax[mem[3]] + ph)
        c = torch.cos(arg)
        waves += c.float()
        expect += (math.exp(-float(k @ k) / (2 * SIGMA ** 2)) * c).float()
        del arg, c
    return waves, expect


def _strategy(layout):
    import fftm_amd
    return {"pencil": fftm_amd.strategy_4d_pencil_pencil, "slab": fftm_amd.strategy_4d_slab_slab,
            "xzwy": fftm_amd.strategy_4d_slab_native_xzwy}[layout]()


@pytest.mark.parametrize("n,layout", [(64, "pencil"), (64, "slab"), (64, "xzwy"), (128, "pencil"),
                                      (128, "xzwy")])
def test_4d_gaussian_lowpass_fullsize(gpu, n, layout):
    import torch
    import fftm_amd
    N = n ** 4
    print(f"case fullsize 4d gaussian {n}^4 {layout}", flush=True)
    f = fftm_amd.ffts("float32", strategy_4d=_strategy(layout))
    f.init(n, n, n, n, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    waves, expect = _plane_waves(n, layout, seed=21 + n)
    assert tuple(waves.shape) == tuple(f.real_shape())
    y = f.real_array()
    f.filter(waves, y, "gaussian", scale=1.0 / N, sigma=SIGMA)
    torch.cuda.synchronize()
    err = float((y - expect).abs().max() / expect.abs().max())
    print(f"plane waves through the Gaussian filter: rel-max {err:.3e}", flush=True)
    assert err <= TOL_WAVES
    # the same filter as forward, multiply, backward: agrees to rounding
    X = f.complex_array()
    f.forward(waves, X)
    f.spectral_apply(X, X, "gaussian", scale=1.0 / N, sigma=SIGMA)
    y2 = f.real_array()
    f.backward(X, y2)
    d = float((y2 - expect).abs().max() / expect.abs().max())
    print(f"forward + multiply + backward: rel-max {d:.3e}", flush=True)
    assert d <= TOL_WAVES
    del waves, expect, y2
    # round trip of uniform(-1, 1) noise
    g = torch.Generator(device="cuda").manual_seed(7)
    u = torch.rand(tuple(f.real_shape()), generator=g, device="cuda") * 2 - 1
    f.forward(u, X)
    f.backward(X, y)
    rms = float(torch.sqrt(torch.mean((y.double() / N - u.double()) ** 2)))
    print(f"round trip RMS {rms:.3e}", flush=True)
    assert rms <= 10 * EPS32
    f.close()


def _sample_3d(n):
    """sin x + cos(y)/2 - sin(2z)/4 + sin(x + y - z)/8 on [x][y][z], FP32, built slab by slab."""
    import torch
    g = _grid(n)
    out = torch.empty((n, n, n), device="cuda", dtype=torch.float32)
    y = g[None, :, None]
    z = g[None, None, :]
    for x0 in range(0, n, 128):
        x = g[x0:x0 + 128, None, None]
        out[x0:x0 + 128] = (x.sin() + 0.5 * y.cos() - 0.25 * (2 * z).sin() + 0.125 * (x + y - z).sin()).float()
    return out


def test_1024_cubed_fp32_properties(gpu):
    import torch
    import fftm_amd
    n = 1024
    N = n ** 3
    print("case fullsize 1024^3 fp32", flush=True)
    f = fftm_amd.ffts("float32")
    f.init(n, n, n, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    u = _sample_3d(n)
    X = f.complex_array()
    f.forward(u, X)
    bins = {(1, 0, 0): -0.5j * N, (n - 1, 0, 0): 0.5j * N, (0, 1, 0): 0.25 * N, (0, n - 1, 0): 0.25 * N,
            (0, 0, 2): 0.125j * N, (n - 1, n - 1, 1): 1j * N / 16}
    for idx, v in bins.items():
        X[idx] -= v
    dev = float(X.abs().max())
    print(f"closed-form spectrum: max |X - E| / N = {dev / N:.3e}", flush=True)
    assert dev <= 1e-6 * N
    del u
    g = torch.Generator(device="cuda").manual_seed(2025)
    a = torch.rand((n, n, n), generator=g, device="cuda") * 2 - 1
    f.forward(a, X)
    y = f.real_array()
    f.backward(X, y)
    rms = float(torch.sqrt(torch.mean((y.double() / N - a.double()) ** 2)))
    print(f"round trip RMS {rms:.3e}", flush=True)
    assert rms <= 10 * EPS32
    w = torch.full((n // 2 + 1,), 2.0, device="cuda", dtype=torch.float64)
    w[0] = w[-1] = 1.0
    lhs = float(torch.sum(a.double() ** 2)) * N
    rhs = float(torch.sum((X.abs().double() ** 2) * w))
    print(f"Parseval rel {abs(lhs - rhs) / lhs:.3e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * lhs
    f.close()


def _kolmogorov_field(n, seed):
    """uniform(-1, 1) plus a k^-5/3 random-phase field of half its RMS (the mixed-precision
    workload's input), FP32 on the GPU; the random-phase part is made by torch.fft."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    u = torch.rand((n, n, n), generator=g, device="cuda", dtype=torch.float32) * 2 - 1
    k = torch.fft.fftfreq(n, d=1.0 / n, device="cuda")
    kz = torch.arange(n // 2 + 1, device="cuda", dtype=torch.float32)
    K = torch.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = torch.where(K > 0, K.clamp(min=1.0) ** (-5.0 / 3.0), torch.zeros_like(K))
    del K
    ph = torch.rand(amp.shape, generator=g, device="cuda") * (2 * math.pi)
    spec = torch.polar(amp, ph)
    del amp, ph
    v = torch.fft.irfftn(spec, s=(n, n, n))
    del spec
    v *= 0.5 / float(v.std())
    u += v
    del v
    torch.cuda.empty_cache()
    return u


def test_1024_cubed_fp16_storage_vs_fp32(gpu):
    import torch
    import fftm_amd
    n = 1024
    N = n ** 3
    print("case fullsize 1024^3 fp16 storage", flush=True)
    u = _kolmogorov_field(n, seed=5)
    res = {}
    for storage in ("fp32", "fp16"):
        f = fftm_amd.ffts("float32")
        f.init(n, n, n, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="",
                                                           spectrum_storage=storage))
        X = f.complex_array()
        f.forward(u, X)
        y = f.real_array()
        f.backward(X, y)
        res[storage] = (X, y)
        f.close()
    X32, y32 = res["fp32"]
    X16, y16 = res["fp16"]
    l2 = float(torch.linalg.vector_norm(X16 - X32) / torch.linalg.vector_norm(X32))
    mx = float((X16 - X32).abs().max() / X32.abs().max())
    print(f"FP16 storage vs FP32: forward rel-L2 {l2:.3e} rel-max {mx:.3e}", flush=True)
    assert l2 <= TOL_H16_FWD and mx <= 4 * TOL_H16_FWD
    rt = float(torch.linalg.vector_norm(y16 / N - u) / torch.linalg.vector_norm(u))
    rt32 = float(torch.linalg.vector_norm(y32 / N - u) / torch.linalg.vector_norm(u))
    print(f"round trip rel-L2: FP16 storage {rt:.3e}, FP32 {rt32:.3e}", flush=True)
    assert rt <= TOL_H16_RT and rt32 <= 10 * EPS32
