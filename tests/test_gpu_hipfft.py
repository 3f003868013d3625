"""Our transforms against the reference's own backend on the same inputs: the system
hipFFT (rocFFT), reached through torch.fft on ROCm.  This is the comparison the
reference's ffts makes when built for HIP (hipfft_wrap.h:303-417); the FP64 oracle
stays the accuracy reference, this pins "the same answer as the reference's library"."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    import torch
    a = a.to(torch.complex128) if a.is_complex() else a.double()
    b = b.to(torch.complex128) if b.is_complex() else b.double()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("dims", [(64, 48, 80), (128, 128, 128), (96, 160, 384), (12, 10, 8)])
def test_3d_vs_hipfft(gpu, dt, dims):
    import torch
    import fftm_amd
    f = fftm_amd.ffts(dt)
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    g = torch.Generator(device="cuda").manual_seed(7)
    u = torch.rand(dims, generator=g, device="cuda", dtype=f.real) * 2 - 1
    X = f.complex_array()
    f.forward(u, X)
    Xh = torch.fft.rfftn(u)
    tol = 2e-6 if dt == "float32" else 1e-13
    e_f = _rel(X, Xh)
    y = f.real_array()
    f.backward(X, y)
    yh = torch.fft.irfftn(X, s=dims) * float(np.prod(dims))
    e_b = _rel(y, yh)
    print(f"[hipfft] 3d {dims} {dt}: forward {e_f:.2e} backward {e_b:.2e}", flush=True)
    assert e_f < tol and e_b < tol
    f.close()


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_4d_pencil_vs_hipfft(gpu, dt):
    import torch
    import fftm_amd
    nx, ny, nz, nw = 32, 16, 24, 40
    f = fftm_amd.ffts(dt, strategy_4d=fftm_amd.strategy_4d_pencil_pencil())
    f.init(nx, ny, nz, nw, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    g = torch.Generator(device="cuda").manual_seed(11)
    mem = torch.rand((nw, nz, ny, nx), generator=g, device="cuda", dtype=f.real) * 2 - 1  # [w][z][y][x]
    X = f.complex_array()                                                                  # [y][z][wh][x]
    f.forward(mem, X)
    logical = mem.permute(3, 2, 1, 0)                                                      # (x, y, z, w)
    Xh = torch.fft.rfftn(logical).permute(1, 2, 3, 0)                                      # [y][z][wh][x]
    tol = 2e-6 if dt == "float32" else 1e-13
    e = _rel(X, Xh.contiguous())
    print(f"[hipfft] 4d pencil {dt}: forward {e:.2e}", flush=True)
    assert e < tol
    f.close()
