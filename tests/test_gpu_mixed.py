"""FP32 transforms with FP16-stored intermediate spectra (no reference counterpart).

The half spectrum between passes is stored as FP16 pairs with one power-of-two
exponent per stored line, chosen so that the line's largest component lands in
[2^14, 2^15).  Each storage rounds with unit roundoff 2^-11, so the forward spectrum
(two roundings) stays within a few 1e-4 relative L2 of the FP64 oracle; the plain
FP32 path is the accuracy reference.  Scaling the input by a power of two scales every
stored exponent by the same amount, so results scale bitwise exactly -- that pins the
exponent bookkeeping independently of the FP16 rounding."""

import numpy as np
import pytest

import oracle
from helpers import from_gpu

pytestmark = pytest.mark.gpu

# FP16 storage: 2 roundings (forward) / 4 (round trip) of unit roundoff 2^-11 each
TOL_FWD_L2 = 1.5e-3
TOL_RT_L2 = 3e-3


def _ffts(dims, storage="fp16"):
    import fftm_amd
    f = fftm_amd.ffts("float32")
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="",
                                                     spectrum_storage=storage))
    return f


@pytest.mark.parametrize("dims", [(128, 128, 128), (256, 128, 512), (128, 256, 256), (2048, 128, 128)])
def test_h16_matches_oracle(gpu, dims):
    import torch
    print(f"case h16 3d {dims}", flush=True)
    x = oracle.fill_random_real_3d(*dims, 0x1234ABCDEF987654, np.float32)
    f = _ffts(dims)
    u = torch.from_numpy(x).cuda()
    X = f.complex_array()
    f.forward(u, X)
    Xo = oracle.ffts3d_forward(x.astype(np.float64))
    l2, mx = oracle.rel_errors(from_gpu(X), Xo)
    print(f"[h16] fwd {dims}: rel_l2={l2:.3e} rel_max={mx:.3e}", flush=True)
    assert l2 < TOL_FWD_L2 and mx < 4 * TOL_FWD_L2
    y = f.real_array()
    f.backward(X, y)
    yr = from_gpu(y).astype(np.float64) / np.prod(dims)
    rt = float(np.linalg.norm(yr - x) / np.linalg.norm(x))
    print(f"[h16] round trip {dims}: rel_l2={rt:.3e}", flush=True)
    assert rt < TOL_RT_L2
    # the plain FP32 path on the same input is an order of magnitude more accurate
    g = _ffts(dims, "fp32")
    X32 = g.complex_array()
    g.forward(u, X32)
    l2_32, _ = oracle.rel_errors(from_gpu(X32), Xo)
    assert l2_32 < l2 / 10
    f.close()
    g.close()


@pytest.mark.parametrize("shift", [20, -30])
def test_h16_power_of_two_scaling_is_exact(gpu, shift):
    """x * 2^s gives spectrum * 2^s bitwise: the per-line exponents absorb the scale
    (2^-30 puts every value below the FP16 normal range without them)."""
    import torch
    dims = (128, 128, 256)
    x = oracle.fill_random_real_3d(*dims, 77, np.float32)
    f = _ffts(dims)
    u = torch.from_numpy(x).cuda()
    us = u * (2.0 ** shift)
    X, Xs = f.complex_array(), f.complex_array()
    f.forward(u, X)
    f.forward(us, Xs)
    assert torch.equal(Xs, X * (2.0 ** shift))
    y, ys = f.real_array(), f.real_array()
    f.backward(X, y)
    f.backward(Xs, ys)
    assert torch.equal(ys, y * (2.0 ** shift))
    f.close()


def test_h16_wide_dynamic_range_spectrum(gpu):
    """uniform(-1,1) plus a k^-5/3 random-phase field: the per-line exponents keep the
    small high-wavenumber lines in FP16's normal range."""
    import torch
    dims = (256, 256, 256)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, size=dims)
    kx = np.fft.fftfreq(dims[0]) * dims[0]
    kz = np.arange(dims[2] // 2 + 1)
    K = np.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = np.where(K > 0, np.power(np.maximum(K, 1.0), -5.0 / 3.0), 0.0)
    phase = np.exp(2j * np.pi * rng.uniform(size=amp.shape))
    x = x + np.fft.irfftn(amp * phase * 1e4, s=dims)
    x32 = x.astype(np.float32)
    f = _ffts(dims)
    X = f.complex_array()
    f.forward(torch.from_numpy(x32).cuda(), X)
    Xo = oracle.ffts3d_forward(x32.astype(np.float64))
    l2, mx = oracle.rel_errors(from_gpu(X), Xo)
    print(f"[h16] k^-5/3 field {dims}: rel_l2={l2:.3e} rel_max={mx:.3e}", flush=True)
    assert l2 < TOL_FWD_L2
    f.close()


def test_h16_variants_bitwise_and_autotune(gpu):
    import torch
    import fftm_amd
    from fftm_amd import _lib
    dims = (256, 256, 256)
    f = _ffts(dims)
    L = _lib.lib()
    u = torch.rand(f.real_shape(), dtype=torch.float32, device="cuda") * 2 - 1
    X0, y0 = f.complex_array(), f.real_array()
    f.forward(u, X0)
    f.backward(X0, y0)
    nf, nb = f.num_passes()
    shapes_ok = {s + 8 * o for s in (0, 2, 6, 7) for o in range(4)}
    for var in sorted(shapes_ok):
        for bw, n in ((0, nf), (1, nb)):
            for i in range(n):
                _lib.check(L.fftm_mi_ffts_set_pass_variant(f._handle, bw, i, var), "set_pass_variant")
        X, y = f.complex_array(), f.real_array()
        f.forward(u, X)
        f.backward(X, y)
        assert torch.equal(X, X0) and torch.equal(y, y0), f"variant {var}"
    assert L.fftm_mi_ffts_set_pass_variant(f._handle, 0, 1, 3) == 16
    assert L.fftm_mi_ffts_set_pass_variant(f._handle, 0, 1, 11) == 16
    chosen = f.autotune(warmup=1, iterations=3)
    assert set(chosen.values()) <= shapes_ok
    X, y = f.complex_array(), f.real_array()
    f.forward(u, X)
    f.backward(X, y)
    assert torch.equal(X, X0) and torch.equal(y, y0)
    f.close()


@pytest.mark.parametrize("dims", [(128, 128, 128), (256, 128, 512)])
def test_h16_line_scales_differ(gpu, dims):
    """z-lines scaled by 2^s(x, y), s in [-5, 5]: every stored line of the first pass
    gets its own exponent, so the y pass, whose column tiles straddle two x planes
    (N/2+1 columns per plane), must give each column the exponents of its own plane.
    Compared with the plain FP32 path at the FP16 forward tolerance: a column given
    the other plane's exponents would be off by a power of two."""
    import torch
    print(f"case h16 line scales {dims}", flush=True)
    nx, ny, _ = dims
    x = oracle.fill_random_real_3d(*dims, 0x5EED, np.float32)
    s = (3 * np.arange(nx)[:, None] + 5 * np.arange(ny)[None, :]) % 11 - 5
    x = (x * np.exp2(s)[:, :, None]).astype(np.float32)
    u = torch.from_numpy(x).cuda()
    f, g = _ffts(dims), _ffts(dims, "fp32")
    X, X32 = f.complex_array(), g.complex_array()
    f.forward(u, X)
    g.forward(u, X32)
    l2 = float(torch.linalg.vector_norm(X - X32) / torch.linalg.vector_norm(X32))
    print(f"[h16] line scales {dims}: fwd rel_l2 vs FP32 {l2:.3e}", flush=True)
    assert l2 < TOL_FWD_L2
    y = f.real_array()
    f.backward(X, y)
    rt = float(torch.linalg.vector_norm(y / np.prod(dims) - u) / torch.linalg.vector_norm(u))
    print(f"[h16] line scales {dims}: round trip rel_l2 {rt:.3e}", flush=True)
    assert rt < TOL_RT_L2
    f.close()
    g.close()


def test_h16_paired_end_passes_bitwise(gpu):
    """The FP16-storage R2C and C2R passes move their FP32 real side as 8-byte pairs when
    the real array is 16-byte aligned and element-wise otherwise: a real array at an
    8-byte offset turns pairing off and must give bitwise the same spectrum and output."""
    import torch
    dims = (128, 128, 256)
    n = int(np.prod(dims))
    print(f"case h16 paired end passes {dims}", flush=True)
    x = oracle.fill_random_real_3d(*dims, 0x1234ABCDEF987654, np.float32)
    f = _ffts(dims)
    a = torch.from_numpy(x).cuda()
    buf = torch.zeros(n + 2, dtype=torch.float32, device="cuda")
    off = buf[2:].view(dims)
    off.copy_(a)
    assert a.data_ptr() % 16 == 0 and off.data_ptr() % 16 == 8
    X1, X2 = f.complex_array(), f.complex_array()
    f.forward(a, X1)
    f.forward(off, X2)
    assert torch.equal(X1, X2)
    y1 = f.real_array()
    f.backward(X1, y1)
    ybuf = torch.zeros(n + 2, dtype=torch.float32, device="cuda")
    y2 = ybuf[2:].view(dims)
    f.backward(X1, y2)
    assert torch.equal(y1, y2)
    f.close()


def test_h16_adjoints_match_fp32(gpu):
    """The adjoints of an FP16-storage handle (the same passes as forward / backward with
    the real-axis weights) against the plain FP32 handle's, at the FP16 tolerance."""
    import torch
    dims = (128, 128, 128)
    print(f"case h16 adjoints {dims}", flush=True)
    f16, f32 = _ffts(dims), _ffts(dims, "fp32")
    g = torch.Generator(device="cpu").manual_seed(11)
    v = (torch.rand(dims, generator=g) * 2 - 1).cuda()
    cs = f32.complex_shape()
    G = torch.complex(torch.randn(cs, generator=g), torch.randn(cs, generator=g)).cuda()
    A16, A32 = f16.complex_array(), f32.complex_array()
    f16.c2r_adjoint(v, A16)
    f32.c2r_adjoint(v, A32)
    l2, _ = oracle.rel_errors(from_gpu(A16), from_gpu(A32).astype(np.complex128))
    print(f"[h16] c2r_adjoint rel_l2 vs fp32 = {l2:.3e}", flush=True)
    assert l2 < TOL_FWD_L2
    B16, B32 = f16.real_array(), f32.real_array()
    f16.r2c_adjoint(G, B16)
    f32.r2c_adjoint(G, B32)
    l2, _ = oracle.rel_errors(from_gpu(B16), from_gpu(B32).astype(np.float64))
    print(f"[h16] r2c_adjoint rel_l2 vs fp32 = {l2:.3e}", flush=True)
    assert l2 < TOL_FWD_L2
    f16.close()
    f32.close()


def test_h16_adjoint_without_work_area_is_refused(gpu):
    """An FP16-storage handle keeps its spectrum in the work area in both directions:
    c2r_adjoint, like forward, refuses to run before one is bound (no launch happens)."""
    import ctypes
    import torch
    from fftm_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    n = (ctypes.c_int64 * 3)(128, 128, 128)
    assert L.fftm_mi_ffts_create(0, n, _lib.F32_H16, 0, ctypes.byref(h)) == 0
    try:
        u = torch.zeros((128, 128, 128), dtype=torch.float32, device="cuda")
        X = torch.zeros((128, 128, 65), dtype=torch.complex64, device="cuda")
        assert L.fftm_mi_ffts_c2r_adjoint(h, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(X.data_ptr()), 1) == \
            _lib.NO_WORKSPACE
        assert L.fftm_mi_ffts_forward(h, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(X.data_ptr()), 1) == \
            _lib.NO_WORKSPACE
    finally:
        L.fftm_mi_ffts_destroy(h)


# hi + lo FP16 storage (FFTM_MI_F32_H16X2): each stored bin keeps hi = fp16(v) and
# lo = fp16((v - hi) 2^11), about 22 significant bits, so the FP32 parity bar applies
TOL_X2_L2 = 2e-6
TOL_X2_MAX = 1e-5


@pytest.mark.parametrize("dims", [(128, 128, 128), (256, 128, 512), (512, 512, 128)])
def test_h16x2_matches_oracle(gpu, dims):
    import torch
    print(f"case h16x2 3d {dims}", flush=True)
    x = oracle.fill_random_real_3d(*dims, 0x1234ABCDEF987654, np.float32)
    f = _ffts(dims, "fp16x2")
    u = torch.from_numpy(x).cuda()
    X = f.complex_array()
    f.forward(u, X)
    Xo = oracle.ffts3d_forward(x.astype(np.float64))
    l2, mx = oracle.rel_errors(from_gpu(X), Xo)
    print(f"[h16x2] fwd {dims}: rel_l2={l2:.3e} rel_max={mx:.3e}", flush=True)
    assert l2 < TOL_X2_L2 and mx < TOL_X2_MAX
    y = f.real_array()
    f.backward(X, y)
    yr = from_gpu(y).astype(np.float64) / np.prod(dims)
    rt = float(np.sqrt(np.mean((yr - x) ** 2)))
    print(f"[h16x2] round trip {dims}: rms={rt:.3e}", flush=True)
    assert rt < 10 * float(np.finfo(np.float32).eps)
    f.close()


@pytest.mark.parametrize("shift", [20, -30])
def test_h16x2_power_of_two_scaling_is_exact(gpu, shift):
    import torch
    dims = (128, 128, 256)
    x = oracle.fill_random_real_3d(*dims, 77, np.float32)
    f = _ffts(dims, "fp16x2")
    u = torch.from_numpy(x).cuda()
    X, Xs = f.complex_array(), f.complex_array()
    f.forward(u, X)
    f.forward(u * (2.0 ** shift), Xs)
    assert torch.equal(Xs, X * (2.0 ** shift))
    f.close()


def test_h16x2_variants_and_limits(gpu):
    """Tile shapes 0, 2 and 6 (each in every tile order) give bitwise the same result; shape 7,
    the alternative radix schedules and lengths above 1024 are not compiled for it."""
    import torch
    import fftm_amd
    from fftm_amd import _lib
    L = _lib.lib()
    dims = (256, 256, 256)
    f = _ffts(dims, "fp16x2")
    u = torch.rand(f.real_shape(), dtype=torch.float32, device="cuda") * 2 - 1
    X0, y0 = f.complex_array(), f.real_array()
    f.forward(u, X0)
    f.backward(X0, y0)
    nf, nb = f.num_passes()
    for var in (0, 2, 6, 8, 10, 14, 16, 18, 22, 24, 26, 30):
        for bw, n in ((0, nf), (1, nb)):
            for i in range(n):
                _lib.check(L.fftm_mi_ffts_set_pass_variant(f._handle, bw, i, var), "set_pass_variant")
        X, y = f.complex_array(), f.real_array()
        f.forward(u, X)
        f.backward(X, y)
        assert torch.equal(X, X0) and torch.equal(y, y0), f"variant {var}"
    assert L.fftm_mi_ffts_set_pass_variant(f._handle, 0, 1, 7) == _lib.NOT_SUPPORTED
    # the tuner (its reference copy of a pass output covers the 8-byte-per-bin work area)
    chosen = f.autotune(warmup=1, iterations=2)
    assert all((v & 7) in (0, 2, 6) and v < 32 for v in chosen.values())
    X, y = f.complex_array(), f.real_array()
    f.forward(u, X)
    f.backward(X, y)
    assert torch.equal(X, X0) and torch.equal(y, y0)
    f.close()
    import ctypes
    ws = ctypes.c_size_t()
    assert L.fftm_mi_ffts_validate(0, _lib.i64_array([2048, 128, 128]), _lib.F32_H16X2, 0, ctypes.byref(ws)) == 16
