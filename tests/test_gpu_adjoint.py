"""Adjoints and spectral operators on the GPU (the differentiation workload):

* dot-product identities <R2C u, G> = <u, R2C^T G>, <C2R X, v> = <X, C2R^T v>
  (the analytic adjoint), every layout, FP64;
* spectral multiply and shell gradient against a numpy restatement;
* the loss L = ||C2R(H * R2C(u)) - y||^2 with a 32-shell radial H: gradients w.r.t. u
  and the shell coefficients from the hand-written adjoints (through the thin autograd
  binding) against FP64 central finite differences on 32^3, step 1e-6, 1e-6 relative.
"""

import math

import numpy as np
import pytest

import oracle
from helpers import from_gpu, to_gpu

pytestmark = pytest.mark.gpu

LAYOUTS = [("3d", (12, 10, 8)), ("3d", (32, 32, 32)), ("pencil", (8, 6, 10, 12)), ("slab", (8, 6, 10, 12)),
           ("xzwy", (16, 8, 8, 6))]


def _make(kind, dims, real="float64"):
    import fftm_amd
    strat = {"pencil": fftm_amd.strategy_4d_pencil_pencil(), "slab": fftm_amd.strategy_4d_slab_slab(),
             "xzwy": fftm_amd.strategy_4d_slab_native_xzwy(), "3d": None}[kind]
    f = fftm_amd.ffts(real, strategy_4d=strat)
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    return f


def _logical_to_spec_mem(kind, a):
    """logical half-spectrum array (x, y, z[, kw]) -> spectrum memory layout of the strategy."""
    if kind in ("3d", "slab"):
        return np.ascontiguousarray(a)
    return np.ascontiguousarray(np.transpose(a, (1, 2, 3, 0)))   # [y][z][wh][x]


def _inner_c(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag))


@pytest.mark.parametrize("kind,dims", LAYOUTS)
def test_adjoint_dot_products(gpu, kind, dims):
    import torch
    f = _make(kind, dims)
    g = np.random.default_rng(1)
    u = to_gpu(g.uniform(-1, 1, f.real_shape()))
    G = to_gpu(g.standard_normal(f.complex_shape()) + 1j * g.standard_normal(f.complex_shape()))
    Fu = f.complex_array()
    f.forward(u, Fu)
    FtG = f.real_array()
    f.r2c_adjoint(G, FtG)
    lhs, rhs = _inner_c(from_gpu(Fu), from_gpu(G)), float(torch.sum(u * FtG))
    print(f"[adjoint] {kind}{dims} R2C: {lhs:.15e} vs {rhs:.15e}")
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0) * math.sqrt(np.prod(dims))
    X = to_gpu(g.standard_normal(f.complex_shape()) + 1j * g.standard_normal(f.complex_shape()))
    v = to_gpu(g.uniform(-1, 1, f.real_shape()))
    CX = f.real_array()
    f.backward(X, CX)
    Ctv = f.complex_array()
    f.c2r_adjoint(v, Ctv)
    lhs, rhs = float(torch.sum(CX * v)), _inner_c(from_gpu(X), from_gpu(Ctv))
    print(f"[adjoint] {kind}{dims} C2R: {lhs:.15e} vs {rhs:.15e}")
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0) * math.sqrt(np.prod(dims))
    # the weighted forms against the oracle: C2R^T v = w R2C(v)
    logical_w = oracle.realaxis_weights_logical(dims)
    if kind == "3d":
        ref = oracle.ffts3d_forward(from_gpu(v)) * logical_w
    elif kind == "slab":
        ref = oracle.ffts4d_forward(from_gpu(v), "slab", *dims) * logical_w
    elif kind == "pencil":
        ref = oracle.ffts4d_forward(from_gpu(v), "pencil", *dims) * _logical_to_spec_mem(kind, logical_w)
    else:
        ref = oracle.native_xzwy_forward(from_gpu(v), *dims) * _logical_to_spec_mem(kind, logical_w)
    rms, mx = oracle.rel_errors(from_gpu(Ctv), ref)
    assert rms < 1e-13 and mx < 1e-13


@pytest.mark.parametrize("kind,dims", LAYOUTS)
@pytest.mark.parametrize("real", ["float32", "float64"])
def test_spectral_multiply_and_shell_gradient(gpu, kind, dims, real):
    import torch
    f = _make(kind, dims, real)
    S = 32
    g = np.random.default_rng(2)
    cdt = np.complex64 if real == "float32" else np.complex128
    X = (g.standard_normal(f.complex_shape()) + 1j * g.standard_normal(f.complex_shape())).astype(cdt)
    coeffs = g.uniform(0.5, 1.5, S).astype(np.float32 if real == "float32" else np.float64)
    shell_mem = _logical_to_spec_mem(kind, oracle.shell_index_logical(dims, S))
    out = f.complex_array()
    f.spectral_multiply(to_gpu(X), out, shell_coeffs=to_gpu(coeffs))
    expect = X * coeffs[shell_mem]
    assert np.array_equal(from_gpu(out), expect.astype(cdt)), "shell filter must be exactly H_k * X_k"
    filt = g.uniform(-1, 1, f.complex_shape()).astype(coeffs.dtype)
    f.spectral_multiply(to_gpu(X), out, filter=to_gpu(filt))
    assert np.array_equal(from_gpu(out), (X * filt).astype(cdt))
    G = (g.standard_normal(f.complex_shape()) + 1j * g.standard_normal(f.complex_shape())).astype(cdt)
    grad = from_gpu(f.shell_gradient(to_gpu(G), to_gpu(X), S)).astype(np.float64)
    v = (G.real.astype(np.float64) * X.real + G.imag.astype(np.float64) * X.imag)
    ref = np.bincount(shell_mem.reshape(-1), weights=v.reshape(-1), minlength=S)
    tol = 1e-6 if real == "float32" else 1e-12
    assert np.max(np.abs(grad - ref)) <= tol * np.max(np.abs(ref))
    del torch


@pytest.mark.parametrize("kind,dims,S", [("3d", (6, 6, 6), 3), ("3d", (20, 20, 20), 10), ("3d", (30, 30, 30), 15),
                                         ("pencil", (6, 6, 6, 6), 14), ("slab", (12, 12, 12, 12), 14),
                                         ("xzwy", (6, 6, 6, 6), 14)])
def test_shell_boundary_bins(gpu, kind, dims, S):
    """Sizes with bins where S |k| / kmax is an integer (a floating floor(sqrt(k^2) S / kmax)
    puts them one shell low): the shell multiply, the shell gradient and the fused shell
    filter follow the exact shell index of the oracle (pinned against isqrt on the CPU)."""
    import torch
    d = len(dims)
    ks = [oracle.wavenumbers(n, a == d - 1) for a, n in enumerate(dims)]
    k2 = sum(gk.astype(np.float64) ** 2 for gk in np.meshgrid(*ks, indexing="ij"))
    kmax2 = float(sum((n // 2) ** 2 for n in dims))
    naive = [np.minimum(np.floor(np.sqrt(k2) * c).astype(np.int64), S - 1)
             for c in ((1.0 / math.sqrt(kmax2)) * S, S / math.sqrt(kmax2))]
    exact = oracle.shell_index_logical(dims, S)
    assert any(np.count_nonzero(v != exact) > 0 for v in naive), "the case must contain boundary bins"
    f = _make(kind, dims, "float64")
    g = np.random.default_rng(5)
    X = g.standard_normal(f.complex_shape()) + 1j * g.standard_normal(f.complex_shape())
    coeffs = np.arange(1, S + 1, dtype=np.float64)          # distinct per shell
    shell_mem = _logical_to_spec_mem(kind, exact)
    out = f.complex_array()
    f.spectral_multiply(to_gpu(X), out, shell_coeffs=to_gpu(coeffs))
    assert np.array_equal(from_gpu(out), X * coeffs[shell_mem])
    G = g.standard_normal(f.complex_shape()) + 1j * g.standard_normal(f.complex_shape())
    grad = from_gpu(f.shell_gradient(to_gpu(G), to_gpu(X), S))
    ref = np.bincount(shell_mem.reshape(-1), weights=(G.real * X.real + G.imag * X.imag).reshape(-1), minlength=S)
    assert np.max(np.abs(grad - ref)) <= 1e-12 * np.max(np.abs(ref))
    # the analytic shell operator of the spectral-apply kernel and the fused filter pass
    f.spectral_apply(to_gpu(X), out, "shell", coeffs=to_gpu(coeffs))
    assert np.array_equal(from_gpu(out), X * coeffs[shell_mem])
    u = g.uniform(-1, 1, f.real_shape())
    y = f.real_array()
    f.filter(to_gpu(u), y, "shell", scale=1.0, coeffs=to_gpu(coeffs))
    Fu = f.complex_array()
    f.forward(to_gpu(u), Fu)
    Fu_h = from_gpu(Fu) * coeffs[shell_mem]
    y2 = f.real_array()
    f.backward(to_gpu(Fu_h), y2)
    assert np.max(np.abs(from_gpu(y) - from_gpu(y2))) <= 1e-12 * np.max(np.abs(from_gpu(y2)))
    f.close()
    del torch


def _loss_setup(n, S=32, seed=3, real="float64"):
    import torch
    f = _make("3d", (n, n, n), real)
    dt = torch.float32 if real == "float32" else torch.float64
    g = torch.Generator(device="cpu").manual_seed(seed)
    u = (torch.rand((n, n, n), generator=g, dtype=torch.float64) * 2 - 1).to(dt).cuda()
    c_true = (torch.rand(S, generator=g, dtype=torch.float64) + 0.5).to(dt).cuda()   # hidden H*
    c = (torch.rand(S, generator=g, dtype=torch.float64) + 0.5).to(dt).cuda()
    from fftm_amd.autograd import spectral_filter
    with torch.no_grad():
        y = spectral_filter(u, c_true, f) / (n ** 3) + 0.01 * torch.randn((n, n, n), generator=g,
                                                                           dtype=torch.float64).to(dt).cuda()
    return f, u, c, y


def test_adjoint_finite_difference_fp64_32cubed(gpu):
    """Directional derivatives of L(u, c) = ||C2R(H(c) R2C(u))/N - y||^2 from the adjoints
    vs FP64 central differences, step 1e-6, 1e-6 relative."""
    import torch
    from fftm_amd.autograd import spectral_filter
    n = 32
    f, u, c, y = _loss_setup(n)

    def L(uu, cc):
        with torch.no_grad():
            return float(torch.sum((spectral_filter(uu, cc, f) / n ** 3 - y) ** 2))

    uu = u.clone().requires_grad_(True)
    cc = c.clone().requires_grad_(True)
    loss = torch.sum((spectral_filter(uu, cc, f) / n ** 3 - y) ** 2)
    loss.backward()
    gu, gc = uu.grad.detach(), cc.grad.detach()
    h = 1e-6
    gen = torch.Generator(device="cpu").manual_seed(11)
    worst = 0.0
    for trial in range(3):
        du = torch.randn(u.shape, generator=gen, dtype=torch.float64).cuda()
        du /= torch.linalg.vector_norm(du)
        dc = torch.randn(c.shape, generator=gen, dtype=torch.float64).cuda()
        dc /= torch.linalg.vector_norm(dc)
        fd_u = (L(u + h * du, c) - L(u - h * du, c)) / (2 * h)
        fd_c = (L(u, c + h * dc) - L(u, c - h * dc)) / (2 * h)
        an_u, an_c = float(torch.sum(gu * du)), float(torch.sum(gc * dc))
        ru, rc = abs(fd_u - an_u) / abs(an_u), abs(fd_c - an_c) / abs(an_c)
        print(f"[fd] trial {trial}: d/du fd={fd_u:.12e} adj={an_u:.12e} rel={ru:.2e}; "
              f"d/dc fd={fd_c:.12e} adj={an_c:.12e} rel={rc:.2e}")
        worst = max(worst, ru, rc)
    assert worst <= 1e-6


def test_adjoint_gradients_fp32_vs_torch_fft_fp64(gpu):
    """The 256^3 FP32 gradients vs the same loss differentiated by torch.fft (hipFFT) in FP64."""
    import torch
    from fftm_amd.autograd import spectral_filter
    n, S = 256, 32
    f, u, c, y = _loss_setup(n, S, real="float32")
    uu = u.clone().requires_grad_(True)
    cc = c.clone().requires_grad_(True)
    loss = torch.sum((spectral_filter(uu, cc, f) / n ** 3 - y) ** 2)
    loss.backward()
    shell = torch.from_numpy(oracle.shell_index_logical((n, n, n), S)).cuda()
    u64 = u.double().requires_grad_(True)
    c64 = c.double().requires_grad_(True)
    H = c64[shell]
    v = torch.fft.irfftn(H * torch.fft.rfftn(u64), s=(n, n, n))
    loss64 = torch.sum((v - y.double()) ** 2)
    loss64.backward()
    eu = float(torch.linalg.vector_norm(uu.grad.double() - u64.grad) / torch.linalg.vector_norm(u64.grad))
    ec = float(torch.linalg.vector_norm(cc.grad.double() - c64.grad) / torch.linalg.vector_norm(c64.grad))
    print(f"[adjoint256] loss fp32={float(loss):.6e} fp64={float(loss64):.6e} rel_l2 grad_u={eu:.2e} grad_c={ec:.2e}")
    assert eu < 1e-4 and ec < 1e-4
