"""The real -> real spectral filter (forward, x H(k), backward in one call).

The last forward pass, the multiply and the first backward pass run as one kernel
along the same lines when both passes transform the same axis; otherwise (lengths
without a compiled kernel) the chain is forward, an elementwise multiply, backward.
Both are checked against the FP64 oracle, the fused chain also bitwise against
forward + spectral_apply + backward, and the operators against closed forms."""

import math
import re

import numpy as np
import pytest

import oracle
from helpers import check_close, from_gpu

pytestmark = pytest.mark.gpu


def _ffts(dims, dt="float32", strategy=None):
    import fftm_amd
    f = fftm_amd.ffts(dt, strategy_4d=strategy)
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    return f


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("dims", [(128, 128, 128), (64, 48, 32), (512, 64, 64)])
def test_poisson_periodic_problem(gpu, dt, dims):
    """-Lap u = f on [0, 2 pi)^3 with f = 3 sin x sin y sin z: u = sin x sin y sin z
    (periodic_poisson_3d_problem, versioned_ffts_3d_test_driver.h:187-209); the
    reference's solve + 1/N scale become one filter call with scale = 1/N."""
    import torch
    nx, ny, nz = dims
    print(f"case poisson {dims} {dt}", flush=True)
    h = [2 * math.pi / n for n in dims]
    x, y, z = np.meshgrid(*(np.arange(n) * hh for n, hh in zip(dims, h)), indexing="ij")
    exact = np.sin(x) * np.sin(y) * np.sin(z)
    rhs = -3.0 * exact  # the reference's rhs: Lap u = rhs, solved with H = -1/|k|^2
    f = _ffts(dims, dt)
    rt = f.real
    u = torch.from_numpy(rhs.astype(np.float32 if dt == "float32" else np.float64)).cuda()
    out = torch.empty_like(u)
    f.filter(u, out, "poisson", scale=1.0 / (nx * ny * nz))
    err = np.abs(from_gpu(out) - exact).max()
    print(f"[poisson] {dims} {dt}: max |u - exact| = {err:.3e}", flush=True)
    assert err < (2e-6 if dt == "float32" else 1e-13)
    f.close()


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_gaussian_plane_waves(gpu, dt):
    """cos(k . x) -> exp(-|k|^2 / (2 sigma^2)) cos(k . x) for integer wave vectors."""
    import torch
    dims = (64, 32, 128)
    L = [np.arange(n) * 2 * math.pi / n for n in dims]
    X, Y, Z = np.meshgrid(*L, indexing="ij")
    waves = [(1, 2, 3), (5, -3, 7), (0, 0, 9), (-20, 11, 40)]
    sigma = 8.0
    src = sum(np.cos(a * X + b * Y + c * Z) for a, b, c in waves)
    exp_out = sum(math.exp(-(a * a + b * b + c * c) / (2 * sigma ** 2)) * np.cos(a * X + b * Y + c * Z)
                  for a, b, c in waves)
    f = _ffts(dims, dt)
    u = torch.from_numpy(src.astype(np.float32 if dt == "float32" else np.float64)).cuda()
    out = torch.empty_like(u)
    f.filter(u, out, "gaussian", scale=1.0 / np.prod(dims), sigma=sigma)
    err = np.abs(from_gpu(out) - exp_out).max()
    print(f"[gaussian] plane waves {dt}: max err = {err:.3e}", flush=True)
    assert err < (5e-6 if dt == "float32" else 1e-12)
    f.close()


@pytest.mark.parametrize("kind,axis", [("poisson", None), ("gaussian", None), ("shell", None), ("derivative", 1),
                                       ("derivative", 0), ("derivative", 2), ("poisson_gradient", 0),
                                       ("bandpass", None)])
@pytest.mark.parametrize("dims", [(128, 128, 128), (256, 64, 32), (30, 20, 24)])
def test_filter_vs_oracle_and_unfused(gpu, kind, axis, dims):
    import torch
    print(f"case filter {kind} axis={axis} {dims}", flush=True)
    x = oracle.fill_random_real_3d(*dims, 0x5EED, np.float32)
    f = _ffts(dims)
    coeffs = None
    cf = None
    if kind == "shell":
        cf = np.random.default_rng(3).uniform(0.2, 2.0, size=32)
        coeffs = torch.from_numpy(cf.astype(np.float32)).cuda()
    sigma = 8.0 if kind == "gaussian" else None
    band = (4.0, 11.5) if kind == "bandpass" else None
    scale = 1.0 / np.prod(dims)
    kw = dict(scale=scale, sigma=sigma, coeffs=coeffs, axis=axis, band=band)
    u = torch.from_numpy(x).cuda()
    out = torch.empty_like(u)
    f.filter(u, out, kind, **kw)
    H = oracle.spectral_operator_logical(dims, kind, scale, sigma or 0.0, cf, axis or 0, band or (0.0, 0.0))
    ref = oracle.filter_logical(x.astype(np.float64), H)
    check_close(from_gpu(out), ref, np.float32, f"filter {kind} {dims}")
    # forward + spectral_apply + backward: the same arithmetic, bitwise
    X = f.complex_array()
    f.forward(u, X)
    f.spectral_apply(X, X, kind, **kw)
    y = f.real_array()
    f.backward(X, y)
    assert torch.equal(y, out)
    # in place (real_in == real_out)
    v = u.clone()
    f.filter(v, v, kind, **kw)
    assert torch.equal(v, out)
    f.close()


@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_derivatives_and_poisson_gradient_closed_form(gpu, dt):
    """The reference's v4 check (versioned_ffts_3d_test_driver.h:640-700): u = sin x sin y sin z,
    rhs = -3 u; du/dx_a through the derivative operator and through the fused Poisson
    gradient of rhs both equal cos on axis a times sin on the others."""
    import torch
    dims = (64, 32, 48)
    L = [np.arange(n) * 2 * math.pi / n for n in dims]
    X, Y, Z = np.meshgrid(*L, indexing="ij")
    s = [np.sin(X), np.sin(Y), np.sin(Z)]
    c = [np.cos(X), np.cos(Y), np.cos(Z)]
    u = s[0] * s[1] * s[2]
    npdt = np.float32 if dt == "float32" else np.float64
    f = _ffts(dims, dt)
    ug = torch.from_numpy(u.astype(npdt)).cuda()
    rg = torch.from_numpy((-3.0 * u).astype(npdt)).cuda()
    out = torch.empty_like(ug)
    tol = 2e-6 if dt == "float32" else 1e-12
    # i k_a multiplies the transform's roundoff at wavenumber k by k: bound by k_max = n_a / 2
    for a in range(3):
        exact = np.prod([c[b] if b == a else s[b] for b in range(3)], axis=0)
        f.filter(ug, out, "derivative", scale=1.0 / np.prod(dims), axis=a)
        e1 = np.abs(from_gpu(out) - exact).max()
        f.filter(rg, out, "poisson_gradient", scale=1.0 / np.prod(dims), axis=a)
        e2 = np.abs(from_gpu(out) - exact).max()
        print(f"[derivative] axis {a} {dt}: {e1:.2e}, poisson gradient {e2:.2e}", flush=True)
        assert e1 < tol * dims[a] / 2 and e2 < tol
    f.close()


def test_bandpass_plane_waves(gpu):
    import torch
    dims = (64, 64, 64)
    L = [np.arange(n) * 2 * math.pi / n for n in dims]
    X, Y, Z = np.meshgrid(*L, indexing="ij")
    waves = [(1, 0, 0), (3, 4, 0), (0, 6, 8), (12, 0, 5), (20, 20, 20)]   # |k| = 1, 5, 10, 13, 34.6
    src = sum(np.cos(a * X + b * Y + c * Z) for a, b, c in waves)
    keep = sum(np.cos(a * X + b * Y + c * Z) for a, b, c in waves if 5 <= math.sqrt(a * a + b * b + c * c) <= 13)
    f = _ffts(dims)
    u = torch.from_numpy(src.astype(np.float32)).cuda()
    out = torch.empty_like(u)
    f.filter(u, out, "bandpass", scale=1.0 / np.prod(dims), band=(5.0, 13.0))
    assert np.abs(from_gpu(out) - keep).max() < 5e-6
    f.close()


@pytest.mark.parametrize("strategy,dims", [("pencil", (64, 64, 64, 64)), ("slab", (16, 8, 12, 16)),
                                           ("xzwy", (32, 16, 8, 32)), ("pencil", (16, 8, 12, 16))])
def test_filter_4d(gpu, strategy, dims):
    import torch
    import fftm_amd
    strat = {"pencil": fftm_amd.strategy_4d_pencil_pencil(), "slab": fftm_amd.strategy_4d_slab_slab(),
             "xzwy": fftm_amd.strategy_4d_slab_native_xzwy()}[strategy]
    print(f"case filter 4d {strategy} {dims}", flush=True)
    lg = oracle.fill_random_real_4d_logical(*dims, 0xF00D, np.float32)
    arr = oracle.A0123 if strategy == "pencil" else oracle.A3210
    mem = oracle.to_memory(lg, arr)
    f = _ffts(dims, strategy=strat)
    u = torch.from_numpy(np.ascontiguousarray(mem)).cuda()
    out = torch.empty_like(u)
    scale = 1.0 / np.prod(dims)
    f.filter(u, out, "gaussian", scale=scale, sigma=6.0)
    H = oracle.spectral_operator_logical(dims, "gaussian", scale, 6.0)
    ref = oracle.to_memory(oracle.filter_logical(lg.astype(np.float64), H), arr)
    check_close(from_gpu(out), ref, np.float32, f"filter 4d {strategy} {dims}")
    f.close()


def test_filter_argument_errors(gpu):
    import torch
    import fftm_amd
    f = _ffts((32, 32, 32))
    u = torch.zeros(f.real_shape(), device="cuda")
    with pytest.raises(ValueError):
        f.filter(u, u, "laplace")
    with pytest.raises(fftm_amd.FftmMiError) as e:
        f.filter(u, u, "gaussian", sigma=0.0)
    assert e.value.status == 4
    g = fftm_amd.ffts("float32")
    g.init(32, 32, 32, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="",
                                                          destructive_backward=True))
    with pytest.raises(fftm_amd.FftmMiError) as e:
        g.filter(u, u, "poisson")
    assert e.value.status == 16
    f.close()
    g.close()


@pytest.mark.parametrize("fvar", [0, 2, 4, 28])
def test_filter_pass_tile_shapes_bitwise(gpu, tmp_path, fvar):
    """Every tile shape of the fused filter pass (0, 2, 4, in any tile order) gives bitwise
    the result of forward + spectral_apply + backward.  The shape is installed through a
    tuning cache holding the handle's filter record.  The cache also holds the separate
    passes to radix schedule 0 (variant mod 32), the schedule the filter pass computes
    with: the tuner may keep another schedule for a strided pass, which agrees with
    schedule 0 to within rounding only."""
    import torch
    import fftm_amd
    dims = (512, 64, 32)
    f = fftm_amd.ffts("float32")
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    cache = str(tmp_path / "flt.cache")
    f.autotune(warmup=0, iterations=1, cache_path=cache)
    lines = open(cache).read().splitlines()
    key = [ln.split("=", 1)[0] for ln in lines if ln.endswith(tuple(f"={v}" for v in range(32)))
           and ln.split("=", 1)[0].endswith("_FILTER")]
    assert len(key) == 1, key
    def rec(ln):
        k, _, v = ln.partition("=")
        if k == key[0]:
            return f"{k}={fvar}"
        if re.search(r"_(FWD|BWD)\d+$", k) and v.isdigit():
            return f"{k}={int(v) % 32}"
        return ln
    with open(cache, "w") as fh:
        fh.write("\n".join(rec(ln) for ln in lines) + "\n")
    assert f.load_tuning(cache) > 0
    g = torch.Generator(device="cuda").manual_seed(fvar)
    u = torch.rand(dims, generator=g, device="cuda") * 2 - 1
    out = f.real_array()
    f.filter(u, out, "gaussian", scale=1.0 / u.numel(), sigma=6.0)
    X = f.complex_array()
    f.forward(u, X)
    f.spectral_apply(X, X, "gaussian", scale=1.0 / u.numel(), sigma=6.0)
    y = f.real_array()
    f.backward(X, y)
    assert torch.equal(y, out), f"filter variant {fvar}"
    f.close()
