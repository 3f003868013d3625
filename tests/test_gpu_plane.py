"""The plane pass (real axis + one complex axis in LDS, one HBM pass) against the FP64
oracle and against the unfused passes, for every ffts layout it serves."""

import numpy as np
import pytest

import oracle
from helpers import check_close, from_gpu, roundtrip_tol, rms_diff, to_gpu

pytestmark = pytest.mark.gpu

CASES = [("3d", (512, 32, 32)), ("3d", (1024, 16, 16)), ("3d", (512, 64, 64)), ("3d", (512, 128, 128)),
         ("pencil", (32, 32, 32, 32)), ("slab", (32, 32, 32, 32)), ("xzwy", (32, 32, 32, 32)),
         ("pencil", (64, 16, 32, 64)), ("slab", (8, 64, 64, 64)), ("xzwy", (64, 32, 16, 64)),
         ("pencil", (128, 8, 64, 128)), ("slab", (4, 128, 128, 128))]


def _make(kind, dims, real, fuse=True):
    import ctypes
    import fftm_amd
    from fftm_amd import _lib
    strat = {"pencil": fftm_amd.strategy_4d_pencil_pencil(), "slab": fftm_amd.strategy_4d_slab_slab(),
             "xzwy": fftm_amd.strategy_4d_slab_native_xzwy(), "3d": None}[kind]
    f = fftm_amd.ffts(real, strategy_4d=strat)
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    if not fuse:  # rebuild the handle without plane fusion
        L = _lib.lib()
        L.fftm_mi_ffts_destroy(f._handle)
        layout = {"3d": 0, "pencil": 1, "slab": 2, "xzwy": 3}[kind]
        h = ctypes.c_void_p()
        _lib.check(L.fftm_mi_ffts_create(layout, _lib.i64_array(dims), 0 if real == "float32" else 1, 2,
                                         ctypes.byref(h)), "create")
        f._handle = h
        if f._work is not None:
            _lib.check(L.fftm_mi_ffts_set_work_area(h, f._work.data_ptr()), "work")
    return f


def _oracle_fwd(kind, x, dims):
    if kind == "3d":
        return oracle.ffts3d_forward(x)
    if kind == "xzwy":
        return oracle.native_xzwy_forward(x, *dims)
    return oracle.ffts4d_forward(x, kind, *dims)


def _oracle_bwd(kind, X, dims):
    if kind == "3d":
        return oracle.ffts3d_backward(X, dims[2])
    if kind == "xzwy":
        return oracle.native_xzwy_backward(X, *dims)
    return oracle.ffts4d_backward(X, kind, *dims)


@pytest.mark.parametrize("real", ["float32", "float64"])
@pytest.mark.parametrize("kind,dims", CASES)
def test_plane_pass_parity(gpu, real, kind, dims):
    dt = np.float32 if real == "float32" else np.float64
    f = _make(kind, dims, real)
    nf, nb = f.num_passes()
    print(f"case plane {kind} {dims} {real}: passes fwd={nf} bwd={nb}", flush=True)
    assert nf == len(dims) - 1 and nb == len(dims) - 1, "plane fusion expected"
    x = np.random.default_rng(7).uniform(-1, 1, f.real_shape()).astype(dt)
    X = f.complex_array()
    f.forward(to_gpu(x), X)
    Xo = _oracle_fwd(kind, x.astype(np.float64), dims)
    check_close(from_gpu(X), Xo, dt, f"plane fwd {kind}{dims}")
    y = f.real_array()
    f.backward(X, y)
    check_close(from_gpu(y), _oracle_bwd(kind, Xo, dims), dt, f"plane bwd {kind}{dims}")
    assert rms_diff(from_gpu(y) / np.prod(dims), x) <= roundtrip_tol(dt)
    # the unfused transform agrees to rounding
    g = _make(kind, dims, real, fuse=False)
    assert g.num_passes() == (len(dims), len(dims))
    X2 = g.complex_array()
    g.forward(to_gpu(x), X2)
    check_close(from_gpu(X2), from_gpu(X).astype(np.complex128), dt, "fused vs unfused")


@pytest.mark.parametrize("kind,dims", [("3d", (512, 32, 32)), ("pencil", (32, 32, 32, 32)),
                                       ("slab", (32, 32, 32, 32)), ("xzwy", (32, 32, 32, 32))])
def test_plane_pass_adjoints(gpu, kind, dims):
    import torch
    f = _make(kind, dims, "float64")
    g = np.random.default_rng(3)
    u = to_gpu(g.uniform(-1, 1, f.real_shape()))
    G = to_gpu(g.standard_normal(f.complex_shape()) + 1j * g.standard_normal(f.complex_shape()))
    Fu = f.complex_array()
    f.forward(u, Fu)
    FtG = f.real_array()
    f.r2c_adjoint(G, FtG)
    a = float(np.sum(from_gpu(Fu).real * from_gpu(G).real + from_gpu(Fu).imag * from_gpu(G).imag))
    b = float(torch.sum(u * FtG))
    assert abs(a - b) <= 1e-11 * abs(a)
    v = to_gpu(g.uniform(-1, 1, f.real_shape()))
    X = to_gpu(g.standard_normal(f.complex_shape()) + 1j * g.standard_normal(f.complex_shape()))
    CX = f.real_array()
    f.backward(X, CX)
    Ctv = f.complex_array()
    f.c2r_adjoint(v, Ctv)
    a = float(torch.sum(CX * v))
    b = float(np.sum(from_gpu(X).real * from_gpu(Ctv).real + from_gpu(X).imag * from_gpu(Ctv).imag))
    assert abs(a - b) <= 1e-11 * abs(a)


# shape 1 runs the register-resident plane kernel; every layout (pencil: B contiguous,
# paired real rows; 3D / slab / xzwy: A contiguous), and plane counts that leave the
# last workgroup partly empty
RX_CASES = [("pencil", "float32", (16, 15, 35, 16)), ("pencil", "float32", (32, 32, 32, 32)),
            ("pencil", "float32", (64, 27, 21, 64)), ("pencil", "float32", (128, 8, 64, 128)),
            ("pencil", "float64", (64, 27, 21, 64)), ("pencil", "float64", (128, 8, 64, 128)),
            ("3d", "float32", (525, 64, 64)), ("3d", "float64", (512, 32, 32)), ("3d", "float32", (512, 128, 128)),
            ("slab", "float32", (8, 64, 64, 64)), ("slab", "float64", (4, 128, 128, 128)),
            ("xzwy", "float32", (64, 32, 16, 64)), ("xzwy", "float64", (32, 32, 32, 32))]


@pytest.mark.parametrize("kind,real,dims", RX_CASES)
def test_plane_register_kernel_bitwise(gpu, kind, real, dims):
    """Plane-pass variants 1 (register-resident kernel), 8, 9 (round-robin tile order),
    16, 17, 24 and 25 (chunked tile orders) reproduce variant 0 bitwise: forward,
    backward and both adjoints (weight modes 1, 2)."""
    import torch
    from fftm_amd import _lib
    L = _lib.lib()
    f = _make(kind, dims, real)
    nf, nb = f.num_passes()
    print(f"case plane-rx {kind} {dims} {real}", flush=True)
    assert nf == len(dims) - 1 and nb == len(dims) - 1, "plane fusion expected"
    dt = np.float32 if real == "float32" else np.float64
    ct = np.complex64 if real == "float32" else np.complex128
    g = np.random.default_rng(11)
    x = g.uniform(-1, 1, f.real_shape()).astype(dt)
    u = to_gpu(x)
    S = to_gpu((g.standard_normal(f.complex_shape()) + 1j * g.standard_normal(f.complex_shape())).astype(ct))

    def run():
        X = f.complex_array()
        f.forward(u, X)
        y = f.real_array()
        f.backward(S, y)
        A = f.real_array()
        f.r2c_adjoint(S, A)
        B = f.complex_array()
        f.c2r_adjoint(u, B)
        return X, y, A, B

    def set_all(var):
        for bw, n in ((0, nf), (1, nb)):
            for i in range(n):
                _lib.check(L.fftm_mi_ffts_set_pass_variant(f._handle, bw, i, var), "set_pass_variant")

    set_all(0)
    ref = run()
    check_close(from_gpu(ref[0]), _oracle_fwd(kind, x.astype(np.float64), dims), dt, f"plane-rx fwd {kind}{dims}")
    for var in (1, 8, 9, 16, 17, 24, 25):
        set_all(var)
        out = run()
        for a, b, name in zip(ref, out, ("forward", "backward", "r2c_adjoint", "c2r_adjoint")):
            assert torch.equal(a, b), f"variant {var}: {name} differs from variant 0"
    f.close()
