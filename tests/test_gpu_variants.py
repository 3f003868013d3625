"""Every variant of every pass (tile shape 0..7, each in the four tile orders) produces
bitwise the same result as the default (the per-column arithmetic depends on neither the
tile shape nor the order the tiles run in), so the tuner can pick any of them without a
parity risk."""

import numpy as np
import pytest

import oracle
from helpers import check_close, from_gpu, to_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", ["float32", "float64"])
@pytest.mark.parametrize("kind,dims", [("3d", (64, 128, 256)), ("3d", (512, 64, 96)), ("pencil", (64, 32, 16, 128)),
                                       ("slab", (32, 64, 128, 64)), ("xzwy", (128, 16, 32, 64)),
                                       ("3d", (1024, 8, 2048)), ("3d", (4096, 4, 64))])
def test_variants_bitwise_equal(gpu, dt, kind, dims):
    import torch
    import fftm_amd
    from fftm_amd import _lib
    strat = {"pencil": fftm_amd.strategy_4d_pencil_pencil(), "slab": fftm_amd.strategy_4d_slab_slab(),
             "xzwy": fftm_amd.strategy_4d_slab_native_xzwy(), "3d": None}[kind]
    f = fftm_amd.ffts(dt, strategy_4d=strat)
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    L = _lib.lib()
    print(f"case variants {kind} {dims} {dt}", flush=True)
    u = torch.rand(f.real_shape(), dtype=f.real, device="cuda") * 2 - 1
    X0, y0 = f.complex_array(), f.real_array()
    f.forward(u, X0)
    f.backward(X0, y0)
    nf, nb = f.num_passes()
    for var in range(1, 32):
        for bw, n in ((0, nf), (1, nb)):
            for i in range(n):
                _lib.check(L.fftm_mi_ffts_set_pass_variant(f._handle, bw, i, var), "set_pass_variant")
        X, y = f.complex_array(), f.real_array()
        f.forward(u, X)
        f.backward(X, y)
        dX = float((X - X0).abs().max() / X0.abs().max())
        dy = float((y - y0).abs().max() / y0.abs().max())
        print(f"variant {var}: forward rel-max diff {dX:.3e}, backward {dy:.3e}", flush=True)
        assert torch.equal(X, X0), f"variant {var} forward differs (rel-max {dX:.3e})"
        assert torch.equal(y, y0), f"variant {var} backward differs (rel-max {dy:.3e})"
    if len(dims) == 3 and np.prod(dims) <= 2 ** 24:
        check_close(from_gpu(X0), oracle.ffts3d_forward(from_gpu(u).astype(np.float64)),
                    np.float32 if dt == "float32" else np.float64, f"variants {dims}")
    f.close()
