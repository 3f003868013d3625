"""Radix schedules as tuning candidates, and the radix-16 stage on the matrix cores.

A strided C2C pass of length 512 or 1024 runs its default schedule (512 = 8 8 8,
1024 = 16 8 8), the alternative schedule 8 16 4 / 8 16 8 (variant + 32), or the same
alternative with its radix-16 stage as FP16 MFMA products with hi + lo operand splits and
FP32 accumulation (variant + 64).  Every candidate is checked against the FP64 oracle at
the FP32 tolerance (and, with FP16-stored spectra, at the FP16-storage tolerance and
against the default schedule), and the tuner keeps only candidates within the FP32
parity tolerance of the default schedule's output."""

import numpy as np
import pytest

import oracle
from helpers import TOL_REL_L2, TOL_REL_MAX, from_gpu, roundtrip_tol

pytestmark = pytest.mark.gpu

SEED = 0x1234ABCDEF987654


def _strided_c2c_passes(f):
    """(backward, index) of the strided C2C passes of length 512 / 1024 of a 3D handle
    (the y and x passes: the forward's last two, the backward's first two)."""
    nf, nb = f.num_passes()
    return [(0, nf - 2), (0, nf - 1), (1, 0), (1, 1)]


def _make(dims, storage="fp32"):
    import fftm_amd
    f = fftm_amd.ffts("float32")
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="",
                                                     spectrum_storage=storage))
    return f


@pytest.mark.parametrize("dims", [(512, 512, 64), (1024, 1024, 16)])
@pytest.mark.parametrize("rs,shape", [(1, 0), (1, 6), (1, 2), (2, 0), (2, 6)])
def test_radix_schedules_fp32_vs_oracle(gpu, dims, rs, shape):
    import torch
    print(f"case radix fp32 dims={dims} rs={rs} shape={shape}", flush=True)
    x = oracle.fill_random_real_3d(*dims, SEED, np.float32)
    Xo = oracle.ffts3d_forward(x.astype(np.float64))
    f = _make(dims)
    for bw, i in _strided_c2c_passes(f):
        f.set_pass_variant(bw, i, shape + 32 * rs)
    u = torch.from_numpy(x).cuda()
    X = f.complex_array()
    f.forward(u, X)
    l2, mx = oracle.rel_errors(from_gpu(X), Xo)
    print(f"[radix] fwd rel_l2={l2:.3e} rel_max={mx:.3e}", flush=True)
    assert l2 <= TOL_REL_L2[np.float32] and mx <= TOL_REL_MAX[np.float32]
    y = f.real_array()
    f.backward(X, y)
    rt = float(np.sqrt(np.mean((from_gpu(y).astype(np.float64) / np.prod(dims) - x) ** 2)))
    print(f"[radix] round trip rms={rt:.3e}", flush=True)
    assert rt <= roundtrip_tol(np.float32)
    f.close()


@pytest.mark.parametrize("rs", [1, 2])
def test_radix_schedules_fp16_storage(gpu, rs):
    """FP16-stored spectra with the alternative schedules on the strided passes: the FP16
    tolerance against the oracle, and close to the default schedule's result (the two
    differ only where a stored value rounds to the other neighbouring FP16 number)."""
    import torch
    dims = (512, 512, 128)
    print(f"case radix fp16 dims={dims} rs={rs}", flush=True)
    x = oracle.fill_random_real_3d(*dims, SEED, np.float32)
    Xo = oracle.ffts3d_forward(x.astype(np.float64))
    u = torch.from_numpy(x).cuda()
    f0 = _make(dims, "fp16")
    X0 = f0.complex_array()
    f0.forward(u, X0)
    f = _make(dims, "fp16")
    for bw, i in _strided_c2c_passes(f):
        f.set_pass_variant(bw, i, 6 + 32 * rs)
    X = f.complex_array()
    f.forward(u, X)
    l2, mx = oracle.rel_errors(from_gpu(X), Xo)
    d2, _ = oracle.rel_errors(from_gpu(X), from_gpu(X0))
    print(f"[radix fp16] fwd rel_l2={l2:.3e} rel_max={mx:.3e}; vs default schedule rel_l2={d2:.3e}", flush=True)
    assert l2 < 1.5e-3 and mx < 6e-3
    assert d2 < 3e-4
    y = f.real_array()
    f.backward(X, y)
    rt = float(np.linalg.norm(from_gpu(y).astype(np.float64) / np.prod(dims) - x) / np.linalg.norm(x))
    assert rt < 3e-3
    f0.close()
    f.close()


def test_radix_candidates_in_the_tuning_cache(gpu, tmp_path):
    import torch
    dims = (512, 512, 64)
    f = _make(dims)
    cache = str(tmp_path / "radix.cache")
    chosen = f.autotune(warmup=1, iterations=3, cache_path=cache)
    rec = {}
    for ln in open(cache).read().splitlines()[1:]:
        if ln and not ln.startswith("#"):
            k, v = ln.split("=", 1)
            rec[k] = v
    # every strided C2C pass recorded candidates of both alternative schedules
    for var in (32, 34, 38, 64, 70):
        keys = [k for k in rec if k.endswith(f"_CANDIDATE_{var:03d}_VALID")]
        assert len(keys) == 4, (var, keys)
        assert all(rec[k] == "1" for k in keys), (var, [rec[k] for k in keys])
    print(f"[radix tuning] chosen {chosen}", flush=True)
    x = oracle.fill_random_real_3d(*dims, SEED, np.float32)
    X = f.complex_array()
    f.forward(torch.from_numpy(x).cuda(), X)
    l2, mx = oracle.rel_errors(from_gpu(X), oracle.ffts3d_forward(x.astype(np.float64)))
    assert l2 <= TOL_REL_L2[np.float32] and mx <= TOL_REL_MAX[np.float32]
    f.close()


def test_radix_schedules_rejected_where_not_compiled(gpu):
    from fftm_amd import _lib
    L = _lib.lib()
    f = _make((512, 512, 64))
    h = f._handle
    # the z-R2C pass (real axis) and the 64-point passes have no alternative schedules
    assert L.fftm_mi_ffts_set_pass_variant(h, 0, 0, 32) == _lib.NOT_SUPPORTED
    assert L.fftm_mi_ffts_set_pass_variant(h, 0, 1, 32 + 1) == _lib.NOT_SUPPORTED   # shape 1
    assert L.fftm_mi_ffts_set_pass_variant(h, 0, 1, 64 + 2) == _lib.NOT_SUPPORTED   # shape 2, matrix cores
    assert L.fftm_mi_ffts_set_pass_variant(h, 0, 1, 32 + 2) == 0                    # shape 2, schedule 1
    assert L.fftm_mi_ffts_set_pass_variant(h, 0, 1, 96) == _lib.INVALID_VALUE       # no schedule 3
    assert L.fftm_mi_ffts_set_pass_variant(h, 0, 1, 64 + 8 * 3 + 6) == 0
    f.close()
