"""fftm_mi_ffts_shell_gradient_noalloc: the shell-coefficient gradient with its partial
sums in the library's static device scratch instead of a buffer it allocates, for callers
that cannot allocate.  Checked through the C ABI against fftm_mi_ffts_shell_gradient on
the same inputs (FP64 accumulation in both; the LDS atomics order the additions
differently from run to run, so the comparison is to a relative 1e-12), on two calls in a
row, the second with a smaller problem and fewer shells, and in FP32 and FP64."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grads(f, G, X, nshells, dtype):
    import torch
    from fftm_amd import _lib
    L = _lib.lib()
    tdt = torch.float32 if dtype == "float32" else torch.float64
    g_ref = torch.full((nshells,), float("nan"), dtype=tdt, device="cuda")
    g_new = torch.full((nshells,), float("nan"), dtype=tdt, device="cuda")
    _lib.check(L.fftm_mi_ffts_shell_gradient(f._handle, G.data_ptr(), X.data_ptr(), g_ref.data_ptr(), nshells, 1),
               "fftm_mi_ffts_shell_gradient")
    _lib.check(L.fftm_mi_ffts_shell_gradient_noalloc(f._handle, G.data_ptr(), X.data_ptr(), g_new.data_ptr(),
                                                     nshells, 1), "fftm_mi_ffts_shell_gradient_noalloc")
    return g_ref.double().cpu().numpy(), g_new.double().cpu().numpy()


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_noalloc_matches_allocating_entry_point(gpu, dtype):
    import torch
    import fftm_amd
    tol = 1e-12 if dtype == "float64" else 1e-6  # FP64 sums of FP32 products, written as FP32
    cases = [((64, 48, 40), 32), ((12, 10, 8), 7)]  # second call: smaller input, fewer shells
    for dims, nshells in cases:
        print(f"case noalloc {dtype} dims={dims} shells={nshells}", flush=True)
        f = fftm_amd.ffts(dtype)
        f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
        g = torch.Generator(device="cuda").manual_seed(sum(dims))
        cdt = torch.complex64 if dtype == "float32" else torch.complex128
        shape = tuple(f.complex_shape())
        X = torch.randn(shape, generator=g, device="cuda", dtype=cdt)
        G = torch.randn(shape, generator=g, device="cuda", dtype=cdt)
        ref, new = _grads(f, G, X, nshells, dtype)
        assert np.all(np.isfinite(new))
        err = np.abs(new - ref).max() / np.abs(ref).max()
        print(f"noalloc vs allocating: rel-max {err:.3e}", flush=True)
        assert err <= tol
        f.close()


def test_noalloc_rejects_oversized_requests(gpu):
    import torch
    import fftm_amd
    from fftm_amd import _lib
    L = _lib.lib()
    f = fftm_amd.ffts("float32")
    f.init(8, 8, 8, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    X = f.complex_array()
    g = torch.empty(2048, device="cuda")
    rc = L.fftm_mi_ffts_shell_gradient_noalloc(f._handle, X.data_ptr(), X.data_ptr(), g.data_ptr(), 2048, 1)
    assert rc == _lib.INVALID_VALUE
    assert "1..1024" in _lib.last_error()
    assert L.fftm_mi_ffts_shell_gradient_noalloc(None, X.data_ptr(), X.data_ptr(), g.data_ptr(), 4, 1) \
        == _lib.INVALID_PLAN
    f.close()
