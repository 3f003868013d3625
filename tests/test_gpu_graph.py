"""The transform, filter and spectral-operator entry points launch kernels only (no
allocation, memset or synchronisation with sync=0), so a caller can capture them in a
HIP graph.  Each graph is replayed several times on new inputs, with eager work and
synchronisation between replays, and every replay is checked bitwise against the
eager result."""

import pytest

pytestmark = pytest.mark.gpu


def test_capture_forward_backward_and_filter(gpu):
    import torch
    import fftm_amd
    dims = (128, 96, 160)
    f = fftm_amd.ffts("float32")
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="", sync=False))
    s = torch.cuda.Stream()
    u = torch.empty(f.real_shape(), device="cuda")
    X, y, z = f.complex_array(), f.real_array(), f.real_array()
    with torch.cuda.stream(s):
        f.set_stream(s)
        u.uniform_(-1, 1)
        f.forward(u, X)
        f.backward(X, y)
        f.filter(u, z, "gaussian", scale=1.0 / (128 * 96 * 160), sigma=6.0)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        f.forward(u, X)
        f.backward(X, y)
        f.filter(u, z, "gaussian", scale=1.0 / (128 * 96 * 160), sigma=6.0)
    e = fftm_amd.ffts("float32")
    e.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    Xe, ye, ze = e.complex_array(), e.real_array(), e.real_array()
    for rep in range(5):
        with torch.cuda.stream(s):
            u.uniform_(-1, 1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        ue = u.clone()
        e.forward(ue, Xe)        # eager, with its own synchronisation
        e.backward(Xe, ye)
        e.filter(ue, ze, "gaussian", scale=1.0 / (128 * 96 * 160), sigma=6.0)
        torch.cuda.synchronize()
        assert torch.equal(X, Xe), f"replay {rep}: forward"
        assert torch.equal(y, ye), f"replay {rep}: backward"
        assert torch.equal(z, ze), f"replay {rep}: filter"
    f.close()
    e.close()
