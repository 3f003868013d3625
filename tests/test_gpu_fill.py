"""The device fill of the benchmark harness's input (fill_random_real_3d_functor /
fill_random_real_4d_functor with splitmix64, source/tests/detail/fft_benchmark_common.h:
201-257), called through the C ABI (fftm_mi_fill_random_real), against the oracle's
restatement: bitwise equal in FP32 and FP64, in the 3D, 4D slab and 4D pencil real
layouts, and with a slab origin (start offsets) equal to the same window of the whole
field."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SEEDS = (0x1234ABCDEF987654, 0xCAFEBABE13579BDF)


def _c_strides(dims):
    s = [1] * len(dims)
    for a in range(len(dims) - 2, -1, -1):
        s[a] = s[a + 1] * dims[a + 1]
    return s


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [(12, 10, 8), (64, 32, 128), (7, 5, 3)])
@pytest.mark.parametrize("seed", SEEDS)
def test_fill_3d_bitwise(gpu, dtype, dims, seed):
    import torch
    import fftm_amd
    print(f"case fill3d dims={dims} dtype={np.dtype(dtype).name} seed={seed:#x}", flush=True)
    t = torch.empty(dims, dtype=torch.float32 if dtype == np.float32 else torch.float64, device="cuda")
    fftm_amd.fill_random_real(t, dims, seed, sync=True)
    expect = oracle.fill_random_real_3d(*dims, seed, dtype)
    got = t.cpu().numpy()
    assert got.tobytes() == expect.tobytes()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout", ["slab", "pencil"])
def test_fill_4d_bitwise(gpu, dtype, layout):
    import torch
    import fftm_amd
    dims = (16, 8, 12, 10)  # logical (x, y, z, w)
    print(f"case fill4d dims={dims} layout={layout} dtype={np.dtype(dtype).name}", flush=True)
    logical = oracle.fill_random_real_4d_logical(*dims, SEEDS[0], dtype)
    if layout == "slab":     # memory [x][y][z][w]
        strides = _c_strides(dims)
        expect = logical
    else:                    # memory [w][z][y][x]
        strides = [1, dims[0], dims[0] * dims[1], dims[0] * dims[1] * dims[2]]
        expect = np.ascontiguousarray(logical.transpose(3, 2, 1, 0))
    t = torch.empty(expect.shape, dtype=torch.float32 if dtype == np.float32 else torch.float64, device="cuda")
    fftm_amd.fill_random_real(t, dims, SEEDS[0], strides, sync=True)
    assert t.cpu().numpy().tobytes() == expect.tobytes()


def test_fill_slab_origin(gpu):
    """A rank's slab [x0:x0+nx] of the global field: fill with start offsets equals the
    same window of the whole field (the functor's start_x/start_y/start_z)."""
    import torch
    import fftm_amd
    whole = (20, 12, 10)
    start = (6, 2, 0)
    dims = (8, 10, 10)
    full = oracle.fill_random_real_3d(*whole, SEEDS[1], np.float64)
    t = torch.empty(dims, dtype=torch.float64, device="cuda")
    fftm_amd.fill_random_real(t, dims, SEEDS[1], start=start, sync=True)
    window = full[6:14, 2:12, 0:10]
    assert t.cpu().numpy().tobytes() == np.ascontiguousarray(window).tobytes()


def test_fill_512_cubed_matches_oracle(gpu):
    """The bench field itself (512^3 FP32, seed of test_benchmark_ffts_3D.cu:83)."""
    import torch
    import fftm_amd
    dims = (512, 512, 512)
    print("case fill3d 512^3 fp32", flush=True)
    t = torch.empty(dims, dtype=torch.float32, device="cuda")
    fftm_amd.fill_random_real(t, dims, SEEDS[0], sync=True)
    expect = oracle.fill_random_real_3d(*dims, SEEDS[0], np.float32)
    assert np.array_equal(t.cpu().numpy(), expect)
