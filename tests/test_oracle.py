"""Pinning of the CPU oracle (no GPU).

Each check is chosen so that a plausible slip in the oracle -- a dropped term, a wrong
sign or index, a transposed operand, a wrong arranger order -- fails one of them:
published known-answer vectors, an independent textbook routine (numpy.fft/pocketfft),
a brute-force long-double DFT, closed-form spectra of the reference's analytic test
fields, and transform invariants.
"""

import math

import numpy as np
import pytest

import oracle

M64 = (1 << 64) - 1


def splitmix64_py(x):
    """Independent restatement (fft_benchmark_common.h:201-207) in Python integers."""
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


# ---------------------------------------------------------------- inputs

def test_splitmix64_published_vectors():
    """SplitMix64 seeded with 0 (Vigna's reference generator) outputs
    e220a8397b1dcdaf, 6e789e6aa1b965f4, 06c45d188009454f, f88bb8a8724c81ec; the
    reference's splitmix64(x) is one step from state x, so output k = splitmix64(k*golden)."""
    golden = 0x9E3779B97F4A7C15
    expected = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    for k, e in enumerate(expected):
        assert oracle.splitmix64((k * golden) & M64) == e


def test_splitmix64_c_equals_python():
    rng = np.random.default_rng(3)
    for v in rng.integers(0, 2 ** 63, size=200, dtype=np.int64):
        v = int(v) * 2 + 1
        assert oracle.splitmix64(v) == splitmix64_py(v)


def test_unit_random_range_and_precision():
    states = [splitmix64_py(i) for i in range(2000)]
    d = np.array([oracle.unit_random(s, np.float64) for s in states])
    f = np.array([oracle.unit_random(s, np.float32) for s in states], dtype=np.float32)
    assert d.min() >= -1.0 and d.max() < 1.0
    # float variant: 2*float(unit) - 1 in float (fft_benchmark_common.h:209-215)
    for s, fv in zip(states[:200], f[:200]):
        bits = splitmix64_py(s)
        unit = np.float32((bits >> 11) * (1.0 / 9007199254740992.0))
        assert fv == np.float32(np.float32(2.0) * unit - np.float32(1.0))
        assert oracle.unit_random(s, np.float64) == 2.0 * ((bits >> 11) * (1.0 / 9007199254740992.0)) - 1.0


def test_fill_random_real_3d_state_formula():
    """state = seed ^ gx*C1 ^ gy*C2 ^ gz*C3 (fft_benchmark_common.h:226-230), memory [x][y][z]."""
    seed = 0x1234ABCDEF987654
    a = oracle.fill_random_real_3d(3, 4, 5, seed, np.float64)
    for (x, y, z) in [(0, 0, 0), (2, 3, 4), (1, 0, 3), (2, 1, 0)]:
        st = seed ^ ((x * 0xD6E8FEB86659FD93) & M64) ^ ((y * 0xA5A3564E27F886A5) & M64) ^ \
            ((z * 0x9E3779B97F4A7C15) & M64)
        bits = splitmix64_py(st)
        assert a[x, y, z] == 2.0 * ((bits >> 11) * (1.0 / 9007199254740992.0)) - 1.0


def test_fill_random_real_4d_state_formula():
    seed = 0xCAFEBABE13579BDF
    a = oracle.fill_random_real_4d_logical(2, 3, 2, 3, seed, np.float64)
    x, y, z, w = 1, 2, 1, 2
    st = seed ^ ((x * 0xD6E8FEB86659FD93) & M64) ^ ((y * 0xA5A3564E27F886A5) & M64) ^ \
        ((z * 0x9E3779B97F4A7C15) & M64) ^ ((w * 0x94D049BB133111EB) & M64)
    bits = splitmix64_py(st)
    assert a[x, y, z, w] == 2.0 * ((bits >> 11) * (1.0 / 9007199254740992.0)) - 1.0


def test_fftw_input_value():
    """((17 i + 13) & 1023 - 511.5) / 512 (fftw_mpi_benchmark.cpp:215-219)."""
    assert oracle.fftw_input_value(0) == (13 - 511.5) / 512
    assert oracle.fftw_input_value(60) == (((17 * 60 + 13) & 1023) - 511.5) / 512


# ---------------------------------------------------------------- 1D engine

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 12, 16, 17, 30, 64, 97, 128, 243, 1000, 1024, 1344])
def test_fft1d_vs_numpy(n):
    g = np.random.default_rng(n)
    x = g.standard_normal(n) + 1j * g.standard_normal(n)
    for sign, ref in ((-1, np.fft.fft(x)), (+1, np.fft.ifft(x) * n)):
        y = oracle.fft1d(x, sign)
        assert np.max(np.abs(y - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


# ---------------------------------------------------------------- dense R2C / C2R

SHAPES = [(8,), (13,), (12, 10, 8), (7, 5, 9), (16, 16, 16), (3, 4, 5, 6), (6, 5, 4, 8), (1, 1, 2), (2, 1, 7)]


@pytest.mark.parametrize("shape", SHAPES)
def test_rfftn_vs_numpy(shape):
    x = np.random.default_rng(len(shape)).uniform(-1, 1, size=shape)
    X = oracle.rfftn(x)
    Xn = np.fft.rfftn(x)
    assert np.max(np.abs(X - Xn)) <= 1e-14 * np.max(np.abs(Xn))
    y = oracle.irfftn(X, shape)
    assert np.max(np.abs(y / x.size - x)) <= 1e-14


@pytest.mark.parametrize("shape", [(4, 6, 5), (3, 3, 4, 4), (5,), (2, 7), (4, 3, 2, 5)])
def test_rfftn_vs_brute_force(shape):
    x = np.random.default_rng(9).uniform(-1, 1, size=shape)
    B = oracle.brute_r2c(x)
    X = oracle.rfftn(x)
    assert np.max(np.abs(X - B)) <= 2e-14 * max(1.0, np.max(np.abs(B)))


def test_brute_force_vs_numpy():
    """The brute-force DFT itself against pocketfft (pins its sign and index order)."""
    x = np.random.default_rng(1).uniform(-1, 1, size=(3, 5, 4))
    assert np.max(np.abs(oracle.brute_r2c(x) - np.fft.rfftn(x))) <= 1e-13


def test_irfftn_non_hermitian_semantics():
    """C2R of an arbitrary half spectrum = numpy.irfftn * N (imag of DC/Nyquist ignored)."""
    g = np.random.default_rng(4)
    for shape in [(6, 8, 10), (5, 7), (9,), (4, 3, 2, 6)]:
        h = shape[:-1] + (shape[-1] // 2 + 1,)
        X = g.standard_normal(h) + 1j * g.standard_normal(h)
        y = oracle.irfftn(X, shape)
        ref = np.fft.irfftn(X, s=shape, axes=tuple(range(len(shape)))) * np.prod(shape)
        assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))


def _sample_3d(nx, ny, nz):
    h = [2 * math.pi / n for n in (nx, ny, nz)]
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    x, y, z = h[0] * i, h[1] * j, h[2] * k
    return np.sin(x) + 0.5 * np.cos(y) - 0.25 * np.sin(2 * z) + 0.125 * np.sin(x + y - z)


@pytest.mark.parametrize("dims", [(12, 10, 8), (16, 16, 16), (7, 5, 9), (32, 24, 20)])
def test_closed_form_spectrum_3d(dims):
    """sample_value of versioned_ffts_3d_test_driver.h:62-67 has a sparse exact spectrum:
    (1,0,0) = -iN/2, (nx-1,0,0) = iN/2, (0,1,0) = (0,ny-1,0) = N/4, (0,0,2) = iN/8,
    (nx-1,ny-1,1) = iN/16, every other bin 0."""
    nx, ny, nz = dims
    N = nx * ny * nz
    X = oracle.ffts3d_forward(_sample_3d(*dims))
    E = np.zeros_like(X)
    E[1, 0, 0] += -0.5j * N
    E[nx - 1, 0, 0] += 0.5j * N
    E[0, 1, 0] += 0.25 * N
    E[0, ny - 1, 0] += 0.25 * N
    E[0, 0, 2] += 0.125j * N
    E[nx - 1, ny - 1, 1] += 1j * N / 16
    assert np.max(np.abs(X - E)) <= 1e-12 * N


@pytest.mark.parametrize("dims", [(12, 10, 8, 10), (8, 6, 6, 8)])
def test_closed_form_spectrum_4d(dims):
    """sample_value of versioned_ffts_4d_test_driver.h:64-70 (w is the halved axis):
    (1,0,0,0) = -iN/2, (nx-1,0,0,0) = iN/2, (0,1,0,0) = (0,ny-1,0,0) = N/4,
    (0,0,2,0) = iN/8, (0,0,nz-2,0) = -iN/8, (0,0,0,1) = N/16, (1,1,nz-1,1) = -iN/32."""
    nx, ny, nz, nw = dims
    N = nx * ny * nz * nw
    h = [2 * math.pi / n for n in dims]
    i, j, k, l = np.meshgrid(*[np.arange(n) for n in dims], indexing="ij")
    x, y, z, w = h[0] * i, h[1] * j, h[2] * k, h[3] * l
    u = np.sin(x) + 0.5 * np.cos(y) - 0.25 * np.sin(2 * z) + 0.125 * np.cos(w) + 0.0625 * np.sin(x + y - z + w)
    X = oracle.rfftn(u)
    E = np.zeros_like(X)
    E[1, 0, 0, 0] += -0.5j * N
    E[nx - 1, 0, 0, 0] += 0.5j * N
    E[0, 1, 0, 0] += 0.25 * N
    E[0, ny - 1, 0, 0] += 0.25 * N
    E[0, 0, 2, 0] += 0.125j * N
    E[0, 0, nz - 2, 0] += -0.125j * N
    E[0, 0, 0, 1] += N / 16
    E[1, 1, nz - 1, 1] += -1j * N / 32
    assert np.max(np.abs(X - E)) <= 1e-12 * N


def test_invariants_parseval_linearity_shift():
    g = np.random.default_rng(5)
    shape = (10, 12, 14)
    a, b = g.uniform(-1, 1, shape), g.uniform(-1, 1, shape)
    A, B = oracle.rfftn(a), oracle.rfftn(b)
    # linearity
    C = oracle.rfftn(2.5 * a - 0.75 * b)
    assert np.max(np.abs(C - (2.5 * A - 0.75 * B))) <= 1e-12
    # Parseval with Hermitian weights on the halved axis (interior bins count twice)
    w = np.full(A.shape[-1], 2.0)
    w[0] = 1.0
    if shape[-1] % 2 == 0:
        w[-1] = 1.0
    lhs = np.sum(a * a)
    rhs = np.sum(w * np.abs(A) ** 2) / a.size
    assert abs(lhs - rhs) <= 1e-12 * lhs
    # circular shift along x by s multiplies by exp(-2 pi i kx s / nx)
    s = 3
    S = oracle.rfftn(np.roll(a, s, axis=0))
    kx = np.arange(shape[0])[:, None, None]
    assert np.max(np.abs(S - A * np.exp(-2j * np.pi * kx * s / shape[0]))) <= 1e-11


# ---------------------------------------------------------------- plans

def test_plan_exec_strided_layouts_vs_numpy():
    """Advanced layouts: batch-contiguous strided lines, padded embeds, batch > 1."""
    g = np.random.default_rng(6)
    n, batch = 12, 5
    x = g.uniform(-1, 1, (n, batch))              # line stride = batch, dist 1
    out = oracle.plan_exec(oracle.R2C, [n], [1], batch, 1, [1], batch, 1, batch, x.reshape(-1))
    assert np.allclose(out.reshape(n // 2 + 1, batch), np.fft.rfft(x, axis=0), atol=1e-13)
    # 2D R2C with padded output rows (onembed > extent) and idist/odist
    nx, ny = 6, 8
    xin = g.uniform(-1, 1, (3, nx, ny))
    out = oracle.plan_exec(oracle.R2C, [nx, ny], [nx, ny], 1, nx * ny, [nx, 7], 1, nx * 7, 3, xin.reshape(-1),
                           out=np.zeros(3 * nx * 7, dtype=np.complex128))
    got = out.reshape(3, nx, 7)[:, :, :ny // 2 + 1]
    assert np.allclose(got, np.fft.rfftn(xin, axes=(1, 2)), atol=1e-13)
    # C2CB with istride 2 (every other complex element)
    c = g.standard_normal((4, 2 * 10)) + 1j * g.standard_normal((4, 2 * 10))
    out = oracle.plan_exec(oracle.C2CB, [10], [10], 2, 20, [10], 1, 10, 4, c.reshape(-1))
    assert np.allclose(out.reshape(4, 10), np.fft.ifft(c[:, ::2], axis=1) * 10, atol=1e-12)


# ---------------------------------------------------------------- ffts 4D layouts

@pytest.mark.parametrize("dims", [(4, 6, 8, 10), (5, 3, 4, 6), (8, 8, 8, 8)])
def test_ffts4d_pipelines_vs_numpy(dims):
    """The stage-by-stage ffts 4D restatement (plans + arranger transposes) equals numpy's
    rfftn of the logical field, read out in each strategy's memory layout."""
    nx, ny, nz, nw = dims
    u = oracle.fill_random_real_4d_logical(nx, ny, nz, nw, 0x1234ABCDEF987654, np.float64)
    U = np.fft.rfftn(u)  # logical (x, y, z, kw)
    pencil_in = oracle.to_memory(u, oracle.A0123)
    assert pencil_in.shape == (nw, nz, ny, nx)
    Xp = oracle.ffts4d_forward(pencil_in, "pencil", *dims)
    assert np.allclose(Xp, np.transpose(U, (1, 2, 3, 0)), atol=1e-11)        # [y][z][wh][x]
    slab_in = oracle.to_memory(u, oracle.A3210)
    assert slab_in.shape == (nx, ny, nz, nw)
    Xs = oracle.ffts4d_forward(slab_in, "slab", *dims)
    assert np.allclose(Xs, U, atol=1e-11)                                     # [x][y][z][wh]
    Xn = oracle.native_xzwy_forward(slab_in, *dims)
    assert np.allclose(Xn, np.transpose(U, (1, 2, 3, 0)), atol=1e-11)
    N = u.size
    assert np.allclose(oracle.ffts4d_backward(Xp, "pencil", *dims) / N, pencil_in, atol=1e-13)
    assert np.allclose(oracle.ffts4d_backward(Xs, "slab", *dims) / N, slab_in, atol=1e-13)
    assert np.allclose(oracle.native_xzwy_backward(Xn, *dims) / N, slab_in, atol=1e-13)


def test_arranger_memory_orders():
    """custom_index_fast_arranger<a,b,c,d>: first listed logical index is fastest
    (ffts.hpp:116-155; array_arrangers.h:29-57); stage arrays of the pencil strategy."""
    lg = np.arange(2 * 3 * 4 * 5).reshape(2, 3, 4, 5)     # logical (x, y, z, w)
    m = oracle.to_memory(lg, oracle.A0123)
    assert m.shape == (5, 4, 3, 2) and m[4, 3, 2, 1] == lg[1, 2, 3, 4]
    m = oracle.to_memory(lg, oracle.A0213)                # logical (x, z, w, y) stage2 -> [y][w][z][x]
    assert m.shape == (5, 3, 4, 2)
    assert np.array_equal(oracle.from_memory(m, oracle.A0213), lg)
    # direct_transpose <0,1,3,2>: out(x, y, w, z) = in(x, y, z, w)
    t = oracle.direct_transpose(lg, (0, 1, 3, 2))
    assert t[1, 2, 4, 3] == lg[1, 2, 3, 4]


def test_staged_transposes_are_memory_permutations():
    """The memory-array permutations the staged GPU path uses equal the reference's
    arranger transposes (direct_transpose_4d.h:42-120) for the pencil strategy."""
    nx, ny, nz, nwh = 3, 4, 5, 6
    st0_log = np.arange(nx * ny * nz * nwh).reshape(nx, ny, nz, nwh)     # xyzw
    st0 = oracle.to_memory(st0_log, oracle.A0123)                       # [wh][z][y][x]
    st1 = oracle.to_memory(oracle.direct_transpose(st0_log, (0, 1, 3, 2)), oracle.A0123)
    assert np.array_equal(st1, np.transpose(st0, (1, 0, 2, 3)))          # [z][wh][y][x]
    st1_log = oracle.from_memory(st1, oracle.A0123)
    st2 = oracle.to_memory(oracle.direct_transpose(st1_log, (0, 3, 2, 1)), oracle.A0213)
    assert np.array_equal(st2, np.transpose(st1, (2, 0, 1, 3)))          # [y][z][wh][x]
    st2_log = oracle.from_memory(st2, oracle.A0213)
    out = oracle.to_memory(oracle.direct_transpose(st2_log, (3, 1, 2, 0)), oracle.A3210)
    assert np.array_equal(out, st2)                                      # identity in memory
    back1 = oracle.to_memory(oracle.direct_transpose(st2_log, (0, 3, 2, 1)), oracle.A0123)
    assert np.array_equal(back1, np.transpose(st2, (1, 2, 0, 3)))
    # slab: both transposes are identities in memory
    s0_log = st0_log
    s0 = oracle.to_memory(s0_log, oracle.A3210)
    s1 = oracle.to_memory(oracle.direct_transpose(s0_log, (2, 3, 0, 1)), oracle.A1032)
    assert np.array_equal(s1, s0)
    s1_log = oracle.from_memory(s1, oracle.A1032)
    o = oracle.to_memory(oracle.direct_transpose(s1_log, (3, 0, 1, 2)), oracle.A2103)
    assert np.array_equal(o, s0)


def test_rel_errors_metric():
    e = np.array([1.0, -2.0, 3.0])
    a = e + np.array([0.0, 1e-3, 0.0])
    rms, mx = oracle.rel_errors(a, e)
    assert abs(rms - 1e-3 / math.sqrt(14)) < 1e-15 and abs(mx - 1e-3 / 3) < 1e-15


def test_oracle_spectral_operators_closed_forms():
    """spectral_operator_logical + filter_logical against closed forms: the periodic
    Poisson problem and its gradient (versioned_ffts_3d_test_driver.h:187-209), first
    derivatives of a sine product, Gaussian- and band-filtered plane waves."""
    import math
    dims = (16, 12, 20)
    N = float(np.prod(dims))
    L = [np.arange(n) * 2 * math.pi / n for n in dims]
    X, Y, Z = np.meshgrid(*L, indexing="ij")
    s = [np.sin(X), np.sin(Y), np.sin(Z)]
    c = [np.cos(X), np.cos(Y), np.cos(Z)]
    u = s[0] * s[1] * s[2]
    rhs = -3.0 * u
    sol = oracle.filter_logical(rhs, oracle.spectral_operator_logical(dims, "poisson", 1.0 / N))
    assert np.abs(sol - u).max() < 1e-12
    for a in range(3):
        exact = np.prod([c[b] if b == a else s[b] for b in range(3)], axis=0)
        d = oracle.filter_logical(u, oracle.spectral_operator_logical(dims, "derivative", 1.0 / N, axis=a))
        g = oracle.filter_logical(rhs, oracle.spectral_operator_logical(dims, "poisson_gradient", 1.0 / N, axis=a))
        assert np.abs(d - exact).max() < 1e-12
        assert np.abs(g - exact).max() < 1e-12
    waves = [(1, 2, 3), (0, 5, 4), (7, 0, 9)]
    src = sum(np.cos(a * X + b * Y + cc * Z) for a, b, cc in waves)
    g = oracle.filter_logical(src, oracle.spectral_operator_logical(dims, "gaussian", 1.0 / N, sigma=3.0))
    ge = sum(math.exp(-(a * a + b * b + cc * cc) / 18.0) * np.cos(a * X + b * Y + cc * Z) for a, b, cc in waves)
    assert np.abs(g - ge).max() < 1e-12
    bp = oracle.filter_logical(src, oracle.spectral_operator_logical(dims, "bandpass", 1.0 / N, band=(4.0, 7.0)))
    be = np.cos(5 * Y + 4 * Z)  # |k| = 6.4 only
    assert np.abs(bp - be).max() < 1e-12


# ---------------------------------------------------------------- wavenumbers, radial shells

def test_wavenumber_map_vs_numpy_fftfreq():
    """k = i <= n/2 ? i : i - n with the Nyquist bin positive (versioned_ffts_3d_test_driver.h:
    140-142): numpy's fftfreq(n) * n, with its negative Nyquist flipped; the halved axis is
    0..n/2."""
    for n in (1, 2, 3, 7, 8, 12, 33, 64):
        ref = np.rint(np.fft.fftfreq(n) * n).astype(np.int64)
        if n % 2 == 0:
            ref[n // 2] = n // 2
        assert np.array_equal(oracle.wavenumbers(n, False), ref)
        assert np.array_equal(oracle.wavenumbers(n, True), np.arange(n // 2 + 1))


def _shell_isqrt(dims, nshells):
    """Radial shell index from Python's exact integer square root: for integers a, b > 0,
    floor(sqrt(a / b)) = isqrt(a // b) (s^2 <= a / b iff s^2 <= floor(a / b)), so
    floor(S |k| / kmax) = isqrt(S^2 |k|^2 // kmax^2), clamped to S - 1, with
    kmax^2 = sum (n_a / 2)^2 and |k|^2 over the reference wavenumber map."""
    d = len(dims)
    ks = []
    for a, n in enumerate(dims):
        if a == d - 1:
            ks.append(list(range(n // 2 + 1)))
        else:
            ks.append([i if i <= n // 2 else i - n for i in range(n)])
    kmax2 = sum((n // 2) ** 2 for n in dims)
    grids = np.meshgrid(*[np.array(k, dtype=np.int64) for k in ks], indexing="ij")
    k2 = sum(g * g for g in grids)
    if kmax2 == 0:  # all-ones sizes: the only bin is k = 0
        return np.zeros(k2.shape, dtype=np.int64)
    table = {int(v): min(math.isqrt(nshells * nshells * int(v) // kmax2), nshells - 1)
             for v in np.unique(k2)}
    return np.vectorize(table.__getitem__, otypes=[np.int64])(k2)


def test_shell_index_hand_enumerated():
    """A 4x4x4 cube with 4 shells by hand: kmax^2 = 12, shell = floor(4 |k| / sqrt 12),
    so |k|^2 = 0 -> 0, 1 -> 1, 2 -> 1, 3 -> 2 (exactly 4*sqrt3/sqrt12 = 2, the boundary),
    4 -> 2, 5 -> 2, 6 -> 2, 8 -> 3, 9 -> 3, 12 -> 3 (clamped from 4).  The 6^3 cube with 3
    shells has a boundary bin: kmax^2 = 27, |k|^2 = 3 gives 3 sqrt3 / sqrt27 = 1 exactly,
    shell 1 (a floating floor(sqrt(3) * 3 / sqrt(27)) gives 0)."""
    sh = oracle.shell_index_logical((4, 4, 4), 4)
    expect = {0: 0, 1: 1, 2: 1, 3: 2, 4: 2, 5: 2, 6: 2, 8: 3, 9: 3, 12: 3}
    k = [np.array([0, 1, 2, -1]), np.array([0, 1, 2, -1]), np.array([0, 1, 2])]
    seen = set()
    for i in range(4):
        for j in range(4):
            for l in range(3):
                k2 = int(k[0][i] ** 2 + k[1][j] ** 2 + k[2][l] ** 2)
                assert sh[i, j, l] == expect[k2], (i, j, l, k2)
                seen.add(k2)
    assert seen == set(expect)
    # 4D, 2x2x2x2 with 2 shells: kmax^2 = 4, shell = floor(2 |k| / 2) = floor(|k|) clamped to 1
    sh4 = oracle.shell_index_logical((2, 2, 2, 2), 2)
    assert sh4[0, 0, 0, 0] == 0 and sh4[1, 0, 0, 0] == 1 and sh4[1, 1, 1, 1] == 1
    sh6 = oracle.shell_index_logical((6, 6, 6), 3)
    assert sh6[1, 1, 1] == 1 and sh6[1, 0, 0] == 0 and sh6[2, 2, 2] == 2 and sh6[3, 3, 3] == 2


@pytest.mark.parametrize("dims,nshells", [
    ((4, 4, 4), 4), ((32, 32, 32), 32), ((64, 64, 64), 32), ((256, 256, 256), 32),
    ((12, 10, 8), 32), ((12, 10, 8), 7), ((16, 8, 10, 12), 32), ((8, 6, 10, 12), 32),
    ((32, 32, 32, 32), 32), ((20, 20, 20), 10), ((6, 6, 6), 3), ((30, 30, 30), 15),
    ((10, 10, 10, 10), 5), ((1, 1, 1), 4),
])
def test_shell_index_exact_integer_arithmetic(dims, nshells):
    """oracle.shell_index_logical equals the isqrt answer at every bin, boundary bins (where
    S^2 |k|^2 / kmax^2 is a perfect square) included; (20, 20, 20) with 10 shells has 20 of
    them, (30, 30, 30) with 15 has 68."""
    got = oracle.shell_index_logical(dims, nshells)
    exp = _shell_isqrt(dims, nshells)
    assert got.shape == exp.shape
    bad = int(np.count_nonzero(got != exp))
    assert bad == 0, f"{bad} bins differ"
