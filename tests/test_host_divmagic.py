"""The pass kernels divide a column index by its batch extents with a multiply and a shift
(div_magic in fftm_amd/csrc/mi_host.hip, div_col in mi_pass.hpp): for 1 <= d < 2^31,
l = ceil(log2 d), m = ceil(2^(31 + l) / d), every n < 2^31 gives n // d = (n * m) >> (31 + l)
with n * m < 2^63.  This checks those constants against exact integer division over the
divisors the transforms use (batch extents up to 4096, products of two of them) and the
extreme ones, at the dividends where a rounding error would show first (multiples of d and
their neighbours, the top of the range)."""

import random

import pytest


def div_magic(d):
    l = 0
    while (1 << l) < d:
        l += 1
    sh = 31 + l
    m = ((1 << sh) + d - 1) // d
    return m, sh


def dividends(d, rng):
    top = (1 << 31) - 1
    out = {0, 1, top, top - 1, d - 1, d, d + 1}
    q = top // d
    for k in (1, 2, q - 1, q):
        if k > 0:
            out.update({k * d - 1, k * d, k * d + 1})
    out.update(rng.randrange(1 << 31) for _ in range(64))
    return [n for n in out if 0 <= n <= top]


def divisors():
    ds = set(range(1, 4098)) | {257, 513, 1025, 2049, 65, 129, 33, 17}
    ds |= {a * b for a in (17, 33, 65, 129, 257, 513, 1025) for b in (16, 32, 64, 128, 256, 512, 1024)}
    ds |= {(1 << 31) - 1, (1 << 30) + 1, 1 << 30, (1 << 20) + 1, 3 ** 19}
    return sorted(ds)


@pytest.mark.parametrize("chunk", range(4))
def test_multiply_shift_division_is_exact(chunk):
    rng = random.Random(1234 + chunk)
    ds = divisors()[chunk::4]
    for d in ds:
        m, sh = div_magic(d)
        assert m <= 1 << 32
        for n in dividends(d, rng):
            assert n * m < 1 << 63
            assert (n * m) >> sh == n // d, (n, d, m, sh)


def test_a_shift_one_short_fails():
    """The bound is tight enough that a constant with one bit less precision is wrong for
    some divisor: the test above would catch such a change."""
    bad = 0
    for d in (3, 7, 257, 513, 1025, 4097, 129 * 512):
        l = (d - 1).bit_length()
        sh = 30 + l
        m = ((1 << sh) + d - 1) // d
        top = (1 << 31) - 1
        for n in (top, top - 1, (top // d) * d - 1):
            if (n * m) >> sh != n // d:
                bad += 1
    assert bad > 0
