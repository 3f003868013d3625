"""Tuning with the reference's autotune cache protocol (source/fftm_autotune.hpp:126-169,
fftm_autotune_measure.hpp:293-400): every variant of every pass, and of the fused filter
pass, is timed; the winners are bitwise equal to the defaults; the KEY=VALUE cache holds
the schema, signature, per-candidate and winner records; and a new handle applies the
cache, by load_tuning or through FFTM_MI_AUTOTUNE_CACHE at creation."""

import os

import pytest

pytestmark = pytest.mark.gpu


def _read_cache(path):
    rec = {}
    with open(path) as fh:
        lines = fh.read().splitlines()
    assert lines[0].startswith("# FFTM C++ autotune cache")
    for ln in lines[1:]:
        if ln and not ln.startswith("#"):
            k, v = ln.split("=", 1)
            rec[k] = v
    return rec


def _make(dims, dt):
    import fftm_amd
    f = fftm_amd.ffts(dt)
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    return f


@pytest.mark.parametrize("dt,dims", [("float32", (128, 64, 256)), ("float64", (64, 64, 128))])
def test_autotune_cache_round_trip(gpu, tmp_path, dt, dims):
    import torch
    from fftm_amd import _lib
    L = _lib.lib()
    print(f"case tuning {dims} {dt}", flush=True)
    f = _make(dims, dt)
    u = torch.rand(f.real_shape(), dtype=f.real, device="cuda") * 2 - 1
    X0, y0 = f.complex_array(), f.real_array()
    f.forward(u, X0)
    f.backward(X0, y0)
    g0 = f.real_array()
    f.filter(u, g0, "poisson", scale=-1.0 / u.numel())
    # variant numbers: shapes 0..7 plus 8 * the tile order (0..3); anything else is rejected
    assert L.fftm_mi_ffts_set_pass_variant(f._handle, 0, 0, 32) != 0
    assert L.fftm_mi_ffts_set_pass_variant(f._handle, 0, 0, -1) != 0
    cache = str(tmp_path / "tile.cache")
    chosen = f.autotune(warmup=1, iterations=3, cache_path=cache)
    nf, nb = f.num_passes()
    assert len(chosen) == nf + nb
    assert all(0 <= v < 32 for v in chosen.values())
    X, y = f.complex_array(), f.real_array()
    f.forward(u, X)
    f.backward(X, y)
    assert torch.equal(X, X0) and torch.equal(y, y0), "tuned variants changed the result"
    g = f.real_array()
    f.filter(u, g, "poisson", scale=-1.0 / u.numel())
    assert torch.equal(g, g0), "the tuned filter variant changed the result"

    rec = _read_cache(cache)
    assert rec["FFTM_AUTOTUNE_SCHEMA"] == "2"
    assert rec["FFTM_AUTOTUNE_HARDWARE_SIGNATURE"].startswith("v2:")
    winners = {k: v for k, v in rec.items() if k.startswith("FFTM_MI_TILE_") and "_CANDIDATE_" not in k
               and not k.endswith("_WINNER_MEDIAN_MS")}
    # one record per pass, plus the fused filter pass (these layouts fuse it)
    assert len(winners) == nf + nb + 1
    assert any(k.endswith("_FILTER") for k in winners)
    for k, v in winners.items():
        assert rec[k + "_CANDIDATE_%03d_VALID" % int(v)] == "1"
        assert float(rec[k + "_WINNER_MEDIAN_MS"]) > 0
    # the default shape of every pass kernel was measured in every tile order
    for k in winners:
        assert k + "_CANDIDATE_000_VALID" in rec
    assert any("_CANDIDATE_008_VALID" in k for k in rec)

    # a fresh handle: defaults until the cache is loaded, then the tuned variants
    g2 = _make(dims, dt)
    assert g2.load_tuning(cache) == nf + nb + 1
    assert g2.pass_variants() == chosen
    g2.close()
    os.environ["FFTM_MI_AUTOTUNE_CACHE"] = cache
    try:
        h = _make(dims, dt)
    finally:
        del os.environ["FFTM_MI_AUTOTUNE_CACHE"]
    assert h.pass_variants() == chosen
    X2, y2 = h.complex_array(), h.real_array()
    h.forward(u, X2)
    h.backward(X2, y2)
    assert torch.equal(X2, X0) and torch.equal(y2, y0)
    g3 = h.real_array()
    h.filter(u, g3, "poisson", scale=-1.0 / u.numel())
    assert torch.equal(g3, g0)
    h.close()
    f.close()


def test_cache_records_are_keyed_by_creation_flags(gpu, tmp_path):
    """Tuning records are per pass index, and creation flags change the pass list (a plane-
    fused 32^4 pencil handle has 3 forward passes, an unfused one 4): a cache written by the
    fused handle applies nothing to an unfused handle of the same size, and everything to
    another fused one."""
    import ctypes
    import torch
    import fftm_amd
    from fftm_amd import _lib
    L = _lib.lib()
    dims = (32, 32, 32, 32)
    print(f"case tuning flags {dims}", flush=True)
    f = fftm_amd.ffts("float32", strategy_4d=fftm_amd.strategy_4d_pencil_pencil())
    f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
    u = torch.rand(f.real_shape(), dtype=f.real, device="cuda") * 2 - 1
    cache = str(tmp_path / "flags.cache")
    chosen = f.autotune(warmup=1, iterations=2, cache_path=cache)
    nf, nb = f.num_passes()
    assert len(chosen) == nf + nb
    del u
    n = _lib.i64_array(dims)
    for flags, expect in ((_lib.FLAG_NO_PLANE_FUSION, 0), (0, None)):
        h = ctypes.c_void_p()
        _lib.check(L.fftm_mi_ffts_create(_lib.LAYOUT_4D_PENCIL, n, _lib.F32, flags, ctypes.byref(h)), "create")
        try:
            fp, bp = ctypes.c_int32(), ctypes.c_int32()
            _lib.check(L.fftm_mi_ffts_num_passes(h, ctypes.byref(fp), ctypes.byref(bp)), "num_passes")
            applied = ctypes.c_int32(-1)
            _lib.check(L.fftm_mi_ffts_load_tuning(h, cache.encode(), ctypes.byref(applied)), "load_tuning")
            if flags:
                assert fp.value == nf + 1, "the unfused handle has one more forward pass"
                assert applied.value == 0
            else:
                assert applied.value >= nf + nb
        finally:
            L.fftm_mi_ffts_destroy(h)
    f.close()
