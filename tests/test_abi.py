"""C-ABI checks that need no GPU: the library loads, exports every symbol the header
declares, and validates descriptors on the host with hipfftResult-numbered errors."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fftm_mi.h")


def _lib():
    from fftm_amd import _lib
    return _lib


def _declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fftm_mi_\w+)\s*\(", txt)))


def test_header_declares_expected_entry_points():
    syms = _declared_symbols()
    for s in ["fftm_mi_plan_create", "fftm_mi_plan_exec", "fftm_mi_plan_destroy", "fftm_mi_plan_get_work_size",
              "fftm_mi_plan_set_work_area", "fftm_mi_plan_set_stream", "fftm_mi_transpose_4d",
              "fftm_mi_ffts_forward", "fftm_mi_ffts_backward", "fftm_mi_last_error"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    L = _lib().lib()
    for s in _declared_symbols():
        assert hasattr(L, s), f"{s} declared in include/fftm_mi.h but not exported"
        assert s in _lib().SIGNATURES, f"{s} has no ctypes signature"


def test_nm_exports_match_header():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", _lib().LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fftm_mi_\w+)", out))
    assert set(_declared_symbols()) <= exported


def _desc(**kw):
    L = _lib()
    d = L.PlanDesc()
    d.dir = kw.get("dir", L.R2C)
    d.precision = kw.get("precision", L.F32)
    n = kw.get("n", [16, 16, 16])
    d.rank = kw.get("rank", len(n))
    d.default_layout = kw.get("default_layout", 1)
    for k, v in enumerate(n[:3]):
        d.n[k] = v
    for k, v in enumerate(kw.get("inembed", [])):
        d.inembed[k] = v
    for k, v in enumerate(kw.get("onembed", [])):
        d.onembed[k] = v
    d.istride = kw.get("istride", 1)
    d.idist = kw.get("idist", 1)
    d.ostride = kw.get("ostride", 1)
    d.odist = kw.get("odist", 1)
    d.batch = kw.get("batch", 1)
    d.flags = kw.get("flags", 0)
    return d


def _validate(d):
    L = _lib()
    ws = ctypes.c_size_t(12345)
    rc = L.lib().fftm_mi_plan_validate(ctypes.byref(d), ctypes.byref(ws))
    return rc, ws.value, L.last_error()


@pytest.mark.parametrize("n,prec,dirn,expected_work", [
    ([512, 512, 512], 0, 1, 0),
    ([512, 512, 512], 0, 2, 512 * 512 * 257 * 8),
    ([128, 128, 128], 1, 2, 128 * 128 * 65 * 16),
    ([12, 10, 8], 1, 1, 0),
    ([64], 0, 2, 0),                      # rank-1 C2R needs no work area
    ([64, 64], 0, 3, 0),
])
def test_validate_ok_and_work_size(n, prec, dirn, expected_work):
    rc, ws, err = _validate(_desc(n=n, precision=prec, dir=dirn))
    assert rc == 0, err
    assert ws == expected_work


def test_validate_overwrite_flag_drops_work():
    rc, ws, _ = _validate(_desc(n=[32, 32, 32], dir=2, flags=1))
    assert rc == 0 and ws == 0


@pytest.mark.parametrize("kw,code", [
    (dict(rank=0), 4),
    (dict(rank=4, n=[4, 4, 4]), 4),
    (dict(n=[0, 4, 4]), 8),
    (dict(batch=0), 4),
    (dict(dir=7), 3),
    (dict(precision=5), 3),
    (dict(n=[16, 16, 4099]), 16),           # prime factor above 13
    (dict(n=[8192]), 16),                   # beyond the single-pass length
    (dict(default_layout=0, n=[8, 8], inembed=[8, 4], onembed=[8, 5], idist=32, odist=40, batch=2), 4),
    (dict(default_layout=0, n=[8], inembed=[8], onembed=[8], istride=0), 4),
])
def test_validate_errors(kw, code):
    rc, _, err = _validate(_desc(**kw))
    assert rc == code, err
    assert err  # a message accompanies every failure


def test_ffts_validate_layouts():
    L = _lib()
    ws = ctypes.c_size_t()
    for layout, n, work in [(0, [512, 512, 512], 512 * 512 * 257 * 8), (1, [64, 64, 64, 64], 64 * 64 * 64 * 33 * 8),
                            (2, [12, 10, 8, 10], 12 * 10 * 8 * 6 * 8), (3, [16, 8, 8, 16], 16 * 8 * 8 * 9 * 8)]:
        rc = L.lib().fftm_mi_ffts_validate(layout, L.i64_array(n), 0, 0, ctypes.byref(ws))
        assert rc == 0, L.last_error()
        assert ws.value == work
    rc = L.lib().fftm_mi_ffts_validate(9, L.i64_array([4, 4, 4, 4]), 0, 0, ctypes.byref(ws))
    assert rc == 4
    rc = L.lib().fftm_mi_ffts_validate(1, L.i64_array([4, 4, 4, 4]), 0, 1, ctypes.byref(ws))
    assert rc == 0 and ws.value == 0


def test_null_handles_are_rejected():
    L = _lib().lib()
    assert L.fftm_mi_plan_destroy(None) == 1
    assert L.fftm_mi_plan_exec(None, 1, None, None, 0) == 1
    assert L.fftm_mi_ffts_forward(None, None, None, 0) == 1
    assert L.fftm_mi_plan_validate(None, None) == 4


def test_transpose_argument_validation():
    L = _lib()
    dims = L.i64_array([2, 3, 4, 5])
    rc = L.lib().fftm_mi_transpose_4d(8, dims, L.i32_array([0, 1, 1, 3]), 16, 32, None, 0)
    assert rc == 4 and "permutation" in L.last_error()
    rc = L.lib().fftm_mi_transpose_4d(6, dims, L.i32_array([0, 1, 2, 3]), 16, 32, None, 0)
    assert rc == 4
    rc = L.lib().fftm_mi_transpose_4d(8, dims, L.i32_array([0, 1, 2, 3]), 16, 16, None, 0)
    assert rc == 16


def test_python_mirror_logic_errors_without_gpu():
    import fftm_amd
    f = fftm_amd.ffts("float64")
    with pytest.raises(fftm_amd.LogicError, match="call init"):
        f.forward(None, None)
    with pytest.raises(fftm_amd.LogicError, match="call init"):
        f.dimension()
    with pytest.raises(TypeError):
        fftm_amd.ffts("float16")


def test_version():
    assert _lib().lib().fftm_mi_version() >= 100


def fnv1a64_py(text: bytes) -> str:
    """detail/fftm_autotune_hardware.inc:87-98 restated: offset basis 14695981039346656037, prime 1099511628211."""
    h = 14695981039346656037
    for c in text:
        h ^= c
        h = (h * 1099511628211) & ((1 << 64) - 1)
    return f"{h:016x}"


@pytest.mark.parametrize("text,expected", [(b"", "cbf29ce484222325"), (b"a", "af63dc4c8601ec8c"),
                                           (b"foobar", "85944171f73967e8")])
def test_fnv1a64_published_vectors(text, expected):
    """FNV-1a 64 test vectors (Fowler/Noll/Vo reference suite)."""
    L = _lib()
    buf = ctypes.create_string_buffer(17)
    assert L.lib().fftm_mi_fnv1a64(text, buf) == 0
    assert buf.value.decode() == expected == fnv1a64_py(text)


def test_fnv1a64_canonical_string():
    L = _lib()
    canon = b"version=2;device_classes=gfx950;device_models=AMD_Instinct_MI355X;compute_units=256"
    buf = ctypes.create_string_buffer(17)
    assert L.lib().fftm_mi_fnv1a64(canon, buf) == 0
    assert buf.value.decode() == fnv1a64_py(canon)


def test_hardware_signature_needs_device():
    L = _lib()
    buf = ctypes.create_string_buffer(64)
    rc = L.lib().fftm_mi_hardware_signature(buf, 64)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        assert rc == 11
    else:
        assert rc == 0 and buf.value.startswith(b"v2:")


def test_ffts_validate_fp16_storage():
    L = _lib()
    ws = ctypes.c_size_t()
    al = lambda b: (b + 255) // 256 * 256  # noqa: E731
    n0, n1, n2 = 256, 128, 512
    nzh = n2 // 2 + 1
    rc = L.lib().fftm_mi_ffts_validate(0, L.i64_array([n0, n1, n2]), L.F32_H16, 0, ctypes.byref(ws))
    assert rc == 0, L.last_error()
    # FP16 bins, then the two exponent arrays, each padded by 16 bytes (a strided pass
    # reads a tile's exponents as whole dwords, up to 4 bytes past the last one)
    assert ws.value == al(4 * n0 * n1 * nzh) + al(max(n0 * n1, n1 * nzh) + 16) + al(n0 * nzh + 16)
    # hi + lo FP16 storage: 8 bytes per stored bin
    rc = L.lib().fftm_mi_ffts_validate(0, L.i64_array([n0, n1, n2]), L.F32_H16X2, 0, ctypes.byref(ws))
    assert rc == 0, L.last_error()
    assert ws.value == al(8 * n0 * n1 * nzh) + al(max(n0 * n1, n1 * nzh) + 16) + al(n0 * nzh + 16)
    # 4D layouts and lengths outside 128..2048 are not implemented for FP16 storage
    assert L.lib().fftm_mi_ffts_validate(1, L.i64_array([128] * 4), L.F32_H16, 0, ctypes.byref(ws)) == 16
    assert L.lib().fftm_mi_ffts_validate(0, L.i64_array([128, 96, 128]), L.F32_H16, 0, ctypes.byref(ws)) == 16
    assert L.lib().fftm_mi_ffts_validate(0, L.i64_array([64, 128, 128]), L.F32_H16, 0, ctypes.byref(ws)) == 16


def test_header_is_plain_c(tmp_path):
    """include/fftm_mi.h compiles as C99 (no C++ in the declarations) and a C program
    links against the library and calls the host-only entry points."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "abi.c"
    src.write_text(
        '#include "fftm_mi.h"\n'
        '#include <stdio.h>\n'
        'int main(void) {\n'
        '    fftm_mi_plan_desc d = {0};\n'
        '    size_t ws = 0;\n'
        '    int64_t n[3] = {512, 512, 512};\n'
        '    fftm_mi_spectral_op op = {0};\n'
        '    (void)op;\n'
        '    d.dir = 1; d.precision = FFTM_MI_F32; d.rank = 3; d.default_layout = 1;\n'
        '    d.n[0] = d.n[1] = d.n[2] = 64; d.batch = 1;\n'
        '    if (fftm_mi_plan_validate(&d, &ws) != 0) return 1;\n'
        '    if (fftm_mi_ffts_validate(FFTM_MI_LAYOUT_3D, n, FFTM_MI_F32, 0, &ws) != 0) return 2;\n'
        '    printf("%d %zu\\n", fftm_mi_version(), ws);\n'
        '    return 0;\n'
        '}\n')
    exe = tmp_path / "abi"
    lib_dir = os.path.dirname(_lib().LIB_PATH)
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe), "-L", lib_dir, "-lfftm_mi", "-Wl,-rpath," + lib_dir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split()[1] == str(512 * 512 * 257 * 8)


def test_fill_random_real_rejects_bad_arguments_on_the_host():
    """fftm_mi_fill_random_real validates before any launch (dst is never touched)."""
    L = _lib()
    lib = L.lib()
    n = L.i64_array([4, 4, 4])
    s = L.i64_array([16, 4, 1])
    dummy = 0x1000
    assert lib.fftm_mi_fill_random_real(7, 3, n, s, None, 1, dummy, None, 1) == L.INVALID_TYPE
    assert lib.fftm_mi_fill_random_real(L.F32, 5, n, s, None, 1, dummy, None, 1) == L.INVALID_VALUE
    assert lib.fftm_mi_fill_random_real(L.F32, 3, n, s, L.i64_array([-1, 0, 0]), 1, dummy, None, 1) \
        == L.INVALID_VALUE
    assert lib.fftm_mi_fill_random_real(L.F32, 3, L.i64_array([4, 0, 4]), s, None, 1, dummy, None, 1) \
        == L.INVALID_SIZE
    assert lib.fftm_mi_fill_random_real(L.F32, 3, n, s, None, 1, None, None, 1) == L.INVALID_VALUE
    assert "fill_random_real" in L.last_error()
