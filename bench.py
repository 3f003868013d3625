#!/usr/bin/env python3
"""Benchmark of the ffts R2C+C2R hot path on one MI355X.

One step = one forward R2C + one backward C2R (unnormalised) of a 512^3 FP32 field
already resident in HBM: the headline workload of the metric
"R2C+C2R GFLOP/s & HBM GB/s at 512^3/64^4".  Protocol after
source/tests/test_benchmark_ffts_3D.cu:65-121: W untimed warmup pairs, then K timed
pairs bracketed by device synchronisation.

Prints ONE JSON line.  ``value`` is GFLOP/s with the 2.5 N log2 N per-R2C convention
(5 N log2 N per pair).  Also reported:
  * ``roofline``: the dominant pass kernel's algorithmic bytes per launch over its
    average duration (HIP events on the transform's stream over the timed region),
    against 8 TB/s;
  * ``hbm_gbps_alg``: B_alg = 2 [(4N + 8N_c) + (d-1) 16 N_c] bytes per pair / pair time;
  * max and RMS relative errors of the timed transform (forward spectrum) against an
    FP64 transform of the same input bytes (pocketfft on the host; this library's FP64
    path on the GPU at 1024^3), and of the round trip;
  * the 64^4 pencil-layout 4D pair (GFLOP/s, HBM GB/s);
  * ``cpu_baseline``: the reference's CPU benchmark protocol (FFTW harness: 3 warmup +
    10 timed pairs, forward and backward timed apart) with pocketfft standing in for the
    absent FFTW, on this host's cores, on the same 512^3 field.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
SEED = 0x1234ABCDEF987654  # test_benchmark_ffts_3D.cu:83


def _flops_pair(N):
    return 5.0 * N * math.log2(N)


def _b_alg(dims, elem_real):
    N = int(np.prod(dims))
    Nc = N // dims[-1] * (dims[-1] // 2 + 1)
    d = len(dims)
    return 2 * ((elem_real * N + 2 * elem_real * Nc) + (d - 1) * 4 * elem_real * Nc)


TRAFFIC_PATH = os.path.join(ROOT, "profiles", "traffic.json")


def _events(n, stream):
    """n timing events, each recorded once on stream and waited for: torch creates the HIP
    event at its first record, which would otherwise fall inside a timed region."""
    import torch
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
    for e in ev:
        e.record(stream)
    torch.cuda.synchronize()
    return ev


def source_stamp():
    """Digest of the kernel sources (fftm_amd/csrc/*.hip, *.hpp): a traffic figure measured
    on another build of the kernels is stale."""
    import hashlib
    csrc = os.path.join(ROOT, "fftm_amd", "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(csrc)):
        if name.endswith((".hip", ".hpp")):
            h.update(name.encode())
            with open(os.path.join(csrc, name), "rb") as f:
                h.update(f.read())
    return h.hexdigest()[:16]


def _traffic_summary(kernel_key):
    """(HBM bytes per launch measured with rocprofv3 --pmc, stamp of the build they were
    measured on); the bytes are None when the committed summary lacks the pass or was
    measured on other kernel sources (stale)."""
    try:
        with open(TRAFFIC_PATH) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None, None
    stamp = data.get("source_stamp")
    v = data.get("passes", {}).get(kernel_key)
    if v is None or stamp != source_stamp():
        return None, stamp
    return float(v["bytes_per_launch"]), stamp


def write_traffic(trace_csv, fetch_csv, write_csv, passes=6, pairs=30, size=512):
    """profiles/traffic.json from the rocprofv3 runs of bench.py --size S --no-cpu --no-check
    --no-4d --no-extra --calibrate-copy --tuning-cache C (trace, --pmc FETCH_SIZE, --pmc
    WRITE_SIZE): the last pairs * passes k_pass dispatches are the timed loops, pass =
    position mod 6; the counters are scaled by their reading for the 1 GiB copy (read 2^30 B,
    written 2^30 B).  Entries are keyed by size; those of other sizes already in the file
    are kept when they were measured on the same kernel sources."""
    import csv

    def rows(path, counter=None):
        out = []
        with open(path) as f:
            for r in csv.DictReader(f):
                if counter is not None and r.get("Counter_Name") != counter:
                    continue
                out.append(r)
        return out

    def per_dispatch(path, counter):
        acc = {}
        for r in rows(path, counter):
            acc.setdefault(int(r["Dispatch_Id"]), [r["Kernel_Name"], 0.0])[1] += float(r["Counter_Value"])
        return [acc[k] for k in sorted(acc)]

    def calib(seq):
        vals = [v for name, v in seq if "copyBuffer" in name]
        return (1 << 30) / (max(vals) * 1024.0) if vals else 1.0

    fetch, write = per_dispatch(fetch_csv, "FETCH_SIZE"), per_dispatch(write_csv, "WRITE_SIZE")
    fr, fw = calib(fetch), calib(write)
    fk = [v * 1024.0 * fr for name, v in fetch if "k_pass" in name][-pairs * passes:]
    wk = [v * 1024.0 * fw for name, v in write if "k_pass" in name][-pairs * passes:]
    tr = [r for r in rows(trace_csv) if "k_pass" in r["Kernel_Name"]][-pairs * passes:]
    names = [f"ffts3d_{size}_f32_{'bwd' if k >= 3 else 'fwd'}{k % 3}" for k in range(passes)]
    stamp = source_stamp()
    out = {"source_stamp": stamp, "fetch_scale": fr, "write_scale": fw, "passes": {}}
    try:
        with open(TRAFFIC_PATH) as f:
            old = json.load(f)
        if old.get("source_stamp") == stamp:
            out["passes"].update({k: v for k, v in old.get("passes", {}).items() if k not in names})
    except (OSError, ValueError):
        pass
    for k, name in enumerate(names):
        rd = [fk[i] for i in range(k, len(fk), passes)]
        wr = [wk[i] for i in range(k, len(wk), passes)]
        us = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in tr[k::passes]]
        out["passes"][name] = {"bytes_per_launch": round(float(np.mean(rd)) + float(np.mean(wr)), 1),
                               "read": round(float(np.mean(rd)), 1), "write": round(float(np.mean(wr)), 1),
                               "avg_us": round(float(np.mean(us)), 3), "kernel": tr[k]["Kernel_Name"][:120]}
    with open(TRAFFIC_PATH, "w") as f:
        json.dump(out, f, indent=1)
    return out


def fill_field(u, dims, pencil=False):
    """The benchmark harness's splitmix64 field (fill_random_real_{3d,4d}_functor,
    fft_benchmark_common.h:201-257) written by the library's device fill into u, whose
    memory layout is the ffts real layout: [x][y][z] (3D), [x][y][z][w] (4D slab) or
    [w][z][y][x] (4D pencil)."""
    import fftm_amd
    nd = len(dims)
    strides = [1] * nd
    if pencil:       # memory [w][z][y][x]: x fastest
        for a in range(1, nd):
            strides[a] = strides[a - 1] * dims[a - 1]
    else:
        for a in range(nd - 2, -1, -1):
            strides[a] = strides[a + 1] * dims[a + 1]
    fftm_amd.fill_random_real(u, dims, SEED, strides, sync=True)
    return u


def fp64_spectrum(x_mem, pencil=False):
    """FP64 reference spectrum (pocketfft via scipy.fft, all host threads) of a real field in
    an ffts memory layout, returned in that layout's spectrum order: [x][y][z/2+1] for
    [x][y][z]; for the 4D pencil layout real [w][z][y][x] -> spectrum [y][z][w/2+1][x]."""
    import scipy.fft as sfft
    w = _host_threads()
    x = np.asarray(x_mem, dtype=np.float64)
    if not pencil:
        return sfft.rfftn(x, workers=w)
    logical = x.transpose(3, 2, 1, 0)                       # u[x][y][z][w]
    return np.ascontiguousarray(sfft.rfftn(logical, workers=w).transpose(1, 2, 3, 0))


def rel_errors(actual, expected):
    """(||a - e||_2 / ||e||_2, max|a - e| / max|e|)."""
    a = np.asarray(actual).astype(np.complex128)
    e = np.asarray(expected).astype(np.complex128)
    d = np.abs(a - e)
    ae = np.abs(e)
    return float(np.linalg.norm(d) / np.linalg.norm(ae)), float(d.max() / ae.max())


TUNING_CACHE = None  # --tuning-cache PATH


def timing_stats(samples):
    """Per-pair statistics in ms: mean and population standard deviation as the reference's
    compute_timing_statistics (fft_benchmark_common.h:41-61), plus median, min and p95."""
    a = np.asarray(samples, dtype=np.float64)
    if a.size == 0:
        return {}
    return {"mean": round(float(a.mean()), 4), "stddev": round(float(a.std()), 4),
            "median": round(float(np.median(a)), 4), "min": round(float(a.min()), 4),
            "p95": round(float(np.percentile(a, 95)), 4), "samples": int(a.size)}


def tune_handle(f, warmup, iterations):
    """Tune f's passes, or load them from TUNING_CACHE when it already holds this problem
    (the reference's KEY=VALUE cache: runs that must see the same variants share it)."""
    if TUNING_CACHE is None:
        return f.autotune(warmup=warmup, iterations=iterations)
    if os.path.exists(TUNING_CACHE) and f.load_tuning(TUNING_CACHE) > 0:
        return f.pass_variants()
    return f.autotune(warmup=warmup, iterations=iterations, cache_path=TUNING_CACHE)


def run_gpu(dims, strategy, steps, warmup, check, tune=True, storage="fp32", u_dev=None, destructive=False,
            radix_schedule=None):
    """Time the forward + backward pair.  check: True (FP64 pocketfft on the host, same
    input bytes), "gpu64" (this library's FP64 transform of the same field, for sizes where
    the host transform takes too long) or False.  u_dev: a device input (else the
    splitmix64 field, filled on the device)."""
    import torch
    import fftm_amd
    from fftm_amd import _lib
    import ctypes

    dt = np.float32
    f = fftm_amd.ffts("float32", strategy_4d=strategy)
    opts = fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="", execution="fused", sync=False,
                                      spectrum_storage=storage, destructive_backward=destructive)
    f.init(*dims, options=opts)
    stream = torch.cuda.current_stream()
    f.set_stream(stream)
    h = f._handle
    L = _lib.lib()
    # plan-time tuning, outside the timed region (the reference's autotune protocol:
    # every variant of every pass timed, fastest bitwise-valid one kept); with
    # --tuning-cache, a cache that already holds this problem is applied instead
    variants = tune_handle(f, warmup=2, iterations=7) if tune else f.pass_variants()
    if radix_schedule is not None:
        # the strided C2C passes (the y and x passes of a 3D transform) in the given radix
        # schedule, keeping their tuned tile order (tile shape 0 or 6)
        nf_, nb_ = f.num_passes()
        for key in ((0, nf_ - 2), (0, nf_ - 1), (1, 0), (1, 1)):
            v = variants[key]
            shape = v & 7 if (v & 7) in (0, 6) else 6
            f.set_pass_variant(key[0], key[1], shape + (v & 24) + 32 * radix_schedule)
        variants = f.pass_variants()

    # input: splitmix64 uniform[-1,1) (fill_random_real_{3d,4d}_functor) generated on the
    # device by the library's fill kernel, resident in HBM
    pencil = len(dims) == 4 and strategy.family.value == "pencil_pencil"
    if u_dev is not None:
        u = u_dev
    else:
        u = f.real_array()
        fill_field(u, dims, pencil)
    X = f.complex_array()
    y = f.real_array()
    nf, nb = f.num_passes()
    passes = [(0, i) for i in range(nf)] + [(1, i) for i in range(nb)]
    pbytes = {}
    for bw, i in passes:
        b = ctypes.c_int64()
        _lib.check(L.fftm_mi_ffts_pass_bytes(h, bw, i, ctypes.byref(b)), "pass_bytes")
        pbytes[(bw, i)] = int(b.value)

    def pair():
        f.forward(u, X)
        f.backward(X, y)

    def run_one(bw, i):
        src, dst = (u, X) if bw == 0 else (X, y)
        _lib.check(L.fftm_mi_ffts_run_pass(h, bw, i, src.data_ptr(), dst.data_ptr(), 0), "run_pass")

    # every timing event is created before the warmup, so that neither an event creation
    # nor the synchronization after it leaves the GPU idle between the warmup and the
    # timed region
    ev_first = [_events(len(passes) + 1, stream) for _ in range(5)]
    ev_pp = [_events(len(passes) + 1, stream) for _ in range(steps)]
    ev0 = _events(steps, stream)
    ev1 = _events(steps, stream)
    pev = _events(steps + 1, stream)

    # warmup: W pairs, then pairs back to back until 0.3 s have passed.  The GPU's clocks
    # rise under sustained load: right after an idle gap of a second or more, the same
    # pairs run 2-3 % slower for their first few tens of milliseconds.
    tw = time.perf_counter()
    for _ in range(warmup):
        pair()
    torch.cuda.synchronize()
    while time.perf_counter() - tw < 0.3:
        for _ in range(8):
            pair()
        torch.cuda.synchronize()

    def per_pass(ev, stat=np.mean):
        # pairs pass by pass, a HIP event on the transform's stream after every pass
        for s in range(len(ev)):
            ev[s][0].record(stream)
            for k, (bw, i) in enumerate(passes):
                run_one(bw, i)
                ev[s][k + 1].record(stream)
        torch.cuda.synchronize()
        return {p: float(stat([ev[s][k].elapsed_time(ev[s][k + 1]) for s in range(len(ev))]))
                for k, p in enumerate(passes)}

    # the dominant pass kernel, from five untimed instrumented pairs by the median
    first = per_pass(ev_first, np.median)
    dom = max(passes, key=lambda p: first[p])

    # timed region: K pairs, pass by pass, with HIP events on the transform's stream
    # only around the dominant kernel (an event between every pass adds gaps)
    def timed_pairs():
        for s in range(steps):
            for bw, i in passes:
                if (bw, i) == dom:
                    ev0[s].record(stream)
                    run_one(bw, i)
                    ev1[s].record(stream)
                else:
                    run_one(bw, i)

    # the same K pairs once untimed first: measured with HIP events around the whole run,
    # the first run of this sequence took 1.205 ms per 512^3 pair, the runs right after it
    # 1.181 ms (10.43 against 10.15 ms at 1024^3), after the same warmup
    timed_pairs()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timed_pairs()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pair_s = (t1 - t0) / steps
    dom_ms = float(np.mean([ev0[s].elapsed_time(ev1[s]) for s in range(steps)]))
    # per-pass breakdown (untimed, events between all passes)
    per_pass_ms = per_pass(ev_pp)
    # also the plain calls (forward + backward API), with a HIP event between consecutive
    # pairs only: the reference's protocol (events around forward + backward,
    # test_benchmark_ffts_3D.cu:65-98) for the per-pair statistics
    t2 = time.perf_counter()
    pev[0].record(stream)
    for k in range(steps):
        pair()
        pev[k + 1].record(stream)
    torch.cuda.synchronize()
    api_pair_s = (time.perf_counter() - t2) / steps
    pair_stats = timing_stats([pev[k].elapsed_time(pev[k + 1]) for k in range(steps)])

    res = {"pair_s": pair_s, "api_pair_s": api_pair_s, "per_pass_ms": per_pass_ms, "pass_bytes": pbytes,
           "dom": dom, "dom_ms": dom_ms, "pair_stats_ms": pair_stats,
           "passes": passes, "variants": {f"{'bwd' if b else 'fwd'}{i}": v for (b, i), v in variants.items()}}
    if check == "gpu64":
        pair()
        torch.cuda.synchronize()
        ud = u.double()
        yr = y.double() / float(np.prod(dims))
        res["roundtrip_rms"] = float(torch.sqrt(torch.mean((yr - ud) ** 2)))
        res["roundtrip_rel_rms"] = float(torch.linalg.vector_norm(yr - ud) / torch.linalg.vector_norm(ud))
        del yr
        f.forward(u, X)
        g = fftm_amd.ffts("float64")
        g.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key=""))
        X64 = g.complex_array()
        g.forward(ud, X64)
        del ud
        dX = X.to(torch.complex128) - X64
        res["fwd_rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(dX) / torch.linalg.vector_norm(X64))
        res["fwd_rel_max_vs_fp64"] = float(dX.abs().max() / X64.abs().max())
        res["reference"] = "this library's FP64 transform of the same field on the GPU"
        del dX, X64
        g.close()
    elif check:
        pair()
        torch.cuda.synchronize()
        x_host = u.cpu().numpy().astype(np.float64)
        yr = y.cpu().numpy().astype(np.float64) / float(np.prod(dims))
        d = yr - x_host
        res["roundtrip_rms"] = float(np.sqrt(np.mean(d * d)))
        res["roundtrip_max"] = float(np.max(np.abs(d)))
        res["roundtrip_rel_rms"] = float(np.linalg.norm(d) / np.linalg.norm(x_host))
        del yr, d
        f.forward(u, X)
        torch.cuda.synchronize()
        rms, mx = rel_errors(X.cpu().numpy(), fp64_spectrum(x_host, pencil))
        res["fwd_rel_l2_vs_fp64"] = rms
        res["fwd_rel_max_vs_fp64"] = mx
        res["reference"] = "pocketfft (scipy.fft) FP64 on the host, same input bytes"
        del x_host
    res["bytes_per_pair"] = int(sum(pbytes.values()))
    f.close()
    del u, X, y
    torch.cuda.empty_cache()
    return res


def kolmogorov_field(n, seed=5):
    """uniform(-1,1) plus a k^(-5/3) random-phase field of similar RMS, FP32 on the GPU
    (synthetic input of the mixed-precision workload)."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    u = torch.rand((n, n, n), generator=g, device="cuda", dtype=torch.float32) * 2 - 1
    k = torch.fft.fftfreq(n, d=1.0 / n, device="cuda")
    kz = torch.arange(n // 2 + 1, device="cuda", dtype=torch.float32)
    K = torch.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = torch.where(K > 0, K.clamp(min=1.0) ** (-5.0 / 3.0), torch.zeros_like(K))
    del K
    ph = torch.rand(amp.shape, generator=g, device="cuda") * (2 * math.pi)
    spec = torch.polar(amp, ph)
    del amp, ph
    v = torch.fft.irfftn(spec, s=(n, n, n))
    del spec
    v *= (0.5 / float(v.std()))
    u += v
    del v
    torch.cuda.empty_cache()
    return u


def run_poisson(n, steps, warmup, seed=11):
    """Periodic -Lap u = f on [0, 2 pi)^3 at n^3 FP32: f = 64 random Fourier modes,
    |k| <= 32, N(0,1) amplitudes, random phases, no zero mode; exact u known mode by mode.
    The fused filter (5 HBM passes) against forward + spectral multiply + backward."""
    import torch
    import fftm_amd
    rng = np.random.default_rng(seed)
    modes = []
    while len(modes) < 64:
        kv = rng.integers(-32, 33, size=3)
        if 0 < kv @ kv <= 32 * 32:
            modes.append(kv)
    amps = rng.standard_normal(64)
    phases = rng.uniform(0, 2 * math.pi, 64)
    xs = torch.arange(n, device="cuda", dtype=torch.float64) * (2 * math.pi / n)
    rhs = torch.zeros((n, n, n), device="cuda", dtype=torch.float64)
    exact = torch.zeros_like(rhs)
    for kv, a, ph in zip(modes, amps, phases):
        arg = (kv[0] * xs[:, None, None] + kv[1] * xs[None, :, None]) + (kv[2] * xs[None, None, :] + ph)
        c = torch.cos(arg)
        rhs += float(a) * c
        exact += float(a) / float(kv @ kv) * c
        del arg, c
    f32 = rhs.float()
    del rhs
    f = fftm_amd.ffts("float32")
    f.init(n, n, n, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="", sync=False))
    stream = torch.cuda.current_stream()
    f.set_stream(stream)
    tune_handle(f, warmup=2, iterations=5)
    out = torch.empty_like(f32)
    X = f.complex_array()
    # the operator is the reference functor's H = -scale/|k|^2 (it solves Lap u = f);
    # -Lap u = f needs u_hat = f_hat/|k|^2, i.e. scale = -1/N (1/N normalises the pair)
    scale = -1.0 / float(n) ** 3

    def fused():
        f.filter(f32, out, "poisson", scale=scale)

    def unfused():
        f.forward(f32, X)
        f.spectral_apply(X, X, "poisson", scale=scale)
        f.backward(X, out)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_unf = timed(unfused)
    t_fus = timed(fused)
    fused()
    torch.cuda.synchronize()
    err = float((out.double() - exact).abs().max())
    umax = float(exact.abs().max())
    f.close()
    del f32, out, X, exact
    torch.cuda.empty_cache()
    return {"workload": f"periodic Poisson solve {n}^3 FP32: R2C, x(-1/|k|^2), C2R, x(1/N); f = 64 random modes "
                        f"|k|<=32, N(0,1) amplitudes",
            "ms_fused_filter": round(t_fus * 1e3, 4), "ms_forward_multiply_backward": round(t_unf * 1e3, 4),
            "hbm_passes_fused": 5, "hbm_passes_unfused": 7,
            "max_abs_err_vs_exact": err, "rel_max_err_vs_exact": err / umax}


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def _graph_pair_ms(f, u, X, y, steps):
    """One forward + backward pair captured in a HIP graph (with sync off the entry points
    only launch kernels) and replayed `steps` times on a side stream; ms per replay from
    HIP events around the replays.  For small transforms this removes the per-call host
    launch cost that the plain calls pay."""
    import torch
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    f.set_stream(s)
    with torch.cuda.stream(s):
        f.forward(u, X)
        f.backward(X, y)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        f.forward(u, X)
        f.backward(X, y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(s):
        for _ in range(3):
            g.replay()
        e0.record(s)
        for _ in range(steps):
            g.replay()
        e1.record(s)
    torch.cuda.synchronize()
    f.set_stream(torch.cuda.current_stream())
    return e0.elapsed_time(e1) / steps


def run_config1(steps, warmup):
    """128^3 round trip on the splitmix64 field in FP64 (the reference's CPU-oracle case)
    and FP32, with errors against FP64 pocketfft of the same input.  Plain calls, and the same pair
    replayed from a HIP graph (the errors are those of the replayed outputs)."""
    import torch
    import fftm_amd
    dims = (128, 128, 128)
    out = {"workload": "ffts 3D R2C + C2R round trip 128^3, splitmix64 uniform(-1,1) field, normalised by 1/N",
           "reference": "pocketfft (scipy.fft) FP64 on the host, same input bytes"}
    for dt in ("float64", "float32"):
        f = fftm_amd.ffts(dt)
        f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="", sync=False))
        f.set_stream(torch.cuda.current_stream())
        tune_handle(f, warmup=2, iterations=7)
        u = fill_field(f.real_array(), dims)
        x = u.cpu().numpy()
        Xo = fp64_spectrum(x)
        X, y = f.complex_array(), f.real_array()
        t = _timed(lambda: (f.forward(u, X), f.backward(X, y)), steps, warmup)
        X.zero_()
        y.zero_()
        tg = _graph_pair_ms(f, u, X, y, max(steps, 20))
        rms, mx = rel_errors(X.cpu().numpy(), Xo)
        d = y.cpu().numpy().astype(np.float64) / float(np.prod(dims)) - x.astype(np.float64)
        out[dt] = {"ms_per_pair": round(t * 1e3, 4), "ms_per_pair_hip_graph": round(tg, 4),
                   "fwd_rel_l2_vs_fp64": rms, "fwd_rel_max_vs_fp64": mx,
                   "roundtrip_rms": float(np.sqrt(np.mean(d * d)))}
        f.close()
    return out


def run_gaussian_4d(steps, warmup):
    """Spatio-temporal Gaussian low-pass (sigma_k = 8) on 64^4 and 128^4 FP32, pencil layout:
    uniform(-1,1) noise plus 16 random plane waves, one fused filter call (scale 1/N).
    Accuracy on the plane waves alone against the closed form."""
    import torch
    import fftm_amd
    out = {"workload": "4D Gaussian low-pass sigma_k=8, pencil layout, fused ffts filter (forward, xH, backward)"}
    rng = np.random.default_rng(21)
    for n in (64, 128):
        dims = (n, n, n, n)
        f = fftm_amd.ffts("float32", strategy_4d=fftm_amd.strategy_4d_pencil_pencil())
        f.init(*dims, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="", sync=False))
        f.set_stream(torch.cuda.current_stream())
        tune_handle(f, warmup=1, iterations=3)
        # real layout [w][z][y][x]
        ar = torch.arange(n, device="cuda", dtype=torch.float32) * (2 * math.pi / n)
        waves = torch.zeros((n, n, n, n), device="cuda")
        expect = torch.zeros_like(waves)
        for _ in range(16):
            k = rng.integers(-12, 13, size=4)
            ph = float(rng.uniform(0, 2 * math.pi))
            arg = (float(k[3]) * ar[:, None, None, None] + float(k[2]) * ar[None, :, None, None]) + \
                  (float(k[1]) * ar[None, None, :, None] + float(k[0]) * ar[None, None, None, :] + ph)
            c = torch.cos(arg)
            waves += c
            expect += math.exp(-float(k @ k) / (2 * 64.0)) * c
            del arg, c
        y = torch.empty_like(waves)
        f.filter(waves, y, "gaussian", scale=1.0 / float(n) ** 4, sigma=8.0)
        torch.cuda.synchronize()
        err = float((y - expect).abs().max() / expect.abs().max())
        u = waves + (torch.rand_like(waves) * 2 - 1)
        del waves, expect
        t = _timed(lambda: f.filter(u, y, "gaussian", scale=1.0 / float(n) ** 4, sigma=8.0), steps, warmup)
        N = n ** 4
        Nc = N // n * (n // 2 + 1)
        balg = 2 * ((4 * N + 8 * Nc) + 3 * 16 * Nc)
        out[f"{n}^4"] = {"ms_filter": round(t * 1e3, 4), "gflops_equiv": round(5 * N * math.log2(N) / t / 1e9, 1),
                         "hbm_gbps_vs_unfused_pair_bytes": round(balg / t / 1e9, 1),
                         "plane_waves_rel_max_err": err}
        f.close()
        del u, y
        torch.cuda.empty_cache()
    return out


def run_adjoint(n, steps, warmup, nshells=32):
    """Loss L = ||C2R(H(c) R2C(u)) - y||^2 on n^3 FP32 with 32 radial shell coefficients;
    time of the loss and of its gradients w.r.t. u and c through the hand-written adjoints."""
    import torch
    import fftm_amd
    from fftm_amd.autograd import spectral_filter
    f = fftm_amd.ffts("float32")
    f.init(n, n, n, options=fftm_amd.ffts_init_options(profiling_key="", memory_profiling_key="", sync=False))
    f.set_stream(torch.cuda.current_stream())
    g = torch.Generator(device="cuda").manual_seed(3)
    u = (torch.rand((n, n, n), generator=g, device="cuda") * 2 - 1).requires_grad_(True)
    c = (torch.rand(nshells, generator=g, device="cuda") + 0.5).requires_grad_(True)
    target = torch.rand((n, n, n), generator=g, device="cuda")

    def step():
        u.grad = None
        c.grad = None
        v = spectral_filter(u, c, f)
        loss = ((v - target) ** 2).sum()
        loss.backward()
        return loss

    t_all = _timed(step, steps, warmup)
    with torch.no_grad():
        t_fwd = _timed(lambda: spectral_filter(u, c, f), steps, warmup)
    f.close()
    return {"workload": f"L = ||C2R(H(c) R2C(u)) - y||^2, {n}^3 FP32, {nshells} shell coefficients; "
                        "gradients w.r.t. u and c via the C-ABI adjoints (forward R2C recomputed)",
            "ms_loss_forward": round(t_fwd * 1e3, 4), "ms_loss_and_gradients": round(t_all * 1e3, 4)}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads():
    """Threads this process may use: the affinity mask, capped by OMP_NUM_THREADS when set
    (the GPU box grants a 16-thread share of a larger host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def run_cpu_baseline(dims, pairs=10, warmup=3):
    """The reference's CPU baseline protocol (cpu_fftw_mpi/fftw_mpi_benchmark.cpp:249-276,
    621-698): out-of-place R2C then C2R of one rank's whole field, `warmup` untimed pairs, then
    `pairs` timed pairs with forward and backward timed separately; mean / population stddev /
    median / min / max / p95 per pair, and the round-trip relative L2 error after 1/N.
    FFTW is not installed here, so the transform is pocketfft (scipy.fft.rfftn / irfftn with
    one worker per host thread), in FP64 (the reference harness's precision, the `value`) and
    FP32.  The input is the splitmix64 field of the GPU measurement (same seed)."""
    import scipy
    import scipy.fft as sfft
    import oracle
    threads = _host_threads()
    N = int(np.prod(dims))
    out = {}
    for dt in (np.float64, np.float32):
        x = oracle.fill_random_real_3d(*dims, SEED, dt)
        X = sfft.rfftn(x, workers=threads)           # plan creation outside the timed loop
        y = sfft.irfftn(X, s=dims, workers=threads)
        fwd, bwd = [], []
        for k in range(warmup + pairs):
            t0 = time.perf_counter()
            X = sfft.rfftn(x, workers=threads)
            t1 = time.perf_counter()
            y = sfft.irfftn(X, s=dims, workers=threads)   # pocketfft's irfftn includes the 1/N
            t2 = time.perf_counter()
            if k >= warmup:
                fwd.append((t1 - t0) * 1e3)
                bwd.append((t2 - t1) * 1e3)
        err = float(np.linalg.norm((y - x).astype(np.float64)) / np.linalg.norm(x.astype(np.float64)))
        pair = [a + b for a, b in zip(fwd, bwd)]
        stats = timing_stats(pair)
        stats["max"] = round(float(np.max(pair)), 4)
        out["f64" if dt == np.float64 else "f32"] = {
            "gflops": round(_flops_pair(N) / (stats["mean"] * 1e-3) / 1e9, 2),
            "pair_ms": stats, "forward_ms": timing_stats(fwd), "backward_ms": timing_stats(bwd),
            "roundtrip_rel_l2": err}
        del x, X, y
    f64 = out["f64"]
    return {"value": f64["gflops"], "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"pocketfft (scipy {scipy.__version__} scipy.fft.rfftn/irfftn, workers={threads}; FFTW "
                      f"unavailable) on the same {dims[0]}x{dims[1]}x{dims[2]} splitmix64 field, FP64 and FP32, "
                      f"{warmup} warmup + {pairs} timed R2C+C2R pairs each (the FFTW harness protocol); value = "
                      f"FP64 mean {f64['pair_ms']['mean']:.1f} ms per pair",
            "cpu_model": _cpu_model(), "ms_per_pair": f64["pair_ms"]["mean"], **out}


def _progress(msg):
    """One progress line per finished leg, on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-4d", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shapes (skip plan-time tuning)")
    ap.add_argument("--tuning-cache", default=None,
                    help="autotune cache file: applied where it holds the problem, else written by the tuner")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 256^3, FP16-storage and Poisson secondary measurements")
    ap.add_argument("--no-big", dest="big", action="store_false",
                    help="skip the 1024^3 FP32 vs FP16-storage comparison (the mixed-precision workload)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy 5 times (counter calibration)")
    ap.add_argument("--write-traffic", nargs=3, metavar=("TRACE_CSV", "FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic.json from rocprofv3 kernel-trace / FETCH_SIZE / WRITE_SIZE csv files "
                         "of a run with --no-cpu --no-check --no-4d --no-extra --calibrate-copy, then exit")
    args = ap.parse_args()
    if args.write_traffic:
        print(json.dumps(write_traffic(*args.write_traffic, pairs=3 * args.steps, size=args.size), indent=1))
        return
    global TUNING_CACHE
    TUNING_CACHE = args.tuning_cache
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (the transform runs on one device); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import torch
    import fftm_amd
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.calibrate_copy:
        nbytes = 1 << 30
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes_read": nbytes, "calibration_copy_bytes_written": nbytes,
                          "launches": 5}), file=sys.stderr, flush=True)
        del a, b
        torch.cuda.empty_cache()

    n = args.size
    dims = (n, n, n)
    N = n ** 3
    r3 = run_gpu(dims, fftm_amd.strategy_4d_pencil_pencil(), args.steps, args.warmup, not args.no_check,
                 not args.no_tune)
    pair_s = r3["pair_s"]
    _progress(f"{n}^3 FP32 pair {pair_s * 1e3:.4f} ms")
    gflops = _flops_pair(N) / pair_s / 1e9
    balg = _b_alg(dims, 4)
    # dominant pass kernel: its average duration over the timed region (HIP events)
    dom = r3["dom"]
    dom_ms = r3["dom_ms"]
    dom_bytes = r3["pass_bytes"][dom]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    dom_name = f"{'backward' if dom[0] else 'forward'} pass {dom[1]}"
    traffic, traffic_stamp = _traffic_summary(f"ffts3d_{n}_f32_{'bwd' if dom[0] else 'fwd'}{dom[1]}")

    out = {
        "metric": "R2C+C2R GFLOP/s & HBM GB/s at 512^3/64^4, mixed vs FP32 at ≤1e-5 RMS err",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(pair_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: splitmix64 uniform[-1,1) field, seed {SEED:#x} (fill_random_real_3d_functor)",
        "config": {"workload": f"ffts 3D R2C forward + C2R inverse pair, {n}^3 float32, fused passes",
                   "nx": n, "ny": n, "nz": n, "flop_convention": "2.5 N log2 N per R2C (5 N log2 N per pair)"},
        "hbm_gbps_alg": round(balg / pair_s / 1e9, 1),
        "b_alg_bytes": balg,
        "api_ms_per_pair": round(r3["api_pair_s"] * 1e3, 4),
        "api_pair_timing_ms": r3["pair_stats_ms"],
        "per_pass_ms": {f"{'bwd' if p[0] else 'fwd'}{p[1]}": round(v, 4) for p, v in r3["per_pass_ms"].items()},
        "pass_variants": r3["variants"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": dom_name, "alg_bytes_per_launch": dom_bytes, "avg_ms": round(dom_ms, 4),
                     "traffic_source": {"file": "profiles/traffic.json", "measured_on": traffic_stamp,
                                        "this_build": source_stamp(),
                                        "stale": traffic_stamp is not None and traffic_stamp != source_stamp()}},
    }
    if not args.no_check:
        out["errors_vs_fp64"] = {k: r3[k] for k in ("fwd_rel_l2_vs_fp64", "fwd_rel_max_vs_fp64",
                                                           "roundtrip_rms", "roundtrip_max")}
    if not args.no_4d:
        d4 = (64, 64, 64, 64)
        r4 = run_gpu(d4, fftm_amd.strategy_4d_pencil_pencil(), args.steps, args.warmup, not args.no_check,
                     not args.no_tune)
        N4 = 64 ** 4
        _progress(f"64^4 pencil pair {r4['api_pair_s'] * 1e3:.4f} ms")
        out["secondary_64^4"] = {
            "workload": "ffts 4D pencil layout R2C+C2R pair 64^4 float32 (real [w][z][y][x] -> [y][z][w/2+1][x])",
            # plain API pairs; events between the ~25 us passes add launch gaps
            "gflops": round(_flops_pair(N4) / r4["api_pair_s"] / 1e9, 2),
            "ms_per_pair": round(r4["api_pair_s"] * 1e3, 4),
            "ms_per_pair_timed_region": round(r4["pair_s"] * 1e3, 4),
            "pair_timing_ms": r4["pair_stats_ms"],
            "hbm_gbps_alg": round(_b_alg(d4, 4) / r4["api_pair_s"] / 1e9, 1),
            "per_pass_ms": {f"{'bwd' if p[0] else 'fwd'}{p[1]}": round(v, 4) for p, v in r4["per_pass_ms"].items()},
            "pass_variants": r4["variants"],
        }
        if not args.no_check:
            out["secondary_64^4"]["errors_vs_fp64"] = {
                k: r4[k] for k in ("fwd_rel_l2_vs_fp64", "fwd_rel_max_vs_fp64", "roundtrip_rms", "roundtrip_max")}
    if not args.no_extra:
        r256 = run_gpu((256, 256, 256), fftm_amd.strategy_4d_pencil_pencil(), args.steps, args.warmup, False,
                       not args.no_tune)
        out["secondary_256^3"] = {"gflops": round(_flops_pair(256 ** 3) / r256["api_pair_s"] / 1e9, 2),
                                  "ms_per_pair": round(r256["api_pair_s"] * 1e3, 4),
                                  "ms_per_pair_timed_region": round(r256["pair_s"] * 1e3, 4),
                                  "hbm_gbps_alg": round(_b_alg((256,) * 3, 4) / r256["api_pair_s"] / 1e9, 1)}
        # the backward allowed to overwrite its input (hipfftExecC2R's contract, which the
        # reference's 3D backward inherits, ffts.hpp:363-368): every backward pass in place
        rd = run_gpu(dims, fftm_amd.strategy_4d_pencil_pencil(), args.steps, args.warmup, False, not args.no_tune,
                     destructive=True)
        out["destructive_backward_512^3"] = {
            "workload": f"ffts 3D pair {n}^3 FP32, backward overwrites its input spectrum (no work area)",
            "ms_per_pair": round(rd["pair_s"] * 1e3, 4),
            "gflops": round(_flops_pair(N) / rd["pair_s"] / 1e9, 2),
            "hbm_gbps_alg": round(balg / rd["pair_s"] / 1e9, 1),
            "per_pass_ms": {f"{'bwd' if p[0] else 'fwd'}{p[1]}": round(v, 4) for p, v in rd["per_pass_ms"].items()},
            "pass_variants": rd["variants"]}
        # FP16-stored spectra against the plain FP32 pair: same field, same harness
        rm = run_gpu(dims, None, args.steps, args.warmup, not args.no_check, not args.no_tune, storage="fp16")
        mp = {"workload": f"ffts 3D pair {n}^3, FP32 arithmetic, FP16 half spectra between passes "
                          "(per-line power-of-two exponents)",
              "ms_per_pair": round(rm["pair_s"] * 1e3, 4),
              "gflops": round(gflops * r3["pair_s"] / rm["pair_s"], 2),
              "speedup_vs_fp32": round(r3["pair_s"] / rm["pair_s"], 3),
              "bytes_per_pair": rm["bytes_per_pair"], "bytes_per_pair_fp32": r3["bytes_per_pair"],
              "hbm_gbps": round(rm["bytes_per_pair"] / rm["pair_s"] / 1e9, 1),
              "per_pass_ms": {f"{'bwd' if p[0] else 'fwd'}{p[1]}": round(v, 4) for p, v in rm["per_pass_ms"].items()},
              "pass_variants": rm["variants"]}
        for k in ("fwd_rel_l2_vs_fp64", "fwd_rel_max_vs_fp64", "roundtrip_rel_rms"):
            if k in rm:
                mp[k] = rm[k]
        if "fwd_rel_l2_vs_fp64" in rm:
            mp["meets_1e-5_rms"] = bool(rm["fwd_rel_l2_vs_fp64"] <= 1e-5)
        out["mixed_precision"] = mp

        def leg(r, what):
            d = {"workload": what, "ms_per_pair": round(r["pair_s"] * 1e3, 4),
                 "speedup_vs_fp32": round(r3["pair_s"] / r["pair_s"], 3), "bytes_per_pair": r["bytes_per_pair"],
                 "per_pass_ms": {f"{'bwd' if p[0] else 'fwd'}{p[1]}": round(v, 4) for p, v in r["per_pass_ms"].items()},
                 "pass_variants": r["variants"]}
            for k in ("fwd_rel_l2_vs_fp64", "fwd_rel_max_vs_fp64", "roundtrip_rel_rms"):
                if k in r:
                    d[k] = r[k]
            if "fwd_rel_l2_vs_fp64" in r:
                d["meets_1e-5_rms"] = bool(r["fwd_rel_l2_vs_fp64"] <= 1e-5)
            return d
        # hi + lo FP16 storage: FP32-level accuracy at FP32's bytes
        rx = run_gpu(dims, None, args.steps, args.warmup, not args.no_check, not args.no_tune, storage="fp16x2")
        mp["hi_lo_fp16_storage"] = leg(rx, f"{n}^3 pair, half spectra stored as hi + lo FP16 pairs "
                                           "(8 bytes per bin, per-line exponents)")
        # the radix-16 stage of the strided passes as FP16 MFMA products (hi + lo operands,
        # FP32 accumulation), with FP32 and with FP16-stored spectra
        for storage in ("fp32", "fp16"):
            rr = run_gpu(dims, None, args.steps, args.warmup, not args.no_check, not args.no_tune, storage=storage,
                         radix_schedule=2)
            mp[f"mfma_radix16_{storage}_storage"] = leg(
                rr, f"{n}^3 pair, {storage} spectra, y and x passes in the 8 16 4 schedule with the radix-16 stage "
                    "on the matrix cores (v_mfma_f32_32x32x16_f16, hi + lo operands, FP32 accumulation)")
        _progress("hi + lo storage and matrix-core radix-16 legs done")
        _progress(f"256^3 and {n}^3 FP16-storage legs done")
        if args.big:
            # 1024^3: uniform(-1,1) + k^-5/3 field; FP64 reference from this library on the GPU
            import torch
            ub = kolmogorov_field(1024)
            big = {}
            for storage in ("fp32", "fp16", "fp16x2"):
                rb = run_gpu((1024,) * 3, None, max(3, args.steps // 4), 2, "gpu64", not args.no_tune,
                             storage=storage, u_dev=ub)
                big[storage] = {"ms_per_pair": round(rb["pair_s"] * 1e3, 4),
                                "gflops": round(_flops_pair(1024 ** 3) / rb["pair_s"] / 1e9, 2),
                                "hbm_gbps": round(rb["bytes_per_pair"] / rb["pair_s"] / 1e9, 1),
                                **{k: rb[k] for k in ("fwd_rel_l2_vs_fp64", "fwd_rel_max_vs_fp64",
                                                      "roundtrip_rel_rms") if k in rb}}
            _progress("1024^3 FP32 and FP16-storage legs done")
            big["speedup_fp16_vs_fp32"] = round(big["fp32"]["ms_per_pair"] / big["fp16"]["ms_per_pair"], 3)
            big["speedup_fp16x2_vs_fp32"] = round(big["fp32"]["ms_per_pair"] / big["fp16x2"]["ms_per_pair"], 3)
            big["input"] = "uniform(-1,1) + k^(-5/3) random-phase field (half the RMS), 1024^3 FP32"
            big["reference"] = "this library's FP64 transform of the same field on the GPU"
            out["mixed_precision"]["1024^3"] = big
            del ub
            torch.cuda.empty_cache()
        out["poisson_512^3"] = run_poisson(512, args.steps, args.warmup)
        _progress("Poisson 512^3 done")
        out["config1_128^3"] = run_config1(args.steps, args.warmup)
        out["gaussian_4d"] = run_gaussian_4d(max(3, args.steps // 2), 2)
        _progress("128^3 round trip and 4D Gaussian filter done")
        out["adjoint_256^3"] = run_adjoint(256, max(3, args.steps // 2), 2)
    else:
        out["mixed_precision"] = None
    if not args.no_cpu:
        _progress("GPU legs done; CPU baseline (pocketfft, FP64 and FP32)")
        out["cpu_baseline"] = run_cpu_baseline(dims)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
