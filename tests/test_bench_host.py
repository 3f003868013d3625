"""Host-side pieces of bench.py that need no GPU: the per-pair timing statistics follow
the reference's compute_timing_statistics (source/tests/detail/fft_benchmark_common.h:
41-61: arithmetic mean, population standard deviation), and `--gpus` above 1 exits with
a message."""

import importlib.util
import os
import statistics
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_timing_stats_match_reference_definitions():
    samples = [1.3, 1.25, 1.41, 1.33, 1.29, 1.36, 1.31, 1.5, 1.28, 1.3]
    st = _bench().timing_stats(samples)
    assert abs(st["mean"] - statistics.fmean(samples)) < 1e-4
    # population (divide by n), as the reference; the sample stddev (n - 1) would differ
    assert abs(st["stddev"] - statistics.pstdev(samples)) < 1e-4
    assert abs(st["stddev"] - statistics.stdev(samples)) > 1e-3
    assert st["median"] == round(statistics.median(samples), 4)
    assert st["min"] == min(samples) and st["samples"] == len(samples)
    srt = sorted(samples)
    # numpy's default (linear) 95th percentile lies between the two largest samples here
    assert srt[-2] <= st["p95"] <= srt[-1]
    assert _bench().timing_stats([]) == {}


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_cpu_baseline_protocol_small():
    """The CPU baseline leg on a tiny field: forward and backward timed apart, the
    reference's statistics per pair, both precisions, and a round trip at rounding level."""
    import numpy as np
    b = _bench()
    r = b.run_cpu_baseline((16, 12, 10), pairs=4, warmup=1)
    assert r["kind"] == "port" and r["unit"] == "GFLOP/s" and r["cores"] >= 1
    assert "pocketfft" in r["sample"] and "FFTW unavailable" in r["sample"]
    for key, eps in (("f64", np.finfo(np.float64).eps), ("f32", np.finfo(np.float32).eps)):
        leg = r[key]
        assert leg["pair_ms"]["samples"] == 4 and leg["forward_ms"]["samples"] == 4
        assert leg["pair_ms"]["min"] <= leg["pair_ms"]["median"] <= leg["pair_ms"]["max"]
        assert leg["roundtrip_rel_l2"] < 10 * eps
    assert r["value"] == r["f64"]["gflops"] > 0


def test_traffic_figure_is_stamped_with_the_kernel_sources(tmp_path, monkeypatch):
    """roofline.traffic comes from profiles/traffic.json only when the file was measured
    on the current kernel sources; a summary stamped for other sources reads as stale."""
    import json
    b = _bench()
    stamp = b.source_stamp()
    assert len(stamp) == 16 and stamp == b.source_stamp()
    good = {"source_stamp": stamp, "passes": {"ffts3d_512_f32_fwd1": {"bytes_per_launch": 123.0}}}
    path = tmp_path / "traffic.json"
    path.write_text(json.dumps(good))
    monkeypatch.setattr(b, "TRAFFIC_PATH", str(path))
    assert b._traffic_summary("ffts3d_512_f32_fwd1") == (123.0, stamp)
    assert b._traffic_summary("ffts3d_512_f32_fwd9") == (None, stamp)
    good["source_stamp"] = "0" * 16
    path.write_text(json.dumps(good))
    assert b._traffic_summary("ffts3d_512_f32_fwd1") == (None, "0" * 16)


def _traffic_csvs(tmp_path, tag, npass=6, pairs=2, read_kib=100.0, write_kib=50.0):
    """Synthetic rocprofv3 csv files: a 1 GiB copy (read 512 MiB, written 1 GiB as the
    counters report it), then pairs * npass k_pass dispatches."""
    trace = tmp_path / f"{tag}_trace.csv"
    fetch = tmp_path / f"{tag}_fetch.csv"
    write = tmp_path / f"{tag}_write.csv"
    t_rows = ["Kernel_Name,Start_Timestamp,End_Timestamp"]
    f_rows = ["Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value"]
    w_rows = ["Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value"]
    f_rows.append(f"0,__amd_rocclr_copyBuffer,FETCH_SIZE,{512 * 1024}")
    w_rows.append(f"0,__amd_rocclr_copyBuffer,WRITE_SIZE,{1024 * 1024}")
    for d in range(pairs * npass):
        k = d % npass
        t_rows.append(f"k_pass<{k}>,{1000 * d},{1000 * d + 200 + k}")
        f_rows.append(f"{d + 1},k_pass<{k}>,FETCH_SIZE,{read_kib * (k + 1)}")
        w_rows.append(f"{d + 1},k_pass<{k}>,WRITE_SIZE,{write_kib * (k + 1)}")
    trace.write_text("\n".join(t_rows) + "\n")
    fetch.write_text("\n".join(f_rows) + "\n")
    write.write_text("\n".join(w_rows) + "\n")
    return str(trace), str(fetch), str(write)


def test_traffic_entries_are_keyed_by_size_and_merged_per_build(tmp_path, monkeypatch):
    """write_traffic keys the passes by transform size: a 1024^3 run adds its entries next
    to the 512^3 ones measured on the same kernel sources instead of replacing them, and
    the counters are scaled by the copy calibration (FETCH x 2, WRITE x 1)."""
    import json
    b = _bench()
    path = tmp_path / "traffic.json"
    monkeypatch.setattr(b, "TRAFFIC_PATH", str(path))
    b.write_traffic(*_traffic_csvs(tmp_path, "a"), pairs=2, size=512)
    b.write_traffic(*_traffic_csvs(tmp_path, "b", read_kib=200.0), pairs=2, size=1024)
    data = json.loads(path.read_text())
    assert data["source_stamp"] == b.source_stamp()
    p512, p1024 = data["passes"]["ffts3d_512_f32_fwd1"], data["passes"]["ffts3d_1024_f32_fwd1"]
    assert p512["read"] == 2 * 100.0 * 2 * 1024 and p512["write"] == 50.0 * 2 * 1024
    assert p1024["read"] == 2 * 200.0 * 2 * 1024
    assert p512["avg_us"] == 0.201
    assert b._traffic_summary("ffts3d_512_f32_fwd1")[0] == p512["bytes_per_launch"]
    # entries measured on other kernel sources are dropped, not merged
    data["source_stamp"] = "0" * 16
    path.write_text(json.dumps(data))
    b.write_traffic(*_traffic_csvs(tmp_path, "c"), pairs=2, size=1024)
    assert sorted(json.loads(path.read_text())["passes"]) == [f"ffts3d_1024_f32_{d}{k}" for d in ("bwd", "fwd")
                                                                for k in range(3)]
