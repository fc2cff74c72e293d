"""The profile summariser on synthetic rocpd databases: kernel durations,
the copy calibration of FETCH_SIZE / WRITE_SIZE and the per-primary traffic
and instruction figures the benchmark's roofline reads."""

import json
import sqlite3

import pytest

from ggems_amd import profiling as P


def _db(path, kernels=(), counters=(), top=()):
    path.mkdir(parents=True, exist_ok=True)
    con = sqlite3.connect(str(path / "run_results.db"))
    con.execute("create table kernels (name text, duration integer)")
    con.execute("create table counters_collection (dispatch_id integer, grid_size integer, value real, "
                "counter_name text, kernel_name text)")
    con.execute("create table top_kernels (name text, total_calls integer, total_duration real, average real, "
                "percentage real)")
    con.executemany("insert into kernels values (?, ?)", kernels)
    con.executemany("insert into counters_collection values (?, ?, ?, ?, ?)", counters)
    con.executemany("insert into top_kernels values (?, ?, ?, ?, ?)", top)
    con.commit()
    con.close()
    return str(path)


KERNEL = "void (anonymous namespace)::transport_kernel<3, true, true, false, false, true>(RunParams)"
COPY = "__amd_rocclr_copyBuffer"


def _runs(tmp_path):
    mib = 256
    copy_bytes = mib << 20
    kt = _db(tmp_path / "kt", kernels=[(KERNEL, 2_000_000_000), (KERNEL, 3_000_000_000), (COPY, 50_000)],
             top=[(KERNEL, 2, 5.0e6, 2.5e6, 99.9), (COPY, 1, 50.0, 50.0, 0.1)])
    # the copy reads half of its bytes according to FETCH_SIZE (factor 2) and
    # all of them according to WRITE_SIZE (factor 1); a small copy is ignored
    fetch = _db(tmp_path / "fetch", counters=[
        (1, 1 << 20, copy_bytes / 2048.0, "FETCH_SIZE", COPY),
        (2, 64, 1.0, "FETCH_SIZE", COPY),
        (3, 512, 1000.0, "FETCH_SIZE", KERNEL)])
    write = _db(tmp_path / "write", counters=[
        (1, 1 << 20, copy_bytes / 1024.0, "WRITE_SIZE", COPY),
        (3, 512, 250.0, "WRITE_SIZE", KERNEL)])
    sq = _db(tmp_path / "sq", counters=[(3, 512, 6.0e6, "SQ_INSTS_VALU", KERNEL),
                                        (3, 512, 1.0e6, "SQ_INSTS_SALU", KERNEL),
                                        (3, 512, 300.0, "TCC_HIT_sum", KERNEL),
                                        (3, 512, 100.0, "TCC_MISS_sum", KERNEL)])
    return kt, fetch, write, sq, mib


def test_summary_figures(tmp_path):
    kt, fetch, write, sq, mib = _runs(tmp_path)
    s = P.summarize(kt, fetch, write, sq, copy_mb=mib, photons=1000, kernel="transport_kernel", steps=2.0e6)
    assert s["calls"] == 2 and s["avg_ns"] == 2.5e9
    assert s["fetch_factor"] == pytest.approx(2.0) and s["write_factor"] == pytest.approx(1.0)
    assert s["fetch_bytes"] == pytest.approx(1000 * 1024 * 2.0)
    assert s["write_bytes"] == pytest.approx(250 * 1024)
    assert s["traffic_bytes_per_primary"] == pytest.approx((2048000 + 256000) / 1000)
    assert s["l2_hit_rate"] == pytest.approx(0.75)
    text = P.format_summary(s, ["a command"])
    assert "traffic per primary: 2304.0 B" in text and "3.000 per voxel step" in text and "a command" in text


def test_trace_only_summary(tmp_path):
    kt, _, _, _, _ = _runs(tmp_path)
    s = P.summarize(kt, kernel="transport_kernel")
    assert "traffic_bytes" not in s
    assert "copy calibration" not in P.format_summary(s)


def test_json_record_carries_the_source_digest(tmp_path, monkeypatch):
    kt, fetch, write, sq, mib = _runs(tmp_path)
    out = tmp_path / "rec.json"
    monkeypatch.setattr("sys.argv", ["profiling", "--kt", kt, "--fetch", fetch, "--write", write, "--sq", sq,
                                     "--copy-mb", str(mib), "--photons", "1000", "--steps", "2000000",
                                     "--json", str(out)])
    P.main()
    rec = json.loads(out.read_text())
    assert rec["source_sha256"] == P.source_digest()
    assert rec["valu_instr_per_step"] == pytest.approx(3.0) and rec["salu_instr_per_step"] == pytest.approx(0.5)


def test_benchmark_byte_model_and_multi_gpu_refusal():
    import subprocess
    import sys

    import bench
    st = {"primaries": 1000, "voxel_steps": 286_000, "interactions": 2_500, "detector_interactions": 2}
    b, s, i, d = bench.algorithmic_bytes_per_primary(st)
    assert (s, i, d) == (286.0, 2.5, 0.002)
    assert b == pytest.approx(286.0 + 2.5 * 40 + 0.002 * 8 + 48)
    # one GPU per box: more than one exits with a message before touching a device
    r = subprocess.run([sys.executable, bench.__file__, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
