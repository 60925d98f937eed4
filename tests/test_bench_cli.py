"""bench.py's command line on the CPU: the benchmark is one GPU, so --gpus N
with N > 1 exits (status 2) with a message before touching any device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_bench_lists_the_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0
    for w in ("d2h-120", "c2v-24", "c2v-180", "c2v-batch32", "c2v-100-gradient", "d2h-120-16roots"):
        assert w in r.stdout, w


def test_summarize_gathers_joins_trace_counters_and_launch_list(tmp_path):
    """--summarize-gathers on a synthetic trace: two marker sections (a sigma
    call with one expansion, one pack and one GEMM; one sort), a calibration
    copy; the GB/s it prints follow from the bytes and durations given."""
    import csv
    import json
    mk = "void dpd::k_precond(int, double)"
    seq = [(mk, 0), ("void dpd::k_gather_code<double, double, 4>(x)", 10_000), ("void dpd::k_pack_code<2, double>(y)", 4_000),
           ("void dpd::k_gemm_pk<2, 128, 64>(z)", 50_000), (mk, 0), ("void dpd::k_gather_map(w)", 0), (mk, 0),
           ("void dpd::k_gather_map(w)", 20_000), (mk, 0), ("copy_kernel", 100_000)]
    tr = tmp_path / "kernel_trace.csv"
    with open(tr, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        t = 1_000
        for name, dur in seq:
            w.writerow([name, t, t + dur])
            t += dur + 100
    # FETCH_SIZE reads half the bytes fetched (scale 2), WRITE_SIZE all of them
    fetched = [0, 1e6, 1e6, 4e6, 0, 8e6, 0, 8e6, 0, 2**29]
    written = [0, 2e6, 1e6, 1e6, 0, 8e6, 0, 8e6, 0, 2**29]
    for name, vals, div in (("fetch.csv", fetched, 2048.0), ("write.csv", written, 1024.0)):
        with open(tmp_path / name, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp",
                        "End_Timestamp"])
            for i, ((kname, dur), v) in enumerate(zip(seq, vals)):
                w.writerow([i + 1, kname, "FETCH_SIZE" if name == "fetch.csv" else "WRITE_SIZE", v / div, 0, dur])
    meta = {"precision": "bf16x3", "workload": "d2h-120", "copy_bytes": 2**29,
            "sigma_records": [["t:x1", "gather", 1, 3e6], ["t:bf16x3:t2", "gemm", 1, 9e6], ["t:pack", "gather", 1, 2e6]],
            "sorts": [["sstgenso '3142'", 16e6]]}
    pj = tmp_path / "probe.json"
    pj.write_text(json.dumps(meta))
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize-gathers", str(tr),
                        str(tmp_path / "fetch.csv"), str(tmp_path / "write.csv"), str(pj)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "scale 2.000" in out and "scale 1.000" in out
    # expansion: 3 MB in 10 us = 300 GB/s; traffic 1 + 2 MB = 300 GB/s
    assert "| expansions | 1 | 10.0 | 3.0 | 300 | 0.04 | 3.0 | 300 | 0.04 |" in out
    # pack: 2 MB in 4 us = 500 GB/s; traffic 2 MB
    assert "| operand packing | 1 | 4.0 | 2.0 | 500 | 0.06 | 2.0 | 500 | 0.06 |" in out
    # the sort: the second call (20 us), 16 MB -> 800 GB/s, traffic 16 MB
    assert "sstgenso '3142' (`dpd::k_gather_map`) | 1 | 20.0 | 16.0 | 800 | 0.10 | 16.0 | 800 | 0.10 |" in out


def test_summarize_gemms_derives_busy_share_and_counted_flops(tmp_path):
    """--summarize-gemms on a synthetic trace and SQ counter run: one GEMM launch
    of 10 us with 1 GFLOP (100 TFLOP/s, 1.27 of the FP64 peak is impossible, so
    FP32: 0.64), a split-K reduce that is not a GEMM launch, busy cycles of half
    the SIMD-time and MOPS that count 2 GFLOP."""
    import csv
    import json
    mk = "void dpd::k_precond(int, double)"
    seq = [(mk, 0), ("void dpd::k_pack_code<1, double>(y)", 3_000), ("void dpd::k_gemm_pk<1, 128, 64>(z)", 10_000),
           ("void dpd::k_splitk_reduce(r)", 1_000), (mk, 0)]
    tr = tmp_path / "kernel_trace.csv"
    with open(tr, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        t = 1_000
        for name, dur in seq:
            w.writerow([name, t, t + dur])
            t += dur + 100
    busy = 0.5 * 1024 * 10e-6 * 2.4e9
    with open(tmp_path / "pmc.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp",
                    "End_Timestamp"])
        for i, (kname, dur) in enumerate(seq):
            w.writerow([i + 1, kname, "SQ_VALU_MFMA_BUSY_CYCLES", busy if "gemm" in kname else 0, 0, dur])
            w.writerow([i + 1, kname, "SQ_INSTS_VALU_MFMA_MOPS_F32", 2e9 / 512 if "gemm" in kname else 0, 0, dur])
    meta = {"precision": "fp32", "workload": "d2h-120",
            "sigma_records": [["t:x1", "gather", 1, 1e6, 0, 0], ["t:fp32:t2", "gemm", 1, 5e6, 1e9, 1]]}
    pj = tmp_path / "probe.json"
    pj.write_text(json.dumps(meta))
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--summarize-gemms", str(tr),
                        str(tmp_path / "pmc.csv"), str(pj)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "| t:fp32:t2 | `dpd::k_gemm_pk<1, 128, 64>` | 10.0 | 1.00 | 100.0 | 0.64 | 0.50 | 200.0 |" in r.stdout
