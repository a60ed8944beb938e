"""Host logic of bench.py (no GPU): the one-GPU refusal and the rocprofv3
summariser that cuts the timed region out of a kernel trace by the clock
window the bench logs, groups the GEMM instantiations and averages counters."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def _write(path, header, rows):
    with open(path, "w") as f:
        f.write(",".join(header) + "\n")
        for r in rows:
            f.write(",".join(str(x) for x in r) + "\n")


def test_summarize_rocprof_timed_window_and_grouping(tmp_path, capsys):
    d = tmp_path / "prof"
    d.mkdir()
    f64 = "void peps::gemm_kernel<false, 0, false, true, 4, 4>(long, long)"
    f64b = "void peps::gemm_kernel<false, 0, true, true, 2, 2>(long, long)"
    tiled = "void peps::gemm_f32_tiled<float, false, false>(long, long)"
    svd = "void peps::jacobi_svd_reg_batched_kernel<false, 16, 10>(peps::JacobiItem const*, int)"
    _write(d / "c3_kernel_stats.csv", ["Name", "Calls", "TotalDurationNs"],
           [[f'"{f64}"', 3, 3000], [f'"{f64b}"', 1, 1000], [f'"{svd}"', 2, 8000]])
    # timed window [1000, 2000] ns of CLOCK_BOOTTIME; one dispatch of each
    # kind lies outside it
    trace = [[f'"{f64}"', 1100, 1200], [f'"{f64b}"', 1300, 1350], [f'"{f64}"', 500, 600],
             [f'"{tiled}"', 1400, 1500], [f'"{svd}"', 1500, 1900], [f'"{svd}"', 2100, 2500]]
    _write(d / "c3_kernel_trace.csv", ["Kernel_Name", "Start_Timestamp", "End_Timestamp"], trace)
    _write(d / "c3_counter_collection.csv",
           ["Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"],
           [[f'"{f64}"', "FETCH_SIZE", 10.0, 1100, 1200], [f'"{f64b}"', "FETCH_SIZE", 30.0, 1300, 1350],
            [f'"{f64}"', "FETCH_SIZE", 999.0, 500, 600]])
    log = tmp_path / "bench.log"
    log.write_text("[bench] timed region clock_boottime_ns=1000..2000 clock_monotonic_ns=1..2\n")
    bench.summarize_rocprof(str(d), str(log))
    out = capsys.readouterr().out
    assert "timed region: 4 dispatches, wall 0.000 s, kernel time 0.000 s" in out
    # both f64 instantiations are one group; the dispatch before the window is dropped
    timed = out.split("timed region:")[1].split("counters")[0]
    g = [l for l in timed.splitlines() if "gemm_kernel PREC=0" in l][0]
    assert "calls        2" in g and "total      0.000 ms" in g
    assert any("gemm_f32_tiled" in l for l in timed.splitlines())
    s = [l for l in timed.splitlines() if "jacobi_svd_reg_batched_kernel" in l][0]
    assert "calls        1" in s
    # counters: mean over the two in-window dispatches only
    c = [l for l in out.split("counters")[1].splitlines() if "FETCH_SIZE" in l][0]
    assert "dispatches       2" in c and "mean 2.0000e+01 KiB" in c
