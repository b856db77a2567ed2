"""bench.py's summaries of rocprofv3 output (the figures quoted in profiles/):
per-kernel averages of a kernel trace, the working and the last launches of one
kernel, and HBM bytes per launch scaled by a copy kernel's counters, on small
synthetic CSV files with known answers."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout


def _trace(path):
    # kernel A: 4 launches of 10 us then one near-empty launch of 1 us (not working),
    # kernel B: 2 launches of 3 us; timestamps in ns
    rows, t = [], 0
    for name, dur in [("void dfm::k_spmv<3, float>(int)", 10000)] * 4 + [("void dfm::k_spmv<3, float>(int)", 1000)] + \
            [("dfm::k_fin(double*)", 3000)] * 2:
        rows.append({"Kernel_Name": name, "Start_Timestamp": t, "End_Timestamp": t + dur})
        t += dur + 500
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        w.writerows(rows)


def test_summarize_trace_stdout_and_file(tmp_path):
    tr = tmp_path / "trace.csv"
    _trace(tr)
    s = _bench("--summarize-trace", str(tr), "--kernel", "k_spmv<3, float>", "--tail", "3")
    lines = s.splitlines()
    spmv = [ln for ln in lines if ln.rstrip().endswith("k_spmv<3, float>") and not ln.startswith("#")][0].split()
    # share, calls, average, min / max
    assert spmv[1] == "5" and float(spmv[2]) == 8.2 and float(spmv[3]) == 1.0 and float(spmv[5]) == 10.0
    assert abs(float(spmv[0].rstrip("%")) - 100 * 41 / 47) < 0.01
    assert any("5 launches, average 8.20 us; the 4 that did work (>= half the median) average 10.00 us" in ln
               for ln in lines)
    # the last three: 10, 10, 1 -> 7.00; working 10.00
    assert any("the last 3 launches" in ln and "average 7.00 us; the 2 that did work average 10.00 us" in ln
               for ln in lines)
    out = tmp_path / "sum.txt"
    _bench("--summarize-trace", str(tr), "--kernel", "k_spmv<3, float>", "--tail", "3", "--out", str(out))
    assert out.read_text().strip() == s.strip()


def _pmc(path, counter, kernel_vals, copy_vals):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        d = 0
        for name, vals in (("dfm::k_spmv", kernel_vals), ("dfm::k_copy4", copy_vals)):
            for per_dispatch in vals:
                d += 1
                for v in per_dispatch:  # one row per XCD / instance, summed per dispatch
                    w.writerow({"Dispatch_Id": d, "Kernel_Name": name, "Counter_Name": counter, "Counter_Value": v})


def test_summarize_pmc_scales_by_copy(tmp_path):
    f, w = tmp_path / "fetch.csv", tmp_path / "write.csv"
    # copy of 1 MiB: FETCH 512 units (2048 B each), WRITE 1024 units (1024 B each)
    _pmc(f, "FETCH_SIZE", [[100, 100], [150, 150]], [[256, 256]])
    _pmc(w, "WRITE_SIZE", [[30, 20], [50, 0]], [[1000, 24]])
    s = _bench("--summarize-pmc", str(f), str(w), "--kernel", "k_spmv", "--copy-mib", "1",
               "--out", str(tmp_path / "t.txt"))
    d = json.loads(s.strip().splitlines()[-1])
    assert d["fetch_counter"] == 250.0 and d["write_counter"] == 50.0
    assert d["fetch_scale"] == 2048.0 and d["write_scale"] == 1024.0
    assert d["bytes_per_launch"] == 250 * 2048 + 50 * 1024
    assert d["dispatches"] == [2, 2]
