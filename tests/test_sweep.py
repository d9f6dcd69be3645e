"""The harness columns of scripts/tests.sh process() (:11-33) as sweep.py computes them,
on hand-written rows (CPU only)."""
import importlib.util

from conftest import PKG

spec = importlib.util.spec_from_file_location("sweep", PKG / "sweep.py")
sweep = importlib.util.module_from_spec(spec)
spec.loader.exec_module(sweep)


def test_bc_formatting():
    assert sweep.bc6(1, 3) == "0.333333"
    assert sweep.bc6(10, 4) == "2.500000"
    assert sweep.bc6(2, 3) == "0.666666"  # truncated, as bc with scale=6
    assert sweep.bc6(0, 5) == "0"
    assert sweep.bc6(1, 0) == ""


def test_process_columns():
    rows = ["1024,4,1,1,1024,1024,0.010000,0.004000,0.000000",
            "2048,16,1,4,1024,4096,0.040000,0.016000,0.020000"]
    out = sweep.process(rows, {1024: 1.5, 2048: 12.0})
    assert out[0] == sweep.HEADER
    assert out[1] == "1024,4,1,1,1024,1024,0.010000,0.004000,0.000000,150.000000,375.000000,,0.146484,0.366210"
    assert out[2] == "2048,16,1,4,1024,4096,0.040000,0.016000,0.020000,300.000000,750.000000,600.000000,0.073242,0.183105"


def test_reference_sweeps_listed():
    assert set(sweep.SWEEPS) == {"test0", "testA1", "testA2", "testB", "testC", "testD"}
    runs, key = sweep.SWEEPS["testA2"]
    assert (4, 2048, 32, 8, 8) in runs and key == "threads"


def test_gathered_csv_reads_like_the_reference_plots(tmp_path):
    """scripts/plots.py:183-208 reads a gathered sweep file with pandas and uses the
    columns matrix_size, n_proc, total_threads, time_cuda, time_cuda_gpu and time_cublas;
    the file sweep.py writes (with the process() columns) serves those reads unchanged."""
    import pandas as pd
    rows = ["1024,4,1,1,1024,1024,0.010000,0.004000,0.020000",
            "2048,16,1,4,1024,4096,0.040000,0.016000,0.050000"]
    lines = sweep.process(rows, {1024: 1.5, 2048: 12.0})
    path = tmp_path / "testB.csv"
    path.write_text("\n".join(lines) + "\n")
    results = pd.read_csv(path)
    total_threads = results["n_proc"] * results["total_threads"]
    assert list(total_threads) == [4096, 65536]
    assert list(results["matrix_size"]) == [1024, 2048]
    t1 = 1.5
    speedup = t1 / results["time_cuda"]
    assert abs(speedup[0] - 150.0) < 1e-9
    assert abs(results["time_cuda_gpu"][1] - 0.016) < 1e-12 and abs(results["time_cublas"][1] - 0.05) < 1e-12
    assert abs(results["speedup"][0] - 150.0) < 1e-9 and abs(results["efficiency_kernel"][1] - 0.183105) < 1e-9


def test_run_iterative_writes_the_harness_baseline_file(tmp_path):
    """--run-iterative: bin/iterative.out once per distinct size, "n,seconds" lines in
    csv/iterative.csv, which read_iterative() then feeds to process()."""
    out = sweep.run_iterative([48, 16, 48, 32], tmp_path)
    assert out == tmp_path / "csv" / "iterative.csv"
    lines = out.read_text().split()
    assert [ln.split(",")[0] for ln in lines] == ["16", "32", "48"]
    times = sweep.read_iterative(out)
    assert sorted(times) == [16, 32, 48] and all(t >= 0.0 for t in times.values())
    row = "48,4,1,1,1024,1024,0.001,0.0005,0.002"
    got = sweep.process([row], {48: 1.0})[1].split(",")
    assert got[9:12] == ["1000.000000", "2000.000000", "500.000000"]
