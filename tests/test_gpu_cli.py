"""main.out end to end on the GPU: the reference's positional contract, the CSV row in
the format of src/utils.c:24-30, and the --check verification of C."""
import re
import subprocess

import pytest

from conftest import phpc

pytestmark = pytest.mark.gpu

ROW = re.compile(r"^(\d+),(\d+),(\d+),(\d+),(\d+),(\d+),(\d+\.\d{6}),(\d+\.\d{6}),(\d+\.\d{6})\n$")


def run(tmp_path, args):
    (tmp_path / "csv").mkdir(exist_ok=True)
    r = subprocess.run([str(phpc.BIN_PATH)] + args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(args, r.returncode, r.stderr, flush=True)
    return r


@pytest.mark.parametrize("args,N,P,T,GW,GH", [
    (["256", "32", "2", "2", "cli", "--ranks", "4", "--dist", "uniform", "--check", "64"], 256, 4, 32, 2, 2),
    (["512", "16", "1", "1", "ramp", "--check", "32"], 512, 1, 16, 1, 1),
    (["301", "8", "4", "1", "un", "--ranks", "9", "--uneven", "--mode", "chained", "--check", "64"], 301, 9, 8, 4, 1),
    (["384", "32", "1", "1", "sp", "--ranks", "4", "--dist", "wide", "--mode", "split", "--check", "64",
      "--no-vendor"], 384, 4, 32, 1, 1),
    (["256", "32", "1", "1", "st", "--ranks", "16", "--mode", "stepwise", "--check", "16"], 256, 16, 32, 1, 1),
])
def test_cli_row_and_check(tmp_path, args, N, P, T, GW, GH):
    r = run(tmp_path, args)
    assert r.returncode == 0, r.stderr
    assert "check:" in r.stderr and r.stderr.rstrip().endswith("ok")
    name = args[4]
    f = tmp_path / "csv" / f"{name}_N{N}_T{P}_G1_TW{T}_GW{GW}_GH{GH}.csv"
    text = f.read_text()
    m = ROW.match(text)
    assert m, text
    vals = [int(x) for x in m.groups()[:6]]
    assert vals == [N, P, 1, GW * GH, T * T, GW * GH * T * T]
    assert float(m.group(7)) > 0 and float(m.group(8)) > 0
    assert (float(m.group(9)) > 0) == ("--no-vendor" not in args)


def test_sweep_gathers_rows_and_process_columns(tmp_path):
    """sweep.py runs the reference harness's testC sweep shape and appends process()'s
    columns from a CPU-baseline file."""
    import importlib.util
    from conftest import PKG
    spec = importlib.util.spec_from_file_location("sweep", PKG / "sweep.py")
    sw = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(sw)
    runs = [(4, 64, 4, 1, 1), (4, 128, 8, 1, 1), (1, 96, 32, 2, 2)]
    it = {64: 0.001, 128: 0.008, 96: 0.003}
    out = sw.sweep(runs, "size", "tc", tmp_path, it, ["--no-vendor"])
    lines = out.read_text().splitlines()
    assert lines[0] == sw.HEADER
    assert [int(l.split(",")[0]) for l in lines[1:]] == [64, 96, 128]
    for l in lines[1:]:
        f = l.split(",")
        assert len(f) == 14 and float(f[9]) > 0 and f[11] == ""  # no vendor pass: empty speedup_cublas
    assert not list((tmp_path / "csv").glob("tc_*.csv"))
