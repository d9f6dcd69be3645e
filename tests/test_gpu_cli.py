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
