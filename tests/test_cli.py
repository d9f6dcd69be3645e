"""main.out argument contract and its failure paths (src/main.c:26-52), on the CPU."""
import os
import subprocess

import pytest

from conftest import phpc


def run(args, cwd, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([str(phpc.BIN_PATH)] + args, cwd=cwd, capture_output=True, text=True, env=e,
                          timeout=120)


def test_usage_on_wrong_argc(tmp_path):
    r = run(["64", "32", "1"], tmp_path)
    assert r.returncode != 0
    assert "Usage:" in r.stderr and "<matrix_size> <tile_width> <grid_width> <grid_height> <test_name>" in r.stderr


def test_divisibility_check(tmp_path):
    r = run(["9", "32", "1", "1", "t", "--ranks", "4"], tmp_path)
    assert r.returncode != 0
    assert "must be divisible by process grid dimensions (2 x 2)" in r.stderr


def test_unknown_option(tmp_path):
    r = run(["16", "32", "1", "1", "t", "--bogus"], tmp_path)
    assert r.returncode != 0


def test_no_visible_device_fails_cleanly(tmp_path):
    (tmp_path / "csv").mkdir()
    r = run(["16", "32", "1", "1", "t"], tmp_path, env={"HIP_VISIBLE_DEVICES": "-1", "ROCR_VISIBLE_DEVICES": "-1"})
    assert r.returncode != 0
    assert "Error" in r.stderr
    assert not list((tmp_path / "csv").glob("*.csv"))


def test_bad_mode(tmp_path):
    r = run(["16", "32", "1", "1", "t", "--mode", "bogus"], tmp_path)
    assert r.returncode != 0
    assert "--mode must be" in r.stderr


ITER = phpc.BIN_PATH.parent / "iterative.out"


@pytest.mark.parametrize("dist,n", [("ramp", 1), ("ramp", 40), ("uniform", 33), ("wide", 48)])
def test_iterative_out_matches_oracle(tmp_path, dist, n):
    """bin/iterative.out, the harness's CPU path (src/iterative.c:15-46): prints "n,seconds"
    and its C is bit-identical to the oracle's restatement of the reference loop on the
    same generated inputs (which the device fill also produces)."""
    import numpy as np
    import oracle as O
    out = tmp_path / "c.bin"
    args = [str(ITER), str(n), "--out", str(out)]
    if dist != "ramp":
        args += ["--dist", dist, "--seed", "5"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    size, secs = r.stdout.strip().split(",")
    assert int(size) == n and float(secs) >= 0.0
    gen = {"ramp": lambda s, seed: O.ramp(s), "uniform": O.uniform, "wide": O.wide}[dist]
    A, B = gen((n, n), 5), gen((n, n), 6)
    want = O.iterative(A, B, np.zeros((n, n)))
    got = np.fromfile(out, dtype=np.float64).reshape(n, n)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    if dist == "ramp":  # the closed form of the reference's own input
        assert np.array_equal(got, O.ramp_closed_matrix(n))


def test_iterative_out_usage_and_errors(tmp_path):
    r = subprocess.run([str(ITER)], capture_output=True, text=True, timeout=30)
    assert r.returncode != 0 and "Usage:" in r.stderr
    for bad in (["8", "--dist", "normal"], ["8", "--bogus"], ["0"]):
        r = subprocess.run([str(ITER)] + bad, capture_output=True, text=True, timeout=30)
        assert r.returncode != 0 and "Error" in r.stderr, bad
