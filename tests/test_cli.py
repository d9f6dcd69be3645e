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
