"""Benchmark sweeps in the shape of the reference's harness (scripts/tests.sh:3-33).

Each run is `main.out N T GW GH <name>_N<N>_T<P>_G1_TW<T>_GW<GW>_GH<GH> --ranks P`, the
command line of tests.sh run() (:3-9) with `mpirun -n P` replaced by the emulated rank
count. The rows of one sweep are gathered and sorted into csv/<name>.csv, and, given a
file of CPU baseline times (`n,seconds` lines, the output of the reference's
iterative.out), the derived columns of tests.sh process() (:11-33) are appended:
speedup = t_iterative / time, speedup_kernel = t_iterative / kernel_time,
speedup_cublas = t_iterative / cublas_time, efficiency = speedup / total_threads,
efficiency_kernel = speedup_kernel / total_threads, printed with six decimals.

    python sweep.py --sweep testB [--iterative csv/iterative.csv] [--extra "--mode chained"]
    python sweep.py --sweep testB --run-iterative      # writes csv/iterative.csv first
    python sweep.py --runs "4 512 32 1 1,16 1024 32 2 2" --name mine

--run-iterative produces the CPU baseline file the way the harness's sequential block does
(scripts/tests.sh:36-44): bin/iterative.out once per matrix size of the sweep, its "n,seconds"
lines gathered into csv/iterative.csv.
"""
from __future__ import annotations

import argparse
import shlex
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
MAIN = HERE / "bin" / "main.out"
ITERATIVE = HERE / "bin" / "iterative.out"

# (P, N, T, GW, GH) per run and the sort key of the gathered file (scripts/tests.sh:45-100)
SWEEPS = {
    "test0": ([(4, n, 32, 1, 1) for n in (32, 64, 128, 256, 512, 1024, 2048, 4096)], "size"),
    "testA1": ([(p, 2048, 32, 1, 1) for p in (2, 4, 8, 16)], "proc"),
    "testA2": ([(4, 2048, t, 1, 1) for t in (4, 8, 16, 32)] + [(4, 2048, 32, g, g) for g in (2, 4, 8)], "threads"),
    "testB": ([(1, 1024, 32, 1, 1), (4, 2048, 32, 1, 1), (16, 4096, 32, 1, 1)], "size"),
    "testC": ([(4, 64, 4, 1, 1), (4, 128, 8, 1, 1), (4, 256, 16, 1, 1), (4, 512, 32, 1, 1), (4, 1024, 32, 2, 2),
               (4, 2048, 32, 4, 4)], "size"),
    "testD": ([(1, 1024, 32, 1, 1), (4, 2048, 32, 1, 1), (16, 4096, 32, 1, 1)], "size"),
}

HEADER = ("matrix_size,n_proc,n_gpu,n_block,n_thread_per_block,total_threads,time_cuda,time_cuda_gpu,time_cublas,"
          "speedup,speedup_kernel,speedup_cublas,efficiency,efficiency_kernel")


def run_one(P, N, T, GW, GH, name, workdir, extra):
    """One main.out run; returns the CSV row it wrote."""
    fname = f"{name}_N{N}_T{P}_G1_TW{T}_GW{GW}_GH{GH}"
    cmd = [str(MAIN), str(N), str(T), str(GW), str(GH), fname, "--ranks", str(P)] + extra
    r = subprocess.run(cmd, cwd=workdir, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed ({r.returncode}): {r.stderr.strip()}")
    out = Path(workdir) / "csv" / f"{fname}_N{N}_T{P}_G1_TW{T}_GW{GW}_GH{GH}.csv"
    row = out.read_text().strip()
    out.unlink()
    return row


def bc6(num, den):
    """`bc -l <<< "scale=6; num / den"` as tests.sh prints it: truncated to six decimals,
    a leading zero added by its sed step, '0' for zero and nothing when den is zero."""
    if den == 0:
        return ""
    q = num / den
    v = int(abs(q) * 10 ** 6) / 10 ** 6
    if v == 0:
        return "0"
    return ("-" if q < 0 else "") + f"{v:.6f}"


def process(rows, iterative):
    """The derived columns of scripts/tests.sh process() for each 9-column row."""
    out = [HEADER]
    for row in rows:
        f = row.split(",")
        size, tot = int(f[0]), int(f[5])
        time, kernel, cublas = float(f[6]), float(f[7]), float(f[8])
        base = iterative.get(size)
        if base is None:
            raise KeyError(f"no iterative time for matrix size {size}")
        sp, spk, spc = bc6(base, time), bc6(base, kernel), bc6(base, cublas)
        eff = bc6(float(sp), tot) if sp else ""
        effk = bc6(float(spk), tot) if spk else ""
        out.append(",".join(f[:9] + [sp, spk, spc, eff, effk]))
    return out


def read_iterative(path):
    times = {}
    for line in Path(path).read_text().splitlines():
        if line.strip():
            n, secs = line.split(",")[:2]
            times[int(n)] = float(secs)
    return times


def run_iterative(sizes, workdir=".", exe=ITERATIVE):
    """bin/iterative.out for each size, its "n,seconds" lines into csv/iterative.csv (the
    harness's sequential block, scripts/tests.sh:36-44); returns that file's path."""
    (Path(workdir) / "csv").mkdir(parents=True, exist_ok=True)
    lines = []
    for n in sorted(set(sizes)):
        r = subprocess.run([str(exe), str(n)], cwd=workdir, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{exe} {n} failed ({r.returncode}): {r.stderr.strip()}")
        lines.append(r.stdout.strip())
    out = Path(workdir) / "csv" / "iterative.csv"
    out.write_text("\n".join(lines) + "\n")
    return out


def sweep(runs, sort_key, name, workdir=".", iterative=None, extra=()):
    (Path(workdir) / "csv").mkdir(parents=True, exist_ok=True)
    rows = [run_one(*r, name, workdir, list(extra)) for r in runs]
    col = {"size": 0, "proc": 1, "threads": 5}[sort_key]
    rows.sort(key=lambda r: (int(r.split(",")[col]), r))
    lines = process(rows, iterative) if iterative is not None else rows
    out = Path(workdir) / "csv" / f"{name}.csv"
    out.write_text("\n".join(lines) + "\n")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--sweep", choices=sorted(SWEEPS), help="one of the reference harness's sweeps")
    ap.add_argument("--runs", help="comma-separated 'P N T GW GH' runs (instead of --sweep)")
    ap.add_argument("--name", help="output name (default: the sweep's)")
    ap.add_argument("--iterative", help="CPU baseline times, 'n,seconds' per line")
    ap.add_argument("--run-iterative", action="store_true",
                    help="first run bin/iterative.out for every size of the sweep into csv/iterative.csv")
    ap.add_argument("--extra", default="", help="extra main.out options, e.g. '--mode chained --dist uniform'")
    ap.add_argument("--workdir", default=".")
    a = ap.parse_args(argv)
    if not MAIN.exists():
        sys.exit(f"{MAIN} is not built (make -C {HERE})")
    if a.sweep:
        runs, key = SWEEPS[a.sweep]
        name = a.name or a.sweep
    elif a.runs:
        runs = [tuple(int(x) for x in r.split()) for r in a.runs.split(",")]
        key, name = "size", a.name or "sweep"
    else:
        ap.error("give --sweep or --runs")
    if a.run_iterative:
        a.iterative = str(run_iterative([r[1] for r in runs], a.workdir))
    it = read_iterative(a.iterative) if a.iterative else None
    out = sweep(runs, key, name, a.workdir, it, shlex.split(a.extra))
    print(out)


if __name__ == "__main__":
    main()
