"""Benchmark: FP64 SUMMA on one MI355X (N = 16384, emulated 4 x 4 grid, k-panel 256).

One step = one SUMMA pass C += A*B over the 64 panels for all 16 emulated ranks, run by
the plain FP64-MFMA path (one persistent grouped launch). Inputs are uniform[-1,1) from
fixed seeds and resident in HBM before the timed region starts. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import importlib.util
import json
import os
import shutil
import statistics
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "hpc-multigpu-matrixmult_amd"
PROFILES = ROOT / "profiles"

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak (AMD public spec)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_phpc():
    spec = importlib.util.spec_from_file_location("phpc", PKG / "phpc.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["phpc"] = mod
    spec.loader.exec_module(mod)
    mod.lib()
    return mod


def cpu_baseline():
    """The reference's CPU path (src/iterative.c), timed on one host core at N = 1024."""
    ref = ROOT / "oracle" / "_ref" / "iterative.out"
    port = ROOT / "oracle" / "_build" / "iterative_O0.out"
    exe, kind = (ref, "reference") if ref.exists() else (port, "port")
    if not exe.exists():
        return None
    n, runs = 1024, 3
    cmd = [str(exe), str(n)]
    if shutil.which("taskset"):
        cmd = ["taskset", "-c", "0"] + cmd
    secs = []
    for _ in range(runs):
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        if out.returncode != 0:
            return None
        secs.append(float(out.stdout.strip().split(",")[1]))
    t = statistics.median(secs)
    tflops = 2.0 * n ** 3 / t / 1e12
    return {
        "value": tflops,
        "unit": "TFLOP/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"{exe.relative_to(ROOT)} n={n} (ramp inputs, gcc -O0 as the reference Makefile), "
                   f"median of {runs} runs = {t:.3f} s on 1 pinned core of {os.cpu_count()}; "
                   f"N=16384 would take ~{t * 16 ** 3 / 3600:.1f} h (O(N^3) extrapolation)"),
    }


def traffic_from_profiles(workload):
    f = PROFILES / "traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        return d.get(workload)
    except Exception:
        return None


def grouped_bench(phpc, seed=21):
    """(a) N = 10000 FP64 on an emulated 3 x 3 grid (rank blocks 3334/3333/3333), panel
    N / lcm = 3333, and (b) 512 independent FP32 GEMMs with M, N, K log-uniform in
    [64, 4096], each set in one persistent launch replayed from a captured hipGraph, with
    the tile configuration chosen by the autotuner."""
    import numpy as np
    import torch

    out = {}
    s = torch.cuda.Stream()
    # (a)
    N, p, q = 10000, 3, 3
    A = torch.empty((N, N), dtype=torch.float64, device="cuda")
    B = torch.empty_like(A)
    phpc.fill_device(A, phpc.DIST_UNIFORM, seed)
    phpc.fill_device(B, phpc.DIST_UNIFORM, seed + 1)
    C = torch.zeros_like(A)
    starts = [0, 3334, 6667, 10000]
    items = [(A[starts[i]:starts[i + 1]], B[:, starts[j]:starts[j + 1]],
              C[starts[i]:starts[i + 1], starts[j]:starts[j + 1]], N // 3) for i in range(p) for j in range(q)]
    plan = phpc.Plan(phpc.F64, items, autotune=True, stream=s)
    C.zero_()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        plan.graph_launch(s)  # capture + first replay
    s.synchronize()
    C.zero_()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 3
    with torch.cuda.stream(s):
        e0.record(s)
        for _ in range(reps):
            plan.graph_launch(s)
        e1.record(s)
    s.synchronize()
    ms = e0.elapsed_time(e1) / reps
    C.zero_()
    with torch.cuda.stream(s):
        plan.graph_launch(s)
    s.synchronize()
    rng = np.random.default_rng(seed)
    rows = np.concatenate([rng.integers(0, N, 2048), [3333, 3334, 6666, 6667, 9999]])
    cols = np.concatenate([rng.integers(0, N, 2048), [3333, 3334, 6666, 6667, 9999]])
    cw, nw = phpc.check_sampled(A, B, C, rows, cols)
    info = plan.info()
    out["fp64_n10000_grid3x3"] = {"tflops": round(2.0 * N ** 3 / (ms * 1e-3) / 1e12, 3), "ms": round(ms, 3),
                                  "blocks": "3334/3333/3333", "panel": N // 3, "config": info["config"],
                                  "tiles": info["tiles"], "err_cw_max_sampled": cw, "err_nw_sampled": nw,
                                  "launch": "one persistent panel-grouped launch for all 9 rank blocks and all "
                                            "k panels (3333, 3333, 3333, 1), hipGraph replay"}
    plan.close()
    del A, B, C, items
    torch.cuda.empty_cache()
    # (b)
    rng = np.random.default_rng(seed)
    shapes = np.rint(np.exp(rng.uniform(np.log(64), np.log(4096), size=(512, 3)))).astype(int)
    gemms = []
    flops = 0.0
    for i, (m, n, k) in enumerate(shapes):
        a = torch.empty((m, k), dtype=torch.float32, device="cuda")
        b = torch.empty((k, n), dtype=torch.float32, device="cuda")
        phpc.fill_device(a, phpc.DIST_UNIFORM, 10000 + 2 * i)
        phpc.fill_device(b, phpc.DIST_UNIFORM, 10001 + 2 * i)
        gemms.append((a, b, torch.zeros((m, n), dtype=torch.float32, device="cuda")))
        flops += 2.0 * m * n * k
    plan = phpc.Plan(phpc.F32, gemms, autotune=True, stream=s)
    for g in gemms:
        g[2].zero_()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        plan.graph_launch(s)
    s.synchronize()
    with torch.cuda.stream(s):
        e0.record(s)
        for _ in range(reps):
            plan.graph_launch(s)
        e1.record(s)
    s.synchronize()
    ms = e0.elapsed_time(e1) / reps
    for g in gemms:
        g[2].zero_()
    with torch.cuda.stream(s):
        plan.graph_launch(s)
    s.synchronize()
    # accuracy: 16 GEMMs x 64 entries against float64 dot products of the FP32 inputs
    worst = 0.0
    for gi in rng.choice(512, 16, replace=False):
        a, b, c = gemms[gi]
        m, k = a.shape
        n = b.shape[1]
        r = rng.integers(0, m, 64)
        cc = rng.integers(0, n, 64)
        ar = a[r].double().cpu().numpy()
        bc = b[:, cc].double().cpu().numpy()
        ref = np.einsum("ik,ki->i", ar, bc)
        scale = np.einsum("ik,ki->i", np.abs(ar), np.abs(bc))
        got = c[r, cc].double().cpu().numpy()
        worst = max(worst, float(np.max(np.abs(got - ref) / np.maximum(scale, 1e-300))))
    info = plan.info()
    out["fp32_512_gemms"] = {"tflops": round(flops / (ms * 1e-3) / 1e12, 3), "ms": round(ms, 3),
                             "flops": flops, "config": info["config"], "tiles": info["tiles"],
                             "tuned_ms": round(info["tuned_ms"], 3), "err_cw_max_sampled": worst,
                             "fp32_mfma_util": round(flops / (ms * 1e-3) / 157.3e12, 4),
                             "launch": ("one batched kernel transposing every A + one persistent "
                                        "transposed-A kernel for all 512 GEMMs (B read in place)"
                                        if info["config"] == 3 else
                                        "one persistent kernel for all 512 GEMMs") +
                                       ", items in decreasing k, hipGraph replay"}
    plan.close()
    del gemms
    torch.cuda.empty_cache()
    return out


def split_sweep(phpc, sizes=(4096, 8192), seed=31):
    """Split-precision FP64 GEMM at each slice / modulus count: error against the
    double-double product on sampled entries and FP64-equivalent TFLOP/s."""
    import numpy as np
    import torch

    out = []
    for N in sizes:
        A = torch.empty((N, N), dtype=torch.float64, device="cuda")
        B = torch.empty_like(A)
        C = torch.zeros_like(A)
        rng = np.random.default_rng(seed)
        rows, cols = rng.integers(0, N, 2048), rng.integers(0, N, 2048)
        for dist, tag in ((phpc.DIST_UNIFORM, "uniform"), (phpc.DIST_WIDE, "wide_exponent")):
            phpc.fill_device(A, dist, seed)
            phpc.fill_device(B, dist, seed + 1)
            runs = [(phpc.SPLIT_CRT, 0)] + [(phpc.SPLIT_CRT, L) for L in (8, 10, 12, 14, 16)] + \
                   [(phpc.SPLIT_SLICES, s_) for s_ in (3, 4, 5, 6, 7, 8)]
            for method, count in runs:
                C.zero_()
                info = phpc.dgemm_split_device(A, B, C, slices=count, target=1e-14, method=method)
                torch.cuda.synchronize()
                cw, nw = phpc.check_sampled(A, B, C, rows, cols)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(2):
                    phpc.dgemm_split_device(A, B, C, slices=count, target=1e-14, method=method)
                e1.record()
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / 2
                out.append({"N": N, "inputs": tag, "method": "crt" if method == phpc.SPLIT_CRT else "slices",
                            "requested": count if count else "auto", "count": info.slices,
                            "int8_gemms": info.products, "int_bits": info.int_bits, "err_cw_max_sampled": cw,
                            "error_estimate": info.error_estimate,
                            "tflops_fp64_equiv": round(2.0 * N ** 3 / (ms * 1e-3) / 1e12, 3), "ms": round(ms, 3)})
        del A, B, C
        torch.cuda.empty_cache()
    phpc.lib().phpc_split_set_method(phpc.SPLIT_CRT)
    return out


def fp32_accurate_bench(phpc, sizes=(4096, 8192, 16384), seed=41):
    """FP32 inputs, uniform[-1,1): the FP32 matrix-core GEMM and the FP32-accurate 3 x BF16
    GEMM (six BF16 slice products, FP32 accumulation). Error against float64 dot products
    of the same FP32 inputs on sampled entries, normalised by |A||B|."""
    import numpy as np
    import torch

    out = []
    for N in sizes:
        A = torch.empty((N, N), dtype=torch.float32, device="cuda")
        B = torch.empty_like(A)
        C = torch.zeros_like(A)
        phpc.fill_device(A, phpc.DIST_UNIFORM, seed)
        phpc.fill_device(B, phpc.DIST_UNIFORM, seed + 1)
        rng = np.random.default_rng(seed)
        r = rng.integers(0, N, 256)
        c = rng.integers(0, N, 256)
        ar = A[r].double().cpu().numpy()
        bc = B[:, c].double().cpu().numpy()
        ref = np.einsum("ik,ki->i", ar, bc)
        scale = np.einsum("ik,ki->i", np.abs(ar), np.abs(bc))
        row = {"N": N}
        for method, tag in ((phpc.SGEMM_FP32, "fp32_mfma"), (phpc.SGEMM_BF16X3, "bf16x3")):
            C.zero_()
            phpc.sgemm_device(A, B, C, method=method)
            torch.cuda.synchronize()
            err = float(np.max(np.abs(C[r, c].double().cpu().numpy() - ref) / scale))
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 3
            e0.record()
            for _ in range(reps):
                phpc.sgemm_device(A, B, C, method=method)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            tf = 2.0 * N ** 3 / (ms * 1e-3) / 1e12
            row[tag] = {"tflops": round(tf, 3), "ms": round(ms, 3), "err_cw_max_sampled": err,
                        "mfma_util": round(tf * (6 if method == phpc.SGEMM_BF16X3 else 1) /
                                           (2500.0 if method == phpc.SGEMM_BF16X3 else 157.3), 4)}
        row["bf16x3_vs_fp32"] = round(row["bf16x3"]["tflops"] / row["fp32_mfma"]["tflops"], 3)
        out.append(row)
        del A, B, C
        torch.cuda.empty_cache()
    return out


def poisson_run(srv, submit, n_req, rate, load, acc, flops, iso, rng):
    """Submit n_req requests at Poisson arrival times (rate per second) from this thread;
    a second thread waits for them in order."""
    import threading

    import numpy as np

    arr = np.cumsum(rng.exponential(1.0 / rate, n_req))
    tickets = [None] * n_req
    got = [None] * n_req
    ready = threading.Semaphore(0)

    def waiter():
        for i in range(n_req):
            ready.acquire()
            got[i] = srv.wait(tickets[i])

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for i in range(n_req):
        dt = arr[i] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        tickets[i] = submit(i)
        ready.release()
    th.join()
    span = time.perf_counter() - t0
    lat = np.array([g["latency_s"] for g in got])
    slow = lat / iso
    cls = {}
    for a, tag in ((0, "fp64"), (1, "fp32")):
        sel = acc == a
        cls[tag] = {"p50_ms": round(float(np.percentile(lat[sel], 50)) * 1e3, 3),
                    "p99_ms": round(float(np.percentile(lat[sel], 99)) * 1e3, 3)}
    return {
        "load": load, "rate_req_per_s": round(rate, 2),
        "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
        "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "max_ms": round(float(lat.max()) * 1e3, 3),
        "p99_vs_isolated_p99": round(float(np.percentile(lat, 99) / np.percentile(iso, 99)), 3),
        "slowdown_p50": round(float(np.percentile(slow, 50)), 3),
        "slowdown_p99": round(float(np.percentile(slow, 99)), 3),
        "tflops": round(flops.sum() / span / 1e12, 3),
        "mean_batch": round(float(np.mean([g["batch_size"] for g in got])), 2),
        "by_class": cls,
    }


def serving_bench(phpc, n_req, seed=11, slots=3, batch_max=32, split_min=4096, loads=(0.3, 0.7, 0.9)):
    """GEMM serving stream: square sizes 60% from [256,1024], 30% from [1024,4096], 10% from
    [4096,8192]; accuracy class 50% FP64-level, 50% FP32-level; operands generated on the
    device from per-request seeds; admission under a 200 GB budget. Measures each
    request alone (isolated latency), the saturated rate (all requests at once), then
    Poisson arrivals at `loads` x the saturated rate under each scheduling policy."""
    import numpy as np

    rng = np.random.default_rng(seed)
    u = rng.random(n_req)
    size = np.where(u < 0.6, rng.integers(256, 1025, n_req),
                    np.where(u < 0.9, rng.integers(1024, 4097, n_req), rng.integers(4096, 8193, n_req)))
    acc = np.where(rng.random(n_req) < 0.5, phpc.ACC_FP64, phpc.ACC_FP32)
    flops = 2.0 * size.astype(np.float64) ** 3
    reqs = [(int(acc[i]), int(size[i]), 1000 + 2 * i) for i in range(n_req)]
    srv = phpc.Server(200 * 10 ** 9, slots=slots, max_dim=8192, batch_max=batch_max)
    srv.set_split_min(split_min)

    def submit(i):
        a, n, sd = reqs[i]
        return srv.submit(a, n, n, n, phpc.DIST_UNIFORM, sd)

    try:
        for i in range(min(8, n_req)):  # first-use set-up of every kernel family
            srv.wait(submit(i))
        iso = np.empty(n_req)
        for i in range(n_req):
            iso[i] = srv.wait(submit(i))["latency_s"]
        t0 = time.perf_counter()
        ts = [submit(i) for i in range(n_req)]
        res = [srv.wait(t) for t in ts]
        t_sat = time.perf_counter() - t0
        sat_rate = n_req / t_sat
        _, peak = srv.memory()
        out = {
            "requests": n_req, "slots": slots, "batch_max": batch_max, "budget_bytes": 200 * 10 ** 9,
            "fp64_split_min_dim": split_min,
            "mix": "sizes 60% [256,1024], 30% [1024,4096], 10% [4096,8192]; 50% FP64-level, 50% FP32-level",
            "isolated_ms": {"p50": round(float(np.percentile(iso, 50)) * 1e3, 3),
                            "p99": round(float(np.percentile(iso, 99)) * 1e3, 3),
                            "sum_s": round(float(iso.sum()), 3)},
            "saturated": {"req_per_s": round(sat_rate, 2), "tflops": round(flops.sum() / t_sat / 1e12, 3),
                          "mean_batch": round(float(np.mean([r["batch_size"] for r in res])), 2),
                          "peak_operand_bytes": peak},
            "poisson": {},
        }
        for policy, aging, tag in ((phpc.SCHED_FIFO, 0.0, "fifo"), (phpc.SCHED_SJF, 0.0, "sjf"),
                                   (phpc.SCHED_SJF, 8.0, "sjf_aging8"), (phpc.SCHED_LANES, 0.0, "lanes")):
            srv.set_policy(policy, aging)
            out["poisson"][tag] = []
            for load in loads:
                out["poisson"][tag].append(poisson_run(srv, submit, n_req, load * sat_rate, load, acc, flops, iso,
                                                       np.random.default_rng(seed + int(load * 100))))
                r = out["poisson"][tag][-1]
                log(f"serving {tag} load {load}: p50 {r['p50_ms']} ms, p99 {r['p99_ms']} ms, "
                    f"p99/iso p99 {r['p99_vs_isolated_p99']}")
        return out
    finally:
        srv.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--grid", default="4x4")
    ap.add_argument("--panel", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-vendor", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-grouped", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--only", choices=["summa", "split", "fp32acc", "grouped", "sweep", "serve"],
                    help="run one section alone and print its JSON (for profiler runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device-to-device copy three times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build emulates the process grid on a single "
              f"MI355X and has no multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_probe:
        n = (1 << 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8, "copies": 3}))
        return

    phpc = load_phpc()
    if args.only:
        if args.only == "summa":
            # the headline CHAINED SUMMA at N on the --grid / --panel emulated ranks, one
            # warm-up and `steps` timed calls
            p, q = (int(x) for x in args.grid.split("x"))
            A = torch.empty((args.n, args.n), dtype=torch.float64, device="cuda:0")
            B = torch.empty_like(A)
            C = torch.zeros_like(A)
            phpc.fill_device(A, phpc.DIST_UNIFORM, 1)
            phpc.fill_device(B, phpc.DIST_UNIFORM, 2)
            sp = phpc.stream_ptr(torch.cuda.current_stream())
            for i in range(1 + args.steps):
                if i == 1:
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                phpc.check(phpc.lib().phpc_summa_device(p, q, args.panel, phpc.SUMMA_CHAINED, phpc.dptr(A),
                                                         phpc.dptr(B), phpc.dptr(C), args.n, 0, sp))
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / max(1, args.steps)
            res = {"mode": "chained", "n": args.n, "grid": args.grid, "panel": args.panel,
                   "tflops": round(2.0 * args.n ** 3 / dt / 1e12, 2)}
        elif args.only == "split":
            # the split CRT GEMM at N (uniform inputs), one warm-up and steps timed calls
            A = torch.empty((args.n, args.n), dtype=torch.float64, device="cuda:0")
            B = torch.empty_like(A)
            C = torch.zeros_like(A)
            phpc.fill_device(A, phpc.DIST_UNIFORM, 1)
            phpc.fill_device(B, phpc.DIST_UNIFORM, 2)
            info = None
            t0 = time.perf_counter()
            for _ in range(1 + args.steps):
                info = phpc.dgemm_split_device(A, B, C, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
            torch.cuda.synchronize()
            res = {"moduli": info.slices, "int_bits": info.int_bits, "calls": 1 + args.steps,
                   "wall_s": round(time.perf_counter() - t0, 3)}
        elif args.only == "fp32acc":
            res = fp32_accurate_bench(phpc)
        elif args.only == "grouped":
            res = grouped_bench(phpc)
        elif args.only == "sweep":
            res = split_sweep(phpc)
        else:
            res = serving_bench(phpc, args.serve_requests)
        print(json.dumps({"only": args.only, "result": res}))
        return
    N, panel = args.n, args.panel
    p, q = (int(x) for x in args.grid.split("x"))
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sp = phpc.stream_ptr(stream)
    lib = phpc.lib()

    A = torch.empty((N, N), dtype=torch.float64, device=dev)
    B = torch.empty_like(A)
    C = torch.zeros_like(A)
    phpc.fill_device(A, phpc.DIST_UNIFORM, 1)
    phpc.fill_device(B, phpc.DIST_UNIFORM, 2)

    def run_mode(mode):
        rc = lib.phpc_summa_device(p, q, panel, mode, phpc.dptr(A), phpc.dptr(B), phpc.dptr(C), N, 0, sp)
        if rc:
            phpc.check(rc)

    def timed(mode, steps, warmup):
        for _ in range(warmup):
            run_mode(mode)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t_start = torch.cuda.Event(enable_timing=True)
        t_end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        t_start.record(stream)
        for i in range(steps):
            ev[i][0].record(stream)
            run_mode(mode)
            ev[i][1].record(stream)
        t_end.record(stream)
        torch.cuda.synchronize()
        wall_s = time.perf_counter() - w0
        return t_start.elapsed_time(t_end) / steps, statistics.mean(a.elapsed_time(b) for a, b in ev), wall_s

    flops = 2.0 * N ** 3
    # primary: CHAINED schedule (one FMA chain per output over the 64 panels)
    ms_per_step, kernel_ms, wall = timed(phpc.SUMMA_CHAINED, args.steps, args.warmup)
    value = flops / (ms_per_step * 1e-3) / 1e12
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    # the reference's per-panel grouping (C rounded after every panel), same launch shape
    exact_ms, _, _ = timed(phpc.SUMMA_FUSED, max(2, args.steps // 2), 1)
    step = lambda: run_mode(phpc.SUMMA_CHAINED)  # noqa: E731

    # accuracy of the timed path: one clean pass, sampled double-double check on the device
    C.zero_()
    step()
    torch.cuda.synchronize()
    rng = np.random.default_rng(3)
    rows = rng.integers(0, N, 4096)
    cols = rng.integers(0, N, 4096)
    err_cw, err_nw = phpc.check_sampled(A, B, C, rows, cols)

    extra = {}
    # split-precision FP64 (CRT / Ozaki scheme II on int8 MFMA), same N, both input sets
    split = {"method": "crt (Ozaki scheme II): int8 residue GEMMs, Garner reconstruction",
             "target_err_cw": 1e-14}
    for dist, tag in ((phpc.DIST_UNIFORM, "uniform"), (phpc.DIST_WIDE, "wide_exponent")):
        phpc.fill_device(A, dist, 1)
        phpc.fill_device(B, dist, 2)
        info = [None]

        def run_split():
            info[0] = phpc.dgemm_split_device(A, B, C, slices=0, target=1e-14, method=phpc.SPLIT_CRT)

        run_split()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(2, args.steps // 2)
        e0.record(stream)
        for _ in range(reps):
            run_split()
        e1.record(stream)
        torch.cuda.synchronize()
        t_ms = e0.elapsed_time(e1) / reps
        C.zero_()
        run_split()
        torch.cuda.synchronize()
        cw, nw = phpc.check_sampled(A, B, C, rows, cols)
        i = info[0]
        tf = flops / (t_ms * 1e-3) / 1e12
        split[tag] = {"tflops_fp64_equiv": round(tf, 3), "ms": round(t_ms, 3), "vs_plain": round(tf / value, 3),
                      "moduli": i.slices, "int_bits": i.int_bits, "error_estimate": i.error_estimate,
                      "err_cw_max_sampled": cw, "err_nw_sampled": nw,
                      "int8_mfma_util": round(i.products * flops / (t_ms * 1e-3) / 5.0e15, 4)}
    extra["split"] = split
    phpc.fill_device(A, phpc.DIST_UNIFORM, 1)
    phpc.fill_device(B, phpc.DIST_UNIFORM, 2)
    if not args.no_vendor:
        # rocBLAS comparator on the same SUMMA schedule and as one plain DGEMM
        C.zero_()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        rc = lib.phpc_summa_device(p, q, panel, phpc.SUMMA_VENDOR, phpc.dptr(A), phpc.dptr(B), phpc.dptr(C), N, 0, sp)
        e1.record(stream)
        torch.cuda.synchronize()
        if rc == 0:
            extra["vendor_summa_tflops"] = round(flops / (e0.elapsed_time(e1) * 1e-3) / 1e12, 3)
        C.zero_()
        phpc.dgemm_vendor_device(A, B, C)
        torch.cuda.synchronize()
        e0.record(stream)
        phpc.dgemm_vendor_device(A, B, C)
        e1.record(stream)
        torch.cuda.synchronize()
        extra["vendor_dgemm_tflops"] = round(flops / (e0.elapsed_time(e1) * 1e-3) / 1e12, 3)

    del A, B, C
    torch.cuda.empty_cache()
    if not args.no_grouped:
        extra["grouped"] = grouped_bench(phpc)
    if not args.no_sweep:
        extra["split_sweep"] = split_sweep(phpc)
        extra["fp32_accurate"] = fp32_accurate_bench(phpc)
    if not args.no_serve:
        extra["serving"] = serving_bench(phpc, args.serve_requests)

    workload = f"summa_fp64_n{N}_grid{p}x{q}_panel{panel}"
    out = {
        "metric": "FP64-accurate TFLOP/s (split vs plain) at N=16384; p50/p99 req latency",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: A, B uniform[-1,1) from SplitMix64 seeds 1, 2, resident in HBM",
        "config": {
            "workload": workload,
            "N": N,
            "grid": f"{p}x{q}",
            "panel": panel,
            "path": "plain FP64 MFMA, SUMMA CHAINED schedule: A row blocks transposed once per pass, "
                    "then one persistent launch, 128x128 tiles fed HBM->LDS directly, one FMA chain per "
                    "output over the 64 panels",
            "flops_per_step": flops,
        },
        "accuracy": {"err_cw_max_sampled": err_cw, "err_nw_sampled": err_nw, "samples": 4096,
                     "reference": "double-double on the device"},
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": traffic_from_profiles(workload),
            "kernel": "f64_at_128x128 (transposed-A, direct-to-LDS, 2 workgroups per CU)",
            "kernel_ms_avg": round(kernel_ms, 3),
            "timed_unit": "one SUMMA pass: 4 A row-block transposes (~1% of the pass) + the GEMM launch",
        },
        "wall_s_timed": round(wall, 3),
        "summa_panel_exact_tflops": round(flops / (exact_ms * 1e-3) / 1e12, 3),
    }
    out.update(extra)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
