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


def _cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n=1024, seed=1, runs=3):
    """The reference's CPU path (src/iterative.c:8-13) on the first workload's inputs: N = 1024
    FP64, A and B uniform[-1,1) from seeds 1 and 2, timed on one pinned host core. The -O0
    build is the reference's own flags (Makefile:9) and gives `value`; -O2 is an extra line.
    C is checked bit for bit against the oracle's restatement of the same loop and against
    the double-double product."""
    import tempfile

    import numpy as np

    sys.path.insert(0, str(ROOT))
    import oracle as O

    ref = {lvl: ROOT / "oracle" / "_ref" / f"iterative_ref_{lvl}.out" for lvl in ("O0", "O2")}
    port = {lvl: ROOT / "oracle" / "_build" / f"iterative_{lvl}.out" for lvl in ("O0", "O2")}
    kind = "reference" if all(p.exists() for p in ref.values()) else "port"
    exes = ref if kind == "reference" else port
    if not all(p.exists() for p in exes.values()):
        return None
    pin = ["taskset", "-c", "0"] if shutil.which("taskset") else []
    lines, c_out = {}, None
    with tempfile.TemporaryDirectory() as td:
        for lvl, exe in exes.items():
            secs = []
            for r in range(runs):
                out_file = os.path.join(td, f"c_{lvl}.bin")
                cmd = [str(exe), str(n), str(seed), out_file] if kind == "reference" else [str(exe), str(n), "u"]
                res = subprocess.run(pin + cmd, capture_output=True, text=True, timeout=600)
                if res.returncode != 0:
                    return None
                secs.append(float(res.stdout.strip().split(",")[1]))
                if kind == "reference" and r == 0:
                    c = np.fromfile(out_file, dtype=np.float64).reshape(n, n)
                    if c_out is None:
                        c_out = c
                    elif not np.array_equal(c.view(np.uint64), c_out.view(np.uint64)):
                        c_out = None
                        lines["mismatch"] = "O0 and O2 builds differ"
            t = statistics.median(secs)
            lines[lvl] = {"seconds_median": round(t, 4), "seconds_runs": [round(x, 4) for x in secs],
                          "tflops": 2.0 * n ** 3 / t / 1e12, "gflops": round(2.0 * n ** 3 / t / 1e9, 4),
                          "build": f"gcc -{lvl} -Wall (reference Makefile:9 flags{'' if lvl == 'O0' else ' + -O2'})"}
    checked = None
    if c_out is not None:
        A, B = O.uniform((n, n), seed), O.uniform((n, n), seed + 1)
        want = O.iterative(A, B, np.zeros((n, n)))
        bit_exact = bool(np.array_equal(c_out.view(np.uint64), want.view(np.uint64)))
        cw, _ = O.err_full(c_out, O.gemm_dd(A, B), A, B)
        checked = {"bit_exact_vs_oracle_iterative": bit_exact, "err_cw_vs_double_double": cw,
                   "bound_n_u": n * 2.0 ** -53}
    t0 = lines["O0"]["seconds_median"]
    return {
        "value": lines["O0"]["tflops"],
        "unit": "TFLOP/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"src/iterative.c loop at N={n}, A/B uniform[-1,1) seeds {seed}/{seed + 1}, "
                   f"gcc -O0 as the reference Makefile, median of {runs} runs = {t0:.3f} s on 1 pinned core "
                   f"({os.cpu_count()} visible, {_cpu_model()}); N=16384 would take "
                   f"~{t0 * 16 ** 3 / 3600:.1f} h (O(N^3) extrapolation)"),
        "lines": lines,
        "checked": checked,
        "host": {"nproc": os.cpu_count(), "cpu_model": _cpu_model(), "pinned_core": 0 if pin else None},
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
    [64, 4096], both as an FP32 plan (FP32 matrix cores) and as an FP32-accurate 3 x BF16
    plan, each set replayed from a captured hipGraph, with the tile configuration chosen by
    the autotuner."""
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
    for key, dtype in (("fp32_512_gemms", phpc.F32), ("fp32_512_gemms_bf16x3", phpc.F32X3)):
        plan = phpc.Plan(dtype, gemms, autotune=True, stream=s)
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
        if dtype == phpc.F32X3:
            launch = ("one batched kernel cutting every A and B into BF16 slice planes + one persistent "
                      "BF16 level launch (three k segments per item) for all 512 GEMMs"
                      if info["config"] >= 4 else "one persistent 3 x BF16 kernel for all 512 GEMMs")
        else:
            launch = ("one batched kernel transposing every A + one persistent transposed-A kernel for all "
                      "512 GEMMs (B read in place)" if info["config"] == 3 else
                      "one persistent kernel for all 512 GEMMs")
        out[key] = {"tflops": round(flops / (ms * 1e-3) / 1e12, 3), "ms": round(ms, 3), "flops": flops,
                    "config": info["config"], "tiles": info["tiles"], "tuned_ms": round(info["tuned_ms"], 3),
                    "err_cw_max_sampled": worst, "fp32_mfma_util": round(flops / (ms * 1e-3) / 157.3e12, 4),
                    "launch": launch + ", items in decreasing k, hipGraph replay"}
        plan.close()
    del gemms
    torch.cuda.empty_cache()
    return out


def clock_during(run, seconds=1.5):
    """Median graphics clock (MHz) that `rocm-smi -g` reports while `run` is called back to
    back, untimed, for about `seconds`; None when rocm-smi is absent or reports no clock."""
    import re
    import threading
    import torch

    exe = shutil.which("rocm-smi")
    if exe is None:
        return None
    mhz, stop = [], [False]

    def sample():
        while not stop[0]:
            try:
                txt = subprocess.run([exe, "-g", "--json"], capture_output=True, text=True, timeout=10).stdout
                m = re.search(r"\((\d+)\s*Mhz\)", txt)
                if m:
                    mhz.append(int(m.group(1)))
            except (OSError, subprocess.SubprocessError):
                pass
            time.sleep(0.2)

    run()
    torch.cuda.synchronize()
    th = threading.Thread(target=sample, daemon=True)
    th.start()
    t0 = time.time()
    while time.time() - t0 < seconds or (len(mhz) < 3 and time.time() - t0 < 4 * seconds):
        run()
        torch.cuda.synchronize()
    stop[0] = True
    th.join()
    if not mhz:
        return None
    return {"mhz_median": statistics.median(mhz), "mhz_min": min(mhz), "mhz_max": max(mhz), "samples": len(mhz),
            "source": "rocm-smi -g every 0.2 s while the same work ran back to back, untimed"}


def split_sweep(phpc, sizes=(4096, 8192, 16384), seed=31):
    """Split-precision FP64 GEMM at each slice / modulus count (CRT moduli, int8 slices,
    FP8-E4M3 slices): error against the double-double product on sampled entries and
    FP64-equivalent TFLOP/s."""
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
                   [(phpc.SPLIT_SLICES, s_) for s_ in (3, 4, 5, 6, 7, 8)] + \
                   [(phpc.SPLIT_FP8, s_) for s_ in (0, 3, 4, 5, 6, 7, 8)]
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
                out.append({"N": N, "inputs": tag,
                            "method": {phpc.SPLIT_CRT: "crt", phpc.SPLIT_SLICES: "slices"}.get(method, "fp8"),
                            "requested": count if count else "auto", "count": info.slices,
                            "gemms": info.products, "int_bits": info.int_bits, "err_cw_max_sampled": cw,
                            "error_estimate": info.error_estimate,
                            "tflops_fp64_equiv": round(2.0 * N ** 3 / (ms * 1e-3) / 1e12, 3), "ms": round(ms, 3)})
        del A, B, C
        torch.cuda.empty_cache()
    phpc.lib().phpc_split_set_method(phpc.SPLIT_CRT)
    return out


def fp32_accurate_bench(phpc, sizes=(4096, 8192, 16384), seed=41, clock=True):
    """FP32 inputs, uniform[-1,1): the FP32 matrix-core GEMM and the FP32-accurate 3 x BF16
    GEMM (six BF16 slice products, FP32 accumulation). Error against float64 dot products
    of the same FP32 inputs on sampled entries, normalised by |A||B|. At the largest N, the
    graphics clock each path holds (clock_during), unless clock is False."""
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
            if clock and N == max(sizes):
                row[tag]["clock_mhz"] = clock_during(lambda: phpc.sgemm_device(A, B, C, method=method))
        row["bf16x3_vs_fp32"] = round(row["bf16x3"]["tflops"] / row["fp32_mfma"]["tflops"], 3)
        out.append(row)
        del A, B, C
        torch.cuda.empty_cache()
    return out


def poisson_run(srv, submit, n_req, rate, load, acc, flops, iso, rng, size=None):
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
    # the requests that make the tail: what they are and where their time went
    tail = []
    for i in np.argsort(lat)[::-1][:12]:
        g = got[i]
        tail.append({"n": int(size[i]) if size is not None else None, "fp64": int(acc[i] == 0),
                     "lat_ms": round(g["latency_s"] * 1e3, 2), "queue_ms": round(g["queue_s"] * 1e3, 2),
                     "compute_ms": round(g["compute_s"] * 1e3, 2), "gemm_ms": round(g["gemm_s"] * 1e3, 2),
                     "iso_ms": round(float(iso[i]) * 1e3, 2), "batch": g["batch_size"]})
    return {
        "tail": tail,
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


SERVE_POLICIES = (("fifo", 0, 0.0), ("sjf", 1, 0.0), ("sjf_aging8", 1, 8.0), ("lanes", 2, 0.0))


def serving_bench(phpc, n_req, seed=11, slots=3, batch_max=32, split_min=4096, loads=(0.3, 0.7, 0.9),
                  policies=SERVE_POLICIES, default_policy="sjf", big_limit=1):
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
    srv.set_big_limit(big_limit)
    srv.set_policy(dict((t, p) for t, p, _ in policies)[default_policy],
                   dict((t, a) for t, _, a in policies)[default_policy])

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
            "default_policy": default_policy, "big_limit": big_limit,
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
        for tag, policy, aging in policies:
            srv.set_policy(policy, aging)
            out["poisson"][tag] = []
            for load in loads:
                out["poisson"][tag].append(poisson_run(srv, submit, n_req, load * sat_rate, load, acc, flops, iso,
                                                       np.random.default_rng(seed + int(load * 100)), size))
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
    ap.add_argument("--no-clock", action="store_true",
                    help="skip sampling the graphics clock (rocm-smi) beside the plain and split GEMMs")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-grouped", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--only", choices=["summa", "split", "fp32acc", "grouped", "sweep", "serve"],
                    help="run one section alone and print its JSON (for profiler runs)")
    ap.add_argument("--mode", choices=["chained", "fused"], default="chained",
                    help="SUMMA schedule of --only summa")
    ap.add_argument("--split-method", choices=["crt", "slices", "fp8"], default="crt",
                    help="method of --only split (CRT moduli, int8 slices or FP8-E4M3 slices)")
    ap.add_argument("--split-count", type=int, default=0,
                    help="moduli or slices of --only split (0: chosen for err_cw 1e-14)")
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
                mode = phpc.SUMMA_CHAINED if args.mode == "chained" else phpc.SUMMA_FUSED
                phpc.check(phpc.lib().phpc_summa_device(p, q, args.panel, mode, phpc.dptr(A),
                                                         phpc.dptr(B), phpc.dptr(C), args.n, 0, sp))
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / max(1, args.steps)
            res = {"mode": args.mode, "n": args.n, "grid": args.grid, "panel": args.panel,
                   "tflops": round(2.0 * args.n ** 3 / dt / 1e12, 2)}
        elif args.only == "split":
            # the split GEMM at N (uniform inputs), one warm-up and steps timed calls
            method = {"crt": phpc.SPLIT_CRT, "slices": phpc.SPLIT_SLICES, "fp8": phpc.SPLIT_FP8}[args.split_method]
            A = torch.empty((args.n, args.n), dtype=torch.float64, device="cuda:0")
            B = torch.empty_like(A)
            C = torch.zeros_like(A)
            phpc.fill_device(A, phpc.DIST_UNIFORM, 1)
            phpc.fill_device(B, phpc.DIST_UNIFORM, 2)
            info = None
            t0 = time.perf_counter()
            for _ in range(1 + args.steps):
                info = phpc.dgemm_split_device(A, B, C, slices=args.split_count, target=1e-14, method=method)
            torch.cuda.synchronize()
            phpc.lib().phpc_split_set_method(phpc.SPLIT_CRT)
            res = {"method": args.split_method, "count": info.slices, "gemms": info.products,
                   "int_bits": info.int_bits, "calls": 1 + args.steps, "wall_s": round(time.perf_counter() - t0, 3)}
        elif args.only == "fp32acc":
            res = fp32_accurate_bench(phpc, clock=not args.no_clock)
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
        """Whole-step time (events around the K steps on the launch stream, and the wall
        clock), and the average duration of the GEMM kernel launch alone (the library's
        events around that launch on the same stream)."""
        for _ in range(warmup):
            run_mode(mode)
        torch.cuda.synchronize()
        t_start = torch.cuda.Event(enable_timing=True)
        t_end = torch.cuda.Event(enable_timing=True)
        phpc.kernel_timing(True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        t_start.record(stream)
        for i in range(steps):
            run_mode(mode)
        t_end.record(stream)
        torch.cuda.synchronize()
        wall_s = time.perf_counter() - w0
        k_ms, k_n = phpc.kernel_timing_read()
        phpc.kernel_timing(False)
        return t_start.elapsed_time(t_end) / steps, k_ms / max(1, k_n), k_n, wall_s

    flops = 2.0 * N ** 3
    # primary: CHAINED schedule (one FMA chain per output over the 64 panels)
    ms_per_step, kernel_ms, k_launches, wall = timed(phpc.SUMMA_CHAINED, args.steps, args.warmup)
    value = flops / (ms_per_step * 1e-3) / 1e12
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    # the reference's per-panel grouping (C rounded after every panel), same launch shape
    exact_ms, exact_kernel_ms, _, _ = timed(phpc.SUMMA_FUSED, max(2, args.steps // 2), 1)
    exact_tf = flops / (exact_ms * 1e-3) / 1e12
    exact_kernel_tf = flops / (exact_kernel_ms * 1e-3) / 1e12

    # accuracy of the timed path: one clean pass, sampled double-double check on the device
    C.zero_()
    run_mode(phpc.SUMMA_CHAINED)
    torch.cuda.synchronize()
    rng = np.random.default_rng(3)
    rows = rng.integers(0, N, 4096)
    cols = rng.integers(0, N, 4096)
    err_cw, err_nw = phpc.check_sampled(A, B, C, rows, cols)
    plain_clock = clock_during(lambda: run_mode(phpc.SUMMA_CHAINED)) if not args.no_clock else None

    detail = {}
    # split-precision FP64 (CRT / Ozaki scheme II on int8 MFMA), same N, both input sets
    split = {}
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
        phpc.kernel_timing(True)
        e0.record(stream)
        for _ in range(reps):
            run_split()
        e1.record(stream)
        torch.cuda.synchronize()
        lvl_ms, lvl_n = phpc.kernel_timing_read()
        phpc.kernel_timing(False)
        t_ms = e0.elapsed_time(e1) / reps
        C.zero_()
        run_split()
        torch.cuda.synchronize()
        cw, nw = phpc.check_sampled(A, B, C, rows, cols)
        i = info[0]
        tf = flops / (t_ms * 1e-3) / 1e12
        kp = (N + 127) // 128 * 128
        lvl_avg = lvl_ms / max(1, lvl_n)
        split[tag] = {"tflops_fp64_equiv": round(tf, 3), "ms": round(t_ms, 3), "vs_plain": round(tf / value, 3),
                      "moduli": i.slices, "int_bits": i.int_bits, "error_estimate": i.error_estimate,
                      "err_cw_max_sampled": cw, "err_nw_sampled": nw,
                      "int8_kernel_ms_avg": round(lvl_avg, 4), "int8_launches": lvl_n,
                      "int8_kernel_tops": round(2.0 * N * N * kp / (lvl_avg * 1e-3) / 1e12, 2),
                      "int8_share_of_time": round(lvl_ms / reps / t_ms, 4)}
        if tag == "uniform" and not args.no_clock:
            split[tag]["clock"] = clock_during(run_split)
    phpc.fill_device(A, phpc.DIST_UNIFORM, 1)
    phpc.fill_device(B, phpc.DIST_UNIFORM, 2)
    vendor = {}
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
            vendor["summa_tflops"] = round(flops / (e0.elapsed_time(e1) * 1e-3) / 1e12, 3)
        C.zero_()
        phpc.dgemm_vendor_device(A, B, C)
        torch.cuda.synchronize()
        e0.record(stream)
        phpc.dgemm_vendor_device(A, B, C)
        e1.record(stream)
        torch.cuda.synchronize()
        vendor["dgemm_tflops"] = round(flops / (e0.elapsed_time(e1) * 1e-3) / 1e12, 3)

    del A, B, C
    torch.cuda.empty_cache()
    if not args.no_grouped:
        detail["grouped"] = grouped_bench(phpc)
    if not args.no_sweep:
        detail["split_sweep"] = split_sweep(phpc)
        detail["fp32_accurate"] = fp32_accurate_bench(phpc, clock=not args.no_clock)
    if not args.no_serve:
        detail["serving"] = serving_bench(phpc, args.serve_requests)
    # the full detail on a line of its own, before the result line
    print("bench-detail " + json.dumps(detail), flush=True)

    workload = f"summa_fp64_n{N}_grid{p}x{q}_panel{panel}"
    uni, wide = split["uniform"], split["wide_exponent"]
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
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": traffic_from_profiles(workload),
            "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel, "
                              "scaled by a 1 GiB copy), not measured in this run",
            "kernel": "f64_at_128x128 (transposed-A, direct-to-LDS, 2 workgroups per CU)",
            "kernel_ms_avg": round(kernel_ms, 3),
            "kernel_launches_timed": k_launches,
            "pass_ms_avg": round(ms_per_step, 3),
            "timed_unit": "achieved = 2N^3 / kernel_ms_avg (HIP events around the GEMM launch alone, on its "
                          "stream); a step (pass) adds the 4 A row-block transposes",
            "clock_mhz": (plain_clock or {}).get("mhz_median"),
        },
        "accuracy": {"err_cw_max_sampled": err_cw, "err_nw_sampled": err_nw, "samples": 4096,
                     "reference": "double-double on the device"},
        # the metric's parts: split vs plain at N, and the serving tail
        "split_tflops": {"uniform": uni["tflops_fp64_equiv"], "wide_exponent": wide["tflops_fp64_equiv"]},
        "split_vs_plain": {"uniform": uni["vs_plain"], "wide_exponent": wide["vs_plain"]},
        "split_err_cw": {"uniform": uni["err_cw_max_sampled"], "wide_exponent": wide["err_cw_max_sampled"],
                         "target": 1e-14},
        "split_moduli": {"uniform": uni["moduli"], "wide_exponent": wide["moduli"]},
        "split_roofline": {
            "bound": "mfma",
            "kernel": "i8_level_256x256_w16_k128_mi16 (one int8 residue GEMM per modulus, 16x16x64 MFMAs)",
            "achieved": uni["int8_kernel_tops"],
            "peak": 5000.0,
            "unit": "TOP/s",
            "frac": round(uni["int8_kernel_tops"] / 5000.0, 4),
            "kernel_ms_avg": uni["int8_kernel_ms_avg"],
            "ops_per_launch": 2.0 * N * N * ((N + 127) // 128 * 128),
            "share_of_split_time": uni["int8_share_of_time"],
            # the nominal peak assumes 2.4 GHz; under this load the chip holds a lower clock
            "clock_mhz": (uni.get("clock") or {}).get("mhz_median"),
            "peak_at_clock": (round(5000.0 * uni["clock"]["mhz_median"] / 2400.0, 1) if uni.get("clock") else None),
            "frac_at_clock": (round(uni["int8_kernel_tops"] / (5000.0 * uni["clock"]["mhz_median"] / 2400.0), 4)
                              if uni.get("clock") else None),
        },
        "split": split,
        "summa_panel_exact": {
            "tflops": round(exact_tf, 3),
            "roofline": {"bound": "mfma", "achieved": round(exact_kernel_tf, 3), "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(exact_kernel_tf / FP64_PEAK_TFLOPS, 4),
                         "kernel": "f64_at_128x128_w16_mp_k32 (16 waves, 32-deep slabs; per-panel sums added "
                                   "to the C carried in registers)",
                         "kernel_ms_avg": round(exact_kernel_ms, 3), "pass_ms_avg": round(exact_ms, 3)},
            "arithmetic": "bit-exact to the reference's per-panel grouping (src/phpc_summa.c:63-95)",
        },
        "vendor": vendor,
        "wall_s_timed": round(wall, 3),
    }
    if "serving" in detail:
        srv = detail["serving"]
        default = srv.get("default_policy", "fifo")
        at = {tag: {str(r["load"]): r["p99_vs_isolated_p99"] for r in runs} for tag, runs in srv["poisson"].items()}
        r70 = next(r for r in srv["poisson"][default] if abs(r["load"] - 0.7) < 1e-9)
        out["serving_p99_ratio_at_0.7"] = r70["p99_vs_isolated_p99"]
        out["serving"] = {
            "default_policy": default,
            "p99_vs_isolated_p99": at,
            "at_0.7": {k: r70[k] for k in ("p50_ms", "p99_ms", "rate_req_per_s", "tflops", "mean_batch",
                                            "slowdown_p99")},
            "isolated_ms": srv["isolated_ms"],
            "saturated": {k: srv["saturated"][k] for k in ("req_per_s", "tflops", "mean_batch")},
            "requests": srv["requests"], "budget_bytes": srv["budget_bytes"],
        }
    if "grouped" in detail:
        g = detail["grouped"]
        out["grouped"] = {
            "fp64_n10000_grid3x3": {"tflops": g["fp64_n10000_grid3x3"]["tflops"],
                                    "frac_fp64_peak": round(g["fp64_n10000_grid3x3"]["tflops"] / FP64_PEAK_TFLOPS, 4),
                                    "err_cw": g["fp64_n10000_grid3x3"]["err_cw_max_sampled"]},
            "fp32_512_gemms": {"tflops": g["fp32_512_gemms"]["tflops"],
                               "frac_fp32_peak": g["fp32_512_gemms"]["fp32_mfma_util"],
                               "err_cw": g["fp32_512_gemms"]["err_cw_max_sampled"]},
            "fp32_512_gemms_bf16x3": {"tflops": g["fp32_512_gemms_bf16x3"]["tflops"],
                                      "config": g["fp32_512_gemms_bf16x3"]["config"],
                                      "err_cw": g["fp32_512_gemms_bf16x3"]["err_cw_max_sampled"]},
        }
    if "fp32_accurate" in detail:
        out["fp32_accurate"] = {str(r["N"]): {"fp32_mfma_tflops": r["fp32_mfma"]["tflops"],
                                              "bf16x3_tflops": r["bf16x3"]["tflops"],
                                              "bf16x3_err_cw": r["bf16x3"]["err_cw_max_sampled"],
                                              **({"clock_mhz": [(r[t]["clock_mhz"] or {}).get("mhz_median")
                                                                for t in ("fp32_mfma", "bf16x3")]}
                                                 if "clock_mhz" in r["bf16x3"] else {})}
                                for r in detail["fp32_accurate"]}
    if "split_sweep" in detail:
        # error against split count at each N (CRT moduli / slices): err_cw and TFLOP/s
        sw = {}
        for r in detail["split_sweep"]:
            if r["method"] == "fp8" and r["N"] != max(x["N"] for x in detail["split_sweep"]):
                continue  # the line stays under 8 KB: FP8 rows at the largest N only (all on bench-detail)
            key = f'{r["N"]}_{r["inputs"][:4]}_{r["method"]}'
            sw.setdefault(key, []).append([r["count"], float(f'{r["err_cw_max_sampled"]:.2e}'),
                                           round(r["tflops_fp64_equiv"], 1)])
        out["split_sweep"] = {"columns": ["count", "err_cw", "tflops"], "rows": sw}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
