"""Parity at the benchmark workloads' own sizes, through the C ABI:

- SUMMA at N = 16384 on the emulated 4 x 4 grid, panel 256 (the timed workload): the
  CHAINED schedule and the panel-exact FUSED schedule, sampled entries bit for bit
  against the oracle's per-entry restatement of the reference arithmetic, and CHAINED
  over the whole matrix against the register-staged kernel;
- the grouped N = 10000 3 x 3 plan (rank blocks 3334/3333/3333), sampled bit for bit;
- the split-precision GEMM at N = 8192 and 16384 on both input sets, sampled against the
  double-double product, and its FP32-accurate 3 x BF16 variant at the same sizes within
  its bound;
- 512 FP32 GEMMs with M, N, K log-uniform in [64, 4096] in one plan, every GEMM within
  the FP32 bound, hipGraph replays equal to the eager launch; the same GEMMs as a 3 x BF16
  plan on slice planes within the 3 x BF16 bound;
- a Poisson serving stream with the 60/30/10 size law up to 8192, FP64/FP32 classes 50/50
  and a 200 GB budget: every result equals the same request through the one-shot entry
  points bit for bit and meets its class's bound.
"""
import numpy as np
import pytest

from conftest import phpc
import oracle as O

pytestmark = pytest.mark.gpu


def i64(t):
    import torch
    return t.view(torch.int64)


def sample_entries(N, count, seed):
    """Random entries plus a sweep over both diagonals and the last row and column (the
    rank-block and tile edges are on these)."""
    rng = np.random.default_rng(seed)
    d = np.arange(0, N, max(1, N // 256))
    rows = np.concatenate([rng.integers(0, N, count), d, N - 1 - d, np.full(8, N - 1), rng.integers(0, N, 8)])
    cols = np.concatenate([rng.integers(0, N, count), d, d, rng.integers(0, N, 8), np.full(8, N - 1)])
    return rows.astype(np.int64), cols.astype(np.int64)


@pytest.mark.parametrize("mode", ["chained", "fused"])
@pytest.mark.parametrize("dist", [phpc.DIST_UNIFORM, phpc.DIST_WIDE])
def test_summa_n16384_sampled_bit_exact(mode, dist):
    """The timed workload (N = 16384, 4 x 4, panel 256): sampled entries equal the oracle's
    restatement bit for bit, CHAINED as one FMA chain over all k (one local-GEMM call,
    src/phpc_gemm.cu:32,50-55), FUSED with each of the 64 panels summed from zero and added
    to C in panel order (src/phpc_summa.c:63-95); both within the FP64 bound."""
    import torch
    N, panel = 16384, 256
    print(f"summa n16384 {mode} dist={dist}", flush=True)
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, dist, 1)
    phpc.fill_device(dB, dist, 2)
    dC = torch.zeros_like(dA)
    m = phpc.SUMMA_CHAINED if mode == "chained" else phpc.SUMMA_FUSED
    phpc.summa_device(dA, dB, dC, 4, 4, panel=panel, mode=m)
    torch.cuda.synchronize()
    rows, cols = sample_entries(N, 2048, 5)
    got = dC[torch.from_numpy(rows).cuda(), torch.from_numpy(cols).cuda()].cpu().numpy()
    A, B = dA.cpu().numpy(), dB.cpu().numpy()
    want = O.summa_entries(A, B, rows, cols, panel=0 if mode == "chained" else panel)
    bad = np.flatnonzero(got.view(np.uint64) != want.view(np.uint64))
    assert bad.size == 0, (bad.size, rows[bad[:4]], cols[bad[:4]])
    cw, _ = phpc.check_sampled(dA, dB, dC, rows[:1024], cols[:1024])
    print(f"  {got.size} entries bit-exact, err_cw={cw:.2e}", flush=True)
    assert cw <= 1e-14


def test_chained_n16384_matches_register_staged_full_matrix():
    """The timed CHAINED launch (transposed-A kernel) and the register-staged 256x128 kernel
    (configuration 6) agree bit for bit over the whole N = 16384 matrix."""
    import torch
    N = 16384
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, phpc.DIST_UNIFORM, 1)
    phpc.fill_device(dB, phpc.DIST_UNIFORM, 2)
    dC1 = torch.zeros_like(dA)
    phpc.summa_device(dA, dB, dC1, 4, 4, panel=256, mode=phpc.SUMMA_CHAINED)
    dC2 = torch.zeros_like(dA)
    phpc.set_launch_policy(0, 6, 0)
    try:
        phpc.summa_device(dA, dB, dC2, 4, 4, panel=256, mode=phpc.SUMMA_CHAINED)
        torch.cuda.synchronize()
    finally:
        phpc.set_launch_policy(0, -1, 0)
    assert torch.equal(i64(dC1), i64(dC2))


def test_grouped_n10000_grid3x3_sampled_bit_exact():
    """N = 10000 on an emulated 3 x 3 grid (blocks 3334/3333/3333, which the reference cannot
    run) as one panel-grouped plan replayed from a hipGraph: sampled entries, including every
    block edge, equal the oracle's panel-grouped restatement (panels 3333, 3333, 3333, 1)."""
    import torch
    N, p = 10000, 3
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, phpc.DIST_UNIFORM, 21)
    phpc.fill_device(dB, phpc.DIST_UNIFORM, 22)
    dC = torch.zeros_like(dA)
    st = [0, 3334, 6667, 10000]
    items = [(dA[st[i]:st[i + 1]], dB[:, st[j]:st[j + 1]], dC[st[i]:st[i + 1], st[j]:st[j + 1]], N // 3)
             for i in range(p) for j in range(p)]
    s = torch.cuda.Stream()
    plan = phpc.Plan(phpc.F64, items, autotune=True, stream=s)
    dC.zero_()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        plan.graph_launch(s)
    s.synchronize()
    edges = np.array([0, 3333, 3334, 6666, 6667, 9999])
    rows, cols = sample_entries(N, 2048, 9)
    rows = np.concatenate([rows, np.repeat(edges, edges.size)])
    cols = np.concatenate([cols, np.tile(edges, edges.size)])
    got = dC[torch.from_numpy(rows).cuda(), torch.from_numpy(cols).cuda()].cpu().numpy()
    want = O.summa_entries(dA.cpu().numpy(), dB.cpu().numpy(), rows, cols, panel=N // 3)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    plan.close()


@pytest.mark.parametrize("N", [8192, 16384])
@pytest.mark.parametrize("dist", [phpc.DIST_UNIFORM, phpc.DIST_WIDE])
def test_split_crt_config_sizes(N, dist):
    """The split-precision CRT GEMM at the sizes its throughput is quoted on: sampled entries
    within err_cw 1e-14 of the double-double product, and a second call gives the same bits."""
    import torch
    print(f"split crt N={N} dist={dist}", flush=True)
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, dist, 1)
    phpc.fill_device(dB, dist, 2)
    dC = torch.zeros_like(dA)
    info = phpc.dgemm_split_device(dA, dB, dC, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    rows, cols = sample_entries(N, 2048, 4)
    cw, nw = phpc.check_sampled(dA, dB, dC, rows, cols)
    print(f"  moduli={info.slices} p={info.int_bits} err_cw={cw:.2e} estimate={info.error_estimate:.2e}",
          flush=True)
    assert cw <= 1e-14 and info.error_estimate <= 1e-14
    dC2 = torch.zeros_like(dA)
    phpc.dgemm_split_device(dA, dB, dC2, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    assert torch.equal(i64(dC), i64(dC2))


@pytest.mark.parametrize("dist", [phpc.DIST_UNIFORM, phpc.DIST_WIDE])
def test_split_fp8_slices_config_size(dist):
    """The FP8-E4M3 slice rows of the split-count table at N = 8192: the automatic count
    reaches err_cw 1e-14 on sampled entries against the double-double product, and a
    second call gives the same bits."""
    import torch
    N = 8192
    print(f"split fp8 N={N} dist={dist}", flush=True)
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, dist, 1)
    phpc.fill_device(dB, dist, 2)
    dC = torch.zeros_like(dA)
    dC2 = torch.zeros_like(dA)
    try:
        info = phpc.dgemm_split_device(dA, dB, dC, slices=0, target=1e-14, method=phpc.SPLIT_FP8)
        phpc.dgemm_split_device(dA, dB, dC2, slices=0, target=1e-14, method=phpc.SPLIT_FP8)
        torch.cuda.synchronize()
    finally:
        phpc.lib().phpc_split_set_method(phpc.SPLIT_CRT)
    rows, cols = sample_entries(N, 2048, 5)
    cw, _ = phpc.check_sampled(dA, dB, dC, rows, cols)
    print(f"  slices={info.slices} err_cw={cw:.2e} estimate={info.error_estimate:.2e}", flush=True)
    assert cw <= 1e-14 and info.error_estimate <= 1e-14
    assert torch.equal(i64(dC), i64(dC2))


def f32_sampled_ok(A, B, C, count, seed, C0=None, extra=0.0):
    """C (FP32, on the device) against float64 dot products of the FP32 inputs on `count`
    sampled entries, within ((k + 2) 2^-24 + extra) (|C0| + |A||B|); extra = 2^-23 for the
    3 x BF16 path's dropped slice terms. Computed on the device with elementwise products
    (FP32 x FP32 is exact in FP64)."""
    import torch
    m, k = A.shape
    n = B.shape[1]
    g = torch.Generator(device="cpu").manual_seed(seed)
    r = torch.randint(0, m, (count,), generator=g).cuda()
    c = torch.randint(0, n, (count,), generator=g).cuda()
    a = A[r].double()
    b = B[:, c].t().double()
    ref = (a * b).sum(1)
    mag = (a.abs() * b.abs()).sum(1)
    if C0 is not None:
        ref = ref + C0[r, c].double()
        mag = mag + C0[r, c].double().abs()
    err = (C[r, c].double() - ref).abs()
    bound = ((k + 2) * 2.0 ** -24 + extra) * mag
    return bool(torch.all(err <= bound)), float((err / mag.clamp_min(1e-300)).max())


@pytest.mark.parametrize("N", [8192, 16384])
def test_fp32_accurate_bf16x3_config_sizes(N):
    """The FP32-accurate 3 x BF16 variant of the split-precision workload at its benchmark
    sizes, FP32 uniform inputs: 512 sampled entries within the 3 x BF16 bound
    ((k + 2) 2^-24 + 2^-23) (|C0| + |A||B|), C not zero, and a second call on the same
    inputs gives the same bits."""
    import torch
    print(f"bf16x3 N={N}", flush=True)
    A = torch.empty((N, N), dtype=torch.float32, device="cuda")
    B = torch.empty_like(A)
    C0 = torch.empty_like(A)
    phpc.fill_device(A, phpc.DIST_UNIFORM, 41)
    phpc.fill_device(B, phpc.DIST_UNIFORM, 42)
    phpc.fill_device(C0, phpc.DIST_UNIFORM, 43)
    C = C0.clone()
    phpc.sgemm_device(A, B, C, method=phpc.SGEMM_BF16X3)
    C2 = C0.clone()
    phpc.sgemm_device(A, B, C2, method=phpc.SGEMM_BF16X3)
    torch.cuda.synchronize()
    ok, rel = f32_sampled_ok(A, B, C, 512, N, C0=C0, extra=2.0 ** -23)
    print(f"  worst err_cw {rel:.2e}", flush=True)
    assert ok
    assert torch.equal(C.view(torch.int32), C2.view(torch.int32))


def test_grouped_fp32_512_gemms_config_law():
    """512 FP32 GEMMs with M, N, K = rint(exp(U[ln 64, ln 4096])) in one autotuned plan:
    every GEMM within the FP32 bound on sampled entries, and three hipGraph replays, with
    eager work and synchronisations between them, equal the eager launch bit for bit."""
    import torch
    rng = np.random.default_rng(21)
    shapes = np.rint(np.exp(rng.uniform(np.log(64), np.log(4096), size=(512, 3)))).astype(int)
    gemms = []
    for i, (m, n, k) in enumerate(shapes):
        a = torch.empty((m, k), dtype=torch.float32, device="cuda")
        b = torch.empty((k, n), dtype=torch.float32, device="cuda")
        phpc.fill_device(a, phpc.DIST_UNIFORM, 10000 + 2 * i)
        phpc.fill_device(b, phpc.DIST_UNIFORM, 10001 + 2 * i)
        gemms.append((a, b, torch.zeros((m, n), dtype=torch.float32, device="cuda")))
    s = torch.cuda.Stream()
    plan = phpc.Plan(phpc.F32, gemms, autotune=True, stream=s)
    for g in gemms:
        g[2].zero_()
    torch.cuda.synchronize()
    plan.launch(s)
    s.synchronize()
    print(f"  plan {plan.info()}", flush=True)
    worst = 0.0
    for i, (a, b, c) in enumerate(gemms):
        ok, rel = f32_sampled_ok(a, b, c, 32, i)
        worst = max(worst, rel)
        assert ok, (i, tuple(shapes[i]))
    print(f"  512 GEMMs in bound, worst err_cw {worst:.2e}", flush=True)
    eager = [g[2].clone() for g in gemms]
    for rep in range(3):
        for g in gemms:
            g[2].zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            plan.graph_launch(s)
        s.synchronize()
        _ = torch.ones(1024, device="cuda").sum().item()  # eager work between replays
        for i, g in enumerate(gemms):
            assert torch.equal(g[2].view(torch.int32), eager[i].view(torch.int32)), (rep, i)
    plan.close()


def test_grouped_fp32_512_gemms_bf16x3_slice_planes():
    """The same 512 FP32 GEMMs as an FP32-accurate 3 x BF16 plan on slice planes (one cut
    of every operand, one grouped BF16 level launch): every GEMM within the 3 x BF16 bound
    ((k + 2) 2^-24 + 2^-23) (|A||B|) on sampled entries."""
    import torch
    rng = np.random.default_rng(21)
    shapes = np.rint(np.exp(rng.uniform(np.log(64), np.log(4096), size=(512, 3)))).astype(int)
    gemms = []
    for i, (m, n, k) in enumerate(shapes):
        a = torch.empty((m, k), dtype=torch.float32, device="cuda")
        b = torch.empty((k, n), dtype=torch.float32, device="cuda")
        phpc.fill_device(a, phpc.DIST_UNIFORM, 10000 + 2 * i)
        phpc.fill_device(b, phpc.DIST_UNIFORM, 10001 + 2 * i)
        gemms.append((a, b, torch.zeros((m, n), dtype=torch.float32, device="cuda")))
    plan = phpc.Plan(phpc.F32X3, gemms)
    info = plan.info()
    print(f"  plan {info}", flush=True)
    assert info["config"] == 6  # slice planes on 256 x 256 tiles, by work
    plan.launch()
    torch.cuda.synchronize()
    worst = 0.0
    for i, (a, b, c) in enumerate(gemms):
        ok, rel = f32_sampled_ok(a, b, c, 32, i, extra=2.0 ** -23)
        worst = max(worst, rel)
        assert ok, (i, tuple(shapes[i]))
    print(f"  512 GEMMs in bound, worst err_cw {worst:.2e}", flush=True)
    plan.close()


def _one_shot(acc, n, seed, split_min=4096):
    """One request through the non-serving entry points on the same generated operands:
    FP64 class by the split CRT GEMM when n >= split_min, else the plain FP64 kernel; FP32
    class by the 3 x BF16 slice planes when the product fills the GPU, else an FP32 plan."""
    import torch
    dt = torch.float64 if acc == phpc.ACC_FP64 else torch.float32
    A = torch.empty((n, n), dtype=dt, device="cuda")
    B = torch.empty_like(A)
    C = torch.zeros_like(A)
    phpc.fill_device(A, phpc.DIST_UNIFORM, seed)
    phpc.fill_device(B, phpc.DIST_UNIFORM, seed + 1)
    if acc == phpc.ACC_FP64 and n >= split_min:
        phpc.dgemm_split_device(A, B, C, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
    elif acc == phpc.ACC_FP64:
        phpc.dgemm_device(A, B, C)
    elif n * n >= 2048 * 2048 and ((n + 127) // 128) ** 2 >= 256:
        phpc.sgemm_device(A, B, C, method=phpc.SGEMM_BF16X3)
    else:
        plan = phpc.Plan(phpc.F32, [(A, B, C)])
        plan.launch()
        torch.cuda.synchronize()
        plan.close()
    torch.cuda.synchronize()
    return A, B, C


def test_serving_poisson_stream_config_law():
    """150 requests arriving as a Poisson stream: sizes 60% U[256,1024], 30% U[1024,4096],
    10% U[4096,8192], FP64/FP32 classes 50/50, operands generated in HBM under a 200 GB
    budget, three workers with continuous batching. Every result equals the same request
    through the one-shot entry points bit for bit (so it does not depend on what it ran
    beside), and meets its class's bound against the double-double (FP64) or float64 (FP32)
    product on sampled entries."""
    import time

    import torch
    torch.cuda.empty_cache()
    n_req = 150
    rng = np.random.default_rng(11)
    u = rng.random(n_req)
    size = np.where(u < 0.6, rng.integers(256, 1025, n_req),
                    np.where(u < 0.9, rng.integers(1024, 4097, n_req), rng.integers(4096, 8193, n_req)))
    acc = np.where(rng.random(n_req) < 0.5, phpc.ACC_FP64, phpc.ACC_FP32)
    arrivals = np.cumsum(rng.exponential(1.0 / 300.0, n_req))
    srv = phpc.Server(200 * 10 ** 9, slots=3, max_dim=8192, batch_max=32)
    outs, tickets = [], []
    try:
        t0 = time.perf_counter()
        for i in range(n_req):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            n = int(size[i])
            out = np.full((n, n), np.nan, dtype=np.float64 if acc[i] == phpc.ACC_FP64 else np.float32)
            outs.append(out)
            tickets.append(srv.submit(int(acc[i]), n, n, n, phpc.DIST_UNIFORM, 1000 + 2 * i, out))
        res = [srv.wait(t) for t in tickets]
        _, peak = srv.memory()
    finally:
        srv.close()
    assert all(r["status"] == 0 for r in res)
    assert 0 < peak <= 200 * 10 ** 9
    lat = np.array([r["latency_s"] for r in res])
    print(f"  stream: p50 {np.percentile(lat, 50) * 1e3:.2f} ms, p99 {np.percentile(lat, 99) * 1e3:.2f} ms, "
          f"batched {sum(r['batch_size'] > 1 for r in res)} of {n_req}", flush=True)
    for i in range(n_req):
        n = int(size[i])
        A, B, C = _one_shot(int(acc[i]), n, 1000 + 2 * i)
        want = C.cpu().numpy()
        view = np.uint64 if acc[i] == phpc.ACC_FP64 else np.uint32
        assert np.array_equal(outs[i].view(view), want.view(view)), (i, int(acc[i]), n)
        if acc[i] == phpc.ACC_FP64:
            rows, cols = sample_entries(n, 64, i)
            cw, _ = phpc.check_sampled(A, B, C, rows, cols)
            assert cw <= 1e-14, (i, n, cw)
        else:
            bf16x3 = n * n >= 2048 * 2048 and ((n + 127) // 128) ** 2 >= 256
            ok, rel = f32_sampled_ok(A, B, C, 64, i, extra=2.0 ** -23 if bf16x3 else 0.0)
            assert ok, (i, n, rel)
        del A, B, C
        outs[i] = None
