"""Serving loop: every request's result is the same bits whether it runs alone or in a
mix of concurrent requests, and the same bits as the library's one-shot entry point on
the same generated operands; admission never exceeds the memory budget."""
import numpy as np
import pytest

from conftest import phpc
import oracle as O

pytestmark = pytest.mark.gpu

GiB = 1 << 30


def bits(x):
    x = np.ascontiguousarray(x)
    return x.view(np.uint64 if x.dtype == np.float64 else np.uint32)


SPLIT_MIN = 256


def one_shot(acc, m, n, k, dist, seed):
    """The same request through the non-serving entry points (split CRT GEMM for FP64
    requests with min(m, n, k) >= SPLIT_MIN, plain FP64 kernel below, FP32 plan)."""
    import torch
    dt = torch.float64 if acc == phpc.ACC_FP64 else torch.float32
    A = torch.empty((m, k), dtype=dt, device="cuda")
    B = torch.empty((k, n), dtype=dt, device="cuda")
    C = torch.zeros((m, n), dtype=dt, device="cuda")
    phpc.fill_device(A, dist, seed)
    phpc.fill_device(B, dist, seed + 1)
    if acc == phpc.ACC_FP64 and min(m, n, k) >= SPLIT_MIN:
        phpc.dgemm_split_device(A, B, C, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
    elif acc == phpc.ACC_FP64:
        phpc.dgemm_device(A, B, C)
    elif m * n >= 2048 * 2048 and k >= 256 and (m + 127) // 128 * ((n + 127) // 128) >= 256:
        phpc.sgemm_device(A, B, C, method=phpc.SGEMM_BF16X3)
    else:
        plan = phpc.Plan(phpc.F32, [(A, B, C)])
        plan.launch()
        torch.cuda.synchronize()
        plan.close()
    torch.cuda.synchronize()
    return A.cpu().numpy(), B.cpu().numpy(), C.cpu().numpy()


def test_fill_f32_is_rounded_f64_stream():
    import torch
    for dist, gen in ((phpc.DIST_UNIFORM, O.uniform), (phpc.DIST_WIDE, O.wide)):
        x = torch.empty((37, 53), dtype=torch.float32, device="cuda")
        phpc.fill_device(x, dist, 11)
        assert np.array_equal(bits(x.cpu().numpy()), bits(gen((37, 53), 11).astype(np.float32)))


MIX = [
    (phpc.ACC_FP32, 300, 200, 150, 0, 1001),
    (phpc.ACC_FP64, 257, 129, 333, 0, 2001),
    (phpc.ACC_FP32, 64, 64, 64, 1, 3001),
    (phpc.ACC_FP64, 512, 512, 512, 1, 4001),
    (phpc.ACC_FP32, 1000, 17, 999, 0, 5001),
    (phpc.ACC_FP32, 1, 1, 1, 0, 6001),
    (phpc.ACC_FP64, 33, 700, 65, 0, 7001),
    (phpc.ACC_FP32, 512, 768, 256, 0, 8001),
    (phpc.ACC_FP64, 64, 96, 128, 0, 9001),
    (phpc.ACC_FP64, 300, 280, 260, 0, 10001),
    (phpc.ACC_FP32, 2304, 2048, 512, 0, 11001),  # fills the GPU: 3 x BF16 slice planes
]


def out_array(acc, m, n):
    return np.full((m, n), np.nan, dtype=np.float64 if acc == phpc.ACC_FP64 else np.float32)


@pytest.mark.parametrize("policy", [phpc.SCHED_FIFO, phpc.SCHED_SJF, phpc.SCHED_LANES])
def test_alone_vs_mix_vs_one_shot(policy):
    srv = phpc.Server(8 * GiB, slots=3, max_dim=2304, batch_max=8)
    srv.set_split_min(SPLIT_MIN)
    srv.set_policy(policy, 0.0)
    try:
        alone = []
        for acc, m, n, k, dist, seed in MIX:
            out = out_array(acc, m, n)
            r = srv.wait(srv.submit(acc, m, n, k, dist, seed, out))
            assert r["status"] == 0 and r["batch_size"] == 1
            alone.append(out)
        # the same requests submitted together, each twice, interleaved with fillers
        tickets = []
        for rep in range(2):
            for i, (acc, m, n, k, dist, seed) in enumerate(MIX):
                out = out_array(acc, m, n)
                tickets.append((i, out, srv.submit(acc, m, n, k, dist, seed, out)))
                srv.submit(phpc.ACC_FP32, 96, 96, 96, 0, 9000 + i)  # filler, not waited on
        batched = {phpc.ACC_FP32: 0, phpc.ACC_FP64: 0}
        for i, out, t in tickets:
            r = srv.wait(t)
            assert r["status"] == 0
            batched[MIX[i][0]] += r["batch_size"] > 1
            assert np.array_equal(bits(out), bits(alone[i])), MIX[i]
        assert batched[phpc.ACC_FP32] > 0  # some requests shared a launch
        srv.drain()
    finally:
        srv.close()
    for i, (acc, m, n, k, dist, seed) in enumerate(MIX):
        A, B, C = one_shot(acc, m, n, k, dist, seed)
        assert np.array_equal(bits(C), bits(alone[i])), MIX[i]
        A64, B64 = A.astype(np.float64), B.astype(np.float64)
        if m * n > 1 << 20:  # large: sampled entries
            rng = np.random.default_rng(2)
            r, c = rng.integers(0, m, 256), rng.integers(0, n, 256)
            ref = np.einsum("ik,ki->i", A64[r], B64[:, c])
            scale = np.einsum("ik,ki->i", np.abs(A64[r]), np.abs(B64[:, c]))
            got = alone[i][r, c].astype(np.float64)
        else:
            ref = O.gemm_dd(A64, B64)
            scale = np.abs(A64) @ np.abs(B64)
            got = alone[i].astype(np.float64)
        err = np.max(np.abs(got - ref) / np.maximum(scale, 1e-300))
        if acc == phpc.ACC_FP64:
            assert err <= 1e-14, (MIX[i], err)
        else:
            assert err <= (k + 2) * 2.0 ** -24 + 2.0 ** -23, (MIX[i], err)


def test_budget_admission():
    import torch
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    # max_dim 4096: the two slots' workspaces take about 1.9 GiB of the 4 GiB, leaving an
    # operand arena of about 2.1 GiB
    srv = phpc.Server(4 * GiB, slots=2, max_dim=4096, batch_max=4)
    try:
        # a request whose operands alone exceed the budget is refused at submit
        with pytest.raises(RuntimeError, match="error 6"):
            srv.submit(phpc.ACC_FP64, 16384, 16384, 16384)
        # FP64 requests of 384 MiB each (split path) and FP32 ones of 192 MiB (3 x BF16
        # slice planes): at most five of the former fit at once, so admission must wait
        ts = [srv.submit(phpc.ACC_FP64, 4096, 4096, 4096, 0, 50 + i) for i in range(7)]
        ts += [srv.submit(phpc.ACC_FP32, 4096, 4096, 4096, 0, 70 + i) for i in range(3)]
        ts += [srv.submit(phpc.ACC_FP64, 200, 200, 200, 0, 60)]
        for t in ts:
            assert srv.wait(t)["status"] == 0
        in_use, peak = srv.memory()
        assert in_use == 0 and 0 < peak <= 4 * GiB
        # the device memory the server holds (workspaces + arena) stays within the budget:
        # nothing grows on the request path
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        assert free0 - free1 <= 4 * GiB + 64 * 2 ** 20, (free0 - free1) / GiB
    finally:
        srv.close()


def test_requests_wider_than_max_dim_are_refused():
    """The per-slot workspaces are sized for max_dim at creation; a wider request is refused at
    submit instead of growing them outside the budget, and the server keeps serving."""
    srv = phpc.Server(1 * GiB, slots=1, max_dim=256, batch_max=4)
    try:
        for shape in ((257, 8, 8), (8, 257, 8), (8, 8, 257)):
            with pytest.raises(RuntimeError, match="error 6"):
                srv.submit(phpc.ACC_FP32, *shape)
        assert "max_dim" in phpc.lib().phpc_last_error_message().decode()
        t = srv.submit(phpc.ACC_FP64, 256, 256, 256, 0, 5)
        assert srv.wait(t)["status"] == 0
    finally:
        srv.close()


def test_latency_accounting_and_drain():
    srv = phpc.Server(4 * GiB, slots=2, max_dim=512, batch_max=16)
    try:
        rng = np.random.default_rng(3)
        ts = []
        for i in range(40):
            acc = phpc.ACC_FP64 if rng.random() < 0.3 else phpc.ACC_FP32
            m, n, k = (int(x) for x in rng.integers(16, 512, 3))
            ts.append(srv.submit(acc, m, n, k, 0, 100 + 2 * i))
        srv.drain()
        for t in ts:
            r = srv.wait(t)
            assert r["status"] == 0
            assert r["latency_s"] >= r["compute_s"] > 0
            assert abs(r["latency_s"] - r["queue_s"] - r["compute_s"]) < 1e-6
    finally:
        srv.close()
