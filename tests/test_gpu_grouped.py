"""Grouped plans: many GEMMs in one persistent launch (FP64 bit-exact against the
oracle's restatement of the reference kernel; FP32 within the FP32 dot-product bound),
autotuning that leaves C untouched, and hipGraph replays equal to eager launches."""
import numpy as np
import pytest

from conftest import phpc
import oracle as O

pytestmark = pytest.mark.gpu


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def loguniform_shapes(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return [tuple(int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(3)) for _ in range(count)]


def f32_bound(A, B, C0):
    """|fl(C0 + A B) - (C0 + A B)| <= (k + 2) u32 (|C0| + |A||B|), u32 = 2^-24."""
    k = A.shape[1]
    return (k + 2) * 2.0 ** -24 * (np.abs(C0) + np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64)))


@pytest.mark.parametrize("autotune", [False, True])
def test_f32_mixed_shapes_one_launch(autotune):
    import torch
    shapes = loguniform_shapes(48, 8, 700, seed=5 if autotune else 6)
    gemms, host = [], []
    for i, (m, n, k) in enumerate(shapes):
        A = O.uniform((m, k), 100 + i).astype(np.float32)
        B = O.uniform((k, n), 200 + i).astype(np.float32)
        C0 = O.uniform((m, n), 300 + i).astype(np.float32)
        gemms.append((torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), torch.from_numpy(C0.copy()).cuda()))
        host.append((A, B, C0))
    plan = phpc.Plan(phpc.F32, gemms, autotune=autotune)
    torch.cuda.synchronize()
    for (_, _, dC), (_, _, C0) in zip(gemms, host):  # autotuning restores C
        assert np.array_equal(dC.cpu().numpy(), C0)
    plan.launch()
    torch.cuda.synchronize()
    for (_, _, dC), (A, B, C0) in zip(gemms, host):
        ref = C0.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
        assert np.all(np.abs(dC.cpu().numpy().astype(np.float64) - ref) <= f32_bound(A, B, C0))
    info = plan.info()
    assert info["tiles"] > 0 and (info["tuned_ms"] > 0) == autotune


@pytest.mark.parametrize("shapes", [[(512, 768, 256)], [(64, 64, 64)], [(128, 128, 16)],
                                    [(2048, 2048, 256)], [(512, 768, 256), (64, 64, 64), (4, 8, 12)]])
def test_f32_vec4_loads(shapes):
    """Every dimension a multiple of 4 and 16-byte aligned rows: the 128-bit load path."""
    import torch
    gemms, host = [], []
    for i, (m, n, k) in enumerate(shapes):
        A = O.uniform((m, k), 40 + i).astype(np.float32)
        B = O.uniform((k, n), 50 + i).astype(np.float32)
        C0 = O.uniform((m, n), 60 + i).astype(np.float32)
        gemms.append((torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), torch.from_numpy(C0.copy()).cuda()))
        host.append((A, B, C0))
    plan = phpc.Plan(phpc.F32, gemms)
    plan.launch()
    torch.cuda.synchronize()
    for (_, _, dC), (A, B, C0) in zip(gemms, host):
        ref = C0.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
        assert np.all(np.abs(dC.cpu().numpy().astype(np.float64) - ref) <= f32_bound(A, B, C0))


def _f32_plan_with_config(cfg, gemms):
    phpc.set_tuning(phpc.TUNE_F32_PLAN_CONFIG, cfg)
    try:
        return phpc.Plan(phpc.F32, gemms)
    finally:
        phpc.set_tuning(phpc.TUNE_F32_PLAN_CONFIG, -1)


def test_f32_plan_transposed_a_bit_identical_and_replays():
    """FP32 plans on the transposed-A configuration: one batched kernel stages every A
    transposed (an A shared by two items once), then one persistent launch runs all items,
    reading B in place (4 bytes per lane, since some rows are not 16-byte aligned). The
    kernel takes k in the same pairs
    and order as the register-staged configurations, so the results equal configuration 2's
    bit for bit, within the FP32 bound; captured replays (eager work in between) equal eager
    launches."""
    import torch
    shapes = [(300, 257, 200), (128, 128, 16), (1, 1, 1), (77, 4096, 3), (1000, 17, 999), (512, 768, 256),
              (2049, 130, 64), (64, 64, 2)]
    host = []
    for i, (m, n, k) in enumerate(shapes):
        host.append((O.uniform((m, k), 70 + i).astype(np.float32), O.uniform((k, n), 80 + i).astype(np.float32),
                     O.uniform((m, n), 90 + i).astype(np.float32)))
    # one more item sharing the first item's A, with a column-offset (unaligned) view of B
    Bw = O.uniform((200, 70), 99).astype(np.float32)
    host.append((host[0][0], Bw[:, 1:66], O.uniform((300, 65), 98).astype(np.float32)))

    def make():
        dA0 = torch.from_numpy(host[0][0]).cuda()
        g = []
        for i, (A, B, C0) in enumerate(host):
            dA = dA0 if i in (0, len(host) - 1) else torch.from_numpy(A).cuda()
            dB = torch.from_numpy(np.ascontiguousarray(B)).cuda() if i < len(host) - 1 else \
                torch.from_numpy(Bw).cuda()[:, 1:66]
            g.append((dA, dB, torch.from_numpy(C0.copy()).cuda()))
        return g

    ga, gr, ge = make(), make(), make()
    pa = _f32_plan_with_config(3, ga)
    pr = _f32_plan_with_config(2, gr)
    pe = _f32_plan_with_config(3, ge)
    assert pa.info()["config"] == 3 and pr.info()["config"] == 2
    s = torch.cuda.Stream()
    junk = torch.randn(512, 512, device="cuda")
    for rep in range(3):
        with torch.cuda.stream(s):
            pa.graph_launch(s)
        s.synchronize()
        junk = junk @ junk.T / 512.0
        torch.cuda.synchronize()
        pr.launch()
        pe.launch()
        torch.cuda.synchronize()
        for (_, _, c_a), (_, _, c_r), (_, _, c_e) in zip(ga, gr, ge):
            assert torch.equal(c_a.view(torch.int32), c_r.view(torch.int32)), rep
            assert torch.equal(c_a.view(torch.int32), c_e.view(torch.int32)), rep
    for (_, _, dC), (A, B, C0) in zip(ga, host):
        ref = C0.astype(np.float64) + 3 * (A.astype(np.float64) @ B.astype(np.float64))
        # three launches, each within (k + 2) u (|C before| + |A||B|)
        bound = 3 * (A.shape[1] + 2) * 2.0 ** -24 * (np.abs(C0) + 3 * np.abs(A.astype(np.float64)) @
                                                    np.abs(B.astype(np.float64)))
        assert np.all(np.abs(dC.cpu().numpy().astype(np.float64) - ref) <= bound)
    for p in (pa, pr, pe):
        p.close()


def test_f64_grouped_items_bitexact():
    import torch
    shapes = [(130, 70, 90), (1, 1, 1), (257, 3, 33), (64, 64, 64), (300, 129, 17)]
    gemms, host = [], []
    for i, (m, n, k) in enumerate(shapes):
        A, B, C0 = O.uniform((m, k), 10 + i), O.uniform((k, n), 20 + i), O.uniform((m, n), 30 + i)
        gemms.append((torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), torch.from_numpy(C0.copy()).cuda()))
        host.append((A, B, C0))
    plan = phpc.Plan(phpc.F64, gemms, autotune=True)
    plan.launch()
    torch.cuda.synchronize()
    for (_, _, dC), (A, B, C0) in zip(gemms, host):
        assert np.array_equal(bits(dC.cpu().numpy()), bits(O.kernel_emul(A, B, C0.copy())))


def summa_items(dA, dB, dC, N, p, q, panel):
    items = []
    for i in range(p):
        for j in range(q):
            r0, mr = O.block_start(N, p, i), O.block_size(N, p, i)
            c0, nc = O.block_start(N, q, j), O.block_size(N, q, j)
            items.append((dA[r0:r0 + mr], dB[:, c0:c0 + nc], dC[r0:r0 + mr, c0:c0 + nc], panel))
    return items


def test_f64_uneven_rank_blocks_match_summa():
    import torch
    N, p, q = 301, 3, 3
    panel = N // O.find_lcm(p, q)
    A, B, C0 = O.uniform((N, N), 1), O.uniform((N, N), 2), O.uniform((N, N), 3)
    dA, dB, dC = (torch.from_numpy(x.copy()).cuda() for x in (A, B, C0))
    plan = phpc.Plan(phpc.F64, summa_items(dA, dB, dC, N, p, q, panel), autotune=True)
    plan.launch()
    torch.cuda.synchronize()
    assert np.array_equal(bits(dC.cpu().numpy()), bits(O.summa_emul(A, B, C0.copy(), p, q, panel)))


def test_graph_replays_equal_eager():
    """Replay the captured launch several times with eager work and synchronisations in
    between; after every replay C equals the eager result bit for bit (FP64)."""
    import torch
    N = 384
    A, B, C0 = O.uniform((N, N), 4), O.uniform((N, N), 5), O.uniform((N, N), 6)
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dCg, dCe = torch.from_numpy(C0.copy()).cuda(), torch.from_numpy(C0.copy()).cuda()
    s = torch.cuda.Stream()
    pg = phpc.Plan(phpc.F64, summa_items(dA, dB, dCg, N, 2, 2, 32))
    pe = phpc.Plan(phpc.F64, summa_items(dA, dB, dCe, N, 2, 2, 32))
    junk = torch.randn(1000, 1000, device="cuda")
    for rep in range(4):
        with torch.cuda.stream(s):
            pg.graph_launch(s)
        s.synchronize()
        junk = junk @ junk.T / 1000.0  # eager work between replays
        torch.cuda.synchronize()
        pe.launch()
        torch.cuda.synchronize()
        assert torch.equal(dCg.view(torch.int64), dCe.view(torch.int64)), rep


def test_f64_n10000_grid3x3_one_launch():
    """N = 10000 on a 3 x 3 grid: rank blocks 3334/3333/3333 (the reference needs N
    divisible by the grid), panel N / lcm = 3333, one persistent launch per pass."""
    import torch
    N, p, q = 10000, 3, 3
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, phpc.DIST_UNIFORM, 1)
    phpc.fill_device(dB, phpc.DIST_UNIFORM, 2)
    dC = torch.zeros_like(dA)
    s = torch.cuda.Stream()
    plan = phpc.Plan(phpc.F64, summa_items(dA, dB, dC, N, p, q, N // 3), autotune=False)
    with torch.cuda.stream(s):
        plan.graph_launch(s)
    s.synchronize()
    rng = np.random.default_rng(8)
    rows = np.concatenate([rng.integers(0, N, 1024), [3333, 3334, 6666, 6667, 9999]])
    cols = np.concatenate([rng.integers(0, N, 1024), [3333, 3334, 6666, 6667, 9999]])
    cw, _ = phpc.check_sampled(dA, dB, dC, rows, cols)
    assert cw <= 1e-14


def test_f64_plan_transposed_a_graph_replays():
    """Large single-panel FP64 plans take the transposed-A kernel: the plan keeps its own
    transposed copies of A, refreshed by every launch, and captured replays (eager work and
    synchronisations in between) equal eager launches of the same plan bit for bit; the
    result also equals the host-descriptor entry point's."""
    import torch
    N = 4096
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, phpc.DIST_UNIFORM, 11)
    phpc.fill_device(dB, phpc.DIST_UNIFORM, 12)
    dCg, dCe, dCd = torch.zeros_like(dA), torch.zeros_like(dA), torch.zeros_like(dA)
    s = torch.cuda.Stream()
    pg = phpc.Plan(phpc.F64, summa_items(dA, dB, dCg, N, 2, 2, 0))
    pe = phpc.Plan(phpc.F64, summa_items(dA, dB, dCe, N, 2, 2, 0))
    assert pg.info()["config"] == 11
    junk = torch.randn(1000, 1000, device="cuda")
    for rep in range(3):
        with torch.cuda.stream(s):
            pg.graph_launch(s)
        s.synchronize()
        junk = junk @ junk.T / 1000.0
        torch.cuda.synchronize()
        pe.launch()
        phpc.dgemm_device(dA, dB, dCd)
        torch.cuda.synchronize()
        assert torch.equal(dCg.view(torch.int64), dCe.view(torch.int64)), rep
        assert torch.equal(dCg.view(torch.int64), dCd.view(torch.int64)), rep
    pg.close()
    pe.close()


def test_f32_plan_transposed_a_aligned_b_uses_16_byte_loads():
    """Every B 16-byte aligned with a pitch of 4 floats: the 16-byte B load form of the
    transposed-A kernel, bit for bit equal to configuration 2."""
    import torch
    shapes = [(300, 256, 200), (129, 132, 17), (640, 512, 64)]
    host = [(O.uniform((m, k), 30 + i).astype(np.float32), O.uniform((k, n), 40 + i).astype(np.float32),
             O.uniform((m, n), 50 + i).astype(np.float32)) for i, (m, n, k) in enumerate(shapes)]
    outs = []
    for cfg in (3, 2):
        g = [(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), torch.from_numpy(C0.copy()).cuda())
             for A, B, C0 in host]
        plan = _f32_plan_with_config(cfg, g)
        assert plan.info()["config"] == cfg
        plan.launch()
        torch.cuda.synchronize()
        outs.append([c.cpu().numpy() for _, _, c in g])
        plan.close()
    for a, b in zip(*outs):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
