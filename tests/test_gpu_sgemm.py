"""FP32-accurate GEMM from three BF16 slices (and the FP32 matrix-core path) against a
float64 reference of the same FP32 inputs, within the stated bound; exact on small
integers; grouped plans replayed from a hipGraph equal eager launches."""
import numpy as np
import pytest

from conftest import phpc
import oracle as O

pytestmark = pytest.mark.gpu


def bound(A, B, C0):
    k = A.shape[1]
    mag = np.abs(C0).astype(np.float64) + np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    return ((k + 2) * 2.0 ** -24 + 2.0 ** -23) * mag


SHAPES = [(128, 128, 32), (256, 384, 512), (1, 1, 1), (130, 70, 33), (257, 129, 1000), (64, 512, 16),
          (1000, 17, 999), (2048, 2048, 2048)]


@pytest.mark.parametrize("method", [phpc.SGEMM_BF16X3, phpc.SGEMM_FP32])
@pytest.mark.parametrize("m,n,k", SHAPES)
def test_sgemm_within_fp32_bound(method, m, n, k):
    import torch
    A = O.uniform((m, k), 1).astype(np.float32)
    B = O.uniform((k, n), 2).astype(np.float32)
    C0 = O.uniform((m, n), 3).astype(np.float32)
    dC = torch.from_numpy(C0.copy()).cuda()
    phpc.sgemm_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, method=method)
    torch.cuda.synchronize()
    ref = C0.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
    err = np.abs(dC.cpu().numpy().astype(np.float64) - ref)
    assert np.all(err <= bound(A, B, C0)), (m, n, k, float(np.max(err / bound(A, B, C0))))


def _sgemm(A, B, C0, key6, b_view=None):
    import torch
    dB = torch.from_numpy(B).cuda() if b_view is None else b_view
    dC = torch.from_numpy(C0.copy()).cuda()
    phpc.set_tuning(phpc.TUNE_F32_TRANSPOSED_A, key6)
    try:
        phpc.sgemm_device(torch.from_numpy(A).cuda(), dB, dC, method=phpc.SGEMM_FP32)
    finally:
        phpc.set_tuning(phpc.TUNE_F32_TRANSPOSED_A, 1)
    torch.cuda.synchronize()
    return dC.cpu().numpy()


# large enough (>= 2 tiles of 128 x 128 per CU) for the transposed-A kernel; n = 2052 keeps
# B with 16-byte loads, n = 2049 / 2051 (rows not 16-byte aligned) with 4-byte loads
@pytest.mark.parametrize("m,n,k", [(4100, 2052, 300), (4097, 2049, 1000), (4096, 2051, 17), (4224, 4096, 2)])
def test_fp32_transposed_a_bit_identical_to_register_staged(m, n, k):
    """The transposed-A FP32 kernel takes k in the same pairs, in the same order, as the
    register-staged kernel, so the two agree bit for bit; both stay within the FP32 bound."""
    A = O.uniform((m, k), 11).astype(np.float32)
    B = O.uniform((k, n), 12).astype(np.float32)
    C0 = O.uniform((m, n), 13).astype(np.float32)
    got = _sgemm(A, B, C0, 1)
    want = _sgemm(A, B, C0, 0)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (m, n, k)
    ref = C0.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
    assert np.all(np.abs(got.astype(np.float64) - ref) <= bound(A, B, C0))


def test_fp32_transposed_a_on_offset_views():
    """B a column-offset view (base not 16-byte aligned) and A with a padded pitch."""
    import torch
    m, n, k = 4160, 2080, 257
    Ap = O.uniform((m, k + 3), 21).astype(np.float32)
    Bp = O.uniform((k, n + 5), 22).astype(np.float32)
    C0 = O.uniform((m, n), 23).astype(np.float32)
    dA = torch.from_numpy(Ap).cuda()[:, 1:k + 1]
    dB = torch.from_numpy(Bp).cuda()[:, 1:n + 1]
    outs = []
    for key6 in (1, 0):
        dC = torch.from_numpy(C0.copy()).cuda()
        phpc.set_tuning(phpc.TUNE_F32_TRANSPOSED_A, key6)
        try:
            phpc.sgemm_device(dA, dB, dC, method=phpc.SGEMM_FP32)
        finally:
            phpc.set_tuning(phpc.TUNE_F32_TRANSPOSED_A, 1)
        torch.cuda.synchronize()
        outs.append(dC.cpu().numpy())
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    A, B = Ap[:, 1:k + 1], Bp[:, 1:n + 1]
    ref = C0.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
    assert np.all(np.abs(outs[0].astype(np.float64) - ref) <= bound(A, B, C0))


def test_fp32_transposed_a_special_values():
    """Infinities, NaNs and signed zeros through the transposed-A FP32 kernel give the
    register-staged kernel's results: NaNs at the same places, every other entry bit for bit
    (the same k pairs in the same order)."""
    m, n, k = 4100, 2052, 64
    A = O.uniform((m, k), 61).astype(np.float32)
    B = O.uniform((k, n), 62).astype(np.float32)
    C0 = O.uniform((m, n), 63).astype(np.float32)
    A[5, 3] = np.inf
    A[6, 4] = -np.inf
    B[3, 9] = 0.0                # C[5, 9]: inf * 0 = NaN
    B[4, 10] = np.nan
    A[7, :] = -0.0
    C0[7, :] = -0.0
    C0[8, 8] = np.inf
    got, want = _sgemm(A, B, C0, 1), _sgemm(A, B, C0, 0)
    gn, wn = np.isnan(got), np.isnan(want)
    assert np.array_equal(gn, wn) and gn[5, 9] and gn[4, 10]
    assert np.array_equal(np.where(gn, 0, got).view(np.uint32), np.where(wn, 0, want).view(np.uint32))
    assert np.isinf(got[8, 8]) and np.isinf(got[5, 0]) and not np.signbit(got[7, 0])


def test_bf16x3_wide_range_inputs():
    """Entries spread over 2^-40 .. 2^40: each slice keeps its own exponent."""
    import torch
    rng = np.random.default_rng(4)
    m, n, k = 200, 160, 300
    A = np.ldexp(rng.uniform(0.5, 1, (m, k)) * rng.choice([-1, 1], (m, k)), rng.integers(-40, 40, (m, k)))
    B = np.ldexp(rng.uniform(0.5, 1, (k, n)) * rng.choice([-1, 1], (k, n)), rng.integers(-40, 40, (k, n)))
    A, B = A.astype(np.float32), B.astype(np.float32)
    C0 = np.zeros((m, n), np.float32)
    dC = torch.zeros((m, n), dtype=torch.float32, device="cuda")
    phpc.sgemm_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, method=phpc.SGEMM_BF16X3)
    torch.cuda.synchronize()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert np.all(np.abs(dC.cpu().numpy() - ref) <= bound(A, B, C0))


def test_bf16x3_exact_on_small_integers():
    """Integers |x| <= 8 are single BF16 slices; products and sums below 2^24 are exact."""
    import torch
    rng = np.random.default_rng(5)
    m, n, k = 300, 200, 512
    A = rng.integers(-8, 9, (m, k)).astype(np.float32)
    B = rng.integers(-8, 9, (k, n)).astype(np.float32)
    dC = torch.zeros((m, n), dtype=torch.float32, device="cuda")
    phpc.sgemm_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, method=phpc.SGEMM_BF16X3)
    torch.cuda.synchronize()
    want = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
    assert np.array_equal(dC.cpu().numpy(), want)


def test_bf16x3_plan_graph_replays_equal_eager():
    import torch
    shapes = [(300, 200, 150), (512, 768, 256), (64, 64, 64), (1000, 17, 999)]
    gem_g, gem_e, host = [], [], []
    for i, (m, n, k) in enumerate(shapes):
        A = O.uniform((m, k), 10 + i).astype(np.float32)
        B = O.uniform((k, n), 20 + i).astype(np.float32)
        dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        gem_g.append((dA, dB, torch.zeros((m, n), dtype=torch.float32, device="cuda")))
        gem_e.append((dA, dB, torch.zeros((m, n), dtype=torch.float32, device="cuda")))
        host.append((A, B))
    s = torch.cuda.Stream()
    pg = phpc.Plan(phpc.F32X3, gem_g, autotune=True, stream=s)
    # the same shape list: the tuned configuration comes from the cache, so both plans
    # run the same kernels
    pe = phpc.Plan(phpc.F32X3, gem_e, autotune=True)
    assert pg.info()["config"] == pe.info()["config"]
    junk = torch.randn(512, 512, device="cuda")
    for rep in range(3):
        for g in gem_g + gem_e:
            g[2].zero_()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            pg.graph_launch(s)
        s.synchronize()
        junk = junk @ junk.T / 512.0
        pe.launch()
        torch.cuda.synchronize()
        for (_, _, cg), (_, _, ce) in zip(gem_g, gem_e):
            assert torch.equal(cg.view(torch.int32), ce.view(torch.int32)), rep
    for (A, B), (_, _, dC) in zip(host, gem_e):
        ref = A.astype(np.float64) @ B.astype(np.float64)
        assert np.all(np.abs(dC.cpu().numpy() - ref) <= bound(A, B, np.zeros_like(ref)))


@pytest.mark.parametrize("m,n,k", [(2100, 2300, 1000), (2048, 2048, 4096), (4096, 2048, 257)])
def test_bf16x3_slice_planes_within_fp32_bound(m, n, k):
    """Large products take the slice-plane path (BF16 planes cut once, three slice
    levels on the direct-to-LDS pipeline); ragged shapes included."""
    import torch
    A = O.uniform((m, k), 7).astype(np.float32)
    B = O.uniform((k, n), 8).astype(np.float32)
    C0 = O.uniform((m, n), 9).astype(np.float32)
    dC = torch.from_numpy(C0.copy()).cuda()
    phpc.sgemm_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, method=phpc.SGEMM_BF16X3)
    torch.cuda.synchronize()
    rng = np.random.default_rng(1)
    r, c = rng.integers(0, m, 512), rng.integers(0, n, 512)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = C0[r, c].astype(np.float64) + np.einsum("ik,ki->i", A64[r], B64[:, c])
    mag = np.abs(C0[r, c]).astype(np.float64) + np.einsum("ik,ki->i", np.abs(A64[r]), np.abs(B64[:, c]))
    err = np.abs(dC.cpu().numpy()[r, c].astype(np.float64) - ref)
    assert np.all(err <= ((k + 2) * 2.0 ** -24 + 2.0 ** -23) * mag)


def test_bf16x3_slice_planes_exact_on_small_integers():
    import torch
    rng = np.random.default_rng(6)
    m, n, k = 2048, 2304, 640
    A = rng.integers(-8, 9, (m, k)).astype(np.float32)
    B = rng.integers(-8, 9, (k, n)).astype(np.float32)
    dC = torch.zeros((m, n), dtype=torch.float32, device="cuda")
    phpc.sgemm_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, method=phpc.SGEMM_BF16X3)
    torch.cuda.synchronize()
    want = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
    assert np.array_equal(dC.cpu().numpy(), want)


def _x3_plan(items, cfg, stream=None):
    """A 3 x BF16 plan pinned to configuration cfg (tuning key 8)."""
    phpc.set_tuning(8, cfg)
    try:
        plan = phpc.Plan(phpc.F32X3, items, stream=stream)
    finally:
        phpc.set_tuning(8, -1)
    assert plan.info()["config"] == cfg
    return plan


@pytest.mark.parametrize("cfg", [4, 5, 6])
def test_bf16x3_plan_slice_planes_equal_single_path(cfg):
    """A 3 x BF16 plan on slice planes (one batched cut of every distinct operand, then one
    grouped level launch over every item's tiles) gives each item the bits of the
    single-product slice-plane path (phpc_sgemm_device with BF16X3, which takes the planes
    at m n >= 2048^2 and k >= 256): ragged shapes, k not a multiple of 64, one A shared by
    two items, C not zero."""
    import torch
    A0 = torch.from_numpy(O.uniform((2100, 1000), 31).astype(np.float32)).cuda()
    ops = [(A0, 1000, 2300), (A0, 1000, 2050)]
    for i, (m, k, n) in enumerate([(2048, 300, 2048), (2500, 257, 1700)]):
        ops.append((torch.from_numpy(O.uniform((m, k), 40 + i).astype(np.float32)).cuda(), k, n))
    items, want = [], []
    for i, (A, k, n) in enumerate(ops):
        B = torch.from_numpy(O.uniform((k, n), 50 + i).astype(np.float32)).cuda()
        C0 = torch.from_numpy(O.uniform((A.shape[0], n), 60 + i).astype(np.float32)).cuda()
        items.append((A, B, C0.clone()))
        Cs = C0.clone()
        phpc.sgemm_device(A, B, Cs, method=phpc.SGEMM_BF16X3)
        want.append(Cs)
    plan = _x3_plan(items, cfg)
    plan.launch()
    torch.cuda.synchronize()
    for i, ((_, _, C), w) in enumerate(zip(items, want)):
        assert torch.equal(C.view(torch.int32), w.view(torch.int32)), i
    plan.close()


SMALL_ITEMS = [(1, 1, 1), (64, 64, 64), (300, 200, 150), (1000, 17, 999), (7, 513, 65), (129, 130, 4100),
               (256, 256, 64), (33, 1, 700)]


@pytest.mark.parametrize("cfg", [4, 6])
def test_bf16x3_plan_slice_planes_small_items(cfg):
    """Small and ragged items (1 x 1 x 1 up to k = 4100) on slice planes: exact on
    small-integer inputs (every product and partial sum an integer below 2^24), within the
    FP32 bound on uniform ones, and hipGraph replays (eager work between them) equal to the
    eager launch bit for bit."""
    import torch
    rng = np.random.default_rng(8)
    host, items = [], []
    for i, (m, n, k) in enumerate(SMALL_ITEMS):
        A = rng.integers(-8, 9, (m, k)).astype(np.float32)
        B = rng.integers(-8, 9, (k, n)).astype(np.float32)
        items.append((torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(),
                      torch.zeros((m, n), dtype=torch.float32, device="cuda")))
        host.append((A, B))
    plan = _x3_plan(items, cfg)
    plan.launch()
    torch.cuda.synchronize()
    for (A, B), (_, _, C) in zip(host, items):
        want = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float32)
        assert np.array_equal(C.cpu().numpy(), want), A.shape
    plan.close()

    host, items = [], []
    for i, (m, n, k) in enumerate(SMALL_ITEMS):
        A = O.uniform((m, k), 70 + i).astype(np.float32)
        B = O.uniform((k, n), 80 + i).astype(np.float32)
        C0 = O.uniform((m, n), 90 + i).astype(np.float32)
        items.append((torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), torch.from_numpy(C0.copy()).cuda()))
        host.append((A, B, C0))
    s = torch.cuda.Stream()
    plan = _x3_plan(items, cfg, stream=s)
    eager = None
    for rep in range(4):
        for (_, _, C0), (_, _, C) in zip(host, items):
            C.copy_(torch.from_numpy(C0))
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            if rep == 0:
                plan.launch(s)
            else:
                plan.graph_launch(s)
        s.synchronize()
        _ = torch.ones(1024, device="cuda").sum().item()  # eager work between replays
        got = [it[2].clone() for it in items]
        if eager is None:
            eager = got
            for (A, B, C0), C in zip(host, got):
                ref = C0.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
                err = np.abs(C.cpu().numpy().astype(np.float64) - ref)
                assert np.all(err <= bound(A, B, C0)), A.shape
        else:
            for i, (g, e) in enumerate(zip(got, eager)):
                assert torch.equal(g.view(torch.int32), e.view(torch.int32)), (rep, i)
    plan.close()



def test_bf16x3_plan_falls_back_when_planes_do_not_fit():
    """A k whose slice-plane rows would pass the level kernel's 32-bit offsets (256 rows x
    3 planes x 2 bytes x kp >= 2^31) makes a plan pinned to slice planes fall back to the
    in-register 3 x BF16 kernel (configuration 2), with a result within the bound."""
    import torch
    k = 1400000  # kp = k; 256 x 6 x kp > 2^31
    A = O.uniform((1, k), 5).astype(np.float32)
    B = O.uniform((k, 3), 6).astype(np.float32)
    C = torch.zeros((1, 3), dtype=torch.float32, device="cuda")
    phpc.set_tuning(8, 6)
    try:
        plan = phpc.Plan(phpc.F32X3, [(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), C)])
    finally:
        phpc.set_tuning(8, -1)
    assert plan.info()["config"] == 2
    plan.launch()
    torch.cuda.synchronize()
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert np.all(np.abs(C.cpu().numpy() - ref) <= bound(A, B, np.zeros_like(ref)))
    plan.close()


def test_bf16x3_plan_slice_planes_strided_operands():
    """Slice-plane plans on operand views: A and B as column windows of wider matrices
    (lda > k with rows not 16-byte aligned, so the cut takes its 4-byte path; ldb > n) and
    C a window too. Each item equals the single-product slice-plane path on the same views
    bit for bit, and nothing outside the C windows changes."""
    import torch
    specs = [(2100, 2300, 1000, 3, 5), (2048, 2049, 300, 1, 7)]
    items, want, guards = [], [], []
    for i, (m, n, k, pa, pb) in enumerate(specs):
        Aw = torch.from_numpy(O.uniform((m, k + pa), 100 + i).astype(np.float32)).cuda()
        Bw = torch.from_numpy(O.uniform((k, n + pb), 110 + i).astype(np.float32)).cuda()
        Cw = torch.from_numpy(O.uniform((m, n + 2), 120 + i).astype(np.float32)).cuda()
        A, B = Aw[:, pa:], Bw[:, :n]
        Cs = Cw.clone()
        phpc.sgemm_device(A, B, Cs[:, 1:n + 1], method=phpc.SGEMM_BF16X3)
        want.append(Cs)
        items.append((A, B, Cw[:, 1:n + 1]))
        guards.append(Cw)
    plan = _x3_plan(items, 6)
    plan.launch()
    torch.cuda.synchronize()
    for i, (Cw, w) in enumerate(zip(guards, want)):
        assert torch.equal(Cw.view(torch.int32), w.view(torch.int32)), i
    plan.close()
