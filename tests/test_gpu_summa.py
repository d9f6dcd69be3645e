"""GPU parity of the emulated SUMMA (src/phpc_summa.c:24-122) against the oracle's
restatement, bit-exact, plus full-size (N = 16384, 4 x 4 grid, panel 256) properties."""
import ctypes

import numpy as np
import pytest

from conftest import phpc
import oracle as O

pytestmark = pytest.mark.gpu


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


CASES = [(1, 1, 256, 0), (2, 2, 256, 0), (4, 4, 512, 0), (4, 2, 512, 0), (3, 3, 301, 0),
         (4, 4, 512, 32), (2, 3, 257, 16), (3, 3, 300, 7)]


@pytest.mark.parametrize("p,q,N,panel", CASES)
def test_summa_device_modes_match_oracle(p, q, N, panel):
    import torch
    print(f"summa p={p} q={q} N={N} panel={panel}", flush=True)
    A, B, C0 = O.uniform((N, N), 1), O.uniform((N, N), 2), O.uniform((N, N), 3)
    want = O.summa_emul(A, B, C0.copy(), p, q, panel)
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    for mode in (phpc.SUMMA_FUSED, phpc.SUMMA_STEPWISE):
        dC = torch.from_numpy(C0.copy()).cuda()
        phpc.summa_device(dA, dB, dC, p, q, panel=panel, mode=mode)
        torch.cuda.synchronize()
        assert np.array_equal(bits(dC.cpu().numpy()), bits(want)), mode


@pytest.mark.parametrize("cfg", [0, 3, 12, 13, 14, 15, 16])
@pytest.mark.parametrize("p,q,N,panel", [(4, 4, 512, 32), (2, 3, 257, 16), (3, 3, 300, 7), (2, 2, 256, 0)])
def test_summa_fused_panel_configs_match_oracle(cfg, p, q, N, panel):
    """FUSED on each panel-grouped tile configuration, the transposed-A one (12) included,
    is bit-exact against the oracle's SUMMA emulation."""
    import torch
    print(f"fused cfg={cfg} p={p} q={q} N={N} panel={panel}", flush=True)
    A, B, C0 = O.uniform((N, N), 21), O.uniform((N, N), 22), O.uniform((N, N), 23)
    want = O.summa_emul(A, B, C0.copy(), p, q, panel)
    dC = torch.from_numpy(C0.copy()).cuda()
    phpc.set_launch_policy(0, cfg, 0)
    try:
        phpc.summa_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, p, q, panel=panel,
                          mode=phpc.SUMMA_FUSED)
        torch.cuda.synchronize()
    finally:
        phpc.set_launch_policy(0, -1, 0)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


@pytest.mark.parametrize("P,N", [(1, 128), (4, 256), (16, 512), (6, 300)])
def test_summa_cuda_host_entry_point(P, N):
    A, B, C0 = O.uniform((N, N), 4), O.uniform((N, N), 5), O.uniform((N, N), 6)
    g = phpc.Grid(nranks=P)
    p, q = g.dims()
    C = C0.copy()
    t = phpc.summa_cuda(g, A, B, C)
    assert phpc.last_error()[0] == 0 and t > 0
    assert np.array_equal(bits(C), bits(O.summa_emul(A, B, C0.copy(), p, q, 0)))


def test_generic_loop_with_gemm_cuda_equals_device_path(lib):
    N = 128
    A, B, C0 = O.uniform((N, N), 7), O.uniform((N, N), 8), O.uniform((N, N), 9)
    g = phpc.Grid(dims=(2, 4), panel=16)
    f = phpc.GEMM_T(("phpc_gemm_cuda", lib))
    C = C0.copy()
    t = ctypes.c_float(0)
    assert lib.phpc_gemm_summa(f, g.h, phpc._ptr(A), phpc._ptr(B), phpc._ptr(C), N, 1, 4, 4, 16,
                               ctypes.byref(t)) == 0
    assert np.array_equal(bits(C), bits(O.summa_emul(A, B, C0.copy(), 2, 4, 16)))
    C2 = C0.copy()
    phpc.summa_cuda(g, A, B, C2)
    assert np.array_equal(bits(C2), bits(C))


def test_summa_ramp_closed_form():
    N = 1024
    A = O.ramp((N, N))
    g = phpc.Grid(nranks=16, panel=0)
    C = np.zeros((N, N))
    phpc.summa_cuda(g, A, A, C)
    assert np.array_equal(C, O.ramp_closed_matrix(N))


def test_summa_wide_exponent_bitwise():
    import torch
    N = 512
    A, B = O.wide((N, N), 31), O.wide((N, N), 32)
    want = O.summa_emul(A, B, np.zeros((N, N)), 4, 4, 64)
    dC = torch.zeros((N, N), dtype=torch.float64, device="cuda")
    phpc.summa_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, 4, 4, panel=64)
    torch.cuda.synchronize()
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


def test_summa_cublas_within_bound():
    N = 256
    A, B = O.uniform((N, N), 1), O.uniform((N, N), 2)
    g = phpc.Grid(nranks=4)
    C = np.zeros((N, N))
    t = phpc.summa_cublas(g, A, B, C)
    assert phpc.last_error()[0] == 0 and t == 0.0
    cw, _ = O.err_full(C, O.gemm_dd(A, B), A, B)
    assert cw <= 1e-14


@pytest.mark.parametrize("dist", [phpc.DIST_UNIFORM, phpc.DIST_WIDE])
def test_summa_full_size_properties(dist):
    """N = 16384 on the emulated 4 x 4 grid with panel 256: the fused and stepwise
    schedules agree bit for bit over the whole matrix, and sampled entries are within
    the FP64 bound of the double-double value."""
    import torch
    N = 16384
    print(f"summa full size dist={dist}", flush=True)
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, dist, 1)
    phpc.fill_device(dB, dist, 2)
    dC1 = torch.zeros_like(dA)
    phpc.summa_device(dA, dB, dC1, 4, 4, panel=256, mode=phpc.SUMMA_FUSED)
    dC2 = torch.zeros_like(dA)
    phpc.summa_device(dA, dB, dC2, 4, 4, panel=256, mode=phpc.SUMMA_STEPWISE)
    torch.cuda.synchronize()
    assert torch.equal(dC1.view(torch.int64), dC2.view(torch.int64))
    del dC2
    rng = np.random.default_rng(3)
    rows = np.concatenate([rng.integers(0, N, 2048), np.arange(0, N, 64)])
    cols = np.concatenate([rng.integers(0, N, 2048), np.arange(N - 1, -1, -64)])
    cw, nw = phpc.check_sampled(dA, dB, dC1, rows, cols)
    print(f"  err_cw={cw:.3e} err_nw={nw:.3e}", flush=True)
    assert cw <= 1e-14


@pytest.mark.parametrize("p,q,N", [(4, 4, 512), (3, 3, 301), (2, 4, 258)])
def test_summa_chained_equals_single_panel_kernel(p, q, N):
    """CHAINED mode: one FMA chain per output over all of k, i.e. the reference kernel's
    arithmetic for one local-GEMM call over the full k."""
    import torch
    A, B, C0 = O.uniform((N, N), 11), O.uniform((N, N), 12), O.uniform((N, N), 13)
    dC = torch.from_numpy(C0.copy()).cuda()
    phpc.summa_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, p, q, panel=16,
                      mode=phpc.SUMMA_CHAINED)
    torch.cuda.synchronize()
    want = O.kernel_emul(A, B, C0.copy())
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


def test_large_single_gemm_uses_wide_tile_and_stays_exact():
    """N = 4096 selects the 256x128 tile: bit-exact against the kernel emulation on
    sampled rows (the emulation of full rows is cheap)."""
    import torch
    N = 4096
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, phpc.DIST_UNIFORM, 5)
    phpc.fill_device(dB, phpc.DIST_UNIFORM, 6)
    dC = torch.zeros_like(dA)
    phpc.dgemm_device(dA, dB, dC)
    torch.cuda.synchronize()
    A = O.uniform((N, N), 5)
    B = O.uniform((N, N), 6)
    rows = [0, 1, 255, 256, 2049, 4095]
    got = dC[rows].cpu().numpy()
    want = O.kernel_emul(np.ascontiguousarray(A[rows]), B, np.zeros((len(rows), N)))
    assert np.array_equal(bits(got), bits(want))


def test_pack_device_byte_exact():
    """The strided -> packed block copy (MPI_Type_vector panel, src/phpc_summa.c:53-59)
    moves bytes unchanged: NaN payloads, signed zeros, infinities, subnormals."""
    import torch
    rng = np.random.default_rng(9)
    src = rng.integers(0, 2 ** 63, size=(70, 90), dtype=np.int64).view(np.float64).copy()
    specials = np.array([np.nan, -0.0, np.inf, -np.inf, 5e-324, -5e-324], dtype=np.float64)
    src.reshape(-1)[:6] = specials
    src.view(np.uint64).reshape(-1)[6] = 0x7FF0000000000123  # signalling NaN payload
    dsrc = torch.from_numpy(src).cuda()
    for (r0, c0, rows, cols) in [(0, 0, 70, 90), (3, 5, 40, 33), (69, 89, 1, 1), (10, 0, 0, 5)]:
        dst = torch.full((max(rows, 1), cols + 3), -1.0, dtype=torch.float64, device="cuda")
        sub = dsrc[r0:, c0:]
        phpc.check(phpc.lib().phpc_pack_device(phpc.dptr(sub), 90, phpc.dptr(dst), cols + 3, rows, cols,
                                               phpc.stream_ptr()))
        torch.cuda.synchronize()
        got = dst.cpu().numpy().view(np.uint64)
        assert np.array_equal(got[:rows, :cols], src[r0:r0 + rows, c0:c0 + cols].view(np.uint64))
        assert np.all(dst.cpu().numpy()[:, cols:] == -1.0)  # padding columns untouched


@pytest.mark.parametrize("p,q,N", [(2, 2, 256), (3, 3, 301), (1, 1, 200)])
def test_summa_split_mode_fp64_accurate(p, q, N):
    import torch
    A, B = O.uniform((N, N), 31), O.uniform((N, N), 32)
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dC = torch.zeros((N, N), dtype=torch.float64, device="cuda")
    phpc.summa_device(dA, dB, dC, p, q, mode=phpc.SUMMA_SPLIT)
    torch.cuda.synchronize()
    cw, _ = O.err_full(dC.cpu().numpy(), O.gemm_dd(A, B), A, B)
    assert cw <= 1e-14


@pytest.mark.parametrize("mode", [phpc.SUMMA_FUSED, phpc.SUMMA_CHAINED, phpc.SUMMA_STEPWISE])
def test_summa_host_bands_match_oracle(mode):
    """Host-pointer SUMMA with the copy/compute pipeline: N = 1000 runs as four row bands
    (the last one partial) streamed through pinned staging; the bits equal the oracle's."""
    N, p, q, panel = 1000, 3, 3, 0
    A, B, C0 = O.uniform((N, N), 41), O.uniform((N, N), 42), O.uniform((N, N), 43)
    C = C0.copy()
    t = ctypes.c_float(0)
    phpc.check(phpc.lib().phpc_summa_host(p, q, panel, mode, A.ctypes.data_as(ctypes.c_void_p),
                                          B.ctypes.data_as(ctypes.c_void_p), C.ctypes.data_as(ctypes.c_void_p),
                                          N, 0, ctypes.byref(t)))
    assert t.value > 0
    if mode == phpc.SUMMA_CHAINED:
        want = O.kernel_emul(A, B, C0.copy())
    else:
        want = O.summa_emul(A, B, C0.copy(), p, q, panel)
    assert np.array_equal(bits(C), bits(want))


def test_gemm_cuda_staged_copies_span_chunks():
    """phpc_gemm_cuda with padded leading dimensions and operands larger than one pinned
    staging slot equals the device-pointer path bit for bit."""
    import torch
    m, k, n = 2304, 2048, 2560
    lda, ldb, ldc = k + 3, n + 5, n + 1
    rng = np.random.default_rng(5)
    Ab = rng.uniform(-1, 1, (m, lda))
    Bb = rng.uniform(-1, 1, (k, ldb))
    Cb = rng.uniform(-1, 1, (m, ldc))
    Ch = Cb.copy()
    phpc.gemm_cuda(Ab, Bb, Ch, m=m, k=k, n=n, lda=lda, ldb=ldb, ldc=ldc)
    assert phpc.last_error()[0] == 0
    dA = torch.from_numpy(np.ascontiguousarray(Ab[:, :k])).cuda()
    dB = torch.from_numpy(np.ascontiguousarray(Bb[:, :n])).cuda()
    dC = torch.from_numpy(np.ascontiguousarray(Cb[:, :n])).cuda()
    phpc.dgemm_device(dA, dB, dC)
    torch.cuda.synchronize()
    assert np.array_equal(bits(Ch[:, :n]), bits(dC.cpu().numpy()))
    assert np.array_equal(bits(Ch[:, n:]), bits(Cb[:, n:]))  # padding columns untouched


def test_chained_transposed_a_matches_register_staged_full_matrix():
    """N = 8192, 4 x 4 grid, CHAINED: the transposed-A kernel (default for large problems)
    and the register-staged 256x128 kernel (configuration 6) give the same C bit for bit
    over the whole matrix, on uniform and wide-exponent inputs."""
    import torch
    N = 8192
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    for dist in (phpc.DIST_UNIFORM, phpc.DIST_WIDE):
        print(f"chained at vs staged dist={dist}", flush=True)
        phpc.fill_device(dA, dist, 31)
        phpc.fill_device(dB, dist, 32)
        dC1 = torch.zeros_like(dA)
        phpc.summa_device(dA, dB, dC1, 4, 4, panel=256, mode=phpc.SUMMA_CHAINED)
        dC2 = torch.zeros_like(dA)
        phpc.set_launch_policy(0, 6, 0)
        try:
            phpc.summa_device(dA, dB, dC2, 4, 4, panel=256, mode=phpc.SUMMA_CHAINED)
            torch.cuda.synchronize()
        finally:
            phpc.set_launch_policy(0, -1, 0)
        assert torch.equal(dC1.view(torch.int64), dC2.view(torch.int64)), dist
        del dC1, dC2


@pytest.mark.parametrize("mode", ["fused", "stepwise", "chained"])
def test_summa_special_values_match_oracle(mode):
    """IEEE special cases through the SUMMA schedules on an uneven 3 x 3 grid: subnormal
    sums, signed zeros, infinities and NaNs come out as in the oracle's restatement (NaN
    positions compared, payloads not)."""
    import torch
    from test_gpu_gemm import special_inputs, same_bits_or_both_nan
    N, p, q, panel = 100, 3, 3, 11
    A, B, C0 = special_inputs(N, N, N, 71)
    if mode == "chained":
        want = O.kernel_emul(A, B, C0.copy())
    else:
        want = O.summa_emul(A, B, C0.copy(), p, q, panel)
    dC = torch.from_numpy(C0.copy()).cuda()
    m = {"fused": phpc.SUMMA_FUSED, "stepwise": phpc.SUMMA_STEPWISE, "chained": phpc.SUMMA_CHAINED}[mode]
    phpc.summa_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, p, q, panel=panel, mode=m)
    torch.cuda.synchronize()
    assert same_bits_or_both_nan(dC.cpu().numpy(), want), mode
