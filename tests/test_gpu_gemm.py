"""GPU parity of the local GEMM (phpc_gemm_cuda and the device API) against the
oracle's restatement of the reference kernel arithmetic (src/phpc_gemm.cu:25-55).
Bit-exact: the FP64 MFMA path runs the same FMA chain per output."""
import ctypes

import numpy as np
import pytest

from conftest import phpc
import oracle as O

pytestmark = pytest.mark.gpu


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


@pytest.fixture(autouse=True)
def default_policy():
    phpc.set_launch_policy(0, -1, 0)
    yield
    phpc.set_launch_policy(0, -1, 0)


SHAPES = [(1, 1, 1), (7, 5, 3), (16, 4, 16), (64, 64, 64), (128, 128, 128), (129, 65, 257),
          (300, 17, 450), (256, 1024, 256), (1000, 333, 77), (513, 1031, 515)]


@pytest.mark.parametrize("m,k,n", SHAPES)
def test_gemm_cuda_matches_reference_kernel_bitwise(m, k, n):
    print(f"gemm_cuda m={m} k={k} n={n}", flush=True)
    A, B, C0 = O.uniform((m, k), 1), O.uniform((k, n), 2), O.uniform((m, n), 3)
    C = C0.copy()
    t = phpc.gemm_cuda(A, B, C)
    assert phpc.last_error()[0] == 0
    want = O.kernel_emul(A, B, C0.copy())
    assert np.array_equal(bits(C), bits(want))
    assert t > 0


def test_gemm_cuda_n1024_uniform_and_error_bound():
    n = 1024
    A, B = O.uniform((n, n), 11), O.uniform((n, n), 12)
    C = np.zeros((n, n))
    phpc.gemm_cuda(A, B, C)
    assert np.array_equal(bits(C), bits(O.kernel_emul(A, B, np.zeros((n, n)))))
    cw, nw = O.err_full(C, O.gemm_dd(A, B), A, B)
    assert cw <= 1e-14, cw


def test_gemm_cuda_ramp_closed_form():
    n = 1024
    A = O.ramp((n, n))
    C = np.zeros((n, n))
    phpc.gemm_cuda(A, A, C)
    assert np.array_equal(C, O.ramp_closed_matrix(n))


def test_gemm_cuda_wide_exponent_bitwise():
    n = 300
    A, B = O.wide((n, n), 21), O.wide((n, n), 22)
    C = np.zeros((n, n))
    phpc.gemm_cuda(A, B, C)
    assert np.array_equal(bits(C), bits(O.kernel_emul(A, B, np.zeros((n, n)))))


def test_gemm_cuda_leading_dimensions():
    A = O.uniform((70, 90), 1)
    B = O.uniform((80, 101), 2)
    C = O.uniform((65, 99), 3)
    want = C.copy()
    O.kernel_emul(A, B, want, m=60, k=75, n=95, lda=90, ldb=101, ldc=99)
    phpc.gemm_cuda(A, B, C, m=60, k=75, n=95, lda=90, ldb=101, ldc=99)
    assert np.array_equal(bits(C), bits(want))


def test_gemm_cuda_odd_leading_dimension_scalar_path():
    A = O.uniform((33, 31), 4)
    B = O.uniform((31, 29), 5)
    C = np.zeros((33, 29))
    phpc.gemm_cuda(A, B, C)
    assert np.array_equal(bits(C), bits(O.kernel_emul(A, B, np.zeros((33, 29)))))


def test_results_invariant_to_grid_and_tile_config():
    m, k, n = 300, 200, 260
    A, B, C0 = O.uniform((m, k), 1), O.uniform((k, n), 2), O.uniform((m, n), 3)
    want = O.kernel_emul(A, B, C0.copy())
    for cfg in range(phpc.num_configs()):
        for honor, (gw, gh) in [(0, (8, 8)), (1, (1, 1)), (1, (3, 5)), (1, (16, 16))]:
            phpc.set_launch_policy(honor, cfg, 0)
            C = C0.copy()
            phpc.gemm_cuda(A, B, C, grid_width=gw, grid_height=gh, block_width=16)
            assert np.array_equal(bits(C), bits(want)), (cfg, honor, gw, gh)


def test_accumulate_semantics():
    A, B, C0 = O.uniform((90, 70), 1), O.uniform((70, 50), 2), O.uniform((90, 50), 3)
    C1 = C0.copy()
    phpc.gemm_cuda(A, B, C1)
    Z = np.zeros((90, 50))
    phpc.gemm_cuda(A, B, Z)
    assert np.array_equal(bits(C1), bits(C0 + Z))


@pytest.mark.parametrize("bw,gw,gh", [(33, 8, 8), (0, 8, 8), (32, 0, 4), (32, 4, 0)])
def test_invalid_launch_leaves_c_unchanged(bw, gw, gh):
    A, B, C0 = O.uniform((8, 8), 1), O.uniform((8, 8), 2), O.uniform((8, 8), 3)
    C = C0.copy()
    phpc.gemm_cuda(A, B, C, block_width=bw, grid_width=gw, grid_height=gh)
    assert phpc.last_error()[0] == phpc.PHPC_ERR_ARG
    assert np.array_equal(bits(C), bits(C0))


def test_zero_k_leaves_c_unchanged():
    A = np.zeros((4, 1))
    B = np.zeros((1, 5))
    C0 = O.uniform((4, 5), 3)
    C = C0.copy()
    phpc.gemm_cuda(A, B, C, k=0)
    assert phpc.last_error()[0] == 0
    assert np.array_equal(bits(C), bits(C0))


def test_vendor_comparator_within_bound():
    n = 512
    A, B = O.uniform((n, n), 1), O.uniform((n, n), 2)
    C = np.zeros((n, n))
    t = phpc.gemm_cublas(A, B, C)
    assert phpc.last_error()[0] == 0 and t == 0.0
    cw, _ = O.err_full(C, O.gemm_dd(A, B), A, B)
    assert cw <= 1e-14


# ---------------- device-resident API ----------------

def test_fill_device_matches_oracle_generators():
    import torch
    for dist, gen in [(phpc.DIST_UNIFORM, O.uniform), (phpc.DIST_WIDE, O.wide)]:
        x = torch.empty((37, 53), dtype=torch.float64, device="cuda")
        phpc.fill_device(x, dist, 77)
        torch.cuda.synchronize()
        assert np.array_equal(bits(x.cpu().numpy()), bits(gen((37, 53), 77)))
    x = torch.empty((5, 7), dtype=torch.float64, device="cuda")
    phpc.fill_device(x, phpc.DIST_RAMP, 0)
    torch.cuda.synchronize()
    assert np.array_equal(x.cpu().numpy(), O.ramp((5, 7)))


@pytest.mark.parametrize("panel", [1, 7, 16, 64, 256])
def test_dgemm_device_panels_match_summa_grouping(panel):
    import torch
    n = 300
    A, B, C0 = O.uniform((n, n), 1), O.uniform((n, n), 2), O.uniform((n, n), 3)
    dA, dB, dC = (torch.from_numpy(x).cuda() for x in (A, B, C0))
    phpc.dgemm_device(dA, dB, dC, panel=panel)
    torch.cuda.synchronize()
    want = O.summa_emul(A, B, C0.copy(), 1, 1, panel)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


def test_check_sampled_matches_oracle():
    import torch
    m, k, n = 200, 700, 150
    A, B, C0 = O.uniform((m, k), 1), O.uniform((k, n), 2), O.uniform((m, n), 3)
    C = O.kernel_emul(A, B, C0.copy())
    rng = np.random.default_rng(0)
    rows = rng.integers(0, m, 500)
    cols = rng.integers(0, n, 500)
    dA, dB, dC0, dC = (torch.from_numpy(x).cuda() for x in (A, B, C0, C))
    cw, nw = phpc.check_sampled(dA, dB, dC, rows, cols, C0=dC0)
    ref = O.gemm_dd_entries(A, B, rows, cols, C0)
    ab = np.array([O.lib().orc_absdot(O._p(A), k, O._p(B), n, int(i), int(j), k) for i, j in zip(rows, cols)])
    d = np.abs(C[rows, cols] - ref)
    want = np.max(np.where(d == 0, 0.0, d / (ab + np.abs(C0[rows, cols]))))
    assert cw == pytest.approx(want, rel=1e-6, abs=1e-30)
    # a planted error is found
    C2 = C.copy()
    C2[rows[3], cols[3]] += 1e-6
    cw2, _ = phpc.check_sampled(dA, dB, torch.from_numpy(C2).cuda(), rows, cols, C0=dC0)
    assert cw2 > 1e-9


@pytest.mark.parametrize("cfg", [10, 11, 12, 13, 14, 15, 16])
@pytest.mark.parametrize("m,k,n", [(257, 131, 199), (300, 200, 260), (33, 17, 1), (1, 64, 3)])
def test_transposed_a_configs_on_ragged_shapes(cfg, m, k, n):
    """The transposed-A configurations (A staged transposed; B staged to an even pitch when
    its rows are not 16-byte aligned) are bit-exact on odd and tiny shapes, through the
    device entry point on strided views whose B rows start at odd column offsets."""
    import torch
    A, B, C0 = O.uniform((m, k), 41), O.uniform((k, n), 42), O.uniform((m, n), 43)
    want = O.kernel_emul(A, B, C0.copy())
    # B as a view with an odd column offset and an odd row stride
    Bbig = torch.zeros((k, n + 3), dtype=torch.float64, device="cuda")
    Bbig[:, 1:1 + n] = torch.from_numpy(B).cuda()
    dB = Bbig[:, 1:1 + n]
    dA = torch.from_numpy(A).cuda()
    dC = torch.from_numpy(C0.copy()).cuda()
    phpc.set_launch_policy(0, cfg, 0)
    try:
        phpc.dgemm_device(dA, dB, dC)
        torch.cuda.synchronize()
    finally:
        phpc.set_launch_policy(0, -1, 0)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


def special_inputs(m, k, n, seed):
    """Uniform operands seeded with IEEE special cases: rows and columns of tiny values whose
    products and partial sums are subnormal, subnormal entries, signed zeros, infinities and
    NaNs in A, B and C0."""
    A, B, C0 = O.uniform((m, k), seed), O.uniform((k, n), seed + 1), O.uniform((m, n), seed + 2)
    A[0, :] *= 1e-160
    B[:, 0] *= 1e-160
    C0[0, 0] = 0.0               # C[0, 0]: sums of 1e-320 products, subnormal throughout
    A[1, :] = 5e-324 * np.arange(1, k + 1)  # subnormal entries; FMAs round in the subnormal range
    C0[1, :] = 0.0
    A[2, :] = -0.0
    C0[2, :] = -0.0              # -0 + (+0 from a chain of -0 products) = +0
    B[:, 3] = -0.0
    A[4, 5] = np.inf
    A[4, 6] = -np.inf
    B[5, 8] = B[6, 8] = 1.0      # C[4, 8]: inf - inf = NaN; the rest of row 4 is +-inf or NaN
    B[7, 9] = np.nan
    C0[6, 10] = np.inf
    C0[7, 11] = np.nan
    C0[8, 12] = 4.9e-324
    return A, B, C0


def same_bits_or_both_nan(got, want):
    gn, wn = np.isnan(got), np.isnan(want)
    return np.array_equal(gn, wn) and np.array_equal(bits(np.where(gn, 0.0, got)), bits(np.where(wn, 0.0, want)))


@pytest.mark.parametrize("cfg", [-1] + list(range(17)))
def test_special_values_match_reference_kernel(cfg):
    """IEEE special cases through every tile configuration: subnormal products and sums are
    kept (no flush to zero), signed zeros, infinities and NaNs propagate as in the reference
    kernel's FMA chain. NaN payloads are not compared."""
    import torch
    m, k, n = 70, 45, 66
    A, B, C0 = special_inputs(m, k, n, 61)
    want = O.kernel_emul(A, B, C0.copy())
    assert np.isnan(want[4, 8]) and np.isnan(want[7, 11]) and 0 < abs(want[0, 0]) < 2.3e-308
    assert np.count_nonzero((np.abs(want[1]) < 2.3e-308) & (want[1] != 0)) > n // 2
    assert not np.signbit(want[2, 0])
    if cfg >= phpc.num_configs():
        pytest.skip("no such configuration")
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dC = torch.from_numpy(C0.copy()).cuda()
    phpc.set_launch_policy(0, cfg, 0)
    try:
        phpc.dgemm_device(dA, dB, dC)
        torch.cuda.synchronize()
    finally:
        phpc.set_launch_policy(0, -1, 0)
    assert same_bits_or_both_nan(dC.cpu().numpy(), want), cfg
    C = C0.copy()
    phpc.gemm_cuda(A, B, C, grid_width=4, grid_height=4, block_width=16)
    assert same_bits_or_both_nan(C, want)


@pytest.mark.parametrize("cfg", [0, 12, 13, 14, 15, 16])
@pytest.mark.parametrize("m,k,n,panel", [(300, 1000, 260, 333), (129, 70, 257, 16), (513, 2050, 130, 1024),
                                         (64, 64, 64, 63), (260, 700, 130, 40)])
def test_panel_grouped_configs_on_ragged_panels(cfg, m, k, n, panel):
    """Panel-grouped work through the device entry point, register-staged (0) and
    transposed-A (12; 13-15 with three or four 16-deep stages or two 32-deep ones, whose
    slabs stop at panel ends that do not fall on a slab boundary, e.g. panels of 40 and
    333): C = ((C0 + s_0) + s_1) + ..., each s_t an FMA chain from zero over
    one panel, bit-exact against per-panel calls of the reference kernel's emulation, on
    ragged shapes and panels that do not divide k."""
    import torch
    A, B, C0 = O.uniform((m, k), 81), O.uniform((k, n), 82), O.uniform((m, n), 83)
    want = C0.copy()
    for k0 in range(0, k, panel):
        k1 = min(k, k0 + panel)
        s = O.kernel_emul(np.ascontiguousarray(A[:, k0:k1]), np.ascontiguousarray(B[k0:k1]), np.zeros((m, n)))
        want = want + s
    dC = torch.from_numpy(C0.copy()).cuda()
    phpc.set_launch_policy(0, cfg, 0)
    try:
        phpc.dgemm_device(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), dC, panel=panel)
        torch.cuda.synchronize()
    finally:
        phpc.set_launch_policy(0, -1, 0)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


def test_kernel_timing_counts_gemm_launches():
    """phpc_kernel_timing brackets each FP64 tile launch with events on its stream: one
    device GEMM is one timed launch of positive duration; nothing is recorded while off."""
    import torch
    N = 1024
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, phpc.DIST_UNIFORM, 1)
    phpc.fill_device(dB, phpc.DIST_UNIFORM, 2)
    dC = torch.zeros_like(dA)
    phpc.kernel_timing(True)
    phpc.dgemm_device(dA, dB, dC)
    ms, n = phpc.kernel_timing_read()
    assert n == 1 and ms > 0.0
    phpc.kernel_timing(False)
    phpc.dgemm_device(dA, dB, dC)
    torch.cuda.synchronize()
    assert phpc.kernel_timing_read() == (0.0, 0)
