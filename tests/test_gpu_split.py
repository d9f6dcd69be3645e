"""GPU parity of the split-precision (int8-slice) FP64 GEMM: bit-exact against the
oracle's restatement of the algorithm at small sizes, FP64-level accuracy against the
double-double product, and full-size sampled accuracy."""
import numpy as np
import pytest

from conftest import phpc
import oracle as O

pytestmark = pytest.mark.gpu


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("m,k,n,s,gen", [(1, 1, 1, 1, "u"), (5, 7, 3, 3, "u"), (64, 64, 64, 7, "u"),
                                         (200, 300, 130, 7, "u"), (257, 129, 255, 4, "w"),
                                         (96, 1000, 80, 8, "w"), (33, 65, 300, 2, "u")])
def test_split_matches_restatement_bitwise(m, k, n, s, gen):
    import torch
    print(f"split m={m} k={k} n={n} s={s} {gen}", flush=True)
    g = O.uniform if gen == "u" else O.wide
    A, B, C0 = g((m, k), 1), g((k, n), 2), O.uniform((m, n), 3)
    dC = dev(C0)
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=s, method=phpc.SPLIT_SLICES)
    torch.cuda.synchronize()
    assert info.slices == s and info.products == s * (s + 1) // 2
    want = O.split_gemm(A, B, C0.copy(), s)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


def test_split_k_pieces_match_restatement():
    """k = 20000 with 8 slices: level 9 spans 8 padded slices (160256 products), more
    than one int32 accumulation can hold, so it runs in two pieces."""
    import torch
    m, k, n, s = 16, 20000, 24, 8
    A, B = O.uniform((m, k), 4), O.uniform((k, n), 5)
    dC = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=s, method=phpc.SPLIT_SLICES)
    torch.cuda.synchronize()
    assert info.launches > s + 6  # at least one level was split
    want = O.split_gemm(A, B, np.zeros((m, n)), s)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


@pytest.mark.parametrize("m,k,n,s,gen", [(1, 1, 1, 1, "u"), (5, 7, 3, 3, "u"), (64, 64, 64, 6, "u"),
                                         (200, 300, 130, 12, "u"), (257, 129, 255, 9, "w"),
                                         (96, 1000, 80, 14, "w"), (600, 333, 520, 13, "u")])
def test_fp8_slices_match_restatement_bitwise(m, k, n, s, gen):
    """Scheme I on FP8-E4M3 slices (4-bit digits, FP32 sums on the FP8 matrix cores) gives
    the restatement's bits: the FP32 sums are exact integers, so they equal the oracle's
    integer level sums, and the FP64 recombination runs in the same order."""
    import torch
    print(f"fp8 m={m} k={k} n={n} s={s} {gen}", flush=True)
    g = O.uniform if gen == "u" else O.wide
    A, B, C0 = g((m, k), 1), g((k, n), 2), O.uniform((m, n), 3)
    dC = dev(C0)
    try:
        info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=s, method=phpc.SPLIT_FP8)
        torch.cuda.synchronize()
    finally:
        phpc.lib().phpc_split_set_method(phpc.SPLIT_CRT)
    assert info.slices == s and info.products == s * (s + 1) // 2 and info.int_bits == 4 * s and info.method == 2
    want = O.split_gemm(A, B, C0.copy(), s, bits=4)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


def test_fp8_large_problem_kernel_matches_small_one():
    """3000 x 3000 outputs (144 tiles of 256 x 256) take the large-problem FP8 kernel; the
    same product in row blocks of 512 takes the small one. Every output's sums and their
    FP64 order are the same, so the bits are too."""
    import torch
    m, k, n, s = 3000, 200, 3000, 5
    dA = torch.empty((m, k), dtype=torch.float64, device="cuda")
    dB = torch.empty((k, n), dtype=torch.float64, device="cuda")
    phpc.fill_device(dA, phpc.DIST_WIDE, 71)
    phpc.fill_device(dB, phpc.DIST_UNIFORM, 72)
    whole = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    blocks = torch.zeros_like(whole)
    try:
        phpc.dgemm_split_device(dA, dB, whole, slices=s, method=phpc.SPLIT_FP8)
        for r0 in range(0, m, 512):
            r1 = min(m, r0 + 512)
            phpc.dgemm_split_device(dA[r0:r1], dB, blocks[r0:r1], slices=s, method=phpc.SPLIT_FP8)
        torch.cuda.synchronize()
    finally:
        phpc.lib().phpc_split_set_method(phpc.SPLIT_CRT)
    assert torch.equal(whole.view(torch.int64), blocks.view(torch.int64))


def test_fp8_slices_k_pieces_match_restatement():
    """k = 20000 with 6 FP8 slices: level 7 spans 6 padded slices (120576 products), more
    than an exact FP32 sum of 4-bit digit products holds (74565), so it runs in pieces of
    three slices."""
    import torch
    m, k, n, s = 16, 20000, 24, 6
    A, B = O.uniform((m, k), 4), O.uniform((k, n), 5)
    dC = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    try:
        info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=s, method=phpc.SPLIT_FP8)
        torch.cuda.synchronize()
    finally:
        phpc.lib().phpc_split_set_method(phpc.SPLIT_CRT)
    assert info.launches > s + 6  # at least one level was split
    want = O.split_gemm(A, B, np.zeros((m, n)), s, bits=4)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


@pytest.mark.parametrize("gen", ["u", "w"])
def test_fp8_slices_auto_reach_fp64_accuracy(gen):
    """With the count left to the error estimate (4-bit digits), the FP8 slices reach the
    FP64 target against the double-double product on sampled entries."""
    import torch
    N = 2048
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    dist = phpc.DIST_UNIFORM if gen == "u" else phpc.DIST_WIDE
    phpc.fill_device(dA, dist, 61)
    phpc.fill_device(dB, dist, 62)
    dC = torch.zeros_like(dA)
    try:
        info = phpc.dgemm_split_device(dA, dB, dC, slices=0, target=1e-14, method=phpc.SPLIT_FP8)
        torch.cuda.synchronize()
    finally:
        phpc.lib().phpc_split_set_method(phpc.SPLIT_CRT)
    rng = np.random.default_rng(63)
    cw, _ = phpc.check_sampled(dA, dB, dC, rng.integers(0, N, 1024), rng.integers(0, N, 1024))
    print(f"  fp8 auto {gen}: slices={info.slices} estimate={info.error_estimate:.2e} err_cw={cw:.2e}", flush=True)
    assert info.error_estimate <= 1e-14 and cw <= 1e-14


def test_split_exact_on_small_integers():
    import torch
    rng = np.random.default_rng(9)
    A = rng.integers(-127, 128, (70, 90)).astype(np.float64)
    B = rng.integers(-127, 128, (90, 50)).astype(np.float64)
    dC = torch.zeros((70, 50), dtype=torch.float64, device="cuda")
    phpc.dgemm_split_device(dev(A), dev(B), dC, slices=1, method=phpc.SPLIT_SLICES)
    torch.cuda.synchronize()
    assert np.array_equal(dC.cpu().numpy(), A @ B)


@pytest.mark.parametrize("gen,method", [("u", 0), ("w", 0), ("u", 1), ("w", 1)])
def test_split_auto_slices_meet_fp64_target(gen, method):
    import torch
    N = 512
    g = O.uniform if gen == "u" else O.wide
    A, B = g((N, N), 11), g((N, N), 12)
    dC = torch.zeros((N, N), dtype=torch.float64, device="cuda")
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=0, target=1e-14, method=method)
    torch.cuda.synchronize()
    cw, _ = O.err_full(dC.cpu().numpy(), O.gemm_dd(A, B), A, B)
    print(f"  {gen}: slices={info.slices} estimate={info.error_estimate:.2e} err_cw={cw:.2e}", flush=True)
    assert cw <= 1e-14
    assert info.error_estimate <= 1e-14


@pytest.mark.parametrize("method", [0, 1])
def test_split_zero_rows_and_columns(method):
    """All-zero rows of A and columns of B (ratio 0, left out of the modulus bound) leave
    their rows and columns of C exactly at C0; the rest meets the FP64 target, including a
    row whose entries are all far below the others' scale."""
    import torch
    m, k, n = 200, 300, 150
    A, B, C0 = O.wide((m, k), 31), O.wide((k, n), 32), O.uniform((m, n), 33)
    A[[0, 7, 199], :] = 0.0
    B[:, [0, 5, 149]] = 0.0
    A[50, :] *= 2.0 ** -600
    dC = dev(C0.copy())
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=0, target=1e-14, method=method)
    torch.cuda.synchronize()
    got = dC.cpu().numpy()
    assert np.array_equal(got[[0, 7, 199], :], C0[[0, 7, 199], :])
    assert np.array_equal(got[:, [0, 5, 149]], C0[:, [0, 5, 149]])
    exact = C0 + O.gemm_dd(A, B)
    mag = np.abs(C0) + np.abs(A) @ np.abs(B)
    assert np.all(np.abs(got - exact) <= 1e-14 * mag), (info.slices, info.int_bits)


@pytest.mark.parametrize("side,bad", [("A", np.inf), ("A", np.nan), ("B", -np.inf), ("B", np.nan)])
@pytest.mark.parametrize("method", [phpc.SPLIT_CRT, phpc.SPLIT_SLICES])
def test_split_rejects_non_finite(side, bad, method):
    """An inf or NaN anywhere in A or B fails the call and leaves C alone (a NaN is checked
    per entry: the column maximum, an fmax, would drop it). 300 x 700 x 260: the bad entry
    of A sits at k = 517, the third of a thread's four loads in the row statistics; that of
    B in the third 256-row block of the column statistics."""
    import torch
    A = np.ones((300, 700))
    B = np.ones((700, 260))
    if side == "A":
        A[201, 517] = bad
    else:
        B[517, 201] = bad
    dC = torch.full((300, 260), 2.0, dtype=torch.float64, device="cuda")
    try:
        with pytest.raises(RuntimeError):
            phpc.dgemm_split_device(dev(A), dev(B), dC, slices=3, method=method)
    finally:
        phpc.lib().phpc_split_set_method(phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    assert torch.all(dC == 2.0)


@pytest.mark.parametrize("dist,method", [(phpc.DIST_UNIFORM, 0), (phpc.DIST_WIDE, 0), (phpc.DIST_UNIFORM, 1),
                                         (phpc.DIST_WIDE, 1)])
def test_split_large_sampled_accuracy(dist, method):
    import torch
    N = 4096
    dA = torch.empty((N, N), dtype=torch.float64, device="cuda")
    dB = torch.empty_like(dA)
    phpc.fill_device(dA, dist, 1)
    phpc.fill_device(dB, dist, 2)
    dC = torch.zeros_like(dA)
    info = phpc.dgemm_split_device(dA, dB, dC, slices=0, target=1e-14, method=method)
    torch.cuda.synchronize()
    rng = np.random.default_rng(4)
    rows, cols = rng.integers(0, N, 2048), rng.integers(0, N, 2048)
    cw, nw = phpc.check_sampled(dA, dB, dC, rows, cols)
    print(f"  dist={dist} slices={info.slices} err_cw={cw:.2e}", flush=True)
    assert cw <= 1e-14


@pytest.mark.parametrize("m,k,n,L,gen", [(1, 1, 1, 3, "u"), (7, 9, 5, 6, "u"), (64, 64, 64, 14, "u"),
                                         (200, 300, 130, 14, "w"), (257, 129, 255, 10, "u"),
                                         (96, 1000, 80, 16, "w"), (33, 65, 300, 4, "u"),
                                         (16, 64, 16, 20, "u"), (8, 3, 8, 20, "w"), (40, 17, 24, 18, "u")])
def test_crt_matches_restatement_bitwise(m, k, n, L, gen):
    import torch
    print(f"crt m={m} k={k} n={n} L={L} {gen}", flush=True)
    g = O.uniform if gen == "u" else O.wide
    A, B, C0 = g((m, k), 5), g((k, n), 6), O.uniform((m, n), 7)
    dC = dev(C0)
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=L, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    assert info.method == 1 and info.slices == L
    # with many moduli and a short k the widest integers that fit would pass 2^72, the
    # residue routine's range; the width is capped at 71 bits there
    assert 1 <= info.int_bits <= 71
    want = O.crt_gemm(A, B, C0.copy(), info.int_bits, L)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


@pytest.mark.parametrize("gen", ["u", "w"])
def test_crt_k_pieces_match_restatement(gen):
    """k = 140000 is more than one int32 accumulation of residue products holds (130944),
    so every residue GEMM runs in two k pieces, the second adding the first's residue before
    it reduces mod m: the result equals the restatement bit for bit."""
    import torch
    m, k, n, L = 6, 140000, 5, 12
    g = O.uniform if gen == "u" else O.wide
    A, B, C0 = g((m, k), 61), g((k, n), 62), O.uniform((m, n), 63)
    dC = dev(C0)
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=L, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    want = O.crt_gemm(A, B, C0.copy(), info.int_bits, L)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173]


def _max_abs_product_bits(A, B, p):
    """log2 of max_ij sum_l |A'_il| |B'_lj| for the scaled integers A' = rint(a 2^(p - e_i)),
    B' = rint(b 2^(p - f_j)), 2^e the power of two just above a row's (column's) largest |entry|."""
    ea = np.frexp(np.abs(A).max(axis=1))[1]
    eb = np.frexp(np.abs(B).max(axis=0))[1]
    Ai = np.rint(np.ldexp(A, (p - ea)[:, None]))
    Bi = np.rint(np.ldexp(B, (p - eb)[None, :]))
    return float(np.log2((np.abs(Ai) @ np.abs(Bi)).max()))


def test_crt_auto_moduli_cover_the_largest_product():
    """The automatic modulus count must leave M = prod m_t above 2 max|C'| for the scaled
    integers actually formed. All entries of one magnitude and sign (row and column ratios of
    about 1) are its worst case, max|C'| ~ 2^(2p) k, and must come back exact; entries spread
    over 2^-20 .. 2^20 get by with fewer bits than the worst case 2p + log2 k + 2."""
    import torch
    m, k, n = 64, 3000, 48
    A = np.full((m, k), 1023.0)
    B = np.full((k, n), 1021.0)
    dC = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    assert np.array_equal(dC.cpu().numpy(), A @ B)
    cum = np.cumsum(np.log2(MODULI))
    assert cum[info.slices - 1] > 1 + _max_abs_product_bits(A, B, info.int_bits)
    Aw, Bw, C0 = O.wide((m, k), 5), O.wide((k, n), 6), O.uniform((m, n), 7)
    dC = dev(C0)
    infw = phpc.dgemm_split_device(dev(Aw), dev(Bw), dC, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    need = 1 + _max_abs_product_bits(Aw, Bw, infw.int_bits)
    worst = 2 * infw.int_bits + np.log2(k) + 2
    print(f"  wide: L={infw.slices} p={infw.int_bits} bits={cum[infw.slices - 1]:.1f} need={need:.1f} "
          f"worst={worst:.1f}", flush=True)
    assert cum[infw.slices - 1] > need
    assert need < worst - 4
    # the same moduli and integer width through the restatement give the same bits
    want = O.crt_gemm(Aw, Bw, C0.copy(), infw.int_bits, infw.slices)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


def _exact_width(X, axis):
    """Largest E - q over the rows (axis=1) or columns (axis=0) of X with a nonzero entry:
    every |x| < 2^E and q is the lowest set bit of any entry, so that 2^(p - E) x is an
    integer for every entry once p >= E - q."""
    f, e = np.frexp(X)
    mant = np.abs(np.ldexp(f, 53)).astype(np.int64)  # odd * 2^(trailing zeros)
    low = np.where(mant != 0, e - 53 + np.log2(mant & -mant).astype(np.int64), 1 << 20)
    mx = np.abs(X).max(axis=axis)
    E = np.frexp(mx)[1]
    return int((E - low.min(axis=axis))[mx > 0].max())


@pytest.mark.parametrize("kind", ["wide", "int"])
def test_crt_auto_uses_the_exact_width(kind):
    """Entries that all fit a narrow width below their row's (column's) largest exponent --
    powers of two over 2^-20 .. 2^20 need 41 bits, integers up to 1000 need 10 -- scale to
    integers exactly at that width, so the automatic rule takes it in place of the wider
    integer the error estimate asks for: the result is the restatement's bits at that width
    and modulus count, and exact (integers) or within two roundings (powers of two) of the
    double-double product."""
    import torch
    m, k, n = 120, 700, 90
    print(f"crt exact width {kind}", flush=True)
    if kind == "wide":
        A, B = O.wide((m, k), 31), O.wide((k, n), 32)
    else:
        rng = np.random.default_rng(34)
        A = rng.integers(-1000, 1001, (m, k)).astype(np.float64)
        B = rng.integers(-1000, 1001, (k, n)).astype(np.float64)
    w = max(_exact_width(A, 1), _exact_width(B, 0))
    assert w == (41 if kind == "wide" else 10)
    dA, dB = dev(A), dev(B)
    dC = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    info = phpc.dgemm_split_device(dA, dB, dC, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    print(f"  p={info.int_bits} moduli={info.slices} estimate={info.error_estimate:.2e}", flush=True)
    assert info.int_bits == w
    want = O.crt_gemm(A, B, np.zeros((m, n)), info.int_bits, info.slices)
    got = dC.cpu().numpy()
    assert np.array_equal(bits(got), bits(want))
    if kind == "int":
        assert np.array_equal(got, A @ B)  # |sums| < 2^53: exact in FP64
    else:
        rows, cols = np.divmod(np.arange(m * n), n)
        cw, _ = phpc.check_sampled(dA, dB, dC, rows, cols)
        assert cw <= 2.0 ** -51


def test_crt_auto_exact_width_with_zero_rows_and_subnormals():
    """The exact width over rows and columns that are all zero (no width), hold one entry,
    or hold subnormal entries (lowest set bits down to 2^-1074): the rule takes the largest
    width of the rest, and the result is the restatement's bits at that width and count."""
    import torch
    rng = np.random.default_rng(37)
    m, k, n = 40, 96, 36
    A = rng.integers(-50, 51, (m, k)).astype(np.float64) * 0.25
    B = rng.integers(-50, 51, (k, n)).astype(np.float64)
    A[3] = 0.0                                        # an all-zero row
    A[5] = 0.0
    A[5, 7] = 3.0                                     # a row with one entry
    A[9] = np.ldexp(rng.integers(-7, 8, k).astype(np.float64), -1070)  # subnormal row
    B[:, 4] = 0.0                                     # an all-zero column
    B[:, 6] = np.ldexp(rng.integers(1, 4, k).astype(np.float64), -1072)  # subnormal column
    w = max(_exact_width(A, 1), _exact_width(B, 0))
    dC = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    print(f"  width={w} p={info.int_bits} moduli={info.slices}", flush=True)
    assert info.int_bits == w
    want = O.crt_gemm(A, B, np.zeros((m, n)), info.int_bits, info.slices)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))


def test_crt_auto_full_mantissas_keep_the_estimate():
    """Uniform entries lie on the 2^-52 grid below 1, an exact width of 52 bits, above the
    estimate's integer width, which the rule keeps: the smallest p >= 8 with
    2 * 2^-p (max row ratio + max column ratio) / sqrt(k) <= target."""
    import torch
    m, k, n = 100, 900, 80
    A, B = O.uniform((m, k), 35), O.uniform((k, n), 36)
    assert max(_exact_width(A, 1), _exact_width(B, 0)) == 52
    dC = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    ra = (k * np.ldexp(1.0, np.frexp(np.abs(A).max(axis=1))[1]) / np.abs(A).sum(axis=1)).max()
    rb = (k * np.ldexp(1.0, np.frexp(np.abs(B).max(axis=0))[1]) / np.abs(B).sum(axis=0)).max()
    p = next(p for p in range(8, 71) if 2 * 2.0 ** -p * (ra + rb) / np.sqrt(k) <= 1e-14)
    assert info.int_bits == p


def test_crt_exact_on_integers():
    """Integer inputs whose product fits the moduli come back exact."""
    import torch
    rng = np.random.default_rng(12)
    A = rng.integers(-1000, 1001, (50, 70)).astype(np.float64)
    B = rng.integers(-1000, 1001, (70, 40)).astype(np.float64)
    dC = torch.zeros((50, 40), dtype=torch.float64, device="cuda")
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=8, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    assert info.int_bits >= 10  # 1000 < 2^10 fits the integer width, so scaling is exact
    assert np.array_equal(dC.cpu().numpy(), A @ B)


def test_int8_level_configs_bit_identical():
    """Every int8 level-kernel configuration computes the same exact integer residue GEMMs,
    so the CRT result is the same bits whichever runs the large-problem path (m = n = 3000
    gives 144 tiles of 256 x 256, enough for the large-problem kernel); the default's
    result is within the FP64 target. k = 900 is not a multiple of the 128-byte slab of
    the large-problem kernels: the planes are padded to whole slabs."""
    import torch
    m, k, n = 3000, 900, 3000
    dA = torch.empty((m, k), dtype=torch.float64, device="cuda")
    dB = torch.empty((k, n), dtype=torch.float64, device="cuda")
    phpc.fill_device(dA, phpc.DIST_WIDE, 41)
    phpc.fill_device(dB, phpc.DIST_UNIFORM, 42)
    results = {}
    cfg = 0
    try:
        while phpc.lib().phpc_split_set_config(cfg) == 0:
            dC = torch.zeros((m, n), dtype=torch.float64, device="cuda")
            phpc.dgemm_split_device(dA, dB, dC, slices=0, target=1e-14, method=phpc.SPLIT_CRT)
            torch.cuda.synchronize()
            results[cfg] = dC
            cfg += 1
    finally:
        phpc.lib().phpc_split_set_config(13)  # the default large-problem kernel
    assert len(results) >= 16
    ref = results[13]
    for c, dC in results.items():
        assert torch.equal(dC.view(torch.int64), ref.view(torch.int64)), c
    rng = np.random.default_rng(6)
    cw, _ = phpc.check_sampled(dA, dB, ref, rng.integers(0, m, 512), rng.integers(0, n, 512))
    assert cw <= 1e-14


@pytest.mark.parametrize("k,L", [(300, 6), (40, 4), (129, 5)])
def test_crt_large_problem_kernel_ragged_k_matches_restatement(k, L):
    """m = n = 2900 (144 tiles of 256 x 256) runs the residue GEMMs on the large-problem
    kernel, which walks k in 128-byte slabs; k = 300, 40 and 129 leave a partial slab that
    the zero padding of the planes must fill. The result equals the oracle's restatement
    bit for bit."""
    import torch
    m = n = 2900
    print(f"crt large-kernel m=n={m} k={k} L={L}", flush=True)
    A, B, C0 = O.uniform((m, k), 51), O.wide((k, n), 52), O.uniform((m, n), 53)
    dC = dev(C0)
    info = phpc.dgemm_split_device(dev(A), dev(B), dC, slices=L, method=phpc.SPLIT_CRT)
    torch.cuda.synchronize()
    want = O.crt_gemm(A, B, C0.copy(), info.int_bits, L)
    assert np.array_equal(bits(dC.cpu().numpy()), bits(want))
