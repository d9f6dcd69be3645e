"""Pins of the CPU oracle (oracle/oracle.c) against values it does not produce itself:
the reference's own iterative.c (built from its sources into oracle/_ref), published
SplitMix64 outputs, closed forms, brute-force integers, hand-computed known answers,
binary128 and exact rational arithmetic."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"


# ---------------- generators ----------------

def test_splitmix64_published_sequence():
    # SplitMix64 seeded with 0 (Steele, Lea, Flood 2014; the reference implementation
    # by Vigna): first outputs of next().
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
            0xF88BB8A8724C81EC, 0x1B39896A51A8749B]
    assert [O.splitmix64_at(0, i) for i in range(5)] == want


def test_uniform_range_and_moments():
    x = O.uniform((200000,), 11)
    assert x.min() >= -1.0 and x.max() < 1.0
    assert abs(x.mean()) < 0.01 and abs(x.var() - 1 / 3) < 0.01
    # the value is 2*(top 53 bits)*2^-53 - 1: every sample is a multiple of 2^-52
    assert np.all(np.ldexp(x + 1.0, 52) == np.floor(np.ldexp(x + 1.0, 52)))


def test_wide_exponents():
    x = O.wide((100000,), 5)
    m, e = np.frexp(np.abs(x))
    assert np.all(m == 0.5)
    ex = e - 1
    assert ex.min() == -20 and ex.max() == 20
    assert len(np.unique(ex)) == 41
    assert abs((x > 0).mean() - 0.5) < 0.01


def test_ramp():
    assert np.array_equal(O.ramp((3, 3)).ravel(), np.arange(9.0))


# ---------------- iterative.c ----------------

@pytest.mark.skipif(not O.ref_iterative_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("n,dist", [(1, "u"), (17, "u"), (64, "u"), (48, "w"), (40, "r")])
def test_iterative_matches_reference_build(n, dist):
    if dist == "u":
        A, B = O.uniform((n, n), 3), O.uniform((n, n), 4)
    elif dist == "w":
        A, B = O.wide((n, n), 3), O.wide((n, n), 4)
    else:
        A, B = O.ramp((n, n)), O.ramp((n, n))
    C0 = O.uniform((n, n), 9)
    mine = O.iterative(A, B, C0.copy())
    ref = O.ref_iterative(A, B, C0.copy())
    assert np.array_equal(mine.view(np.uint64), ref.view(np.uint64))


def test_iterative_matches_golden_from_reference():
    z = np.load(GOLDEN / "ref_iterative_n48.npz")
    C = O.iterative(z["A"], z["B"], np.zeros_like(z["C"]))
    assert np.array_equal(C.view(np.uint64), z["C"].view(np.uint64))


# ---------------- closed forms and brute force ----------------

@pytest.mark.parametrize("N", [1, 2, 3, 7, 16])
def test_ramp_closed_form_vs_integers(N):
    A = np.arange(N * N, dtype=object).reshape(N, N)
    exact = A.dot(A)
    for r in range(N):
        for c in range(N):
            assert O.ramp_closed(N, r, c) == float(exact[r, c])


@pytest.mark.parametrize("N", [5, 64, 130])
def test_ramp_exact_paths(N):
    A = O.ramp((N, N))
    want = O.ramp_closed_matrix(N)
    assert np.array_equal(O.iterative(A, A, np.zeros((N, N))), want)
    assert np.array_equal(O.kernel_emul(A, A, np.zeros((N, N))), want)
    assert np.array_equal(O.summa_emul(A, A, np.zeros((N, N)), 2, 3, 7), want)
    assert np.array_equal(O.gemm_dd(A, A), want)


def test_ramp_golden():
    z = np.load(GOLDEN / "ramp_n64.npz")
    assert np.array_equal(O.ramp_closed_matrix(64), z["C"])


def test_kernel_emul_integer_brute_force():
    rng = np.random.default_rng(1)
    for m, k, n in [(1, 1, 1), (5, 3, 7), (33, 65, 17), (64, 64, 64)]:
        A = rng.integers(-8, 9, (m, k))
        B = rng.integers(-8, 9, (k, n))
        C0 = rng.integers(-100, 101, (m, n))
        want = (A @ B + C0).astype(np.float64)
        got = O.kernel_emul(A.astype(np.float64), B.astype(np.float64), C0.astype(np.float64))
        assert np.array_equal(got, want)


def test_kernel_emul_leading_dimensions():
    # operands embedded in larger arrays: only the m x k / k x n / m x n windows are used
    rng = np.random.default_rng(2)
    A = rng.integers(-4, 5, (6, 11)).astype(np.float64)
    B = rng.integers(-4, 5, (9, 13)).astype(np.float64)
    C = np.full((6, 10), 7.0)
    O.kernel_emul(A, B, C, m=5, k=9, n=8, lda=11, ldb=13, ldc=10)
    want = np.full((6, 10), 7.0)
    want[:5, :8] += A[:5, :9] @ B[:9, :8]
    assert np.array_equal(C, want)


# ---------------- hand-computed known answers ----------------

def test_fma_chain_vs_unfused_known_answer():
    # a0*b0 = 1 + 2^-29 + 2^-60 rounds to 1 + 2^-29; a1*b1 is its exact negative.
    # kernel (fma chain): fma(a1, b1, 1 + 2^-29) = -2^-60 exactly.
    # iterative (round each product, then add): 0.    exact value: 0.
    e = 1.0 + 2.0 ** -30
    A = np.array([[e, e], [0.0, 0.0]])
    B = np.array([[e, 0.0], [-e, 0.0]])
    assert O.kernel_emul(A, B, np.zeros((2, 2)))[0, 0] == -(2.0 ** -60)
    assert O.iterative(A, B, np.zeros((2, 2)))[0, 0] == 0.0
    assert O.gemm_dd(A, B)[0, 0] == 0.0


def test_summa_panel_grouping_known_answer():
    # C0 = 1, panel products s0 = s1 = 2^-53.
    # SUMMA with panel 1: (1 + 2^-53) + 2^-53 -> each add ties to even -> 1.
    # one panel: s = fma(2^-53, 1, 2^-53) = 2^-52, then 1 + 2^-52.
    t = 2.0 ** -53
    A = np.array([[t, t], [0.0, 0.0]])
    B = np.array([[1.0, 0.0], [1.0, 0.0]])
    C0 = np.array([[1.0, 0.0], [0.0, 0.0]])
    assert O.summa_emul(A, B, C0.copy(), 1, 1, panel=1)[0, 0] == 1.0
    assert O.summa_emul(A, B, C0.copy(), 1, 1, panel=2)[0, 0] == 1.0 + 2.0 ** -52
    # the reference default panel N / lcm(p, q) with a 2 x 2 grid on N = 2 is 1
    assert O.summa_emul(A, B, C0.copy(), 2, 2, panel=0)[0, 0] == 1.0


def test_summa_single_panel_equals_kernel():
    A, B, C0 = O.uniform((40, 40), 1), O.uniform((40, 40), 2), O.uniform((40, 40), 3)
    a = O.summa_emul(A, B, C0.copy(), 1, 1, panel=0)
    b = O.kernel_emul(A, B, C0.copy())
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_summa_independent_of_grid_for_fixed_panel():
    # the grid only partitions C; the arithmetic per element depends on the panel width
    A, B = O.uniform((30, 30), 4), O.uniform((30, 30), 5)
    ref = O.summa_emul(A, B, np.zeros((30, 30)), 1, 1, panel=4)
    for p, q in [(2, 2), (3, 2), (5, 3), (4, 4)]:
        got = O.summa_emul(A, B, np.zeros((30, 30)), p, q, panel=4)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


# ---------------- double-double vs independent arithmetic ----------------

def test_dd_cancellation_known_answer():
    A = np.array([[1e16, 1.0, -1e16]])
    B = np.array([[1.0], [1.0], [1.0]])
    assert O.gemm_dd(A, B)[0, 0] == 1.0


def test_dd_vs_exact_rationals():
    rng = np.random.default_rng(5)
    for trial in range(20):
        k = 60
        a = rng.standard_normal(k) * np.exp2(rng.integers(-30, 30, k))
        b = rng.standard_normal(k) * np.exp2(rng.integers(-30, 30, k))
        exact = float(sum(Fraction(x) * Fraction(y) for x, y in zip(a, b)))
        got = O.gemm_dd(a.reshape(1, k), b.reshape(k, 1))[0, 0]
        assert abs(got - exact) <= math.ulp(exact), (trial, got, exact)
        assert abs(O.dot_q(a, b) - exact) <= math.ulp(exact)


def test_dd_entries_match_full():
    A, B, C0 = O.uniform((20, 50), 1), O.uniform((50, 30), 2), O.uniform((20, 30), 3)
    full = O.gemm_dd(A, B, C0)
    rows = np.array([0, 5, 19, 7])
    cols = np.array([0, 29, 3, 11])
    assert np.array_equal(O.gemm_dd_entries(A, B, rows, cols, C0), full[rows, cols])


def test_kernel_emul_within_fp64_bound():
    A, B = O.uniform((37, 300), 8), O.uniform((300, 29), 9)
    C = O.kernel_emul(A, B, np.zeros((37, 29)))
    ref = O.gemm_dd(A, B)
    cw, nw = O.err_full(C, ref, A, B)
    assert 0 < cw <= 300 * 2.0 ** -53


# ---------------- grid helpers ----------------

def test_find_lcm():
    for a in range(1, 30):
        for b in range(1, 30):
            assert O.find_lcm(a, b) == a * b // math.gcd(a, b)


def test_dims_create_known_answers_and_brute_force():
    # MPI standard examples for 2 dimensions: (6) -> (3,2), (7) -> (7,1)
    assert O.dims_create(6) == (3, 2)
    assert O.dims_create(7) == (7, 1)
    assert O.dims_create(16) == (4, 4)
    assert O.dims_create(8) == (4, 2)
    assert O.dims_create(1) == (1, 1)
    for P in range(1, 200):
        d0, d1 = O.dims_create(P)
        pairs = [(P // d, d) for d in range(1, P + 1) if P % d == 0 and P // d >= d]
        best = min(pairs, key=lambda t: t[0] - t[1])
        assert (d0, d1) == best


def test_block_split():
    assert [O.block_size(10000, 3, i) for i in range(3)] == [3334, 3333, 3333]
    for n in (7, 100, 10000):
        for parts in (1, 2, 3, 4, 7):
            sizes = [O.block_size(n, parts, i) for i in range(parts)]
            starts = [O.block_start(n, parts, i) for i in range(parts)]
            assert sum(sizes) == n and max(sizes) - min(sizes) <= 1
            assert starts == [sum(sizes[:i]) for i in range(parts)]


# ---------------- residual and error metrics ----------------

def test_freivalds_detects_planted_error():
    n = 96
    A, B = O.uniform((n, n), 1), O.uniform((n, n), 2)
    C = O.gemm_dd(A, B)
    assert O.freivalds(A, B, C) < 1e-15
    # a single wrong entry moves row 3 of C x by 1e-6 * |x_5|, far above rounding level
    C[3, 5] += 1e-6
    assert O.freivalds(A, B, C) > 1e-11


def test_err_full_planted():
    A, B = O.uniform((8, 8), 1), O.uniform((8, 8), 2)
    ref = O.gemm_dd(A, B)
    C = ref.copy()
    C[2, 3] += 1e-10
    cw, nw = O.err_full(C, ref, A, B)
    ab = np.abs(A) @ np.abs(B)
    assert cw == pytest.approx(1e-10 / ab[2, 3], rel=1e-6)
    assert nw == pytest.approx(1e-10 / np.abs(ref).max(), rel=1e-6)


# ---------------- split-precision restatement ----------------

def test_split_digits_reconstruct():
    rng = np.random.default_rng(7)
    for _ in range(200):
        x = rng.standard_normal() * 2.0 ** rng.integers(-30, 30)
        e = O.exp_of_max(abs(x))
        for s in (1, 3, 7):
            d = O.split_digits(x, e, s)
            assert all(-127 <= t <= 127 for t in d)
            recon = sum(Fraction(t) * Fraction(2) ** (e - 7 * (i + 1)) for i, t in enumerate(d))
            assert abs(Fraction(x) - recon) < Fraction(2) ** (e - 7 * s)
    # exact cases: an integer that fits one 7-bit digit; 2^-9 with e = 0 sits in slice 2
    assert O.split_digits(100.0, 7, 3) == [100, 0, 0]
    assert O.split_digits(2.0 ** -9, 0, 3) == [0, 32, 0]


def test_split_exact_for_small_integers():
    rng = np.random.default_rng(8)
    A = rng.integers(-127, 128, (9, 40)).astype(np.float64)
    B = rng.integers(-127, 128, (40, 11)).astype(np.float64)
    C0 = rng.integers(-50, 50, (9, 11)).astype(np.float64)
    got = O.split_gemm(A, B, C0.copy(), 1)
    assert np.array_equal(got, A @ B + C0)


def test_split_digits_reconstruct_4_bit():
    """4-bit digits (the FP8-E4M3 slices): |d| <= 15, x = 2^e sum d_t 2^-4t within
    2^(e - 4s); exact cases where x fits the digits."""
    rng = np.random.default_rng(17)
    for _ in range(200):
        x = rng.standard_normal() * 2.0 ** rng.integers(-30, 30)
        e = O.exp_of_max(abs(x))
        for s in (1, 3, 7, 13):
            d = O.split_digits(x, e, s, bits=4)
            assert all(-15 <= t <= 15 for t in d)
            recon = sum(Fraction(t) * Fraction(2) ** (e - 4 * (i + 1)) for i, t in enumerate(d))
            assert abs(Fraction(x) - recon) < Fraction(2) ** (e - 4 * s)
    assert O.split_digits(13.0, 4, 3, bits=4) == [13, 0, 0]
    assert O.split_digits(-(2.0 ** -5), 0, 3, bits=4) == [0, -8, 0]
    assert O.split_digits(0.9375 + 0.0625 / 16, 0, 3, bits=4) == [15, 1, 0]


def test_split_4_bit_exact_for_small_integers():
    """Integers up to 15 whose rows and columns reach 8 (exponent 4) are one 4-bit digit:
    one FP8 slice gives the exact product."""
    rng = np.random.default_rng(18)
    A = rng.integers(-15, 16, (9, 40)).astype(np.float64)
    B = rng.integers(-15, 16, (40, 11)).astype(np.float64)
    A[:, 0] = 9.0
    B[0, :] = -12.0
    C0 = rng.integers(-50, 50, (9, 11)).astype(np.float64)
    got = O.split_gemm(A, B, C0.copy(), 1, bits=4)
    assert np.array_equal(got, A @ B + C0)


def test_split_4_bit_error_within_truncation_bound_and_decreasing():
    m, k, n = 24, 300, 20
    for gen in (O.uniform, O.wide):
        A, B = gen((m, k), 1), gen((k, n), 2)
        ref = O.gemm_dd(A, B)
        eA = np.array([O.exp_of_max(np.abs(A[i]).max()) for i in range(m)])
        eB = np.array([O.exp_of_max(np.abs(B[:, j]).max()) for j in range(n)])
        ab = np.abs(A) @ np.abs(B)
        prev = None
        for s in (3, 5, 7, 9, 11, 13, 14):
            C = O.split_gemm(A, B, np.zeros((m, n)), s, bits=4)
            bound = k * (s + 1) * 2.0 ** (-4 * s) * np.ldexp(1.0, eA[:, None] + eB[None, :]) + 8 * 2.0 ** -53 * ab
            assert np.all(np.abs(C - ref) <= bound)
            cw, _ = O.err_full(C, ref, A, B)
            if prev is not None and prev > 1e-15:
                assert cw < prev
            prev = cw
        assert prev <= 1e-15


def test_split_error_within_truncation_bound_and_decreasing():
    m, k, n = 24, 300, 20
    for gen in (O.uniform, O.wide):
        A, B = gen((m, k), 1), gen((k, n), 2)
        ref = O.gemm_dd(A, B)
        eA = np.array([O.exp_of_max(np.abs(A[i]).max()) for i in range(m)])
        eB = np.array([O.exp_of_max(np.abs(B[:, j]).max()) for j in range(n)])
        ab = np.abs(A) @ np.abs(B)
        prev = None
        for s in (2, 3, 4, 5, 6, 7, 8):
            C = O.split_gemm(A, B, np.zeros((m, n)), s)
            # rigorous: each product loses < (s+1) 2^(-7s) 2^(e_i+f_j); plus a few roundings
            bound = k * (s + 1) * 2.0 ** (-7 * s) * np.ldexp(1.0, eA[:, None] + eB[None, :]) + 8 * 2.0 ** -53 * ab
            assert np.all(np.abs(C - ref) <= bound)
            cw, _ = O.err_full(C, ref, A, B)
            if prev is not None and prev > 1e-15:
                assert cw < prev
            prev = cw
        assert prev <= 1e-15


# ---------------- CRT restatement ----------------

def test_crt_digits_reconstruct_exact_integers():
    rng = np.random.default_rng(11)
    for L in (2, 5, 14, 16, 20):
        mods = O.MODULI[:L]
        M = 1
        for mm in mods:
            M *= mm
        for _ in range(50):
            X = int(rng.integers(-(2 ** 62), 2 ** 62)) * int(rng.integers(0, 2 ** 40)) % M - M // 2
            X = max(min(X, M // 2 - 1), -(M // 2) + 1)
            res = []
            for mm in mods:
                r = X % mm
                if r > (mm - 1) // 2:
                    r -= mm
                res.append(r)
            v = O.crt_digits(res)
            val, P = 0, 1
            for t, mm in enumerate(mods):
                val += v[t] * P
                P *= mm
            assert val == X


def test_crt_gemm_within_integerisation_bound():
    m, k, n = 20, 257, 18
    for gen in (O.uniform, O.wide):
        A, B = gen((m, k), 3), gen((k, n), 4)
        ref = O.gemm_dd(A, B)
        eA = np.array([O.exp_of_max(np.abs(A[i]).max()) for i in range(m)])
        eB = np.array([O.exp_of_max(np.abs(B[:, j]).max()) for j in range(n)])
        ab = np.abs(A) @ np.abs(B)
        prev = None
        for p, L in ((10, 4), (20, 7), (30, 9), (40, 12), (48, 14)):
            C = O.crt_gemm(A, B, np.zeros((m, n)), p, L)
            # |A - A'2^(p-e)| <= 2^(e-p-1) per entry (rounding to integers), likewise B
            ea = np.ldexp(1.0, eA - p - 1)[:, None]
            eb = np.ldexp(1.0, eB - p - 1)[None, :]
            bound = (ea * (np.abs(B).sum(axis=0)[None, :]) + eb * (np.abs(A).sum(axis=1)[:, None])
                     + k * ea * eb + 4 * 2.0 ** -53 * ab)
            assert np.all(np.abs(C - ref) <= bound), (p, L)
            cw, _ = O.err_full(C, ref, A, B)
            if prev is not None and prev > 1e-15:
                assert cw < prev
            prev = cw
        assert prev <= 1e-14


@pytest.mark.parametrize("N,p,q,panel", [(24, 1, 1, 0), (30, 2, 3, 7), (33, 3, 3, 0), (40, 4, 4, 5), (17, 2, 2, 17)])
def test_summa_entries_equals_full_emulation(N, p, q, panel):
    """orc_summa_entries, the per-entry form used at sizes the full emulation cannot reach,
    gives every entry of orc_summa_emul (any grid, this panel) and, for one panel, of
    orc_kernel_emul, bit for bit, with and without C0; on wide-exponent inputs, where a
    different panel grouping changes the rounding."""
    A, B, C0 = O.wide((N, N), 1), O.wide((N, N), 2), O.uniform((N, N), 3)
    rr, cc = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    rows, cols = rr.ravel(), cc.ravel()
    w = panel if panel else N // O.find_lcm(p, q)
    want = O.summa_emul(A, B, C0.copy(), p, q, panel)
    got = O.summa_entries(A, B, rows, cols, panel=w, C0=C0).reshape(N, N)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    one = O.summa_entries(A, B, rows, cols, panel=0).reshape(N, N)
    assert np.array_equal(one.view(np.uint64), O.kernel_emul(A, B, np.zeros((N, N))).view(np.uint64))
    # the grouping matters on these inputs: a mistaken single chain would not pass
    if 0 < w < N:
        assert not np.array_equal(one.view(np.uint64), O.summa_emul(A, B, np.zeros((N, N)), p, q, panel).view(np.uint64))


def test_summa_entries_ramp_closed_form():
    N = 64
    A = B = O.ramp((N, N))
    rng = np.random.default_rng(0)
    rows, cols = rng.integers(0, N, 300), rng.integers(0, N, 300)
    got = O.summa_entries(A, B, rows, cols, panel=16)
    assert np.array_equal(got, O.ramp_closed_matrix(N)[rows, cols])
