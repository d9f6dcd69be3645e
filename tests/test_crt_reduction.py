"""The exact-arithmetic steps of the CRT split GEMM's residue and reconstruction kernels
(csrc/split.hip), checked on the CPU with IEEE float32 arithmetic as the device does it:

- the residue GEMM's epilogue: int32 sums reduced mod m in two float32 rounds;
- residues: for every odd modulus and every byte-dot-product sum the kernels can form
  (0 .. 9 * 255^2), q = rint(float32(sum) * float32(1/m)) is the exact nearest quotient,
  so sum - q*m is the symmetric residue; the narrow form (|x| <= 2^46, the 48-bit two's
  complement and its sign byte) gives the same residues;
- Garner digits: with the coefficients stored in the symmetric (int8) range, the digit sums
  stay below 2^18, the same rint reduction is exact there, and the mixed-radix digits
  rebuild every integer in the symmetric range of the moduli product; the plain-FP64
  Horner prefix (top digits while their product is <= 2^53) is exact;
- Horner grouping: two digits a step while the value stays below 2^104 leaves the same
  double-double pair as one digit a step, so the kernel's bits do not depend on it.
"""
import math

import numpy as np

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173]


def sym(x, m):
    r = x % m
    return r - m if r > (m - 1) // 2 else r


def test_residue_rint_reduction_is_exact_over_the_whole_sum_range():
    sums = np.arange(0, 9 * 255 * 255 + 1, dtype=np.int64)
    sf = sums.astype(np.float32)
    for m in MODULI[1:]:
        assert m % 2 == 1
        rmf = np.float32(1.0) / np.float32(m)
        q = np.rint((sf * rmf).astype(np.float32)).astype(np.int64)
        r = sums - q * m
        assert np.all(np.abs(r) <= (m - 1) // 2), m
        assert np.all((sums - r) % m == 0), m
        # the device forms r with one FMA in float32: exact because |q m| < 2^24
        assert int(np.max(np.abs(q * m))) < 2 ** 24


def test_byte_weights_give_the_residue_of_the_integer():
    rng = np.random.default_rng(3)
    for m in MODULI[1:]:
        w = [pow(2, 8 * i, m) for i in range(9)]
        for _ in range(300):
            x = int(rng.integers(0, 2 ** 62)) * int(rng.integers(1, 2 ** 10))
            x %= 2 ** 72
            b = [(x >> (8 * i)) & 0xFF for i in range(9)]
            s = sum(bi * wi for bi, wi in zip(b, w))
            assert s <= 9 * 255 * 255
            assert sym(s, m) == sym(x, m)


def test_narrow_two_complement_form_gives_the_residue():
    """|x| <= 2^46: x = u - s 2^48 with u its 48-bit two's complement and s its sign, so the
    six bytes of u weighted by 2^(8 i) mod m, plus s weighted by (-2^48) mod m, sum to x mod
    m; the sum stays below 2^19 (inside the rint argument above), and u's low byte as int8
    is x's residue mod 256. hi = floor(x 2^-24), lo = x - hi 2^24 are exact in FP64, and
    bits 8..24 of hi (32-bit two's complement) are u's bytes 4-5 and, in bit 16, s."""
    rng = np.random.default_rng(4)
    xs = [0, 1, -1, 2 ** 46, -2 ** 46, 2 ** 46 - 1, 1 - 2 ** 46, 2 ** 24, -2 ** 24, 2 ** 24 - 1, 1 - 2 ** 24, 128, -128,
          -129, 255, -256, 2 ** 32, -2 ** 32, 2 ** 40 + 7, -2 ** 40 - 7]
    xs += [int(v) for v in rng.integers(-2 ** 46, 2 ** 46, 3000, endpoint=True)]
    xs += [int(v) for v in rng.integers(-2 ** 20, 2 ** 20, 1000)]
    for x in xs:
        hd = math.floor(float(x) * 2.0 ** -24)
        lo = float(x) - hd * 2.0 ** 24
        assert -2 ** 22 <= hd <= 2 ** 22 and 0 <= lo < 2 ** 24 and hd * 2 ** 24 + lo == x
        hi32 = int(hd) & 0xFFFFFFFF
        w0 = (int(lo) | (hi32 << 24)) & 0xFFFFFFFF
        w1 = (hi32 >> 8) & 0x1FFFF
        u, sgn = x % 2 ** 48, int(x < 0)
        assert w0 == u & 0xFFFFFFFF and w1 == (u >> 32) | (sgn << 16)
        low = w0 & 0xFF
        assert (low - 256 if low > 127 else low) == sym(x, 256)
        for m in MODULI[1:]:
            c = [pow(2, 8 * i, m) for i in range(6)]
            s = sum(((w0 >> (8 * i)) & 0xFF) * c[i] for i in range(4))
            s += (w1 & 0xFF) * c[4] + ((w1 >> 8) & 0xFF) * c[5] + ((w1 >> 16) & 0xFF) * ((-2 ** 48) % m)
            assert s < 2 ** 19
            assert sym(s, m) == sym(x, m)


def test_epilogue_residue_of_int32_sums():
    """The residue GEMM's epilogue reduces an int32 sum v (|v| <= 128^2 x 130944 + 128 <
    2^31) in two float32 rounds on a = |v|: q1 = rint(float32(a) * float32(1/m)) (below
    2^24, for the kernel's 24-bit multiply), r = a - q1 m (exact), then
    q2 = rint(float32(r) * float32(1/m)), r - q2 m, with v's sign: the symmetric residue for
    every odd modulus; m = 256 takes v's low byte."""
    rng = np.random.default_rng(11)
    vmax = 128 * 128 * 130944 + 128
    assert vmax < 2 ** 31
    v = np.concatenate([rng.integers(-vmax, vmax, 200000, endpoint=True),
                        np.arange(-3000, 3000), np.array([vmax, -vmax, vmax - 1, 1 - vmax, 2 ** 30, -2 ** 30])])
    v = v.astype(np.int64)
    a = np.abs(v)
    for m in MODULI[1:]:
        rmf = np.float32(1.0) / np.float32(m)
        q1 = np.rint((a.astype(np.float32) * rmf).astype(np.float32)).astype(np.int64)
        assert np.all(q1 < 2 ** 24)  # the kernel's 24-bit unsigned multiply
        r = a - q1 * m
        assert np.all(np.abs(r) < 3 * m)
        q2 = np.rint((r.astype(np.float32) * rmf).astype(np.float32)).astype(np.int64)
        res = np.where(v < 0, -(r - q2 * m), r - q2 * m)
        want = v % m
        want = np.where(want > (m - 1) // 2, want - m, want)
        assert np.array_equal(res, want), m
    low = (v & 0xFF).astype(np.int64)
    low = np.where(low > 127, low - 256, low)
    want = v % 256
    assert np.array_equal(low, np.where(want > 127, want - 256, want))


def garner_constants(L):
    invs, q = [0] * L, [[0] * L for _ in range(L)]
    for t in range(L):
        m = MODULI[t]
        prod, pmod = 1, []
        for s in range(t):
            pmod.append(prod)
            prod = prod * MODULI[s] % m
        inv = 1 if t == 0 else pow(prod, -1, m)
        invs[t] = sym(inv, m)
        for s in range(t):
            q[s][t] = sym((m - pmod[s] * inv % m) % m, m)
    return invs, q


def plain_start(L):
    prod, s = MODULI[L - 1], L - 1
    while s > 0 and prod * MODULI[s - 1] <= 2 ** 53:
        prod *= MODULI[s - 1]
        s -= 1
    return s


def test_garner_digits_and_horner_rebuild_every_integer():
    rng = np.random.default_rng(5)
    for L in (8, 12, 13, 14, 15, 16, 20):
        invs, q = garner_constants(L)
        assert all(-128 <= c <= 127 for c in invs) and all(-128 <= c <= 127 for row in q for c in row)
        M = math.prod(MODULI[:L])
        S = plain_start(L)
        assert math.prod(MODULI[S:L]) <= 2 ** 53
        for trial in range(400):
            x = int(rng.integers(-2 ** 62, 2 ** 62)) * int(rng.integers(1, 2 ** 40)) % M
            x = x - M if x >= M // 2 else x
            if trial == 0:
                x = M // 2 - 1  # M is even (m_0 = 256): the symmetric range is [-M/2, M/2)
            if trial == 1:
                x = -(M // 2)
            r = [sym(x, m) for m in MODULI[:L]]
            v = [r[0]]
            for t in range(1, L):
                y = r[t] * invs[t] + sum(v[s] * q[s][t] for s in range(t))
                assert abs(y) < 2 ** 18
                yf = np.float32(y)
                qq = np.rint(np.float32(yf * (np.float32(1.0) / np.float32(MODULI[t]))))
                d = int(y - int(qq) * MODULI[t])
                assert abs(d) <= (MODULI[t] - 1) // 2
                v.append(d)
            # plain FP64 Horner on the top digits is exact
            hi = float(v[L - 1])
            for t in range(L - 2, S - 1, -1):
                hi = math.fma(hi, float(MODULI[t]), float(v[t])) if hasattr(math, "fma") else hi * MODULI[t] + v[t]
            exact = v[L - 1]
            for t in range(L - 2, S - 1, -1):
                exact = exact * MODULI[t] + v[t]
            assert hi == exact
            val = exact
            for t in range(S - 1, -1, -1):
                val = val * MODULI[t] + v[t]
            assert val == x, (L, x)


def _dd_step(hi, lo, mt, vt):
    """One double-double Horner step of the reconstruction kernel, (hi, lo) * mt + vt, in
    IEEE doubles; the fma's exact product error is formed from the integers."""
    ph = hi * mt
    pe = float(int(hi) * int(mt) - int(ph))  # fma(hi, mt, -ph): representable, so exact
    pl = lo * mt + pe
    sh = ph + vt
    bb = sh - ph
    se = (ph - (sh - bb)) + (vt - bb)
    tl = pl + se
    h = sh + tl
    return h, tl - (h - sh)


def exact_floor(L):
    prod, f = 1, L
    while f > 0 and prod * MODULI[f - 1] <= 2 ** 105:
        prod *= MODULI[f - 1]
        f -= 1
    return f


def _horner(v, L, pairs):
    """The reconstruction kernel's Horner over the mixed-radix digits: plain FP64 while the
    prefix is at most 2^53, then double-double; with pairs, two digits a step down to
    exact_floor(L) (one digit a step below it), else one digit a step throughout."""
    S, F = plain_start(L), exact_floor(L) if pairs else L
    hi, lo, t = float(v[L - 1]), 0.0, L - 2
    while pairs and t - 1 >= S:
        mt, vt = MODULI[t] * MODULI[t - 1], v[t] * MODULI[t - 1] + v[t - 1]
        assert mt < 2 ** 16 and abs(vt) < 2 ** 15
        hi = float(int(hi) * mt + vt)  # fma below 2^53: exact
        t -= 2
    while t >= S:
        hi = float(int(hi) * MODULI[t] + v[t])
        t -= 1
    assert abs(int(hi)) <= 2 ** 53
    while t - 1 >= F:
        mt, vt = MODULI[t] * MODULI[t - 1], v[t] * MODULI[t - 1] + v[t - 1]
        hi, lo = _dd_step(hi, lo, float(mt), float(vt))
        t -= 2
    while t >= 0:
        hi, lo = _dd_step(hi, lo, float(MODULI[t]), float(v[t]))
        t -= 1
    return hi, lo


def test_two_digit_horner_steps_match_one_digit_steps():
    """The reconstruction kernel takes the mixed-radix digits two at a time while the value
    stays below 2^104 (multiplier m_t m_{t-1} < 2^16, digit v_t m_{t-1} + v_{t-1} with
    |.| < 2^15). There every step is exact, so the (hi, lo) pair equals the one-digit
    Horner's (hi = the integer rounded once, lo the rest) and the integer itself; the
    digits below exact_floor(L), reached by 14 or more moduli, go one at a time, so the
    result is the one-digit Horner's bits in every case, in and beyond the exact range."""
    rng = np.random.default_rng(8)
    for L in (2, 3, 8, 12, 13, 14, 15, 16, 19, 20):
        invs, q = garner_constants(L)
        M = math.prod(MODULI[:L])
        assert math.prod(MODULI[exact_floor(L):L]) <= 2 ** 105
        for trial in range(300):
            x = int(rng.integers(-2 ** 62, 2 ** 62)) * int(rng.integers(1, 2 ** 62)) * int(rng.integers(1, 2 ** 40)) % M
            x = x - M if x >= M // 2 else x
            if trial == 0:
                x = M // 2 - 1
            if trial == 1:
                x = -(M // 2)
            if trial % 3 == 2 and M > 2 ** 101:  # values inside the exact range
                x = x % 2 ** 100 - 2 ** 99
            r = [sym(x, m) for m in MODULI[:L]]
            v = [r[0]]
            for t in range(1, L):
                y = r[t] * invs[t] + sum(v[s] * q[s][t] for s in range(t))
                v.append(sym(y, MODULI[t]))
            hi2, lo2 = _horner(v, L, pairs=True)
            hi1, lo1 = _horner(v, L, pairs=False)
            assert (hi2, lo2) == (hi1, lo1), (L, x)
            if abs(x) < 2 ** 104:
                assert int(hi2) + int(lo2) == x and hi2 + lo2 == float(x), (L, x)
