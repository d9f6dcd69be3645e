"""The 3 x BF16 cut of FP32 values that the FP32-accurate GEMM uses, restated in numpy:
x = hi + mid + lo exactly, each a BF16 value (round to nearest even), |mid| <= 2^-8 |x|,
|lo| <= 2^-16 |x|. CPU only."""
import numpy as np


def bf16_rne(x):
    u = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def split3(x):
    h = bf16_rne(x)
    r1 = (x - h).astype(np.float32)
    m = bf16_rne(r1)
    r2 = (r1 - m).astype(np.float32)
    lo = bf16_rne(r2)
    return h, m, lo, r2


def test_split_is_exact_and_bounded():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.uniform(-1, 1, 200000),
        np.ldexp(rng.uniform(0.5, 1, 100000) * rng.choice([-1, 1], 100000), rng.integers(-100, 100, 100000)),
        rng.integers(-(2 ** 24), 2 ** 24, 100000).astype(np.float64),
    ]).astype(np.float32)
    h, m, lo, r2 = split3(x)
    for part in (h, m, lo):  # every slice is a BF16 value
        assert np.all((part.view(np.uint32) & 0xFFFF) == 0)
    assert np.array_equal(lo, r2)  # the last remainder needs no rounding
    s = h.astype(np.float64) + m.astype(np.float64) + lo.astype(np.float64)
    assert np.array_equal(s, x.astype(np.float64))
    ax = np.abs(x.astype(np.float64))
    assert np.all(np.abs(m) <= np.ldexp(ax, -8))
    assert np.all(np.abs(lo) <= np.ldexp(ax, -16))


def test_dropped_terms_below_2_pow_minus_23():
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 1, 100000).astype(np.float32)
    b = rng.uniform(-1, 1, 100000).astype(np.float32)
    ha, ma, la, _ = split3(a)
    hb, mb, lb, _ = split3(b)
    f = lambda v: v.astype(np.float64)  # noqa: E731
    kept = f(ha) * f(hb) + f(ha) * f(mb) + f(ma) * f(hb) + f(ha) * f(lb) + f(ma) * f(mb) + f(la) * f(hb)
    exact = f(a) * f(b)
    assert np.all(np.abs(exact - kept) <= 2.0 ** -23 * np.abs(exact) + 1e-300)
