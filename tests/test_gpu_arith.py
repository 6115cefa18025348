"""GPU: the solver's FP64 divide and square root (cfd::ddiv, cfd::dsqrt in physics.h: the
compiler's IEEE sequences without their operand scaling) return exactly the IEEE results, bit
for bit, against numpy's correctly rounded a / b and sqrt and against the compiler's own a / b
and sqrt on the device. The operand classes cover the range the solver's states span and far
beyond it (magnitudes 1e-280..1e280, quotients up to 1e310), exact quotients, zeros, infinities
and NaNs, and each edge of the range in which the omitted scaling is inactive: numerators down
to 2^-969, divisors up to 2^1022, quotients down to 2^-1022, exponent gaps up to 767, roots of
numbers down to 2^-767. Just past those edges (subnormal operands or quotients, numerators
below 2^-969, divisors of 2^1022 and more, gaps of 768 and more, roots below 2^-767) the
scaling acts and the results can differ; the last test shows that they do for subnormal
operands and for roots below 2^-767 (its other classes differed nowhere in these samples)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernels", "libarith_probe.so")


def run(a, b):
    lib = C.CDLL(LIB)
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty((a.size, 4))
    dp = C.POINTER(C.c_double)
    rc = lib.arith_probe(a.ctypes.data_as(dp), b.ctypes.data_as(dp), out.ctypes.data_as(dp), C.c_int(a.size))
    assert rc == 0
    return out


def same(x, y):
    return (x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))


def operands(rng, n, alo, ahi, blo, bhi):
    a = 10.0 ** rng.uniform(alo, ahi, n) * rng.choice([-1.0, 1.0], n)
    b = 10.0 ** rng.uniform(blo, bhi, n) * rng.choice([-1.0, 1.0], n)
    return a, b


@pytest.mark.parametrize("name,alo,ahi,blo,bhi", [
    ("state values", -8, 12, -6, 4),
    ("physical range", -30, 30, -30, 30),
    ("wide", -280, 280, -280, 280),
    ("huge quotients", 200, 300, -100, -10),
])
def test_divide_and_sqrt_are_ieee(gpu, name, alo, ahi, blo, bhi):
    rng = np.random.default_rng([alo + 1000, ahi + 1000, blo + 1000, bhi + 1000])
    a, b = operands(rng, 1 << 20, alo, ahi, blo, bhi)
    out = run(a, b)
    with np.errstate(all="ignore"):
        q, s = a / b, np.sqrt(a)
    print(f"case {name}: {a.size} pairs", flush=True)
    assert same(out[:, 1], q).all() and same(out[:, 3], s).all()  # the compiler's sequences
    assert same(out[:, 0], q).all(), int((~same(out[:, 0], q)).sum())
    assert same(out[:, 2], s).all(), int((~same(out[:, 2], s)).sum())


def test_exact_quotients_and_special_values(gpu):
    rng = np.random.default_rng(5)
    b = 10.0 ** rng.uniform(-5, 5, 1 << 18)
    m = np.floor(rng.uniform(0, 1e6, b.size)) + 1
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 2.0, 0.5, 3.0, 1.7e308, -1e-300, 1e300])
    A = np.concatenate([b * m, np.repeat(sp, sp.size)])
    B = np.concatenate([b, np.tile(sp, sp.size)])
    out = run(A, B)
    with np.errstate(all="ignore"):
        q, s = A / B, np.sqrt(A)
    assert same(out[:, 0], q).all()
    assert same(out[:, 2], s).all()


def mag(rng, n, lo, hi):
    """n values 2^e m with e uniform in [lo, hi), m uniform in [1, 2), random signs."""
    return np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(lo, hi, n)) * rng.choice([-1.0, 1.0], n)


def subnormal(rng, n):
    return rng.integers(1, 1 << 52, n).astype(np.uint64).view(np.float64) * rng.choice([-1.0, 1.0], n)


def test_edges_of_the_unscaled_range(gpu):
    """Operands at the edges of the range in which the compiler's sequences scale nothing."""
    rng = np.random.default_rng(11)
    n = 1 << 16
    cases = {
        "numerators 2^-969..2^-940": (mag(rng, n, -969, -940), mag(rng, n, -40, 40)),
        "divisors 2^1000..2^1022": (mag(rng, n, 990, 1023), mag(rng, n, 1000, 1022)),
        "quotients 2^-1021..2^-899": (mag(rng, n, -900, -800), mag(rng, n, 100, 121)),
        "exponent gaps 568..767": (mag(rng, n, 500, 501), mag(rng, n, -267, -67)),
    }
    for name, (a, b) in cases.items():
        out = run(a, b)
        with np.errstate(all="ignore"):
            q = a / b
        print(f"case {name}: {a.size} pairs", flush=True)
        assert np.isfinite(q).all() and (np.abs(q) >= 2.0 ** -1022).all(), name
        assert same(out[:, 1], q).all(), name
        assert same(out[:, 0], q).all(), (name, int((~same(out[:, 0], q)).sum()))
    x = mag(rng, n, -767, -700)
    out = run(x, np.ones(n))
    assert same(out[:, 3], np.sqrt(x)).all() and same(out[:, 2], np.sqrt(x)).all()


def test_past_the_edges_the_scaling_acts(gpu):
    """Operands the compiler's sequences rescale: those sequences stay IEEE, and the solver's
    unscaled ones differ on some of them (the range the exact arithmetic excludes). The counts
    per class are printed; subnormal operands and roots below 2^-767 are where they differ."""
    rng = np.random.default_rng(12)
    n = 1 << 16
    cases = {
        "subnormal numerator": (subnormal(rng, n), mag(rng, n, -60, 60)),
        "subnormal divisor": (mag(rng, n, -60, 60), subnormal(rng, n)),
        "numerator below 2^-969": (mag(rng, n, -1000, -969), mag(rng, n, -40, 40)),
        "subnormal quotient": (mag(rng, n, -1000, -960), mag(rng, n, 30, 120)),
        "divisor from 2^1022": (mag(rng, n, -40, 40), mag(rng, n, 1022, 1024)),
        "exponent gap from 768": (mag(rng, n, 500, 501), mag(rng, n, -300, -267)),
    }
    differ = 0
    for name, (a, b) in cases.items():
        out = run(a, b)
        with np.errstate(all="ignore"):
            q = a / b
        d = int((~same(out[:, 0], q)).sum())
        print(f"case {name}: {a.size} pairs, {d} unscaled quotients differ from IEEE", flush=True)
        assert same(out[:, 1], q).all(), name
        differ += d
    x = np.concatenate([mag(rng, n, -1000, -767), subnormal(rng, n)])
    out = run(x, np.ones(x.size))
    ds = int((~same(out[:, 2], np.sqrt(x))).sum())
    print(f"roots below 2^-767: {x.size} values, {ds} unscaled roots differ from IEEE", flush=True)
    assert same(out[:, 3], np.sqrt(x)).all()
    assert differ > 0 and ds > 0
