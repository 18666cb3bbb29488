"""The traversal kernels divide by ray-direction components through a
per-ray double reciprocal: q = RN_single(RN_double(a * RN_double(1/b))),
with a single-precision division only when |q| < 2^-126.  This checks on
the CPU that the rule returns the IEEE single quotient bit for bit, on
random bit patterns, traversal-like operands and the cases closest to a
rounding midpoint (integer ratios and their neighbours), and that an exact
subnormal midpoint exists, which is why that range is divided again."""
import numpy as np


def _rule(a, b):
    with np.errstate(all="ignore"):
        q = (a.astype(np.float64) * (1.0 / b.astype(np.float64))).astype(
            np.float32)
        fall = (np.abs(q) < np.float32(2.0 ** -126)) & (a != 0)
        return np.where(fall, a / b, q)


def _same(x, y):
    return (x.view(np.uint32) == y.view(np.uint32)) | (np.isnan(x) &
                                                       np.isnan(y))


def test_rule_equals_single_division():
    rng = np.random.default_rng(5)
    n = 2_000_000
    cases = []
    a = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    cases.append((a.view(np.float32), b.view(np.float32)))
    o = rng.uniform(-700, 700, n).astype(np.float32)
    f = np.floor(rng.uniform(-300, 300, n)).astype(np.float32)
    cases.append((f - o, rng.normal(size=n).astype(np.float32)))
    ai = rng.integers(1, 2 ** 24, n).astype(np.float32)
    bi = rng.integers(1, 2 ** 12, n).astype(np.float32)
    cases += [(ai, bi), (np.nextafter(ai, np.float32(np.inf)), bi),
              (ai * np.float32(2.0 ** 100),
               (bi * np.float32(2.0 ** -60)).astype(np.float32)),
              ((ai * np.float32(2.0 ** -140)).astype(np.float32), bi)]
    specials = np.array([0.0, -0.0, np.inf, -np.inf, 1e-45, -1e-45, 1.0,
                         3.0, 3.4e38], np.float32)
    sa, sb = np.meshgrid(specials, specials)
    cases.append((sa.ravel(), sb.ravel()))
    for a, b in cases:
        with np.errstate(all="ignore"):
            ref = a / b
        assert _same(ref, _rule(a, b)).all()


def test_subnormal_midpoint_needs_the_fallback():
    # 9*2^-150 / 3 = 1.5 * 2^-149 lies exactly between two subnormals
    a = np.array([9 * 2.0 ** -150], np.float32)
    b = np.array([3.0], np.float32)
    with np.errstate(all="ignore"):
        q = (a.astype(np.float64) * (1.0 / b.astype(np.float64))).astype(
            np.float32)
    assert abs(float(q[0])) < 2.0 ** -126  # the rule sends it to a / b
    assert _same(a / b, _rule(a, b)).all()
