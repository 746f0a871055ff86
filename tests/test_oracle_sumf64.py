"""The oracle's exact f64 SUM (binding.sum_f64) pinned on the CPU: the reference's
AggregateSumF64 known answers (kernel/aggregate.rs:187-239), math.fsum (Shewchuk's
correctly rounded summation) on random doubles of every magnitude, and closed forms
where a sequential sum rounds differently (absorbed 1s, intermediate overflow,
cancellation, subnormals, the signs of zero, NaN and the infinities)."""
import math

import numpy as np

from oracle import binding as B


def bits(xs):
    return np.array(xs, np.float64).view(np.uint64)


def test_reference_kats():
    assert B.sum_f64(bits([1.0, 2.0, 3.0, 4.0, 5.0])) == 15.0  # aggregate.rs:196-201
    assert B.sum_f64(bits([float(i) for i in range(100)])) == 4950.0  # :203-226


def test_equals_math_fsum_on_random_doubles():
    rng = np.random.default_rng(9)
    for trial in range(30):
        n = int(rng.integers(0, 3000))
        mant = rng.uniform(-1, 1, n)
        exp = rng.integers(-1070, 1000, n) if trial % 3 == 0 else rng.integers(-40, 40, n)
        xs = np.ldexp(mant, exp)
        got = B.sum_f64(bits(xs), neg_zero_start=False)
        want = math.fsum(xs)
        assert got == want or (got == 0 and want == 0), trial
        if want == 0:
            assert math.copysign(1, got) == 1


def test_closed_forms():
    two53 = float(2 ** 53)
    assert B.sum_f64(bits([two53, 1.0, 1.0])) == two53 + 2  # sequential: 2^53
    assert B.sum_f64(bits([1e308, 1e308, -1e308])) == 1e308  # sequential: inf
    assert B.sum_f64(bits([1e308, 1e308])) == math.inf
    assert B.sum_f64(bits([-1e308, -1e308])) == -math.inf
    assert B.sum_f64(bits([0.1] * 10)) == 1.0  # sequential: 0.9999999999999999
    assert B.sum_f64(bits([5e-324, 5e-324, 5e-324])) == 1.5e-323
    assert B.sum_f64(bits([1e16, 1.0, -1e16])) == 1.0  # sequential: 0.0
    # ties to even at the last place: 2^53 + 1 is a tie between 2^53 and 2^53 + 2
    assert B.sum_f64(bits([two53, 1.0])) == two53
    assert B.sum_f64(bits([two53 + 2, 1.0])) == two53 + 4
    nan = B.sum_f64(bits([1.0, float("nan")]))
    assert math.isnan(nan)
    assert math.isnan(B.sum_f64(bits([math.inf, -math.inf])))
    assert B.sum_f64(bits([math.inf, 1e308, 1e308])) == math.inf
    assert B.sum_f64(bits([-math.inf, 3.0])) == -math.inf


def test_signs_of_zero_and_selection():
    neg = lambda x: math.copysign(1, x) < 0
    assert neg(B.sum_f64(bits([])))              # Iterator::sum starts from -0.0
    assert neg(B.sum_f64(bits([-0.0, -0.0])))
    assert not neg(B.sum_f64(bits([-0.0, 0.0])))
    assert not neg(B.sum_f64(bits([1.5, -1.5])))
    assert not neg(B.sum_f64(bits([-0.0]), neg_zero_start=False))  # 0.0 + -0.0
    xs = np.array([1.0, 2.0, 4.0, 8.0, 16.0])
    sel = np.array([0b10101], np.uint64)
    assert B.sum_f64(bits(xs), sel) == 21.0
