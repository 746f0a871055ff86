"""NUMERIC text encoding of exact int128 results (host code of the library):
against Python's exact integers and decimal arithmetic."""
import ctypes as C
from decimal import ROUND_HALF_UP, Decimal, localcontext

import numpy as np
import pytest

from turbogp_amd import _capi as A
from turbogp_amd.device import GroupAggResult, numeric_text

I128_MIN, I128_MAX = -(1 << 127), (1 << 127) - 1


def words(v):
    u = v % (1 << 128)
    return u & ((1 << 64) - 1), u >> 64


def want_numeric(v, scale):
    with localcontext() as ctx:
        ctx.prec = 80
        return format(Decimal(v).scaleb(-scale), "f") if scale else str(v)


def want_avg(v, n, scale, digits):
    with localcontext() as ctx:
        ctx.prec = 120
        q = (Decimal(v).scaleb(-scale) / Decimal(n)).quantize(Decimal(1).scaleb(-digits),
                                                               rounding=ROUND_HALF_UP)
        s = format(q, "f")
        return s[1:] if s.startswith("-") and set(s[1:]) <= set("0.") else s


def test_known_answers():
    assert numeric_text(*words(12345), 2) == "123.45"
    assert numeric_text(*words(-5), 2) == "-0.05"
    assert numeric_text(*words(0), 0) == "0" and numeric_text(*words(0), 2) == "0.00"
    assert numeric_text(*words(I128_MAX), 0) == "170141183460469231731687303715884105727"
    assert numeric_text(*words(I128_MIN), 0) == "-170141183460469231731687303715884105728"
    assert numeric_text(*words(I128_MIN), 38) == "-1.70141183460469231731687303715884105728"
    buf, ln = C.create_string_buffer(8), C.c_uint64()
    assert A.lib().tgp_int128_to_numeric(*words(10**20), 0, buf, 8, C.byref(ln)) == A.ERR_INVALID
    assert ln.value == 21


def test_random_values_match_python_decimals():
    rng = np.random.default_rng(3)
    for _ in range(3000):
        bits = int(rng.integers(1, 128))
        v = int(rng.integers(0, 2**62)) * (1 << max(0, bits - 62)) + int(rng.integers(0, 1000))
        v = min(v, I128_MAX) * (-1 if rng.random() < 0.5 else 1)
        scale = int(rng.integers(0, 12))
        assert numeric_text(*words(v), scale) == want_numeric(v, scale), (v, scale)


def test_avg_rounds_half_away_from_zero():
    rng = np.random.default_rng(4)
    r = GroupAggResult(np.array([2], np.uint64), np.array([[list(words(12345))]], np.uint64), 0)
    assert r.avg_numeric(0, 2, 4) == ["61.7250"] and r.avg_numeric(0, 2, 2) == ["61.73"]
    for v, n, s, d in ((-5, 2, 0, 0), (5, 2, 0, 0), (1, 3, 0, 4), (-2, 3, 0, 2), (0, 7, 2, 3),
                       (I128_MAX, 1, 0, 0), (I128_MIN, 2**64 - 1, 0, 6), (99, 100, 0, 0),
                       (-1, 3, 2, 2)):
        buf, ln = C.create_string_buffer(96), C.c_uint64()
        assert A.lib().tgp_int128_avg_to_numeric(*words(v), n, s, d, buf, 96, C.byref(ln)) == A.OK
        assert buf.value.decode() == want_avg(v, n, s, d), (v, n, s, d)
    for _ in range(2000):
        v = int(rng.integers(-2**62, 2**62)) * int(rng.integers(1, 2**40))
        n = int(rng.integers(1, 2**63))
        s = int(rng.integers(0, 4))
        d = s + int(rng.integers(0, 8))
        buf, ln = C.create_string_buffer(96), C.c_uint64()
        assert A.lib().tgp_int128_avg_to_numeric(*words(v), n, s, d, buf, 96, C.byref(ln)) == A.OK
        assert buf.value.decode() == want_avg(v, n, s, d), (v, n, s, d)
