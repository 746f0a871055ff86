"""The C-ABI library loads on the CPU and exports every function include/*.h declares."""
import ctypes
import glob
import os
import re

import pytest

from turbogp_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(tgp_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_loads_and_exports_every_declared_symbol():
    lib = _capi.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert declared_functions(), "no declarations found"
    assert missing == []


def test_ctypes_signatures_cover_header():
    assert set(_capi.SIGNATURES) <= set(declared_functions())


def test_abi_version():
    assert _capi.lib().tgp_abi_version() == 1


def test_struct_sizes_match_c_layout():
    # sizes implied by the header's explicit padding
    assert ctypes.sizeof(_capi.Column) == 24
    assert ctypes.sizeof(_capi.Pred) == 24
    assert ctypes.sizeof(_capi.GroupPart) == 24
    assert ctypes.sizeof(_capi.Agg) == 4 * 6 + 8 * 7
    assert ctypes.sizeof(_capi.AggQuery) == 16 + 8 * 24 + 16 + 8 * 24 + 2 * 24 + 6 * 80 + 8


def test_no_device_is_an_error_not_a_crash():
    import torch  # noqa: F401  (only to learn whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = _capi.lib().tgp_ctx_create(0, 0, ctypes.byref(h))
    assert rc == _capi.ERR_NO_DEVICE
    assert b"device" in _capi.lib().tgp_last_error()


def test_plan_validation_needs_no_device():
    # tgp_group_agg_ngroups is pure host logic
    q = _capi.AggQuery()
    q.nparts = 2
    q.parts[0] = _capi.GroupPart(0, 1, 0, 3)
    q.parts[1] = _capi.GroupPart(1, 30, 0, 85)
    n = ctypes.c_uint64()
    assert _capi.lib().tgp_group_agg_ngroups(ctypes.byref(q), ctypes.byref(n)) == 0
    assert n.value == 255


def _explain(spec):
    from turbogp_amd import make_agg_query
    from turbogp_amd import workloads as W

    class FakeCol:  # descriptors only: explain never touches device memory
        def __init__(self, n, dt, addr):
            self.n, self.dt, self.addr = n, dt, addr

        def desc(self):
            return _capi.Column(self.addr, self.n, self.dt, 0)
    names, preds, parts, aggs = spec
    cols = [FakeCol(1000, W.FACT_SCHEMA[nm][1], 0x100000 * (i + 1)) for i, nm in enumerate(names)]
    q = make_agg_query(cols, preds, parts, aggs)
    st, sh = ctypes.c_int32(), ctypes.c_int32()
    _capi.check(_capi.lib().tgp_group_agg_explain(ctypes.byref(q), ctypes.byref(st),
                                                  ctypes.byref(sh)), "explain")
    return st.value, sh.value


def test_benchmark_queries_take_the_specialised_kernels():
    from turbogp_amd import workloads as W
    assert [_explain(W.q1_spec(group=g))[1] for g in ("ab", "a", "b", "month", "")] == [0, 1, 2,
                                                                                        3, 4]
    assert _explain(W.q6_spec())[1] == 5
    # a predicate whose value set is two intervals keeps the generic engine
    names, preds, parts, aggs = W.q6_spec()
    assert _explain((names, [(0, _capi.GE, 731, False)], parts, aggs))[1] == -1
    # strategies of the generic engine
    assert _explain(W.q1_spec(group="ab"))[0] == 1 and _explain(W.q1_spec(group="month"))[0] == 2
