"""The C-ABI library loads on the CPU and exports every function include/*.h declares."""
import ctypes
import glob
import os
import re

import numpy as np
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
    # a specialised form without an ahead-of-time kernel is compiled at run time
    names, preds, parts, aggs = W.q1_spec(group="a")
    assert _explain((names, preds, parts, aggs[:3]))[1] == _capi.SPEC_SHAPE_RUNTIME


def test_batch_and_serve_entry_points_validate_without_a_device():
    L = _capi.lib()
    h = ctypes.c_void_p()
    assert L.tgp_batch_create(None, 16, ctypes.byref(h)) == _capi.ERR_INVALID
    assert L.tgp_batch_set_tuning(None, None) == _capi.ERR_INVALID
    t = _capi.BatchTuning()
    assert L.tgp_batch_get_tuning(None, ctypes.byref(t)) == _capi.ERR_INVALID
    cfg = _capi.ServeConfig(1024, 2, 2, 0, 0, 0, 0, 32, 2, 1, 0, 0)
    assert L.tgp_serve_create(None, ctypes.byref(cfg), ctypes.byref(h)) == _capi.ERR_INVALID
    assert L.tgp_serve_submit(None, None, 0, None) == _capi.ERR_INVALID
    assert L.tgp_serve_destroy(None) == _capi.OK
    assert L.tgp_serve_now_ns() > 0


def test_serve_struct_sizes_match_c_layout():
    assert ctypes.sizeof(_capi.BatchTuning) == 24
    assert ctypes.sizeof(_capi.ServeConfig) == 56
    assert ctypes.sizeof(_capi.ServeTiming) == 40
    assert ctypes.sizeof(_capi.ServeJoin) == 2 * 24 + ctypes.sizeof(_capi.JoinAggQuery)
    assert ctypes.sizeof(_capi.ServeRequest) == 8 + ctypes.sizeof(_capi.AggQuery) + \
        ctypes.sizeof(_capi.ServeJoin)


def test_serve_mix_is_seeded_and_in_range():
    from turbogp_amd import workloads as W
    a = W.serve_mix(500, 1_000_000, 20_000, 3_000_000)
    assert a == W.serve_mix(500, 1_000_000, 20_000, 3_000_000)
    kinds = {k: sum(1 for m in a if m[0] == k) for k in ("q1", "q6", "join", "full")}
    assert 0.6 < (kinds["q1"] + kinds["q6"]) / 500 < 0.8 and 0.18 < kinds["join"] / 500 < 0.32
    for kind, b, e, nb in a:
        limit = 3_000_000 if kind == "join" else 1_000_000
        assert 0 <= b < e <= limit
        assert (kind == "join") == (nb > 0) and nb <= 20_000
    off = W.poisson_offsets(1000, 2000.0)
    assert (np.diff(off) >= 0).all() and 0.3e9 < off[-1] < 0.7e9  # ~1000 / 2000 per s


def test_join_select_distinct_entry_points_validate_without_a_device():
    """The newer entry points reject missing arguments with a status, no device needed."""
    L = _capi.lib()
    h = ctypes.c_void_p()
    keys = (_capi.Column * 2)()
    n = ctypes.c_uint64()
    assert L.tgp_hash_build_multi(None, keys, 2, None, ctypes.byref(h), None) == _capi.ERR_INVALID
    assert L.tgp_hash_join_pairs_sync(None, None, keys, 2, _capi.JOIN_LEFT, 0, 0, None, None, 0,
                                      ctypes.byref(n), None) == _capi.ERR_INVALID
    assert L.tgp_hash_join_mark(None, None, keys, 2, 0, 0, None, None) == _capi.ERR_INVALID
    assert L.tgp_select_rows(None, 0, 10, None, None, 0, 5, None, 5, ctypes.byref(n),
                             None) == _capi.ERR_INVALID
    assert L.tgp_gather(None, None, None, 0, 1, None, None) == _capi.ERR_INVALID
    assert L.tgp_count_distinct(None, None, 0, 0, None, 0, ctypes.byref(n), None) == \
        _capi.ERR_INVALID
    assert L.tgp_tune_set(None, _capi.TUNE_JOIN_PACK_BITS, 3) == _capi.ERR_INVALID
    assert b"tgp_tune_set" in L.tgp_last_error()
