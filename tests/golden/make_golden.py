"""Writes the golden fixtures of the scan / group-aggregate path.

Run from the repository root:  python tests/golden/make_golden.py
Each fixture is computed twice, by the C oracle (oracle/liboracle.so) and by an
independent numpy restatement in this file, and written only if both agree.
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import binding as B  # noqa: E402
from turbogp_amd import workloads as W  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def numpy_q1(cols, cutoff=W.Q1_DATE_CUTOFF):
    """Independent Q1 over numpy int64 (exact: every per-group sum < 2^63 at 1M rows)."""
    m = cols["dt"].astype(np.int64) <= cutoff
    g = cols["flaga"].astype(np.int64) + 3 * cols["flagb"].astype(np.int64)
    price, disc, tax, qty = (cols[k].astype(np.int64) for k in ("price", "disc", "tax", "qty"))
    p1 = price * (100 - disc)
    p2 = p1 * (100 + tax)
    out = []
    for gid in range(6):
        s = m & (g == gid)
        out.append([int(s.sum()), int(qty[s].sum()), int(price[s].sum()), int(p1[s].sum()),
                    int(p2[s].sum())])
    return out


def numpy_q6(cols):
    dt = cols["dt"].astype(np.int64)
    disc, qty, price = (cols[k].astype(np.int64) for k in ("disc", "qty", "price"))
    m = (dt >= 731) & (dt < 1096) & (disc >= 5) & (disc <= 7) & (qty < 24)
    return int(m.sum()), int((price * disc)[m].sum())


def main():
    n = 1_000_000
    cols = W.host_fact_columns(n, B.synth_uniform)
    checks = {k: hashlib.sha256(v.tobytes()).hexdigest() for k, v in cols.items()}

    names, preds, parts, aggs = W.q1_spec()
    q = B.make_query([cols[c] for c in names], preds, parts, aggs)
    cnt, sums, oor = B.group_agg(q, mode=0)
    _, f64, _ = B.group_agg(q, mode=1)
    ref = numpy_q1(cols)
    groups = []
    for gid in range(6):
        row = [int(cnt[gid])] + [int(x) for x in sums[gid]]
        assert row == ref[gid], (gid, row, ref[gid])
        groups.append({"flaga": gid % 3, "flagb": gid // 3, "count": row[0],
                       "sum_qty": str(row[1]), "sum_price": str(row[2]),
                       "sum_disc_price": str(row[3]), "sum_charge": str(row[4]),
                       "ref_f64": [float(x) for x in f64[gid]]})
    assert oor == 0

    names6, preds6, parts6, aggs6 = W.q6_spec()
    q6 = B.make_query([cols[c] for c in names6], preds6, parts6, aggs6)
    c6, s6, _ = B.group_agg(q6, mode=0)
    rc, rs = numpy_q6(cols)
    assert int(c6[0]) == rc and s6[0][0] == rs, (c6, s6, rc, rs)

    fx = {"rows": n, "seed": W.SEED, "column_sha256": checks,
          "q1": {"date_cutoff": W.Q1_DATE_CUTOFF, "groups": groups},
          "q6": {"count": rc, "sum_price_disc": str(rs)}}
    with open(os.path.join(OUT, "fact_1m.json"), "w") as f:
        json.dump(fx, f, indent=1)
    print("wrote fact_1m.json")


if __name__ == "__main__":
    main()
