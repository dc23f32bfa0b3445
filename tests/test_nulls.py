"""NULLs in the oracle (the reference's validity bitmaps, exec/batch.go:113-121)
pinned against numpy brute force written from the SQL rules the reference's
operators follow:

  - a predicate term that reads a NULL is not TRUE: the row fails the
    conjunction (compares merge the operands' validity, exec/expr.go:290-299,
    and the filter keeps TRUE-and-valid rows only);
  - arithmetic over a NULL is NULL; SUM/AVG/MIN/MAX skip NULL arguments and
    are NULL for a group where no row holds a value; COUNT(*) counts every
    row, COUNT(x) the rows where x holds a value (exec/aggregate.go:458-546,
    780-876);
  - a NULL group key is a group of its own (aggregate.go:424-427), reported
    with key -1;
  - zone maps take min/max over the rows that hold a value and count the
    NULLs; a row group with no value has no min/max and never prunes
    (storage/writer.go:205-348, planner/physical.go:653, 688).

A plausible slip (a NULL read as its zeroed value, a NULL counted in a
COUNT(x) or an AVG's divisor, a NULL key folded into code 0) changes one of
the numbers checked here: the NULL rows' values are zero, and 0 passes the
predicates below.
"""
import math

import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen

X = abi.X


def nully_table(n, seed, all_null_rg=False):
    """lineitem with NULLs in a key (flagA), two aggregate arguments (quantity,
    price) and a predicate column (discount); quantity is NULL wherever flagA
    is, so the NULL-key group's SUM(quantity) has no value.  all_null_rg: the
    second row group's prices are all NULL."""
    t = datagen.lineitem(n, seed)
    rng = np.random.default_rng(seed)
    nul = {
        "flagA": rng.random(n) < 0.1,
        "price": rng.random(n) < 0.15,
        "discount": rng.random(n) < 0.05,
    }
    nul["quantity"] = (rng.random(n) < 0.2) | nul["flagA"]
    if all_null_rg and len(t.rg_rows) > 1:
        b = int(t.rg_rows[0])
        nul["price"][b:b + int(t.rg_rows[1])] = True
    for name, m in nul.items():
        i = t.index(name)
        t.columns[i] = datagen.with_nulls(t.columns[i], m)
    return t, {k: ~v for k, v in nul.items()}  # name -> holds a value


CUT = datagen.DATE_MIN + datagen.SHIP_DAYS - 1 - 200


def _fsum(v):
    return math.fsum(v) if len(v) else 0.0


# Each aggregate with its expected value on group mask m, from numpy:
# (builder, expected(d, ok, m) -> (holds a value, value)).
AGGS = [
    (lambda t: abi.count_star(t), lambda d, ok, m: (True, m.sum())),
    (lambda t: abi.agg_count("price", t), lambda d, ok, m: (True, (m & ok["price"]).sum())),
    (lambda t: abi.agg_sum("quantity", t),
     lambda d, ok, m: ((m & ok["quantity"]).any(), d["quantity"][m & ok["quantity"]].sum())),
    (lambda t: abi.agg_sum("price", t),
     lambda d, ok, m: ((m & ok["price"]).any(), _fsum(d["price"][m & ok["price"]]))),
    (lambda t: abi.agg_avg("price", t),
     lambda d, ok, m: ((m & ok["price"]).any(),
                       _fsum(d["price"][m & ok["price"]]) / max(1, (m & ok["price"]).sum()))),
    (lambda t: abi.agg_min("quantity", t),
     lambda d, ok, m: ((m & ok["quantity"]).any(), d["quantity"][m & ok["quantity"]].min(initial=1 << 62))),
    (lambda t: abi.agg_max("price", t),
     lambda d, ok, m: ((m & ok["price"]).any(), d["price"][m & ok["price"]].max(initial=-1e300))),
    (lambda t: abi.agg_sum(X.col("price") * (1 - X.col("discount")), t),
     lambda d, ok, m: ((m & ok["price"]).any(),
                       _fsum(d["price"][m & ok["price"]] * (1 - d["discount"][m & ok["price"]])))),
    (lambda t: abi.agg_avg("discount", t), lambda d, ok, m: (True, _fsum(d["discount"][m]) / m.sum())),
    (lambda t: abi.agg_count(X.col("quantity") * X.col("price"), t),
     lambda d, ok, m: (True, (m & ok["price"] & ok["quantity"]).sum())),
    (lambda t: abi.agg_max("quantity", t),
     lambda d, ok, m: ((m & ok["quantity"]).any(), d["quantity"][m & ok["quantity"]].max(initial=-1))),
]
# two queries of at most VQ_MAX_AGGS aggregates between them cover the list
AGG_SETS = [list(range(0, 8)), [0, 8, 9, 10, 2, 1]]


def nully_query(t, which=0):
    pred = abi.make_pred(t, [("discount", ">=", 0.02), ("shipdate", "<=", CUT)])
    idx = AGG_SETS[which]
    return pred, [t.index("flagA"), t.index("flagB")], [AGGS[i][0](t) for i in idx], idx


def brute_groups(t, ok, keys=True):
    """numpy: the groups of the passing rows in first-occurrence order, as
    (flagA key or -1 for NULL, flagB key, row mask)."""
    d = {c.name: c.data for c in t.columns}
    passing = ok["discount"] & (d["discount"] >= 0.02) & (d["shipdate"] <= CUT)
    if not keys:
        return [(0, 0, passing)], d
    ka = np.where(ok["flagA"], d["flagA"].astype(np.int64), -1)
    kb = d["flagB"].astype(np.int64)
    rows = np.nonzero(passing)[0]
    packed = (ka[rows] + 1) * 4 + kb[rows]
    uniq, first = np.unique(packed, return_index=True)
    groups = []
    for u in uniq[np.argsort(first)]:
        m = np.zeros(t.nrows, dtype=bool)
        m[rows[packed == u]] = True
        groups.append(((int(u) // 4) - 1, int(u) % 4, m))
    return groups, d


def check_against_brute(r, t, ok, idx, keys=True):
    groups, d = brute_groups(t, ok, keys)
    assert r.ngroups == len(groups)
    for g, (ka, kb, m) in enumerate(groups):
        if keys:
            assert (r.keys[g, 0], r.keys[g, 1]) == (ka, kb), g
        assert r.counts[g] == m.sum()
        assert r.first_row[g] == np.nonzero(m)[0][0]
        for j, i in enumerate(idx):
            has, want = AGGS[i][1](d, ok, m)
            assert bool(r.valid[g, j]) == bool(has), (g, i)
            if not has:
                continue
            a = AGGS[i][0](t)
            fl = a.kind == abi.VQ_AGG_AVG or (a.kind != abi.VQ_AGG_COUNT and a.accum == abi.VQ_F64)
            got = r.fvals[g, j] if fl else r.ivals[g, j]
            assert got == want, (g, i, got, want)


@pytest.mark.parametrize("n,seed", [(3000, 1), (70_000, 2)])
@pytest.mark.parametrize("which", [0, 1])
def test_oracle_group_by_with_nulls_matches_numpy(n, seed, which):
    t, ok = nully_table(n, seed)
    pred, keys, aggs, idx = nully_query(t, which)
    r = O.agg_dict(t, pred, keys, aggs, exact=True)
    check_against_brute(r, t, ok, idx)
    g = int(np.nonzero(r.keys[:, 0] == -1)[0][0])  # the NULL key's group is there
    assert r.valid[g, idx.index(2)] == 0  # no quantity under a NULL key
    # the global aggregate (no GROUP BY) of the same rows
    r = O.agg_dict(t, pred, [], aggs, exact=True)
    check_against_brute(r, t, ok, idx, keys=False)


def test_oracle_parallel_agrees_with_serial_on_nulls():
    t, ok = nully_table(200_000, 3)
    pred, keys, aggs, _ = nully_query(t)
    a = O.agg_dict(t, pred, keys, aggs, exact=False)
    b = O.agg_dict(t, pred, keys, aggs, nthreads=4)
    assert a.ngroups == b.ngroups
    np.testing.assert_array_equal(a.keys, b.keys)
    np.testing.assert_array_equal(a.counts, b.counts)
    np.testing.assert_array_equal(a.valid, b.valid)
    np.testing.assert_array_equal(a.ivals, b.ivals)
    np.testing.assert_allclose(a.fvals, b.fvals, rtol=1e-12)


def test_oracle_global_aggregate_all_null_argument():
    """Every price NULL: SUM/AVG/MAX(price) NULL, COUNT(price) 0, COUNT(*) the
    rows."""
    t = datagen.lineitem(5000, 4)
    i = t.index("price")
    t.columns[i] = datagen.with_nulls(t.columns[i], np.ones(t.nrows, dtype=bool))
    aggs = [abi.count_star(t), abi.agg_count("price", t), abi.agg_sum("price", t),
            abi.agg_avg("price", t), abi.agg_max("price", t), abi.agg_sum("quantity", t)]
    r = O.agg_dict(t, None, [], aggs)
    assert r.ngroups == 1
    assert list(r.valid[0]) == [1, 1, 0, 0, 0, 1]
    assert r.ivals[0, 0] == 5000 and r.ivals[0, 1] == 0
    assert r.ivals[0, 5] == t.col("quantity").sum()


def test_oracle_filter_with_nulls_matches_numpy():
    t, ok = nully_table(50_000, 5)
    d = {c.name: c.data for c in t.columns}
    cases = [
        ([("quantity", "<", 1000)], ok["quantity"]),
        ([("flagA", "in", [0, 2])], ok["flagA"] & np.isin(d["flagA"], [0, 2])),
        ([("discount", "<", X.col("tax"))], ok["discount"] & (d["discount"] < d["tax"])),
        ([("price", "!=", 1.5), ("quantity", ">=", 0)], ok["price"] & ok["quantity"]),
        ([("flagB", "==", 1)], d["flagB"] == 1),
    ]
    for terms, want in cases:
        got = O.filter_rows(t, abi.make_pred(t, terms))
        np.testing.assert_array_equal(got, np.nonzero(want)[0].astype(np.uint32), err_msg=str(terms))


def test_oracle_case_when_takes_else_over_null_then():
    """CASE WHEN flagB = 1 THEN price ELSE 0: where the WHEN fails the value
    is the ELSE literal (a value) even where price is NULL; where it holds,
    a NULL price is skipped."""
    t, ok = nully_table(20_000, 6)
    d = {c.name: c.data for c in t.columns}
    aggs = [abi.agg_count(X.case([("flagB", "==", 1)], X.col("price"), 0.0), t),
            abi.agg_sum(X.case([("flagB", "==", 1)], X.col("price"), 0.0), t)]
    r = O.agg_dict(t, None, [], aggs)
    w = d["flagB"] == 1
    assert r.ivals[0, 0] == (~w).sum() + (w & ok["price"]).sum()
    assert r.fvals[0, 1] == math.fsum(d["price"][w & ok["price"]])


def test_oracle_zone_maps_with_nulls():
    t, ok = nully_table(3 * abi.ROW_GROUP_ROWS // 2 if abi.ROW_GROUP_ROWS < 200_000 else 150_000, 7,
                        all_null_rg=True)
    z = O.zone_maps(t)
    nc = len(t.columns)
    base = 0
    for g, n in enumerate(t.rg_rows):
        n = int(n)
        for c, col in enumerate(t.columns):
            zz = z[g * nc + c]
            has = ok.get(col.name, np.ones(t.nrows, dtype=bool))[base:base + n]
            assert zz.null_count == n - has.sum()
            assert zz.has_minmax == int(has.any())
            if has.any() and col.name in ("quantity", "price"):
                v = col.data[base:base + n][has]
                lo, hi = v.min(), v.max()
                if col.type == abi.VQ_F64:
                    lo, hi = np.float64(lo).view(np.int64), np.float64(hi).view(np.int64)
                assert (np.int64(np.uint64(zz.min_bits)), np.int64(np.uint64(zz.max_bits))) == (lo, hi)
        base += n
    if len(t.rg_rows) > 1:
        assert z[1 * nc + t.index("price")].has_minmax == 0
        # a row group with no value never prunes
        keep = O.zone_prune(t, z, abi.make_pred(t, [("price", ">", 1e12)]))
        assert keep[1] == 1 and keep[0] == 0
