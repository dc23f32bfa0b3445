"""Pins the CPU oracle before anything is checked against it.

Each oracle function is checked against something it does not itself
compute: the reference's own known-answer vectors (tests/golden/kats.json,
transcribed from its Go tests), the reference's own AVX2/scalar filter kernel
built from its source (oracle/_ref, where the checkout was present),
math.fsum (a correctly rounded sum), numpy brute force, and algebraic
invariants (the murmur3 finaliser is a bijection whose inverse is fixed by its
constants).

  function                 pinned by
  or_radix_hash            inverse-permutation identity, the spread KAT
  or_radix_bits_for        TestRadixBitsFor table
  or_join_table_capacity   TestJoinTableCapacity properties
  or_exact_sum             math.fsum on hard inputs (cancellation, subnormals)
  or_zone_maps/prune       TestZoneMapPruning KAT, numpy min/max, raw-bit rule
  or_filter                reference AVX2/scalar kernel; numpy predicates
  or_agg_dict              int-key, SelVec, AVG-merge, closed-form KATs;
                           numpy bincount + fsum brute force
  or_join_agg              serial-join / shared-probe / multiplicity /
                           extreme-key KATs; Python dict brute force
  parallel baselines       agree with the serial restatement
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))
M64 = (1 << 64) - 1


def _table(cols, rg_rows):
    return datagen.Table([datagen.Column(n, t, np.ascontiguousarray(d), ds)
                          for n, t, d, ds in cols], np.array(rg_rows, dtype=np.int64))


# ---- radix hash / bits / capacity -------------------------------------------

def _unxorshift(x, s):
    y = x
    for _ in range(64 // s + 1):
        y = x ^ (y >> s)
    return y & M64


def test_radix_hash_inverts():
    """fmix64 = xs33, *C1, xs33, *C2, xs33; undo each step with the modular
    inverses of C1, C2.  A wrong shift or constant breaks the round trip."""
    c1, c2 = 0xFF51AFD7ED558CCD, 0xC4CEB9FE1A85EC53
    i1, i2 = pow(c1, -1, 1 << 64), pow(c2, -1, 1 << 64)
    rng = np.random.default_rng(0)
    keys = [0, 1, -1, 2**63 - 1, -2**63] + [int(k) for k in rng.integers(-2**63, 2**63 - 1, 500)]
    for k in keys:
        h = O.radix_hash(k)
        x = _unxorshift(h, 33)
        x = (x * i2) & M64
        x = _unxorshift(x, 33)
        x = (x * i1) & M64
        x = _unxorshift(x, 33)
        assert x == k & M64
    assert O.radix_hash(0) == 0


def test_radix_hash_spreads_sequential_keys():
    kat = KATS["radix_spread"]
    n, parts, step = kat["num_keys"], kat["num_parts"], kat["step"]
    counts = np.zeros(parts, dtype=np.int64)
    for i in range(n):
        counts[O.radix_hash(i * step) & (parts - 1)] += 1
    want = n // parts
    assert counts.min() >= want // 2 and counts.max() <= want * 2


def test_radix_bits_for_kat():
    for est, bits in KATS["radix_bits_for"]["cases"]:
        assert O.radix_bits_for(est) == bits, est


def test_join_table_capacity_properties():
    for keys in KATS["join_table_capacity_keys"]["keys"]:
        got = O.join_table_capacity(keys)
        assert got >= 8 and got & (got - 1) == 0
        assert keys <= got * 7 // 10
        half = got // 2
        if half >= 8:
            assert keys > half * 7 // 10


# ---- exact summation ----------------------------------------------------------

@pytest.mark.parametrize("case", range(8))
def test_exact_sum_matches_fsum(case):
    rng = np.random.default_rng(case)
    if case == 0:
        x = rng.uniform(0, 1e6, 10_000)
    elif case == 1:  # heavy cancellation
        a = rng.uniform(-1e16, 1e16, 5000)
        x = np.concatenate([a, -a[::-1], rng.uniform(-1, 1, 100)])
    elif case == 2:  # wide exponent range
        x = rng.standard_normal(5000) * 10.0 ** rng.integers(-300, 300, 5000)
    elif case == 3:  # subnormals
        x = rng.standard_normal(3000) * 5e-324 * 1e5
    elif case == 4:  # ties to even around 2^53
        x = np.array([2.0**53, 1.0, 1.0, -2.0**53, 1e-16, 3.0])
    elif case == 5:  # fsum overflows internally here; the exact sum is 1e308
        assert O.exact_sum(np.array([1e308, 1e308, -1e308])) == 1e308
        x = np.array([1e308, -1e308, 1e300, 5e-324])
    elif case == 6:
        x = np.array([])
    else:  # the workload's own per-row values
        t = datagen.lineitem(50_000)
        x = t.col("price") * (1 - t.col("discount")) * (1 + t.col("tax"))
    assert O.exact_sum(x) == math.fsum(x)


def test_exact_sum_specials():
    assert math.isnan(O.exact_sum([1.0, float("nan")]))
    assert O.exact_sum([1.0, float("inf")]) == float("inf")
    assert math.isnan(O.exact_sum([float("inf"), float("-inf")]))


GO_NAN_BITS = 0x7FF8000000000001  # Go's math.NaN() (uvnan)


def test_float_division_by_zero_is_go_nan():
    """x/0 in a float expression is math.NaN() (exec/expr.go:1016-1018):
    the bits 0x7FF8000000000001.  A SUM or AVG over such a row carries those
    bits (an IEEE sum returns its NaN operand), in both summation modes; a
    group with no zero divisor keeps a number."""
    t = datagen.lineitem(20_000, seed=5)
    X = abi.X
    aggs = [abi.agg_sum(X.col("price") / X.col("discount"), t),
            abi.agg_avg(X.col("price") / X.col("discount"), t)]
    zero_b = np.isin(np.arange(2), t.col("flagB")[t.col("discount") == 0])
    assert zero_b.any()
    for exact in (True, False):
        r = O.agg_dict(t, None, [t.index("flagB")], aggs, exact=exact)
        for i in range(r.ngroups):
            b = int(r.keys[i, 0])
            for j in range(2):
                bits = int(r.fvals[i:i + 1, j].view(np.uint64)[0])
                if zero_b[b]:
                    assert bits == GO_NAN_BITS, hex(bits)
                else:
                    assert not math.isnan(r.fvals[i, j])
    # a group whose divisors are all nonzero
    p = abi.make_pred(t, [("discount", ">", 0.0)])
    r = O.agg_dict(t, p, [], aggs, exact=True)
    assert not np.isnan(r.fvals[0]).any()
    nz = t.col("discount") > 0
    assert r.fvals[0, 0] == math.fsum(t.col("price")[nz] / t.col("discount")[nz])


# ---- zone maps and pruning ---------------------------------------------------

def test_zone_pruning_kat():
    kat = KATS["zone_pruning"]
    n = kat["row_groups"] * abi.ROW_GROUP_ROWS
    t = _table([("id", abi.VQ_I64, np.arange(n, dtype=np.int64), 0)],
               [abi.ROW_GROUP_ROWS] * kat["row_groups"])
    z = O.zone_maps(t)
    pred = abi.make_pred(t, [("id", ">=", kat["threshold"])])
    keep = O.zone_prune(t, z, pred)
    assert keep.tolist() == [0, 0, 1, 1]
    rows = O.filter_rows(t, pred, z)
    assert rows.size == kat["rows_read"]


def test_zone_maps_match_numpy():
    t = datagen.lineitem(200_000)
    z = O.zone_maps(t)
    nc = len(t.columns)
    start = 0
    for g, n in enumerate(t.rg_rows):
        for c, col in enumerate(t.columns):
            zz = z[g * nc + c]
            seg = col.data[start:start + n]
            assert zz.has_minmax == 1
            if col.dict_size:
                assert (zz.min_bits, zz.max_bits) == (0, col.dict_size - 1)
            elif col.type == abi.VQ_F64:
                assert zz.min_bits == seg.min().view(np.uint64)
                assert zz.max_bits == seg.max().view(np.uint64)
            elif col.type == abi.VQ_DATE:
                assert zz.min_bits == int(seg.min()) & 0xFFFFFFFF
                assert zz.max_bits == int(seg.max()) & 0xFFFFFFFF
            else:
                assert zz.min_bits == int(seg.min()) & M64
        start += n


def test_zone_prune_raw_bit_rule():
    """Float zones compare IEEE bits as signed int64 (planner/physical.go:708-737),
    so for non-negative floats pruning equals value pruning; NE never prunes;
    pruning never changes the filtered result."""
    t = datagen.lineitem_sorted(400_000, seed=5)
    z = O.zone_maps(t)
    sd = t.col("shipdate")
    for lo in (8100, 9000, 10400):
        p = abi.make_pred(t, [("shipdate", ">=", lo), ("shipdate", "<", lo + 100)])
        keep = O.zone_prune(t, z, p)
        starts = np.concatenate([[0], np.cumsum(t.rg_rows)[:-1]])
        for g, (s, n) in enumerate(zip(starts, t.rg_rows)):
            seg = sd[s:s + n]
            assert keep[g] == int(seg.max() >= lo and seg.min() < lo + 100)
        full = np.nonzero((sd >= lo) & (sd < lo + 100))[0]
        np.testing.assert_array_equal(O.filter_rows(t, p, z), full)
    p = abi.make_pred(t, [("price", ">", 50000.0)])
    keep = O.zone_prune(t, z, p)
    pr = t.col("price")
    starts = np.concatenate([[0], np.cumsum(t.rg_rows)[:-1]])
    for g, (s, n) in enumerate(zip(starts, t.rg_rows)):
        assert keep[g] == int(pr[s:s + n].max() > 50000.0)
    assert O.zone_prune(t, z, abi.make_pred(t, [("shipdate", "!=", 0)])).all()


# ---- filter -------------------------------------------------------------------

@pytest.mark.parametrize("n", KATS["filter_avx2_sizes"]["sizes"])
def test_filter_matches_reference_kernel(n):
    kat = KATS["filter_avx2_sizes"]
    data = O.xorshift_i64(n, kat["seed"])
    t = _table([("x", abi.VQ_I64, data, 0)], datagen.row_groups(n))
    got = O.filter_rows(t, abi.make_pred(t, [("x", ">", kat["threshold"])]))
    np.testing.assert_array_equal(got, np.nonzero(data > kat["threshold"])[0])
    if O.ref_filter_lib() is None:
        pytest.skip("oracle/_ref not built (reference checkout absent at build time)")
    np.testing.assert_array_equal(got, O.ref_filter_gt_i64(data, kat["threshold"], avx2=True))
    np.testing.assert_array_equal(got, O.ref_filter_gt_i64(data, kat["threshold"], avx2=False))


def test_filter_matches_numpy_predicates():
    t = datagen.lineitem(300_000, seed=9)
    d = {c.name: c.data for c in t.columns}
    cases = [
        ([("shipdate", "<=", 10470)], d["shipdate"] <= 10470),
        ([("discount", ">=", 0.05), ("discount", "<=", 0.07), ("quantity", "<", 24)],
         (d["discount"] >= 0.05) & (d["discount"] <= 0.07) & (d["quantity"] < 24)),
        ([("flagA", "!=", 2), ("tax", ">", 0.02), ("price", "==", 2000.0)],
         (d["flagA"] != 2) & (d["tax"] > 0.02) & (d["price"] == 2000.0)),
    ]
    for terms, mask in cases:
        np.testing.assert_array_equal(O.filter_rows(t, abi.make_pred(t, terms)), np.nonzero(mask)[0])


def q12_table(n, seed=9):
    """lineitem plus the columns of a Q12-shape predicate: commit and receipt
    dates around the ship date, and a ship mode code (7 values)."""
    t = datagen.lineitem(n, seed=seed)
    ship = t.columns[t.index("shipdate")].data.astype(np.int64)
    commit = ship + datagen.rand_int(seed, 90, n, -30, 30)
    receipt = ship + datagen.rand_int(seed, 91, n, 1, 30)
    mode = datagen.rand_int(seed, 92, n, 0, 6).astype(np.uint8)
    cols = list(t.columns) + [datagen.Column("commitdate", abi.VQ_DATE, commit.astype(np.int32)),
                              datagen.Column("receiptdate", abi.VQ_DATE, receipt.astype(np.int32)),
                              datagen.Column("shipmode", abi.VQ_U8, mode, 7)]
    return datagen.Table(cols, datagen.row_groups(n))


def test_filter_in_and_column_compares_match_numpy():
    """IN over codes (a mask over [0, 64)) and comparisons of two columns of
    the row (the Q12 shape: shipmode IN (...), commitdate < receiptdate,
    shipdate < commitdate), alone and with literal terms."""
    t = q12_table(200_000)
    d = {c.name: c.data for c in t.columns}
    X = abi.X
    cases = [
        ([("shipmode", "in", [1, 4])], np.isin(d["shipmode"], [1, 4])),
        ([("commitdate", "<", X.col("receiptdate"))], d["commitdate"] < d["receiptdate"]),
        ([("shipmode", "in", [0, 2, 6]), ("commitdate", "<", X.col("receiptdate")),
          ("shipdate", "<", X.col("commitdate")), ("receiptdate", ">=", 9000)],
         np.isin(d["shipmode"], [0, 2, 6]) & (d["commitdate"] < d["receiptdate"])
         & (d["shipdate"] < d["commitdate"]) & (d["receiptdate"] >= 9000)),
        ([("price", ">=", X.col("price"))], np.ones(t.nrows, bool)),
        ([("discount", ">", X.col("tax"))], d["discount"] > d["tax"]),
        ([("flagA", "in", [])], np.zeros(t.nrows, bool)),
        # a string literal absent from the dictionary is code -1 (StringEqExpr,
        # exec/expr.go:1342-1384: no row matches =, every row <>)
        ([("flagA", "=", -1)], np.zeros(t.nrows, bool)),
        ([("flagA", "<>", -1)], np.ones(t.nrows, bool)),
    ]
    for terms, mask in cases:
        np.testing.assert_array_equal(O.filter_rows(t, abi.make_pred(t, terms)), np.nonzero(mask)[0])


def case_aggs(t):
    """Aggregates over CASE WHEN <cond> THEN <fold> ELSE <literal> END:
    Q12's indicator sums (IN, and the NOT IN written as two <> terms),
    guarded float sums and averages, a guard comparing two columns, MIN/MAX
    under a guard.  Seven columns, so a GROUP BY on one more stays within a
    query's eight."""
    X = abi.X
    return [abi.count_star(t),
            abi.agg_sum(X.case([("flagA", "in", [0, 1])], 1, 0), t),
            abi.agg_sum(X.case([("flagA", "<>", 0), ("flagA", "<>", 1)], 1, 0), t),
            abi.agg_sum(X.case([("quantity", "<", 25)], X.col("price"), 0.0), t),
            abi.agg_avg(X.case([("commitdate", "<", X.col("receiptdate"))],
                               X.col("price") * (1 - X.col("discount")), -1.5), t),
            abi.agg_max(X.case([("shipdate", ">=", X.col("commitdate")), ("quantity", ">", 10)],
                               X.col("quantity"), -7), t),
            abi.agg_min(X.case([("discount", ">", 0.05)], X.col("discount") * 2, 1.0), t),
            abi.agg_sum(X.case([("commitdate", ">", X.col("shipdate"))], X.col("quantity") * 3 - 1, 100), t)]


def case_values_numpy(t):
    """The same arguments per row, computed with numpy from the columns."""
    d = {c.name: c.data for c in t.columns}
    q = d["quantity"].astype(np.int64)
    return [
        None,
        np.where(np.isin(d["flagA"], [0, 1]), 1, 0).astype(np.int64),
        np.where((d["flagA"] != 0) & (d["flagA"] != 1), 1, 0).astype(np.int64),
        np.where(q < 25, d["price"], 0.0),
        np.where(d["commitdate"] < d["receiptdate"], d["price"] * (1 - d["discount"]), -1.5),
        np.where((d["shipdate"] >= d["commitdate"]) & (q > 10), q, -7),
        np.where(d["discount"] > 0.05, d["discount"] * 2, 1.0),
        np.where(d["commitdate"] > d["shipdate"], q * 3 - 1, 100),
    ]


def test_case_when_aggregates_match_numpy():
    """CaseExpr with one WHEN and a literal ELSE (exec/expr.go:1386-1435)
    inside SUM/AVG/MIN/MAX, grouped by a dictionary column, against numpy:
    integer results exact, float sums and averages equal to math.fsum of the
    per-row values (correctly rounded, as the oracle's exact sums are)."""
    t = q12_table(60_000, seed=12)
    aggs = case_aggs(t)
    vals = case_values_numpy(t)
    key = t.index("shipmode")
    r = O.agg_dict(t, None, [key], aggs, exact=True)
    codes = t.columns[key].data
    assert r.ngroups == 7
    for g in range(r.ngroups):
        m = codes == r.keys[g, 0]
        assert r.counts[g] == int(m.sum())
        assert int(r.ivals[g, 1]) == int(vals[1][m].sum())
        assert int(r.ivals[g, 2]) == int(vals[2][m].sum())
        assert r.fvals[g, 3] == math.fsum(vals[3][m])
        assert r.fvals[g, 4] == math.fsum(vals[4][m]) / m.sum()
        assert int(r.ivals[g, 5]) == int(vals[5][m].max())
        assert r.fvals[g, 6] == vals[6][m].min()
        assert int(r.ivals[g, 7]) == int(vals[7][m].sum())
    # the indicator sums of Q12 partition the rows
    np.testing.assert_array_equal(r.ivals[:7, 1] + r.ivals[:7, 2], r.counts[:7])


def test_case_when_sum_equals_the_filtered_sum():
    """SUM(CASE WHEN c THEN x ELSE 0 END) over all rows is SUM(x) over the
    rows passing c -- the CASE path against the predicate path of the same
    oracle, for the dictionary, hash and join aggregates."""
    X = abi.X
    t = q12_table(50_000, seed=13)
    cond = [("shipmode", "in", [1, 3, 4]), ("commitdate", "<", X.col("receiptdate"))]
    arg = X.col("price") * (1 - X.col("discount"))
    guarded = [abi.agg_sum(X.case(cond, arg, 0.0), t), abi.agg_sum(X.case(cond, X.col("quantity"), 0), t)]
    plain = [abi.agg_sum(arg, t), abi.agg_sum("quantity", t)]
    key = t.index("flagA")
    a = O.agg_dict(t, None, [key], guarded, exact=True)
    b = O.agg_dict(t, abi.make_pred(t, cond), [key], plain, exact=True)
    for g in range(a.ngroups):
        j = list(b.keys[:, 0]).index(a.keys[g, 0])
        assert a.fvals[g, 0] == b.fvals[j, 0] and a.ivals[g, 1] == b.ivals[j, 1]
    ha = O.agg_hash(t, None, key, guarded, exact=True)
    hb = O.agg_hash(t, abi.make_pred(t, cond), key, plain, exact=True)
    for g in range(ha.ngroups):
        j = list(hb.keys[:, 0]).index(ha.keys[g, 0])
        assert ha.fvals[g, 0] == hb.fvals[j, 0] and ha.ivals[g, 1] == hb.ivals[j, 1]


def q12_join_tables(nb, npr, seed=16):
    return datagen.q12_tables(nb, npr, seed)


def q12_query(b, p, lo=9000, hi=9365):
    """TPC-H Q12's shape (datagen.q_q12: GROUP BY l_shipmode, the two CASE
    WHEN sums over o_orderpriority) plus a build-side MAX and an average under
    a guard comparing a build column with a probe column: eight columns, a
    query's limit."""
    X = abi.X
    v = abi.JoinView(p, b)
    pp, keys, aggs, urgent = datagen.q_q12(b, p, lo, hi)
    aggs = [abi.count_star(v)] + aggs + [
        abi.agg_max("b.orderdate", v),
        abi.agg_avg(X.case([("b.orderdate", "<", X.col("commitdate"))], X.col("quantity"), 0), v)]
    return pp, keys, aggs, urgent


def test_join_group_q12_matches_numpy():
    b, p = q12_join_tables(20_000, 80_000)
    pp, keys, aggs, urgent = q12_query(b, p)
    r = O.join_group(b, 0, None, p, 0, pp, keys, aggs, exact=True)
    # numpy: the unique build key 1 + 4 i is build row i
    d = {c.name: c.data for c in p.columns}
    modes = p.columns[p.index("shipmode")].dictionary
    m = (np.isin(d["shipmode"], [modes.index("MAIL"), modes.index("SHIP")]) & (d["commitdate"] < d["receiptdate"])
         & (d["shipdate"] < d["commitdate"])
         & (d["receiptdate"] >= 9000) & (d["receiptdate"] < 9365))
    rows = np.nonzero(m)[0]
    brow = (d["fk"][rows] - 1) // 4
    prio = b.col("priority")[brow]
    odate = b.col("orderdate")[brow].astype(np.int64)
    mode = d["shipmode"][rows]
    order = [int(x) for x in dict.fromkeys(mode.tolist())]  # first-occurrence order
    assert r.ngroups == len(order) and list(r.keys[:, 0]) == order
    for gi, g in enumerate(order):
        s = mode == g
        assert r.counts[gi] == int(s.sum())
        assert r.first_row[gi] == int(rows[s][0])
        hi = np.isin(prio[s], urgent)
        assert int(r.ivals[gi, 1]) == int(hi.sum()) and int(r.ivals[gi, 2]) == int((~hi).sum())
        assert int(r.ivals[gi, 3]) == int(odate[s].max())
        qv = np.where(odate[s] < d["commitdate"][rows][s], d["quantity"][rows][s], 0)
        assert r.fvals[gi, 4] == math.fsum(qv.astype(np.float64)) / s.sum()


def test_join_group_duplicate_build_keys_pure_python():
    """Build keys repeating (each probe row meets several build rows) and
    INT64_MIN / 0 keys: a pure-Python nested-loop join."""
    rng = np.random.default_rng(17)
    nb, npr = 300, 900
    bkey = rng.integers(0, 60, nb).astype(np.int64)
    bkey[:5] = np.iinfo(np.int64).min
    bval = rng.integers(-1000, 1000, nb).astype(np.int64)
    pkey = rng.integers(0, 70, npr).astype(np.int64)
    pkey[::97] = np.iinfo(np.int64).min
    code = rng.integers(0, 4, npr).astype(np.uint8)
    pval = rng.standard_normal(npr)
    b = _table([("k", abi.VQ_I64, bkey, 0), ("v", abi.VQ_I64, bval, 0)], [nb])
    p = _table([("k", abi.VQ_I64, pkey, 0), ("c", abi.VQ_U8, code, 4), ("x", abi.VQ_F64, pval, 0)], [npr])
    v = abi.JoinView(p, b)
    X = abi.X
    aggs = [abi.count_star(v), abi.agg_sum("b.v", v), abi.agg_sum(X.col("x") * 2, v),
            abi.agg_min(X.case([("b.v", ">", 0)], X.col("b.v"), 5000), v)]
    bp = abi.make_pred(b, [("v", "!=", 7)])
    r = O.join_group(b, 0, bp, p, 0, None, [1], aggs, exact=True)
    groups = {}
    for i in range(npr):
        for j in range(nb):
            if bkey[j] == pkey[i] and bval[j] != 7:
                g = groups.setdefault(int(code[i]), [0, 0, [], 5000, i])
                g[0] += 1
                g[1] += int(bval[j])
                g[2].append(pval[i] * 2)
                g[3] = min(g[3], int(bval[j]) if bval[j] > 0 else 5000)
    order = sorted(groups, key=lambda k: groups[k][4])
    assert list(r.keys[:, 0]) == order
    for gi, k in enumerate(order):
        c, sv, xs, mn, first = groups[k]
        assert (r.counts[gi], r.ivals[gi, 1], r.ivals[gi, 3], r.first_row[gi]) == (c, sv, mn, first)
        assert r.fvals[gi, 2] == math.fsum(xs)


def test_join_group_by_the_join_key_is_the_join_aggregate():
    """GROUP BY the probe's join key over the joined rows is or_join_agg's
    grouping: the two oracle paths agree bit for bit."""
    b, p = q12_join_tables(5_000, 20_000, seed=18)
    bp, pp, aggs = datagen.q_join(b, p)
    aggs = aggs + [abi.count_star(p), abi.agg_max("quantity", p)]
    a = O.join_group(b, 0, bp, p, 0, pp, [p.index("fk")], aggs, exact=True)
    j = O.join_agg(b, 0, bp, p, 0, pp, aggs, exact=True)
    assert a.ngroups == j.ngroups
    for f in ("keys", "counts", "first_row", "ivals", "valid"):
        np.testing.assert_array_equal(getattr(a, f), getattr(j, f))
    np.testing.assert_array_equal(a.fvals.view(np.int64), j.fvals.view(np.int64))


def test_nested_build_pure_python():
    """The nested build (planner/parallel_join.go:341-400) against a
    pure-Python nested-loop evaluation: inner keys repeated (a build row meets
    several inner rows and is that many rows of the outer build side), build
    keys repeated with the same riding columns, unmatched keys on every side."""
    rng = np.random.default_rng(41)
    ni, nb, npr = 120, 400, 1500
    ikey = rng.integers(0, 90, ni).astype(np.int64)
    iflag = rng.integers(0, 3, ni).astype(np.uint8)
    bfk = rng.integers(0, 100, nb).astype(np.int64)
    bkey = rng.integers(0, 300, nb).astype(np.int64)
    bdate = (9000 + (bkey * 7) % 50).astype(np.int32)  # a function of the key
    bval = rng.integers(0, 1000, nb).astype(np.int64)
    pkey = rng.integers(0, 320, npr).astype(np.int64)
    px = rng.standard_normal(npr)
    inner = _table([("k", abi.VQ_I64, ikey, 0), ("f", abi.VQ_U8, iflag, 3)], [ni])
    build = _table([("fk", abi.VQ_I64, bfk, 0), ("k", abi.VQ_I64, bkey, 0), ("d", abi.VQ_DATE, bdate, 0),
                    ("v", abi.VQ_I64, bval, 0)], [nb])
    probe = _table([("k", abi.VQ_I64, pkey, 0), ("x", abi.VQ_F64, px, 0)], [npr])
    ip = abi.make_pred(inner, [("f", "!=", 2)])
    bp = abi.make_pred(build, [("v", "<", 800)])
    aggs = [abi.count_star(probe), abi.agg_sum("x", probe), abi.agg_max("k", probe)]
    r, pay = O.join_agg_nested(inner, 0, ip, build, 0, 1, bp, probe, 0, None, [2], aggs, exact=True)
    nested = [b for b in range(nb) if bval[b] < 800 for i in range(ni) if ikey[i] == bfk[b] and iflag[i] != 2]
    groups = {}
    for p in range(npr):
        for b in nested:
            if bkey[b] == pkey[p]:
                g = groups.setdefault(int(pkey[p]), [0, [], p, int(bdate[b])])
                g[0] += 1
                g[1].append(px[p])
    order = sorted(groups, key=lambda k: groups[k][2])
    assert r.ngroups == len(order) and list(r.keys[:, 0]) == order
    for gi, k in enumerate(order):
        cnt, xs, first, d = groups[k]
        assert (r.counts[gi], r.ivals[gi, 0], r.first_row[gi], r.ivals[gi, 2], pay[gi, 0]) == (cnt, cnt, first, k, d)
        assert r.fvals[gi, 1] == math.fsum(xs)
    # the top 3 by SUM(x): the same groups, ranked
    t3, p3 = O.join_agg_nested(inner, 0, ip, build, 0, 1, bp, probe, 0, None, [2], aggs[1:2], topk=3, exact=True)
    want = sorted(order, key=lambda k: -math.fsum(groups[k][1]))[:3]
    assert list(t3.keys[:, 0]) == want and list(p3[:, 0]) == [groups[k][3] for k in want]


def test_nested_build_with_every_row_matched_once_is_the_join():
    """An inner side that every build row meets exactly once (unique keys, no
    predicate) leaves the build side as it is: or_join_agg's result."""
    c, o, li = datagen.q3_tables(2_000, 20_000, 60_000, seed=42)
    bp, pp, aggs = datagen.q_join(o, li)
    aggs = aggs + [abi.count_star(li)]
    r, _ = O.join_agg_nested(c, 0, None, o, o.index("custkey"), 0, bp, li, 0, pp, [], aggs, exact=True)
    j = O.join_agg(o, 0, bp, li, 0, pp, aggs, exact=True)
    assert r.ngroups == j.ngroups > 100
    for f in ("keys", "counts", "first_row", "ivals"):
        np.testing.assert_array_equal(getattr(r, f), getattr(j, f))
    np.testing.assert_array_equal(r.fvals.view(np.int64), j.fvals.view(np.int64))


# ---- dictionary-code GROUP BY -----------------------------------------------

def test_intkey_remap_kat():
    kat = KATS["intkey_remap"]
    t = _table([("flag", abi.VQ_U8, np.array(kat["flag"], np.uint8), 3),
                ("value", abi.VQ_I64, np.array(kat["value"], np.int64), 0)], kat["row_groups"])
    r = O.agg_dict(t, None, [0], [abi.agg_sum("value", t), abi.count_star(t)])
    got = {kat["dictionary"][k]: [int(r.ivals[i, 0]), int(r.ivals[i, 1])]
           for i, k in enumerate(r.keys[:, 0])}
    assert got == kat["expected"]
    assert [kat["dictionary"][k] for k in r.keys[:, 0]] == ["R", "A", "N"]  # first occurrence


def _selvec_table():
    kat = KATS["selvec_four_paths"]
    return kat, _table([
        ("flag", abi.VQ_U8, np.array(kat["flag"], np.uint8), 2),
        ("region", abi.VQ_U8, np.array(kat["region"], np.uint8), 3),
        ("price", abi.VQ_F64, np.array(kat["price"], np.float64), 0),
        ("qty", abi.VQ_I64, np.array(kat["qty"], np.int64), 0),
        ("sel", abi.VQ_U8, np.array(kat["sel"], np.uint8), 0)], kat["row_groups"])


def test_selvec_four_paths_kat():
    kat, t = _selvec_table()
    sel = abi.make_pred(t, [("sel", "==", 1)])
    r = O.agg_dict(t, sel, [], [abi.agg_sum("price", t), abi.count_star(t),
                                abi.agg_min("price", t), abi.agg_max("price", t)])
    assert [r.fvals[0, 0], r.ivals[0, 1], r.fvals[0, 2], r.fvals[0, 3]] == kat["expected_direct"]
    r = O.agg_dict(t, sel, [0], [abi.agg_sum("price", t), abi.count_star(t)])
    assert r.ngroups == 1 and r.keys[0, 0] == 0 and [r.fvals[0, 0], r.ivals[0, 1]] == [31, 5]
    r = O.agg_dict(t, sel, [1, 0], [abi.agg_sum("price", t), abi.count_star(t)])
    got = {f"{[1, 2, 9][k[0]]}|{'AB'[k[1]]}": [r.fvals[i, 0], r.ivals[i, 1]]
           for i, k in enumerate(r.keys)}
    assert got == kat["expected_by_region_flag"]


def test_avg_merge_kat():
    kat = KATS["avg_merge"]
    t = _table([("v", abi.VQ_F64, np.array(kat["value"]), 0)], kat["row_groups"])
    for nthreads in (0, 2):
        r = O.agg_dict(t, None, [], [abi.agg_avg("v", t)], nthreads=nthreads)
        assert r.fvals[0, 0] == kat["expected_avg"]


def test_closed_form_sum_kat():
    kat = KATS["closed_form_sum"]
    n = kat["row_groups"] * kat["rows_per_group"]
    t = _table([("id", abi.VQ_I64, np.arange(n, dtype=np.int64), 0)],
               [kat["rows_per_group"]] * kat["row_groups"])
    X = abi.X
    p = abi.make_pred(t, [("id", "<", kat["cutoff"])])
    for nthreads in (0, 1, 2, 4, 8):
        r = O.agg_dict(t, p, [], [abi.agg_sum(X.col("id") * 2 + 1, t)], nthreads=nthreads)
        assert r.ivals[0, 0] == kat["expected"]
        rf = O.agg_dict(t, p, [], [abi.agg_sum(X.col("id") * 2.0 + 1, t)], nthreads=nthreads)
        assert rf.fvals[0, 0] == kat["expected"]


def test_empty_global_aggregate():
    """buildEmptyGlobalResult: one row, COUNT 0, SUM/AVG/MIN NULL."""
    t = datagen.lineitem(1000)
    p = abi.make_pred(t, [("quantity", ">", 50)])
    aggs = [abi.count_star(t), abi.agg_sum("price", t), abi.agg_avg("tax", t), abi.agg_min("quantity", t)]
    r = O.agg_dict(t, p, [], aggs)
    assert r.ngroups == 1 and r.ivals[0, 0] == 0 and r.valid[0].tolist() == [1, 0, 0, 0]
    assert O.agg_dict(t, p, [0], aggs).ngroups == 0


def test_groupby_matches_numpy_brute_force():
    t = datagen.lineitem(120_000, seed=4)
    pred, keys, aggs = datagen.q_groupby(t)
    r = O.agg_dict(t, pred, keys, aggs, exact=True)
    d = {c.name: c.data for c in t.columns}
    m = d["shipdate"] <= datagen.DATE_MIN + datagen.SHIP_DAYS - 1 - 90
    price, disc, tax = d["price"][m], d["discount"][m], d["tax"][m]
    dp = price * (1 - disc)
    ch = dp * (1 + tax)
    gid = d["flagA"][m].astype(np.int64) * 2 + d["flagB"][m]
    idx = np.nonzero(m)[0]
    for i in range(r.ngroups):
        s = gid == r.keys[i, 0] * 2 + r.keys[i, 1]
        assert r.counts[i] == s.sum() and r.first_row[i] == idx[s][0]
        assert r.ivals[i, 0] == d["quantity"][m][s].sum()
        assert r.fvals[i, 1] == math.fsum(price[s])
        assert r.fvals[i, 2] == math.fsum(dp[s])
        assert r.fvals[i, 3] == math.fsum(ch[s])
        assert r.fvals[i, 4] == float(d["quantity"][m][s].sum()) / s.sum()
        assert r.fvals[i, 6] == math.fsum(disc[s]) / s.sum()
        assert r.ivals[i, 7] == s.sum()
    # first-occurrence order
    assert np.all(np.diff(r.first_row) > 0)


def test_naive_row_order_sum_is_the_left_fold():
    t = datagen.lineitem(5000, seed=2)
    r = O.agg_dict(t, None, [], [abi.agg_sum("price", t)], exact=False)
    acc = 0.0
    for v in t.col("price"):
        acc += float(v)
    assert r.fvals[0, 0] == acc


def test_parallel_baseline_agrees():
    t = datagen.lineitem(400_000, seed=8)
    pred, keys, aggs = datagen.q_groupby(t)
    ref = O.agg_dict(t, pred, keys, aggs, exact=True)
    for n in (1, 3, 8):
        r = O.agg_dict(t, pred, keys, aggs, nthreads=n)
        np.testing.assert_array_equal(r.keys, ref.keys)
        np.testing.assert_array_equal(r.ivals[:, [0, 7]], ref.ivals[:, [0, 7]])
        np.testing.assert_allclose(r.fvals, ref.fvals, rtol=1e-9)


def test_parallel_filter_baseline_agrees():
    """The morsel-parallel selection (the filter's CPU baseline) gives the
    serial selection vector, ascending, with pruning, for odd row groups."""
    t = datagen.lineitem_sorted(300_001, seed=3)
    t = datagen.Table(t.columns, np.array([1003] * 299 + [300_001 - 1003 * 299], dtype=np.int64))
    for w in range(3):
        pred, _, _ = datagen.q_scan_sum(t, which=w)
        np.testing.assert_array_equal(O.filter_rows(t, pred, nthreads=5), O.filter_rows(t, pred))
    np.testing.assert_array_equal(O.filter_rows(t, None, nthreads=3), np.arange(t.nrows, dtype=np.uint32))


def test_parallel_hash_baseline_agrees():
    """The morsel-parallel restatement of ParallelHashAggregate over one
    integer key (the hash GROUP BY's CPU baseline) against the serial one:
    the same groups in first-occurrence order, exact integers, float sums to
    the reference's own tolerance (1e-9)."""
    t = datagen.keyed_table(300_000, 4000, abi.VQ_I64, 1.1, seed=3)
    pred, key, aggs = datagen.q_keyed(t)
    ref = O.agg_hash(t, pred, key, aggs, exact=True)
    for n in (1, 4):
        r = O.agg_hash(t, pred, key, aggs, nthreads=n)
        assert r.ngroups == ref.ngroups
        for f in ("keys", "counts", "first_row", "ivals", "valid"):
            np.testing.assert_array_equal(getattr(r, f), getattr(ref, f))
        np.testing.assert_allclose(r.fvals, ref.fvals, rtol=1e-9)


def test_packed_keys_up_to_eight_columns():
    """The int-key path packs up to 8 dictionary columns (32 bits each for
    2, 16 for 4, 8 for 8: exec/agg_intkey.go:67-100): the GROUP BY over 2, 3
    and 5 code columns against a numpy brute force over the code tuples, and
    a 5-column GROUP BY whose codes exceed 8 bits is refused."""
    n = 60_000
    rng = np.random.default_rng(5)
    cards = [7, 300, 5, 11, 3]
    cols = [(f"c{i}", abi.VQ_U16, rng.integers(0, c, n).astype(np.uint16), c) if c > 255 else
            (f"c{i}", abi.VQ_U8, rng.integers(0, c, n).astype(np.uint8), c) for i, c in enumerate(cards)]
    cols.append(("x", abi.VQ_I64, rng.integers(-50, 50, n), 0))
    t = _table(cols, [n // 2, n - n // 2])
    aggs = [abi.count_star(t), abi.agg_sum("x", t)]
    for keys in ([0, 1], [0, 2, 3], [0, 2, 3, 4, 0]):
        r = O.agg_dict(t, None, keys, aggs, exact=True)
        tup = np.stack([t.columns[k].data.astype(np.int64) for k in keys], axis=1)
        uniq, first, inv = np.unique(tup, axis=0, return_index=True, return_inverse=True)
        order = np.argsort(first)
        assert r.ngroups == len(uniq)
        np.testing.assert_array_equal(r.keys, uniq[order])
        np.testing.assert_array_equal(r.first_row, first[order])
        counts = np.bincount(inv.ravel(), minlength=len(uniq))
        sums = np.bincount(inv.ravel(), weights=t.col("x"), minlength=len(uniq)).astype(np.int64)
        np.testing.assert_array_equal(r.counts, counts[order])
        np.testing.assert_array_equal(r.ivals[:, 1], sums[order])
    with pytest.raises(RuntimeError):
        O.agg_dict(t, None, [0, 1, 2, 3, 4], aggs)


# ---- join ------------------------------------------------------------------------

def _join_pair(bkeys, pkeys, bpay=None, pval=None):
    bpay = np.arange(len(bkeys), dtype=np.int64) if bpay is None else np.array(bpay, np.int64)
    pval = np.arange(len(pkeys), dtype=np.float64) + 0.25 if pval is None else pval
    b = _table([("k", abi.VQ_I64, np.array(bkeys, np.int64), 0), ("pay", abi.VQ_I64, bpay, 0)],
               [len(bkeys)])
    p = _table([("k", abi.VQ_I64, np.array(pkeys, np.int64), 0), ("x", abi.VQ_F64, pval, 0)],
               [len(pkeys)])
    return b, p


def test_serial_join_kat():
    kat = KATS["serial_join"]
    b, p = _join_pair(kat["build_key"], kat["probe_key"], kat["build_payload"])
    r = O.join_agg(b, 0, None, p, 0, None, [abi.count_star(p)])
    assert r.keys[:, 0].tolist() == kat["expected_group_order"]
    want = {int(k): len(v) for k, v in kat["expected_rows_by_key"].items()}
    assert {int(k): int(c) for k, c in zip(r.keys[:, 0], r.counts)} == want


def test_shared_probe_and_build_count_kats():
    kat = KATS["shared_probe"]
    b, p = _join_pair(kat["build_key"], kat["probe_key"], kat["build_payload"])
    r = O.join_agg(b, 0, None, p, 0, None, [abi.count_star(p)])
    assert {int(k): int(c) for k, c in zip(r.keys[:, 0], r.counts)} == \
        {int(k): len(v) for k, v in kat["expected_rows_by_key"].items()}
    kat = KATS["build_counts"]
    b, p = _join_pair(kat["build_key"], sorted(set(kat["build_key"])))
    r = O.join_agg(b, 0, None, p, 0, None, [abi.count_star(p)])
    assert r.counts.sum() == kat["num_rows"] and r.ngroups == kat["num_keys"]


def test_multiplicity_and_extreme_key_kats():
    mult = {int(k): v for k, v in KATS["multiplicities"]["mult"].items()}
    bkeys = []
    for i in range(17):
        for k in (100, 200, 300):
            if i < mult[k]:
                bkeys.append(k)
    b, p = _join_pair(bkeys, [100, 200, 300])
    r = O.join_agg(b, 0, None, p, 0, None, [abi.count_star(p), abi.agg_sum("x", p)])
    for i, k in enumerate(r.keys[:, 0]):
        assert r.counts[i] == mult[int(k)]
        assert r.fvals[i, 1] == mult[int(k)] * (list(mult).index(int(k)) + 0.25)
    kat = KATS["extreme_keys"]
    b, p = _join_pair(kat["present"], kat["present"] + kat["absent"])
    r = O.join_agg(b, 0, None, p, 0, None, [abi.count_star(p)])
    assert sorted(r.keys[:, 0].tolist()) == sorted(kat["present"]) and (r.counts == 1).all()


def test_join_matches_python_brute_force():
    b = datagen.orders(20_000, seed=6)
    p = datagen.lineitem_probe(b, 80_000, seed=6)
    bp, pp, aggs = datagen.q_join(b, p)
    r = O.join_agg(b, 0, bp, p, 0, pp, aggs, exact=True)
    bk = b.col("key")[b.col("orderdate") < datagen.JOIN_CUTOFF]
    mult = {}
    for k in bk:
        mult[int(k)] = mult.get(int(k), 0) + 1
    m = p.col("shipdate") > datagen.JOIN_CUTOFF
    vals = p.col("price") * (1 - p.col("discount"))
    groups, order = {}, []
    for i in np.nonzero(m)[0]:
        k = int(p.col("fk")[i])
        if k in mult:
            if k not in groups:
                groups[k] = []
                order.append(k)
            groups[k].extend([vals[i]] * mult[k])
    assert r.keys[:, 0].tolist() == order
    for i, k in enumerate(order):
        assert r.counts[i] == len(groups[k]) and r.fvals[i, 0] == math.fsum(groups[k])
    top = O.join_agg(b, 0, bp, p, 0, pp, aggs, topk=10, exact=True)
    ranked = sorted(range(len(order)), key=lambda i: (-r.fvals[i, 0], r.first_row[i]))[:10]
    assert top.keys[:, 0].tolist() == [order[i] for i in ranked]
    par = O.join_agg(b, 0, bp, p, 0, pp, aggs, topk=10, nthreads=4)
    assert par.keys[:, 0].tolist() == top.keys[:, 0].tolist()
    np.testing.assert_allclose(par.fvals, top.fvals, rtol=1e-9)


def test_join_empty_build():
    b, p = _join_pair([1, 2], [1, 2])
    r = O.join_agg(b, 0, abi.make_pred(b, [("pay", "<", 0)]), p, 0, None, [abi.count_star(p)])
    assert r.ngroups == 0


def test_agg_hash_matches_numpy_and_the_dictionary_path():
    """The general-key GROUP BY: first-seen group order, exact counts and
    sums against numpy; on dictionary codes it equals the int-key path."""
    import math
    t = datagen.keyed_table(60_000, 500, abi.VQ_I64, 1.1, seed=3)
    pred, key, aggs = datagen.q_keyed(t)
    r = O.agg_hash(t, pred, key, aggs)
    k = t.col("key")
    keep = t.col("shipdate") <= datagen.DATE_MIN + datagen.SHIP_DAYS - 1 - 90
    kk = k[keep]
    uniq, first = np.unique(kk, return_index=True)
    order = np.argsort(first)
    assert r.ngroups == len(uniq)
    np.testing.assert_array_equal(r.keys[:, 0], uniq[order])
    rows = np.nonzero(keep)[0]
    np.testing.assert_array_equal(r.first_row, rows[first[order]])
    qty = t.col("quantity")[keep]
    val = (t.col("price") * (1 - t.col("discount")))[keep]
    for g, kv in enumerate(uniq[order]):
        m = kk == kv
        assert r.counts[g] == m.sum()
        assert r.ivals[g, 1] == qty[m].sum()
        assert r.fvals[g, 0] == math.fsum(val[m])
    # dictionary codes: the two key paths agree
    li = datagen.lineitem(20_000, seed=5)
    p2, keys2, aggs2 = datagen.q_groupby(li)
    a = O.agg_dict(li, p2, [keys2[0]], aggs2)
    b = O.agg_hash(li, p2, keys2[0], aggs2)
    assert a.ngroups == b.ngroups
    np.testing.assert_array_equal(a.keys, b.keys)
    np.testing.assert_array_equal(a.counts, b.counts)
    np.testing.assert_array_equal(a.fvals.view(np.int64), b.fvals.view(np.int64))
