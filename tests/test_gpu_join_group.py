"""vq_join_group, the GROUP BY over joined rows (TPC-H Q12's shape: GROUP BY
a probe-side dictionary column, CASE WHEN sums over a build-side column),
against the oracle's serial HashJoin -> HashAggregate restatement: keys,
counts, first rows and integer results bit-exact, float sums within 1e-12
relative of the exact oracle (the device's sums are exact too)."""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from test_gpu_parity import assert_agg_equal, same_bits
from test_oracle import q12_join_tables, q12_query
from vexq_amd import abi, datagen, gpu

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[case] join_group {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def test_q12_shape(gpu_ctx):
    _case("q12", build=150_000, probe=600_000)
    b, p = q12_join_tables(150_000, 600_000, seed=21)
    pp, keys, aggs, _ = q12_query(b, p)
    db, dp = gpu_ctx.upload(b), gpu_ctx.upload(p)
    try:
        got = gpu_ctx.join_group(db, 0, None, dp, 0, pp, keys, aggs)
        assert got.ngroups == 2
        assert_agg_equal(got, O.join_group(b, 0, None, p, 0, pp, keys, aggs, exact=True), aggs)
        again = gpu_ctx.join_group(db, 0, None, dp, 0, pp, keys, aggs)
        assert same_bits(got, again)  # match order is free; the result is not
        # a build predicate, a wider date range, two group columns packed
        X = abi.X
        bp = abi.make_pred(b, [("orderdate", "<", 9500)])
        pp2 = abi.make_pred(p, [("commitdate", "<", X.col("receiptdate"))])
        got = gpu_ctx.join_group(db, 0, bp, dp, 0, pp2, [p.index("shipmode")], aggs)
        assert_agg_equal(got, O.join_group(b, 0, bp, p, 0, pp2, [p.index("shipmode")], aggs, exact=True), aggs)
    finally:
        db.close()
        dp.close()


def test_packed_group_columns_and_build_payload(gpu_ctx):
    """GROUP BY two probe code columns (packed as the int-key path packs
    them) over joined rows, aggregates reading build and probe columns."""
    _case("packed")
    b, p = q12_join_tables(40_000, 160_000, seed=22)
    cols = list(p.columns) + [datagen.Column("flag", abi.VQ_U8,
                                             datagen.rand_int(22, 96, p.nrows, 0, 2).astype(np.uint8), 3)]
    p = datagen.Table(cols, p.rg_rows)
    v = abi.JoinView(p, b)
    X = abi.X
    aggs = [abi.count_star(v), abi.agg_sum("b.orderdate", v),
            abi.agg_sum(X.case([("b.priority", "in", [0, 2, 4])], X.col("price"), 0.0), v),
            abi.agg_min("b.priority", v)]
    keys = [p.index("shipmode"), p.index("flag")]
    pp = abi.make_pred(p, [("quantity", "<", 30)])
    db, dp = gpu_ctx.upload(b), gpu_ctx.upload(p)
    try:
        got = gpu_ctx.join_group(db, 0, None, dp, 0, pp, keys, aggs)
        assert got.ngroups == 21
        assert_agg_equal(got, O.join_group(b, 0, None, p, 0, pp, keys, aggs, exact=True), aggs)
    finally:
        db.close()
        dp.close()


@pytest.mark.parametrize("keys", ["signed", "non_negative", "dense"])
def test_duplicate_and_extreme_build_keys(gpu_ctx, keys):
    """Build keys repeating up to dozens of times, probe keys with no match:
    every probe row meets every build row with its key.  signed: INT64_MIN and
    0 keys on both sides (two-array hash table); non_negative: keys below
    2^32 - 1 spread wide (one-word slots, repeated keys chained by
    compare-and-swap), probe keys at and past that bound; dense: keys in a
    range below four times the rows (a direct table, repeated keys chained in
    by its third pass), probe keys on both sides of the range."""
    _case("duplicates", keys=keys)
    rng = np.random.default_rng(23)
    nb, npr = 60_000, 250_000
    if keys == "signed":
        bkey = rng.integers(-500, 3000, nb).astype(np.int64)
        bkey[:40] = np.iinfo(np.int64).min
        bkey[40:90] = 0
        pkey = rng.integers(-600, 3100, npr).astype(np.int64)
        pkey[::131] = np.iinfo(np.int64).min
        pkey[::257] = 0
    elif keys == "dense":
        bkey = rng.integers(100, 3000, nb).astype(np.int64)
        pkey = rng.integers(0, 3200, npr).astype(np.int64)
    else:
        bkey = rng.integers(0, 3000, nb).astype(np.int64)
        bkey[:30] = 0xFFFFFFFE
        pkey = rng.integers(-100, 3100, npr).astype(np.int64)
        pkey[::113] = 0xFFFFFFFE
        pkey[::229] = 0xFFFFFFFF
        pkey[::331] = 1 << 40
    b = datagen.Table([datagen.Column("k", abi.VQ_I64, bkey), datagen.Column("v", abi.VQ_I64,
                       rng.integers(-10**6, 10**6, nb).astype(np.int64)),
                       datagen.Column("w", abi.VQ_F64, rng.standard_normal(nb))], datagen.row_groups(nb))
    p = datagen.Table([datagen.Column("k", abi.VQ_I64, pkey),
                       datagen.Column("c", abi.VQ_U8, rng.integers(0, 6, npr).astype(np.uint8), 6),
                       datagen.Column("x", abi.VQ_F64, rng.standard_normal(npr))], datagen.row_groups(npr))
    v = abi.JoinView(p, b)
    X = abi.X
    aggs = [abi.count_star(v), abi.agg_sum("b.v", v), abi.agg_sum(X.col("b.w") * X.col("x"), v),
            abi.agg_max(X.case([("b.v", ">", X.col("k"))], X.col("b.v"), -1), v), abi.agg_avg("x", v)]
    bp = abi.make_pred(b, [("v", "!=", 12345)])
    db, dp = gpu_ctx.upload(b), gpu_ctx.upload(p)
    try:
        got = gpu_ctx.join_group(db, 0, bp, dp, 0, None, [1], aggs)
        assert_agg_equal(got, O.join_group(b, 0, bp, p, 0, None, [1], aggs, exact=True), aggs)
        # GROUP BY an integer probe column of a few thousand values (the key itself)
        got = gpu_ctx.join_group(db, 0, bp, dp, 0, None, [0], aggs)
        assert_agg_equal(got, O.join_group(b, 0, bp, p, 0, None, [0], aggs, exact=True), aggs)
        # a build predicate that keeps nothing: no joined rows, no groups
        none = abi.make_pred(b, [("v", ">", 10**7)])
        assert gpu_ctx.join_group(db, 0, none, dp, 0, None, [1], aggs).ngroups == 0
    finally:
        db.close()
        dp.close()


def test_by_the_join_key_equals_join_agg(gpu_ctx):
    """GROUP BY the probe's join key over the joined rows is vq_join_agg's
    grouping: the two entry points agree bit for bit."""
    _case("vs_join_agg")
    b, p = q12_join_tables(4_000, 16_000, seed=24)
    bp, pp, aggs = datagen.q_join(b, p)
    aggs = aggs + [abi.count_star(p), abi.agg_max("quantity", p)]
    db, dp = gpu_ctx.upload(b), gpu_ctx.upload(p)
    try:
        a = gpu_ctx.join_group(db, 0, bp, dp, 0, pp, [p.index("fk")], aggs)
        j = gpu_ctx.join_agg(db, 0, bp, dp, 0, pp, aggs)
        assert a.ngroups == j.ngroups > 100
        assert same_bits(a, j)
    finally:
        db.close()
        dp.close()


def test_errors(gpu_ctx):
    _case("errors")
    b, p = q12_join_tables(2_000, 8_000, seed=25)
    v = abi.JoinView(p, b)
    db, dp = gpu_ctx.upload(b), gpu_ctx.upload(p)
    try:
        agg = [abi.agg_sum("b.orderdate", v)]
        bad = abi.agg_array(agg)
        bad[0].arg.f[0].col = abi.VQ_BUILD_COL | 99
        with pytest.raises(gpu.GpuError) as e:
            gpu_ctx.join_group(db, 0, None, dp, 0, None, [p.index("shipmode")], bad)
        assert e.value.code == abi.VQ_E_ARG
        with pytest.raises(gpu.GpuError) as e:  # a float join key
            gpu_ctx.join_group(db, 0, None, dp, p.index("price"), None, [p.index("shipmode")], agg)
        assert e.value.code == abi.VQ_E_UNSUPPORTED
        with pytest.raises(gpu.GpuError) as e:  # a plain integer among several group columns
            gpu_ctx.join_group(db, 0, None, dp, 0, None, [p.index("quantity"), p.index("shipmode")], agg)
        assert e.value.code == abi.VQ_E_UNSUPPORTED
    finally:
        db.close()
        dp.close()
