"""Columns with NULLs on the device (validity bitmaps uploaded beside the
values) against the oracle: the dictionary GROUP BY and its global
aggregates, the filter, zone maps computed at upload, and the serving path.
Every other operator refuses a query that reads a column with NULLs
(VQ_E_UNSUPPORTED, the reference's fall back to its Go operators).

Bars: keys (-1 for a NULL key), counts, first rows, validity flags, integer
results and selection vectors bit-exact; float sums and averages within
1e-12 relative of the oracle's correctly rounded ones."""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen, gpu

from test_gpu_parity import _print_case, assert_agg_equal
from test_nulls import AGG_SETS, nully_query, nully_table

pytestmark = pytest.mark.gpu
X = abi.X


@pytest.mark.parametrize("n", [1, 3000, 70_000, 1_100_000])
@pytest.mark.parametrize("which", [0, 1])
def test_group_by_with_nulls_matches_oracle(gpu_ctx, n, which):
    _print_case("nulls_groupby", n=n, which=which)
    t, _ = nully_table(n, 10 + n % 7, all_null_rg=True)
    pred, keys, aggs, _ = nully_query(t, which)
    dt = gpu_ctx.upload(t)
    try:
        for kc in (keys, keys[:1], []):
            got = gpu_ctx.agg_dict(dt, pred, kc, aggs)
            assert_agg_equal(got, O.agg_dict(t, pred, kc, aggs, exact=True), aggs)
        got = gpu_ctx.agg_dict(dt, None, keys, aggs)
        assert_agg_equal(got, O.agg_dict(t, None, keys, aggs, exact=True), aggs)
    finally:
        dt.close()


def test_zone_maps_with_nulls_match_oracle(gpu_ctx):
    _print_case("nulls_zones")
    t, _ = nully_table(200_000, 5, all_null_rg=True)
    dt = gpu_ctx.upload(t)
    try:
        z = dt.zones()
    finally:
        dt.close()
    want = O.zone_maps(t)
    for i in range(len(t.rg_rows) * len(t.columns)):
        a, b = z[i], want[i]
        assert (a.min_bits, a.max_bits, a.null_count, a.has_minmax) == \
            (b.min_bits, b.max_bits, b.null_count, b.has_minmax), i


@pytest.mark.parametrize("n", [7, 65_537, 500_000])
def test_filter_with_nulls_matches_oracle(gpu_ctx, n):
    _print_case("nulls_filter", n=n)
    t, _ = nully_table(n, 6)
    dt = gpu_ctx.upload(t)
    try:
        for terms in ([("quantity", "<", 1000)], [("flagA", "in", [0, 2])],
                      [("discount", "<", X.col("tax"))],
                      [("price", "!=", 1.5), ("quantity", ">=", 0), ("flagB", "==", 0)],
                      [("shipdate", ">", datagen.DATE_MIN + 500)]):
            p = abi.make_pred(t, terms)
            np.testing.assert_array_equal(gpu_ctx.filter(dt, p), O.filter_rows(t, p), err_msg=str(terms))
    finally:
        dt.close()


def test_shared_scans_with_nulls_run_alone(gpu_ctx):
    """vq_agg_dict_multi over a table with NULLs: each query as it runs alone."""
    _print_case("nulls_multi")
    t, _ = nully_table(300_000, 8)
    qs = [abi.make_pred(t, [("discount", ">=", d), ("shipdate", "<=", datagen.DATE_MIN + 2000)])
          for d in (0.0, 0.03, 0.07)]
    aggs = [abi.count_star(t), abi.agg_sum("price", t), abi.agg_count("quantity", t)]
    dt = gpu_ctx.upload(t)
    try:
        got = gpu_ctx.agg_dict_multi(dt, qs, aggs)
    finally:
        dt.close()
    for p, r in zip(qs, got):
        assert_agg_equal(r, O.agg_dict(t, p, [], aggs, exact=True), aggs)


def test_served_query_with_nulls(gpu_ctx):
    _print_case("nulls_serve")
    t, _ = nully_table(120_000, 9)
    pred, keys, aggs, _ = nully_query(t, 0)
    dt = gpu_ctx.upload(t)
    try:
        tk = [gpu_ctx.submit_agg_dict(dt, pred, keys, aggs) for _ in range(3)]
        outs = [k.wait() for k in tk]
    finally:
        dt.close()
    want = O.agg_dict(t, pred, keys, aggs, exact=True)
    for r in outs:
        assert_agg_equal(r, want, aggs)


def _unsupported(fn):
    with pytest.raises(gpu.GpuError) as e:
        fn()
    assert e.value.code == abi.VQ_E_UNSUPPORTED, str(e.value)


def test_other_operators_refuse_nulls(gpu_ctx, monkeypatch):
    """Queries that read a column with NULLs outside the dictionary GROUP BY
    and the filter: VQ_E_UNSUPPORTED, with no result.  A query on the same
    table that reads no such column runs."""
    _print_case("nulls_refused")
    t, _ = nully_table(50_000, 11)
    b = datagen.orders(20_000, seed=3)
    pr = datagen.lineitem_probe(b, 60_000, seed=3)
    b.columns[1] = datagen.with_nulls(b.columns[1], np.arange(b.nrows) % 5 == 0)
    dt, db, dp = gpu_ctx.upload(t), gpu_ctx.upload(b), gpu_ctx.upload(pr)
    try:
        sq = [abi.agg_sum("quantity", t)]
        _unsupported(lambda: gpu_ctx.agg_hash(dt, None, t.index("quantity"), [abi.count_star(t)]))
        _unsupported(lambda: gpu_ctx.agg_hash(dt, None, t.index("shipdate"), sq))
        # CASE over a column with NULLs
        case = [abi.agg_sum(X.case([("flagB", "==", 1)], X.col("price"), 0.0), t)]
        _unsupported(lambda: gpu_ctx.agg_dict(dt, None, [], case))
        # a join whose build predicate reads a column with NULLs
        pp = abi.make_pred(b, [("orderdate", "<", datagen.JOIN_CUTOFF)])
        _unsupported(lambda: gpu_ctx.join_agg(db, 0, pp, dp, 0, None, [abi.count_star(pr)]))
        # the same join with no predicate on that column runs
        r = gpu_ctx.join_agg(db, 0, None, dp, 0, None, [abi.count_star(pr)])
        assert r.ngroups > 0
        # the generic (not run-time compiled) kernels take no NULLs
        monkeypatch.setenv("VQ_NO_JIT", "1")
        _unsupported(lambda: gpu_ctx.agg_dict(dt, None, [], sq))
        monkeypatch.delenv("VQ_NO_JIT")
        # no column with NULLs read: runs
        ok = [abi.agg_sum("tax", t)]
        r = gpu_ctx.agg_hash(dt, None, t.index("shipdate"), ok)
        assert r.ngroups > 0
    finally:
        dt.close()
        db.close()
        dp.close()
