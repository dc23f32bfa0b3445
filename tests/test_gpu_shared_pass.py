"""Several global aggregates of one shape in one pass (vq_agg_dict_multi, the
serving path's continuous batching of queued scans) against the same queries
run one by one (vq_agg_dict) and against the oracle.

Bars: counts, first rows and integer sums bit-exact; float sums bit-identical
to the query alone when the pass keeps the same row groups as the query does
(each wave folds the same rows in the same order), else within 1e-12
relative; the oracle's exact sums within 1e-12 relative."""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen

from test_gpu_parity import _print_case, assert_agg_equal, same_bits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 5, 16])
def test_shared_scans_equal_alone_unpruned(gpu_ctx, n):
    """Unsorted rows (nothing prunes): every query's words are the same bits
    in the pass as alone."""
    _print_case("shared_unpruned", n=n)
    t = datagen.lineitem(700_000, seed=3)
    qs = [datagen.q_scan_sum(t, seed=11, which=i) for i in range(n)]
    aggs = qs[0][2] + [abi.count_star(t), abi.agg_sum("quantity", t), abi.agg_min("price", t),
                       abi.agg_max("discount", t), abi.agg_avg("tax", t)]
    dt = gpu_ctx.upload(t)
    try:
        got = gpu_ctx.agg_dict_multi(dt, [p for p, _, _ in qs], aggs)
        alone = [gpu_ctx.agg_dict(dt, p, [], aggs) for p, _, _ in qs]
    finally:
        dt.close()
    for i, (p, _, _) in enumerate(qs):
        assert same_bits(got[i], alone[i]), i
        assert_agg_equal(got[i], O.agg_dict(t, p, [], aggs, exact=True), aggs)


def test_shared_scans_pruned_and_empty(gpu_ctx):
    """Rows sorted by shipdate in row groups the windows prune differently:
    the pass reads the union (a query's words within 1e-12 of alone); a
    window past the data (no row: the empty global result) and a query with
    no predicate in the same shape family run alone."""
    _print_case("shared_pruned")
    t = datagen.lineitem_sorted(1_200_000, seed=4)
    qs = [datagen.q_scan_sum(t, seed=21, which=i) for i in range(7)]
    X = abi.X
    none = abi.make_pred(t, [("shipdate", ">=", 0), ("shipdate", "<", 1), ("discount", ">=", 0.0),
                             ("discount", "<=", 1.0), ("quantity", "<", 24)])
    preds = [p for p, _, _ in qs] + [none]
    aggs = qs[0][2] + [abi.count_star(t), abi.agg_avg("price", t)]
    dt = gpu_ctx.upload(t)
    try:
        got = gpu_ctx.agg_dict_multi(dt, preds, aggs)
        alone = [gpu_ctx.agg_dict(dt, p, [], aggs) for p in preds]
        assert got[0].stats.rg_scanned > alone[0].stats.rg_scanned  # the union of the windows
    finally:
        dt.close()
    for i, p in enumerate(preds):
        ref = O.agg_dict(t, p, [], aggs, exact=True)
        assert_agg_equal(got[i], ref, aggs)
        assert_agg_equal(got[i], alone[i], aggs)
    assert got[-1].counts[0] == 0 and not got[-1].valid[0, 0] and got[-1].valid[0, 1]


def test_shared_pass_of_mixed_shapes_runs_one_by_one(gpu_ctx):
    """Predicates of different shapes cannot share a kernel: the call runs
    them one after another, with the same results."""
    t = datagen.lineitem(200_000, seed=5)
    p1, _, aggs = datagen.q_scan_sum(t, seed=1)
    p2 = abi.make_pred(t, [("quantity", "<", 10)])
    dt = gpu_ctx.upload(t)
    try:
        got = gpu_ctx.agg_dict_multi(dt, [p1, p2], aggs)
        alone = [gpu_ctx.agg_dict(dt, p, [], aggs) for p in (p1, p2)]
    finally:
        dt.close()
    assert same_bits(got[0], alone[0]) and same_bits(got[1], alone[1])
