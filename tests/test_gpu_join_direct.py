"""The direct join (build keys whose zone maps bound them to a dense range:
one bit per key of the range, groups indexed by key) against the oracle and
against the partitioned join (VQ_NO_DIRECT=1) on the same inputs.

Bars: keys, counts, first rows and integer aggregates bit-exact, float sums
within 1e-12 relative of the oracle's exact sums and bit-identical to the
partitioned join's (both are exact fixed point rounded once); reruns and the
queries after one that left the region dirty (a repeated build key) give the
same bits.
"""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen

from test_gpu_parity import _print_case, assert_agg_equal, same_bits

pytestmark = pytest.mark.gpu


def _partitioned(monkeypatch, fn):
    monkeypatch.setenv("VQ_NO_DIRECT", "1")
    try:
        return fn()
    finally:
        monkeypatch.delenv("VQ_NO_DIRECT", raising=False)


def _all_aggs(p):
    X = abi.X
    return [abi.agg_sum(X.col("price") * (1 - X.col("discount")), p), abi.count_star(p),
            abi.agg_sum("quantity", p), abi.agg_avg("price", p), abi.agg_min("discount", p),
            abi.agg_max(X.col("quantity") * 2, p), abi.agg_min("quantity", p),
            abi.agg_max("price", p)]


@pytest.mark.parametrize("topk", [0, 10, 16, 40])
def test_direct_join_matches_oracle_and_partitions(gpu_ctx, monkeypatch, topk):
    _print_case("direct_join", topk=topk)
    b = datagen.orders(150_000, 3)
    p = datagen.lineitem_probe(b, 600_000, 3)
    bp, pp, _ = datagen.q_join(b, p)
    aggs = _all_aggs(p)
    db, dp = gpu_ctx.upload(b), gpu_ctx.upload(p)
    try:
        got = gpu_ctx.join_agg(db, 0, bp, dp, 0, pp, aggs, topk=topk)
        again = gpu_ctx.join_agg(db, 0, bp, dp, 0, pp, aggs, topk=topk)
        part = _partitioned(monkeypatch, lambda: gpu_ctx.join_agg(db, 0, bp, dp, 0, pp, aggs, topk=topk))
    finally:
        db.close()
        dp.close()
    assert got.stats.partitions == 0 and part.stats.partitions > 0
    ref = O.join_agg(b, 0, bp, p, 0, pp, aggs, topk=topk, exact=True)
    assert_agg_equal(got, ref, aggs)
    assert same_bits(got, again)
    assert same_bits(got, part)


def _keyed(build_keys, probe_keys, key_type=abi.VQ_I64, seed=1):
    rng = np.random.default_rng(seed)
    nb, npr = build_keys.size, probe_keys.size
    bt = datagen.Table([datagen.Column("k", key_type, build_keys),
                        datagen.Column("v", abi.VQ_I64, rng.integers(0, 100, nb).astype(np.int64))],
                       datagen.row_groups(nb))
    price = np.round(rng.uniform(1, 1000, npr), 2)
    pt = datagen.Table([datagen.Column("k", key_type, probe_keys),
                        datagen.Column("quantity", abi.VQ_I64, rng.integers(-50, 50, npr).astype(np.int64)),
                        datagen.Column("price", abi.VQ_F64, price),
                        datagen.Column("discount", abi.VQ_F64, rng.integers(0, 11, npr) / 100.0)],
                       datagen.row_groups(npr))
    return bt, pt


def test_direct_join_repeated_negative_and_outside_keys(gpu_ctx, monkeypatch):
    """Build keys in [-300, 300] with repeats (a key's build rows multiply its
    probe matches): the direct path counts the passing build keys, finds a
    repeat and runs the join through the partitions (and so the later joins
    over that build side).  Then unique keys from -20 up, probe keys inside
    and outside the range, on the same slot: the direct path, region clean."""
    rng = np.random.default_rng(7)
    bk = rng.integers(-300, 301, 900).astype(np.int64)
    pk = rng.integers(-400, 401, 50_000).astype(np.int64)
    bt, pt = _keyed(bk, pk)
    aggs = _all_aggs(pt)
    bpred = abi.make_pred(bt, [("v", "<", 70)])
    uk = np.arange(-20, 280, 3, dtype=np.int64)
    bt2, pt2 = _keyed(uk, rng.integers(-30, 300, 20_000).astype(np.int64), seed=2)
    aggs2 = _all_aggs(pt2)
    db, dp, db2, dp2 = gpu_ctx.upload(bt), gpu_ctx.upload(pt), gpu_ctx.upload(bt2), gpu_ctx.upload(pt2)
    try:
        got = []
        for topk in (0, 10, 0):
            _print_case("direct_join_repeats", topk=topk)
            got.append(gpu_ctx.join_agg(db, 0, bpred, dp, 0, None, aggs, topk=topk))
            ref = O.join_agg(bt, 0, bpred, pt, 0, None, aggs, topk=topk, exact=True)
            assert_agg_equal(got[-1], ref, aggs)
            assert got[-1].stats.partitions > 0
        assert same_bits(got[0], got[2])
        part = _partitioned(monkeypatch, lambda: gpu_ctx.join_agg(db, 0, bpred, dp, 0, None, aggs))
        assert same_bits(got[0], part)
        for _ in range(2):
            _print_case("direct_join_after_dirty")
            g2 = gpu_ctx.join_agg(db2, 0, None, dp2, 0, None, aggs2)
            assert_agg_equal(g2, O.join_agg(bt2, 0, None, pt2, 0, None, aggs2, exact=True), aggs2)
            assert g2.stats.partitions == 0
    finally:
        for t in (db, dp, db2, dp2):
            t.close()


@pytest.mark.parametrize("key_type", [abi.VQ_U16, abi.VQ_DATE])
def test_direct_join_narrow_keys(gpu_ctx, key_type):
    _print_case("direct_join_narrow", key_type=key_type)
    rng = np.random.default_rng(11)
    lo = 100 if key_type == abi.VQ_U16 else -5000
    bk = rng.permutation(np.arange(lo, lo + 3000)).astype(np.int64)
    pk = rng.integers(lo - 100, lo + 3100, 80_000).astype(np.int64)
    dt = np.uint16 if key_type == abi.VQ_U16 else np.int32
    bt, pt = _keyed(bk.astype(dt), pk.astype(dt), key_type)
    aggs = _all_aggs(pt)
    db, dp = gpu_ctx.upload(bt), gpu_ctx.upload(pt)
    try:
        got = gpu_ctx.join_agg(db, 0, abi.make_pred(bt, [("v", ">=", 10)]), dp, 0, None, aggs, topk=10)
    finally:
        db.close()
        dp.close()
    ref = O.join_agg(bt, 0, abi.make_pred(bt, [("v", ">=", 10)]), pt, 0, None, aggs, topk=10, exact=True)
    assert got.stats.partitions == 0
    assert_agg_equal(got, ref, aggs)


def test_direct_join_no_match_and_empty_build(gpu_ctx):
    """No probe row meets a passing build key, and a build predicate nothing
    passes: no groups, and the next query is unaffected."""
    b = datagen.orders(20_000, 5)
    p = datagen.lineitem_probe(b, 80_000, 5)
    aggs = _all_aggs(p)
    db, dp = gpu_ctx.upload(b), gpu_ctx.upload(p)
    try:
        _print_case("direct_join_empty")
        none = gpu_ctx.join_agg(db, 0, abi.make_pred(b, [("orderdate", "<", 0)]), dp, 0, None, aggs, topk=10)
        nomatch = gpu_ctx.join_agg(db, 0, None, dp, 0, abi.make_pred(p, [("shipdate", "<", 0)]), aggs)
        bp, pp, _ = datagen.q_join(b, p)
        after = gpu_ctx.join_agg(db, 0, bp, dp, 0, pp, aggs, topk=10)
    finally:
        db.close()
        dp.close()
    assert none.ngroups == 0 and nomatch.ngroups == 0
    assert none.stats.partitions == 0
    assert_agg_equal(after, O.join_agg(b, 0, bp, p, 0, pp, aggs, topk=10, exact=True), aggs)
