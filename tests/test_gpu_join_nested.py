"""vq_join_agg_nested, a join whose build side is itself a join (TPC-H Q3's
customer -> orders -> lineitem chain), against the oracle's serial nested
build: keys, counts, first rows, integer results and the columns that ride
with each group bit-exact, float sums within 1e-12 relative of the exact
oracle (the device's sums are exact too)."""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from test_gpu_parity import assert_agg_equal, same_bits
from vexq_amd import abi, datagen, gpu

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[case] join_nested {name} " + " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _run_both(ctx, tabs, dtabs, args, topk, capacity=None):
    c, o, li = tabs
    dc, do, dl = dtabs
    ik, ip, bfk, bk, bp, pk, pp, pay, aggs, _ = args
    got, gpay = ctx.join_agg_nested(dc, ik, ip, do, bfk, bk, bp, dl, pk, pp, pay, aggs, topk=topk,
                                    capacity=capacity)
    want, wpay = O.join_agg_nested(c, ik, ip, o, bfk, bk, bp, li, pk, pp, pay, aggs, topk=topk, exact=True,
                                   capacity=capacity)
    assert_agg_equal(got, want, aggs)
    np.testing.assert_array_equal(gpay, wpay)
    return got, gpay


def test_q3_shape(gpu_ctx):
    _case("q3", customers=50_000, orders=500_000, lineitem=2_000_000)
    tabs = datagen.q3_tables(50_000, 500_000, 2_000_000, seed=51)
    args = datagen.q_q3(*tabs)
    dtabs = [gpu_ctx.upload(t) for t in tabs]
    try:
        top, pay = _run_both(gpu_ctx, tabs, dtabs, args, topk=10)
        assert top.ngroups == 10 and pay.shape == (10, 2)
        full, _ = _run_both(gpu_ctx, tabs, dtabs, args, topk=0)
        assert full.ngroups > 1000
        again, pay2 = gpu_ctx.join_agg_nested(dtabs[0], args[0], args[1], dtabs[1], args[2], args[3], args[4],
                                              dtabs[2], args[5], args[6], args[7], args[8], topk=10)
        assert same_bits(top, again) and np.array_equal(pay, pay2)
    finally:
        for d in dtabs:
            d.close()


def test_inner_keys_repeated(gpu_ctx):
    """Inner keys on several inner rows: each build row is as many rows of
    the outer build side, and the probe's matches multiply."""
    _case("inner_multiplicity")
    rng = np.random.default_rng(52)
    ni, nb, npr = 40_000, 120_000, 500_000
    inner = datagen.Table([datagen.Column("k", abi.VQ_I64, rng.integers(0, 9_000, ni).astype(np.int64)),
                           datagen.Column("f", abi.VQ_U8, rng.integers(0, 3, ni).astype(np.uint8), 3)],
                          datagen.row_groups(ni))
    bkey = rng.integers(0, 60_000, nb).astype(np.int64)
    build = datagen.Table([datagen.Column("fk", abi.VQ_I64, rng.integers(0, 10_000, nb).astype(np.int64)),
                           datagen.Column("k", abi.VQ_I64, bkey),
                           datagen.Column("d", abi.VQ_DATE, (9000 + bkey % 97).astype(np.int32)),
                           datagen.Column("v", abi.VQ_I64, rng.integers(0, 1000, nb).astype(np.int64))],
                          datagen.row_groups(nb))
    probe = datagen.Table([datagen.Column("k", abi.VQ_I64, rng.integers(0, 61_000, npr).astype(np.int64)),
                           datagen.Column("x", abi.VQ_F64, rng.standard_normal(npr))], datagen.row_groups(npr))
    tabs = (inner, build, probe)
    args = (0, abi.make_pred(inner, [("f", "!=", 2)]), 0, 1, abi.make_pred(build, [("v", "<", 900)]), 0, None,
            [2], [abi.count_star(probe), abi.agg_sum("x", probe), abi.agg_min("k", probe)], 0)
    dtabs = [gpu_ctx.upload(t) for t in tabs]
    try:
        _run_both(gpu_ctx, tabs, dtabs, args, topk=0)
        _run_both(gpu_ctx, tabs, dtabs, args[:8] + ([abi.agg_sum("x", probe)], 0), topk=7)
    finally:
        for d in dtabs:
            d.close()


def test_edges(gpu_ctx):
    _case("edges")
    tabs = datagen.q3_tables(3_000, 30_000, 90_000, seed=53)
    c, o, li = tabs
    args = list(datagen.q_q3(*tabs))
    dtabs = [gpu_ctx.upload(t) for t in tabs]
    try:
        dc, do, dl = dtabs
        # no inner row passes: no nested rows, no groups
        none = abi.make_pred(c, [("mktsegment", "=", 99)])
        r, p = gpu_ctx.join_agg_nested(dc, args[0], none, do, args[2], args[3], args[4], dl, args[5], args[6],
                                       args[7], args[8], topk=10)
        assert r.ngroups == 0 and p.shape == (0, 2)
        # every order once (all customers, unique keys): vq_join_agg's result
        bp, pp, aggs = datagen.q_join(o, li)
        r, _ = gpu_ctx.join_agg_nested(dc, 0, None, do, o.index("custkey"), 0, bp, dl, 0, pp, [], aggs)
        j = gpu_ctx.join_agg(do, 0, bp, dl, 0, pp, aggs)
        assert r.ngroups == j.ngroups > 100 and same_bits(r, j)
        # a riding column that is not a function of the key: refused
        with pytest.raises(gpu.GpuError) as e:
            gpu_ctx.join_agg_nested(dc, 0, None, do, o.index("custkey"), o.index("priority"), None, dl,
                                    li.index("quantity"), None, [o.index("orderdate")], aggs)
        assert e.value.code == abi.VQ_E_UNSUPPORTED
        with pytest.raises(RuntimeError):
            O.join_agg_nested(c, 0, None, o, o.index("custkey"), o.index("priority"), None, li,
                              li.index("quantity"), None, [o.index("orderdate")], aggs)
    finally:
        for d in dtabs:
            d.close()
