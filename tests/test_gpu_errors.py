"""Failure paths on the GPU: the device-memory budget (VQ_E_BUDGET, the
signal for the planner to keep the reference's CPU operator) and results
left empty by a failed call."""
import numpy as np
import pytest

from vexq_amd import abi, datagen, gpu

pytestmark = pytest.mark.gpu


def _table_bytes(t):
    w = {abi.VQ_U8: 1, abi.VQ_U16: 2, abi.VQ_DATE: 4, abi.VQ_I64: 8, abi.VQ_F64: 8}
    return sum(((t.nrows * w[c.type] + 255) & ~255) + 256 for c in t.columns)


def test_budget_rejects_table_and_query_arena():
    print("[case] budget", flush=True)
    t = datagen.lineitem(200_000, seed=3)
    with gpu.Context(0, budget_bytes=1 << 20) as ctx:
        with pytest.raises(gpu.GpuError) as e:
            ctx.upload(t)
        assert e.value.code == abi.VQ_E_BUDGET
        assert ctx.bytes_in_use == 0
    # the table fits, a query's scratch arena does not
    with gpu.Context(0, budget_bytes=_table_bytes(t) + 4096) as ctx:
        dt = ctx.upload(t)
        assert ctx.bytes_in_use == _table_bytes(t)
        pred, keys, aggs = datagen.q_groupby(t)
        with pytest.raises(gpu.GpuError) as e:
            ctx.agg_dict(dt, pred, keys, aggs)
        assert e.value.code == abi.VQ_E_BUDGET
        X = abi.X
        hout = abi.AggOut(t.nrows, 1, 2)
        hout.c.ngroups = 5  # a stale count from an earlier use
        with pytest.raises(gpu.GpuError) as e:
            ctx.agg_hash(dt, None, t.index("quantity"),
                         [abi.count_star(t), abi.agg_sum(X.col("price"), t)], out=hout)
        assert e.value.code == abi.VQ_E_BUDGET and hout.ngroups == 0
        assert ctx.bytes_in_use == _table_bytes(t)
        dt.close()
        assert ctx.bytes_in_use == 0


def test_capacity_failure_reports_needed_rows():
    print("[case] capacity", flush=True)
    t = datagen.lineitem(50_000, seed=4)
    with gpu.Context(0) as ctx:
        dt = ctx.upload(t)
        try:
            X = abi.X
            aggs = [abi.count_star(t), abi.agg_sum(X.col("price"), t)]
            with pytest.raises(gpu.GpuError) as e:
                ctx.agg_hash(dt, None, t.index("quantity"), aggs, capacity=10)
            assert e.value.code == abi.VQ_E_ARG and e.value.needed == 50
            got = ctx.agg_hash(dt, None, t.index("quantity"), aggs, capacity=e.value.needed)
            assert got.ngroups == 50 and int(got.counts.sum()) == t.nrows
            pred, keys, daggs = datagen.q_groupby(t)
            with pytest.raises(gpu.GpuError) as e:
                ctx.agg_dict(dt, pred, keys, daggs, capacity=2)
            assert e.value.code == abi.VQ_E_ARG and e.value.needed == 6
        finally:
            dt.close()


def test_dictionary_column_compare_is_unsupported():
    print("[case] dict colcmp", flush=True)
    t = datagen.lineitem(10_000, seed=5)
    X = abi.X
    with gpu.Context(0) as ctx:
        dt = ctx.upload(t)
        try:
            p = abi.make_pred(t, [("flagA", "<", X.col("flagB"))])
            with pytest.raises(gpu.GpuError) as e:
                ctx.filter(dt, p)
            assert e.value.code == abi.VQ_E_UNSUPPORTED
        finally:
            dt.close()
