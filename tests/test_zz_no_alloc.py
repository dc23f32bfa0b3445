"""vq_agg_dict_noalloc, the dictionary GROUP BY that allocates no device
memory, against vq_agg_dict (the entry point it mirrors) through the C ABI:
integers, keys, counts and first rows bit-exact, float64 sums within 1e-12
relative.  Two calls in a row, the second on a smaller input, so the shared
scratch array's leftovers from the first call cannot leak into the second."""
import ctypes as C

import numpy as np
import pytest

from vexq_amd import abi, datagen, gpu

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12


def _noalloc(ctx, dt, pred, keys, aggs, capacity=4096):
    """A direct C-ABI call of vq_agg_dict_noalloc."""
    L = gpu.lib()
    karr = (C.c_int32 * max(1, len(keys)))(*keys)
    arr = abi.agg_array(aggs)
    out = abi.AggOut(capacity, len(keys), len(aggs))
    st = abi.Stats()
    rc = L.vq_agg_dict_noalloc(ctx.h, dt.h, C.byref(pred) if pred is not None else None, len(keys),
                               karr, len(aggs), arr, C.byref(out.c), C.byref(st))
    assert rc == abi.VQ_OK, L.vq_last_error().decode()
    out.stats = st
    return out


def _assert_same(got, want, aggs):
    assert got.ngroups == want.ngroups
    np.testing.assert_array_equal(got.keys, want.keys)
    np.testing.assert_array_equal(got.counts, want.counts)
    np.testing.assert_array_equal(got.first_row, want.first_row)
    np.testing.assert_array_equal(got.valid, want.valid)
    for j, a in enumerate(aggs):
        fl = a.kind == abi.VQ_AGG_AVG or (a.kind != abi.VQ_AGG_COUNT and a.accum == abi.VQ_F64)
        if fl:
            v = want.valid[:, j].astype(bool)
            np.testing.assert_allclose(got.fvals[v, j], want.fvals[v, j], rtol=REL_TOL, atol=0,
                                       equal_nan=True)
        else:
            np.testing.assert_array_equal(got.ivals[:, j], want.ivals[:, j])


def test_noalloc_matches_agg_dict_on_two_calls(gpu_ctx):
    big = datagen.lineitem(1_000_000, seed=21)
    small = datagen.lineitem(70_001, seed=22)
    db, ds = gpu_ctx.upload(big), gpu_ctx.upload(small)
    try:
        for name, t, dt in (("1M", big, db), ("70K", small, ds)):
            print(f"[case] noalloc groupby {name}", flush=True)
            pred, keys, aggs = datagen.q_groupby(t)
            got = _noalloc(gpu_ctx, dt, pred, keys, aggs)
            want = gpu_ctx.agg_dict(dt, pred, keys, aggs)
            assert got.ngroups == 6
            _assert_same(got, want, aggs)
        # the global aggregate, MIN/MAX and an empty result, smaller still
        print("[case] noalloc variants", flush=True)
        X = abi.X
        t, dt = small, ds
        cases = [
            ([], [abi.count_star(t), abi.agg_sum(X.col("price") * X.col("discount"), t),
                  abi.agg_min("price", t), abi.agg_max("quantity", t)],
             abi.make_pred(t, [("quantity", "<", 24)])),
            ([t.index("flagA")], [abi.count_star(t), abi.agg_avg("tax", t)],
             abi.make_pred(t, [("quantity", ">", 99)])),
            ([], [abi.count_star(t), abi.agg_sum("price", t)], abi.make_pred(t, [("quantity", ">", 99)])),
            ([t.index("flagA")], [abi.agg_sum(X.case([("quantity", "<", 25)], X.col("price"), 0.0), t),
                                  abi.agg_sum(X.case([("flagB", "in", [0])], 1, 0), t)], None),
        ]
        for keys, aggs, pred in cases:
            _assert_same(_noalloc(gpu_ctx, dt, pred, keys, aggs), gpu_ctx.agg_dict(dt, pred, keys, aggs), aggs)
    finally:
        db.close()
        ds.close()


def test_noalloc_runs_where_the_arena_does_not_fit():
    """With a budget that holds the table and nothing more, vq_agg_dict cannot
    get its scratch arena; the allocation-free entry point still runs."""
    print("[case] noalloc budget", flush=True)
    t = datagen.lineitem_sorted(300_000, seed=23)
    w = {abi.VQ_U8: 1, abi.VQ_U16: 2, abi.VQ_DATE: 4, abi.VQ_I64: 8, abi.VQ_F64: 8}
    table_bytes = sum(((t.nrows * w[c.type] + 255) & ~255) + 256 for c in t.columns)
    with gpu.Context(0, budget_bytes=table_bytes) as ctx:
        dt = ctx.upload(t)
        try:
            pred, keys, aggs = datagen.q_scan_sum(t)
            with pytest.raises(gpu.GpuError) as e:
                ctx.agg_dict(dt, pred, keys, aggs)
            assert e.value.code == abi.VQ_E_BUDGET
            got = _noalloc(ctx, dt, pred, keys, aggs)
            assert ctx.bytes_in_use == table_bytes
            assert got.stats.rg_scanned < got.stats.rg_total
        finally:
            dt.close()
    with gpu.Context(0) as ctx:
        dt = ctx.upload(t)
        try:
            _assert_same(got, ctx.agg_dict(dt, pred, keys, aggs), aggs)
        finally:
            dt.close()
