"""The seeded generators produce the workload shapes they promise."""
import numpy as np

from vexq_amd import abi, datagen


def test_lineitem_shapes_and_ranges():
    t = datagen.lineitem(200_000)
    assert t.nrows == 200_000 and t.rg_rows.tolist() == [65536, 65536, 65536, 3392]
    a, b = t.col("flagA"), t.col("flagB")
    assert set(np.unique(a)) == {0, 1, 2} and set(np.unique(b)) == {0, 1}
    fa = np.bincount(a) / a.size
    # weights 1/4, 1/4, 1/2 (codes are in first-occurrence order)
    assert sorted(np.round(fa, 2).tolist()) == [0.25, 0.25, 0.5]
    q = t.col("quantity")
    assert q.min() == 1 and q.max() == 50
    p = t.col("price")
    assert np.all(np.abs(p * 100 - np.round(p * 100)) < 1e-6)
    assert np.all(p >= q * 9.0 - 0.01) and np.all(p <= q * 1049.5 + 0.01)
    assert set(np.round(t.col("discount") * 100).astype(int)) == set(range(11))
    assert set(np.round(t.col("tax") * 100).astype(int)) == set(range(9))
    sd = t.col("shipdate")
    assert sd.min() == datagen.DATE_MIN and sd.max() == datagen.DATE_MIN + datagen.SHIP_DAYS - 1


def test_first_occurrence_codes():
    t = datagen.lineitem(1000, seed=3)
    for name in ("flagA", "flagB"):
        c = t.col(name)
        firsts = [int(np.nonzero(c == k)[0][0]) for k in range(c.max() + 1)]
        assert firsts == sorted(firsts)


def test_generators_are_deterministic_and_sliceable():
    a = datagen.lineitem(5000, seed=9)
    b = datagen.lineitem(5000, seed=9)
    for ca, cb in zip(a.columns, b.columns):
        assert np.array_equal(ca.data, cb.data)
    tail = datagen.rand_u64(9, 3, 100, offset=4900)
    assert np.array_equal(tail, datagen.rand_u64(9, 3, 5000)[4900:])


def test_sorted_table_is_sorted():
    t = datagen.lineitem_sorted(100_000)
    assert np.all(np.diff(t.col("shipdate")) >= 0)


def test_join_tables():
    b = datagen.orders(30_000)
    p = datagen.lineitem_probe(b, 120_000)
    assert np.array_equal(b.col("key"), 1 + 4 * np.arange(30_000))
    keys, counts = np.unique(p.col("fk"), return_counts=True)
    assert keys.size == 30_000 and counts.min() >= 1 and counts.max() <= 7
    assert p.nrows == 120_000
    od = dict(zip(b.col("key").tolist(), b.col("orderdate").tolist()))
    delta = p.col("shipdate") - np.array([od[k] for k in p.col("fk").tolist()])
    assert delta.min() >= 1 and delta.max() <= 121


def test_queries_build():
    t = datagen.lineitem(1000)
    pred, keys, aggs = datagen.q_groupby(t)
    assert pred.nterms == 1 and len(aggs) == 8 and keys == [0, 1]
    assert [a.accum for a in aggs] == [abi.VQ_I64] + [abi.VQ_F64] * 6 + [abi.VQ_I64]
    ts = datagen.lineitem_sorted(10_000)
    p, k, a = datagen.q_scan_sum(ts)
    assert p.nterms == 5 and k == [] and a[0].arg.nfactors == 2
