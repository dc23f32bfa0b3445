"""GROUP BY composite keys — dictionary codes with integer columns (the
reference's string-key path, exec/aggregate.go:409-557, which builds each
row's key from every group column) — in the oracle, against numpy: groups
in first-occurrence order, keys reported as the codes and values
themselves, counts, integer sums exact and float sums the correctly rounded
ones.  An integer column packs its values as they are, so values outside
[0, 2^width) are refused (VQ_E_UNSUPPORTED) rather than folded together."""
import math

import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen


def composite_table(n, seed, day_span=60, small_max=9):
    t = datagen.lineitem(n, seed)
    rng = np.random.default_rng(seed)
    t.columns[t.index("shipdate")].data[:] = (datagen.DATE_MIN + rng.integers(0, day_span, n)).astype(np.int32)
    small = rng.integers(0, small_max + 1, n).astype(np.int64)
    t.columns.append(datagen.Column("small", abi.VQ_I64, small))
    return t


def brute(t, key_names, pred_mask=None):
    keys = np.stack([t.col(k).astype(np.int64) for k in key_names], axis=1)
    m0 = np.ones(t.nrows, bool) if pred_mask is None else pred_mask
    rows = np.nonzero(m0)[0]
    uniq, first, inv = np.unique(keys[rows], axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    out = []
    for g in np.argsort(first):
        m = np.zeros(t.nrows, bool)
        m[rows[inv == g]] = True
        out.append((tuple(uniq[g]), m))
    return out


def composite_aggs(t):
    return [abi.count_star(t), abi.agg_sum("quantity", t), abi.agg_sum("price", t), abi.agg_max("discount", t)]


def check(r, t, key_names, pred_mask=None):
    want = brute(t, key_names, pred_mask)
    assert r.ngroups == len(want)
    for i, (k, m) in enumerate(want):
        assert tuple(int(x) for x in r.keys[i]) == k, i
        assert r.counts[i] == m.sum()
        assert r.first_row[i] == np.nonzero(m)[0][0]
        assert r.ivals[i, 0] == m.sum()
        assert r.ivals[i, 1] == t.col("quantity")[m].sum()
        assert r.fvals[i, 2] == math.fsum(t.col("price")[m])
        assert r.fvals[i, 3] == t.col("discount")[m].max()


@pytest.mark.parametrize("keys", [("shipdate", "flagA"), ("flagB", "small", "flagA"), ("small", "shipdate"),
                                  ("flagA", "flagB", "small", "shipdate"),
                                  ("flagA", "flagB", "small", "quantity", "flagA")])
def test_oracle_composite_keys_match_numpy(keys):
    t = composite_table(30_000, 3)
    kc = [t.index(k) for k in keys]
    aggs = composite_aggs(t)
    check(O.agg_dict(t, None, kc, aggs), t, keys)
    pred = abi.make_pred(t, [("discount", ">=", 0.04)])
    check(O.agg_dict(t, pred, kc, aggs), t, keys, t.col("discount") >= 0.04)


def test_oracle_composite_keys_outside_width_are_refused():
    t = composite_table(5000, 4, small_max=70_000)
    kc = [t.index(k) for k in ("flagA", "flagB", "small")]  # 16 bits a column
    with pytest.raises(RuntimeError, match=str(abi.VQ_E_UNSUPPORTED)):
        O.agg_dict(t, None, kc, composite_aggs(t))
    # five columns: 8 bits each, and the days since 1970 need more
    with pytest.raises(RuntimeError, match=str(abi.VQ_E_UNSUPPORTED)):
        O.agg_dict(t, None, [0, 1, t.index("quantity"), 0, t.index("shipdate")], composite_aggs(t))
    t.columns[t.index("small")].data[0] = -1
    with pytest.raises(RuntimeError, match=str(abi.VQ_E_UNSUPPORTED)):
        O.agg_dict(t, None, [t.index("small"), 0], composite_aggs(t))
