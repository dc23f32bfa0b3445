"""MIN/MAX over strings (AccumType TypeString; minMaxString,
exec/aggregate.go:670-730): the oracle against Python's own bytes order, on
a dictionary whose first-occurrence codes disagree with the strings' order
at every turn — prefixes ("a" < "ab" < "abc"), case ("Z" < "a"), the empty
string, multi-byte UTF-8 ("z" < "é"), a byte above 0x7F — with NULLs in
the string column and in the group key.  A MIN of codes instead of strings,
a comparison that stops at the shorter string or compares signed bytes, or a
NULL read as code 0, gives a different string somewhere here."""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen

WORDS = ["mango", "a", "Zebra", "", "abc", "ab", "é", "z", "apple", "\x7f", "ünï", "Apple", "a b", "zz"]


def string_table(n, seed, nulls=True):
    rng = np.random.default_rng(seed)
    # first-occurrence codes: row i's code is its word's first appearance
    words = rng.integers(0, len(WORDS), n)
    names, codes = [], np.empty(n, np.uint8)
    seen = {}
    for i, w in enumerate(words):
        if w not in seen:
            seen[w] = len(names)
            names.append(WORDS[w])
        codes[i] = seen[w]
    grp = rng.integers(0, 5, n).astype(np.uint8)
    val = rng.integers(0, 100, n).astype(np.int64)
    cols = [datagen.Column("name", abi.VQ_U8, codes, len(names), names),
            datagen.Column("grp", abi.VQ_U8, grp, 5, [f"g{k}" for k in range(5)]),
            datagen.Column("val", abi.VQ_I64, val)]
    t = datagen.Table(cols, datagen.row_groups(n, 4096))
    ok = {"name": np.ones(n, bool), "grp": np.ones(n, bool)}
    if nulls:
        ok["name"] = rng.random(n) > 0.1
        ok["grp"] = rng.random(n) > 0.05
        # group 4: every name NULL, so its MIN/MAX are NULL
        ok["name"][grp == 4] = False
        for c in ("name", "grp"):
            i = t.index(c)
            t.columns[i] = datagen.with_nulls(t.columns[i], ~ok[c])
    return t, ok


def string_aggs(t):
    return [abi.agg_min_str("name", t), abi.agg_max_str("name", t), abi.count_star(t),
            abi.agg_sum("val", t)]


def check_brute(r, t, ok, pred_mask=None):
    names = t.columns[t.index("name")].dictionary
    codes = t.col("name").astype(np.int64)
    g = np.where(ok["grp"], t.col("grp").astype(np.int64), -1)
    m0 = np.ones(t.nrows, bool) if pred_mask is None else pred_mask
    assert r.ngroups == len(np.unique(g[m0]))
    for i in range(r.ngroups):
        m = m0 & (g == r.keys[i, 0])
        assert r.counts[i] == m.sum()
        have = m & ok["name"]
        strs = [names[c].encode() for c in codes[have]]
        assert bool(r.valid[i, 0]) == bool(strs) and bool(r.valid[i, 1]) == bool(strs)
        if strs:
            assert names[r.ivals[i, 0]].encode() == min(strs), i
            assert names[r.ivals[i, 1]].encode() == max(strs), i
        assert r.ivals[i, 3] == t.col("val")[m].sum()


def test_dictionary_is_not_in_byte_order():
    t, _ = string_table(2000, 1)
    names = t.columns[0].dictionary
    assert names != sorted(names, key=str.encode)


@pytest.mark.parametrize("n,seed", [(1, 2), (2000, 3), (50_000, 4)])
def test_oracle_string_min_max_matches_python(n, seed):
    t, ok = string_table(n, seed)
    aggs = string_aggs(t)
    check_brute(O.agg_dict(t, None, [1], aggs), t, ok)
    pred = abi.make_pred(t, [("val", "<", 50)])
    check_brute(O.agg_dict(t, pred, [1], aggs), t, ok, t.col("val") < 50)


def test_oracle_string_min_max_parallel_equals_serial():
    t, ok = string_table(300_000, 5)
    aggs = string_aggs(t)
    a = O.agg_dict(t, None, [1], aggs)
    b = O.agg_dict(t, None, [1], aggs, nthreads=4)
    np.testing.assert_array_equal(a.keys, b.keys)
    np.testing.assert_array_equal(a.ivals, b.ivals)
    np.testing.assert_array_equal(a.valid, b.valid)


def test_dict_rank_is_byte_order():
    names = WORDS
    rank = abi.dict_rank(names)
    assert sorted(range(len(names)), key=lambda c: rank[c]) == \
        sorted(range(len(names)), key=lambda c: names[c].encode())
    assert rank[names.index("")] == 0
    assert rank[names.index("a")] < rank[names.index("ab")] < rank[names.index("abc")]
    assert rank[names.index("Zebra")] < rank[names.index("a")]
    assert rank[names.index("zz")] < rank[names.index("é")]
