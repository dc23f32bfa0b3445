"""MIN/MAX over strings on the device against the oracle: the dictionary
GROUP BY and its global aggregates fold each row's rank in byte order (a
device column built from vq_table_dict_order, or from the dictionary a .vxq
table carries) and map the winning rank back to its code.  Bars: keys,
counts, codes and validity bit-exact."""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from oracle import vxq
from vexq_amd import abi, gpu

from test_gpu_parity import _print_case, assert_agg_equal
from test_string_minmax import WORDS, check_brute, string_aggs, string_table

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,seed,nulls", [(1, 2, False), (2000, 3, True), (50_000, 4, True),
                                          (1_200_000, 6, True), (300_000, 7, False)])
def test_string_min_max_matches_oracle(gpu_ctx, n, seed, nulls):
    _print_case("str_minmax", n=n, seed=seed, nulls=nulls)
    t, ok = string_table(n, seed, nulls)
    aggs = string_aggs(t)
    dt = gpu_ctx.upload(t)
    try:
        dt.dict_order(0)
        for keys in ([1], []):
            got = gpu_ctx.agg_dict(dt, None, keys, aggs)
            assert_agg_equal(got, O.agg_dict(t, None, keys, aggs), aggs)
        # two keys: the dense slots when the code space fits them, else the
        # packed-key hash kernels, which take neither NULLs nor string MIN/MAX
        try:
            got = gpu_ctx.agg_dict(dt, None, [1, 0], aggs)
            assert_agg_equal(got, O.agg_dict(t, None, [1, 0], aggs), aggs)
        except gpu.GpuError as e:
            assert e.code == abi.VQ_E_UNSUPPORTED and t.columns[0].dict_size > 1, str(e)
        pred = abi.make_pred(t, [("val", "<", 50)])
        got = gpu_ctx.agg_dict(dt, pred, [1], aggs)
        assert_agg_equal(got, O.agg_dict(t, pred, [1], aggs), aggs)
        check_brute(got, t, ok, t.col("val") < 50)
    finally:
        dt.close()


def test_order_reset_and_errors(gpu_ctx):
    """Without an order: VQ_E_UNSUPPORTED; a rank table that is not a
    permutation: VQ_E_ARG; a second order (reversed) replaces the first, so
    MIN becomes the old MAX; the hash GROUP BY refuses string MIN/MAX."""
    _print_case("str_minmax_errors")
    t, ok = string_table(20_000, 8, nulls=False)
    aggs = string_aggs(t)
    dt = gpu_ctx.upload(t)
    try:
        with pytest.raises(gpu.GpuError) as e:
            gpu_ctx.agg_dict(dt, None, [1], aggs)
        assert e.value.code == abi.VQ_E_UNSUPPORTED
        bad = np.zeros(t.columns[0].dict_size, np.uint32)
        with pytest.raises(gpu.GpuError) as e:
            dt.dict_order(0, bad)
        assert e.value.code == abi.VQ_E_ARG
        rank = abi.dict_rank(t.columns[0].dictionary)
        dt.dict_order(0, rank)
        a = gpu_ctx.agg_dict(dt, None, [1], aggs)
        dt.dict_order(0, len(rank) - 1 - rank)  # the reverse order
        b = gpu_ctx.agg_dict(dt, None, [1], aggs)
        np.testing.assert_array_equal(a.ivals[:, 0], b.ivals[:, 1])
        np.testing.assert_array_equal(a.ivals[:, 1], b.ivals[:, 0])
        with pytest.raises(gpu.GpuError) as e:
            gpu_ctx.agg_hash(dt, None, 2, aggs[:2])
        assert e.value.code == abi.VQ_E_UNSUPPORTED
    finally:
        dt.close()


def test_shared_pass_and_noalloc_with_string_min_max(gpu_ctx):
    _print_case("str_minmax_multi_noalloc")
    # (the allocation-free entry point takes no NULLs: a table without)
    for nulls in (True, False):
        t, ok = string_table(400_000, 9, nulls)
        aggs = string_aggs(t)
        preds = [abi.make_pred(t, [("val", "<", v)]) for v in (10, 50, 90)]
        dt = gpu_ctx.upload(t)
        try:
            if not nulls:
                with pytest.raises(gpu.GpuError) as e:  # no order yet, and this call cannot build one
                    gpu_ctx.agg_dict_noalloc(dt, None, [1], aggs)
                assert e.value.code == abi.VQ_E_UNSUPPORTED
            dt.dict_order(0)
            outs = gpu_ctx.agg_dict_multi(dt, preds, aggs)
            for p, r in zip(preds, outs):
                assert_agg_equal(r, O.agg_dict(t, p, [], aggs), aggs)
            if not nulls:
                r = gpu_ctx.agg_dict_noalloc(dt, None, [1], aggs)
                assert_agg_equal(r, O.agg_dict(t, None, [1], aggs), aggs)
        finally:
            dt.close()


def test_vxq_table_orders_its_own_dictionary(gpu_ctx):
    """A table ingested from a .vxq image carries its dictionaries: the first
    string MIN/MAX query orders them itself."""
    _print_case("str_minmax_vxq")
    rng = np.random.default_rng(11)
    n = 150_000
    words = [WORDS[i] for i in rng.integers(0, len(WORDS), n)]
    grp = [f"g{k}" for k in rng.integers(0, 4, n)]
    img = vxq.write_vxq([("name", vxq.T_STRING), ("grp", vxq.T_STRING)], [words, grp], [65536, 65536, n - 131072])
    vt = gpu_ctx.upload_vxq(img, [0, 1])
    try:
        names, gnames = vt.dictionary(0), vt.dictionary(1)
        from vexq_amd import datagen
        t = datagen.Table([datagen.Column("name", abi.VQ_U8, np.zeros(1, np.uint8), len(names), names),
                           datagen.Column("grp", abi.VQ_U8, np.zeros(1, np.uint8), len(gnames), gnames)],
                          np.array([n], np.int64))
        aggs = [abi.agg_min_str("name", t), abi.agg_max_str("name", t), abi.count_star(t)]
        got = gpu_ctx.agg_dict(vt, None, [1], aggs)
        for i in range(got.ngroups):
            g = gnames[int(got.keys[i, 0])]
            s = [w.encode() for w, gg in zip(words, grp) if gg == g]
            assert names[int(got.ivals[i, 0])].encode() == min(s)
            assert names[int(got.ivals[i, 1])].encode() == max(s)
            assert got.ivals[i, 2] == len(s)
    finally:
        vt.close()
