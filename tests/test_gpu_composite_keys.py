"""Composite GROUP BY keys on the device (dictionary codes with integer
columns, packed into one key and grouped by the hash GROUP BY's kernels)
against the oracle; one integer key through the dictionary GROUP BY's entry
point equals vq_agg_hash; keys whose values do not fit the packed width are
refused.  Bars: keys, counts, first rows and integer sums bit-exact; float
sums within 1e-12 relative of the oracle's correctly rounded ones."""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, gpu

from test_composite_keys import composite_aggs, composite_table
from test_gpu_parity import _print_case, assert_agg_equal

pytestmark = pytest.mark.gpu

KEYSETS = [("shipdate", "flagA"), ("flagB", "small", "flagA"), ("small", "shipdate"),
           ("flagA", "flagB", "small", "shipdate"), ("flagA", "flagB", "small", "quantity", "flagA")]


@pytest.mark.parametrize("n", [1, 30_000, 700_000])
def test_composite_keys_match_oracle(gpu_ctx, n):
    _print_case("composite", n=n)
    t = composite_table(n, 5 + n % 3)
    aggs = composite_aggs(t)
    pred = abi.make_pred(t, [("discount", ">=", 0.04)])
    dt = gpu_ctx.upload(t)
    try:
        for keys in KEYSETS:
            kc = [t.index(k) for k in keys]
            for p in (None, pred):
                got = gpu_ctx.agg_dict(dt, p, kc, aggs, capacity=max(1, n))
                assert_agg_equal(got, O.agg_dict(t, p, kc, aggs, exact=True), aggs)
    finally:
        dt.close()


def test_one_integer_key_is_the_hash_group_by(gpu_ctx):
    _print_case("composite_one_key")
    t = composite_table(200_000, 8)
    aggs = composite_aggs(t)
    dt = gpu_ctx.upload(t)
    try:
        k = t.index("shipdate")
        a = gpu_ctx.agg_dict(dt, None, [k], aggs, capacity=100_000)
        b = gpu_ctx.agg_hash(dt, None, k, aggs)
        np.testing.assert_array_equal(a.keys, b.keys)
        np.testing.assert_array_equal(a.counts, b.counts)
        np.testing.assert_array_equal(a.ivals, b.ivals)
        np.testing.assert_array_equal(a.fvals.view(np.int64), b.fvals.view(np.int64))
    finally:
        dt.close()


def test_composite_keys_outside_width_are_refused(gpu_ctx):
    _print_case("composite_refused")
    t = composite_table(20_000, 9, small_max=70_000)
    aggs = composite_aggs(t)
    dt = gpu_ctx.upload(t)
    try:
        for kc in ([t.index("flagA"), t.index("flagB"), t.index("small")],
                   [0, 1, t.index("quantity"), 0, t.index("shipdate")],
                   [t.index("price"), 0]):
            with pytest.raises(gpu.GpuError) as e:
                gpu_ctx.agg_dict(dt, None, kc, aggs, capacity=20_000)
            assert e.value.code == abi.VQ_E_UNSUPPORTED, str(e.value)
    finally:
        dt.close()
