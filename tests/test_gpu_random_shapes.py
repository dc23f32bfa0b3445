"""Seeded random shapes through the hash GROUP BY and the join (key types,
cardinalities from one key to most rows, Zipf skew, build:probe ratios,
top-k or every group), each against the oracle — the paths, radix bits,
passes and replicas the planner picks vary with the shape."""
import math

import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen

from test_gpu_parity import _print_case, assert_agg_equal, same_bits

pytestmark = pytest.mark.gpu


def _log_uniform(r, lo, hi):
    return int(round(math.exp(math.log(lo) + r.random() * (math.log(hi) - math.log(lo)))))


@pytest.mark.parametrize("case", range(12))
def test_random_hash_groupby(gpu_ctx, case):
    r = np.random.RandomState(1000 + case)
    rows = _log_uniform(r, 1_000, 400_000)
    groups = max(1, min(rows, _log_uniform(r, 1, 50_000)))
    key_type = abi.VQ_U16 if groups <= 65536 and r.random() < 0.4 else abi.VQ_I64
    skew = 1.1 if r.random() < 0.5 else 0.0
    _print_case("random_hash", case=case, rows=rows, groups=groups, key_type=key_type, skew=skew)
    t = datagen.keyed_table(rows, groups, key_type, skew, seed=2000 + case)
    pred, key, aggs = datagen.q_keyed(t)
    dt = gpu_ctx.upload(t)
    try:
        got = gpu_ctx.agg_hash(dt, pred, key, aggs)
        again = gpu_ctx.agg_hash(dt, pred, key, aggs)
    finally:
        dt.close()
    assert_agg_equal(got, O.agg_hash(t, pred, key, aggs, exact=True), aggs)
    assert same_bits(got, again)


@pytest.mark.parametrize("case", range(12))
def test_random_join(gpu_ctx, case):
    r = np.random.RandomState(3000 + case)
    nbuild = _log_uniform(r, 1_000, 150_000)
    ratio = 0.5 + r.random() * 15.5
    skew = 1.1 if r.random() < 0.5 else 0.0
    topk = 10 if r.random() < 0.5 else 0
    _print_case("random_join", case=case, nbuild=nbuild, ratio=round(ratio, 2), skew=skew, topk=topk)
    b, p = datagen.join_instance(nbuild, ratio, skew, seed=4000 + case)
    bp, pp, aggs = datagen.q_join_instance(b, p)
    db, dp = gpu_ctx.upload(b), gpu_ctx.upload(p)
    try:
        got = gpu_ctx.join_agg(db, 0, bp, dp, 0, pp, aggs, topk=topk)
        again = gpu_ctx.join_agg(db, 0, bp, dp, 0, pp, aggs, topk=topk)
    finally:
        db.close()
        dp.close()
    assert_agg_equal(got, O.join_agg(b, 0, bp, p, 0, pp, aggs, topk=topk, exact=True), aggs)
    assert same_bits(got, again)
