"""Batches of independent queries (vq_batch_run): every instance's result is
the one it gets alone — bit for bit — whether the batch runs through a
captured graph (fresh, then replayed), through direct launches on the branch
streams, or after the autotuner has picked each class's radix bits, scatter
buffers and tiles (the accumulation is exact and order-free, so tuning
cannot change a result); a few instances are also checked against the
oracle."""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen

pytestmark = pytest.mark.gpu


def same(a, b):
    n = a.ngroups
    assert b.ngroups == n
    for f in ("keys", "counts", "first_row", "ivals", "valid"):
        np.testing.assert_array_equal(getattr(a, f)[:n], getattr(b, f)[:n])
    np.testing.assert_array_equal(a.fvals[:n].view(np.int64), b.fvals[:n].view(np.int64))


@pytest.fixture(scope="module")
def instances(gpu_ctx):
    specs = datagen.batch_specs(24, seed=5, min_rows=5_000, max_rows=300_000)
    made = []
    for sp in specs:
        if sp["kind"] == "hash":
            t = datagen.keyed_table(sp["rows"], sp["groups"], sp["key_type"], sp["skew"], sp["seed"])
            pred, key, aggs = datagen.q_keyed(t)
            made.append(("hash", (t,), (gpu_ctx.upload(t),), (pred, key, aggs)))
        else:
            b, p = datagen.join_instance(sp["build"], sp["ratio"], sp["skew"], sp["seed"])
            bp, pp, aggs = datagen.q_join_instance(b, p)
            made.append(("join", (b, p), (gpu_ctx.upload(b), gpu_ctx.upload(p)), (bp, pp, aggs)))
    yield made
    for _, _, dts, _ in made:
        for d in dts:
            d.close()


def batch_list(ctx, made):
    out = []
    for kind, _, dts, q in made:
        if kind == "hash":
            out.append(ctx.batch_hash(dts[0], q[0], q[1], q[2]))
        else:
            out.append(ctx.batch_join(dts[0], 0, q[0], dts[1], 0, q[1], q[2]))
    return out


def alone(ctx, made):
    res = []
    for kind, _, dts, q in made:
        if kind == "hash":
            res.append(ctx.agg_hash(dts[0], q[0], q[1], q[2]))
        else:
            res.append(ctx.join_agg(dts[0], 0, q[0], dts[1], 0, q[1], q[2]))
    return res


def test_batch_equals_alone(gpu_ctx, instances):
    print("[case] batch graph / replay / no-graph / tuned", flush=True)
    ref = alone(gpu_ctx, instances)
    for flags in (0, 0, abi.VQ_BATCH_NO_GRAPH, abi.VQ_BATCH_TUNE):
        got = gpu_ctx.batch_run(batch_list(gpu_ctx, instances), flags)
        for a, b in zip(ref, got):
            same(a, b)


def test_batch_instances_match_oracle(gpu_ctx, instances):
    got = gpu_ctx.batch_run(batch_list(gpu_ctx, instances))
    checked = 0
    for (kind, tabs, _, q), g in zip(instances, got):
        if kind == "hash" and tabs[0].nrows <= 150_000:
            ref = O.agg_hash(tabs[0], q[0], q[1], q[2])
        elif kind == "join" and tabs[1].nrows <= 150_000:
            ref = O.join_agg(tabs[0], 0, q[0], tabs[1], 0, q[1], q[2])
        else:
            continue
        assert g.ngroups == ref.ngroups
        np.testing.assert_array_equal(g.keys, ref.keys)
        np.testing.assert_array_equal(g.counts, ref.counts)
        np.testing.assert_array_equal(g.ivals, ref.ivals)
        np.testing.assert_allclose(g.fvals, ref.fvals, rtol=1e-12, atol=0)
        checked += 1
    assert checked >= 4
