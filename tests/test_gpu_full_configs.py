"""The batched and the served workloads at their full size.

Batched (256 independent instances, rows log-uniform in [10K, 4M], half hash
GROUP BYs of 4..1M groups on u16/int64 keys, half joins of build:probe
1:2..1:16, Zipf(1.1) skew on every other one): every instance of one
captured batch against the exact oracle -- keys, counts and integers
bit-exact, float64 sums within 1e-12 relative.

Served (a 60M-row fact table and a 15M-row dimension, 240 queries: 50% range
scans, 30% GROUP BYs, 20% joins + top-10, random windows and cutoffs): every
query's result in the concurrent mix is bitwise the result it gets alone,
and one query of each kind matches the exact oracle."""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, datagen

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12
WORKERS = min(16, len(os.sched_getaffinity(0)))


def assert_matches_oracle(g, ref, aggs, what):
    assert g.ngroups == ref.ngroups, what
    np.testing.assert_array_equal(g.keys, ref.keys, err_msg=what)
    np.testing.assert_array_equal(g.counts, ref.counts, err_msg=what)
    np.testing.assert_array_equal(g.first_row, ref.first_row, err_msg=what)
    np.testing.assert_array_equal(g.valid, ref.valid, err_msg=what)
    for j, a in enumerate(aggs):
        fl = a.kind == abi.VQ_AGG_AVG or (a.kind != abi.VQ_AGG_COUNT and a.accum == abi.VQ_F64)
        if fl:
            v = ref.valid[:, j].astype(bool)
            np.testing.assert_allclose(g.fvals[v, j], ref.fvals[v, j], rtol=REL_TOL, atol=0,
                                       err_msg=f"{what} aggregate {j}")
        else:
            np.testing.assert_array_equal(g.ivals[:, j], ref.ivals[:, j], err_msg=f"{what} aggregate {j}")


def same_bits(a, b, what):
    n = a.ngroups
    assert b.ngroups == n, what
    for f in ("keys", "counts", "first_row", "ivals", "valid"):
        assert np.array_equal(getattr(a, f)[:n], getattr(b, f)[:n]), (what, f)
    assert np.array_equal(a.fvals[:n].view(np.int64), b.fvals[:n].view(np.int64)), what


def _make_instance(sp):
    if sp["kind"] == "hash":
        t = datagen.keyed_table(sp["rows"], sp["groups"], sp["key_type"], sp["skew"], sp["seed"])
        return ("hash", (t,), datagen.q_keyed(t))
    b, p = datagen.join_instance(sp["build"], sp["ratio"], sp["skew"], sp["seed"])
    return ("join", (b, p), datagen.q_join_instance(b, p))


def _oracle_instance(inst):
    kind, tabs, q = inst
    if kind == "hash":
        return O.agg_hash(tabs[0], q[0], q[1], q[2], exact=True)
    return O.join_agg(tabs[0], 0, q[0], tabs[1], 0, q[1], q[2], exact=True)


@pytest.mark.timeout(900)
def test_batch_256_instances_full_size_match_oracle(gpu_ctx):
    specs = datagen.batch_specs(256, max_rows=4_000_000)
    print(f"[case] batch 256 full size: {sum(s['rows'] for s in specs) / 1e6:.1f}M rows", flush=True)
    with cf.ThreadPoolExecutor(WORKERS) as ex:
        made = list(ex.map(_make_instance, specs))
        refs = ex.map(_oracle_instance, made)  # on the host while the GPU runs
        dts = []
        try:
            batch = []
            for kind, tabs, q in made:
                d = [gpu_ctx.upload(x) for x in tabs]
                dts.append(d)
                if kind == "hash":
                    batch.append(gpu_ctx.batch_hash(d[0], q[0], q[1], q[2]))
                else:
                    batch.append(gpu_ctx.batch_join(d[0], 0, q[0], d[1], 0, q[1], q[2]))
            print("[case] batch 256: run", flush=True)
            got = gpu_ctx.batch_run(batch)
            print("[case] batch 256: replay", flush=True)
            again = gpu_ctx.batch_run(batch)
        finally:
            for d in dts:
                for x in d:
                    x.close()
        print("[case] batch 256: oracle", flush=True)
        for i, (inst, g, g2, ref) in enumerate(zip(made, got, again, refs)):
            aggs = inst[2][2]
            what = f"instance {i} ({inst[0]}, {specs[i]['rows']} rows)"
            assert_matches_oracle(g, ref, aggs, what)
            same_bits(g, g2, what)


@pytest.mark.timeout(900)
def test_served_stream_full_size_alone_equals_mix(gpu_ctx):
    print("[case] serve full size: generating 60M x 15M", flush=True)
    dim = datagen.orders(15_000_000)
    fact = datagen.fact_serving(dim, 60_000_000)
    fk = fact.index("fk")
    stream = datagen.serve_stream(fact, dim, 240)
    assert {k for k, _ in stream} == {"scan", "groupby", "join"}
    ddim, dfact = gpu_ctx.upload(dim), gpu_ctx.upload(fact)

    def run(kind, a):
        if kind == "join":
            return gpu_ctx.join_agg(ddim, 0, a[0], dfact, fk, a[1], a[2], topk=10)
        return gpu_ctx.agg_dict(dfact, a[0], a[1], a[2])

    def submit(kind, a):
        if kind == "join":
            return gpu_ctx.submit_join_agg(ddim, 0, a[0], dfact, fk, a[1], a[2], topk=10)
        return gpu_ctx.submit_agg_dict(dfact, a[0], a[1], a[2])

    # one query of each kind against the oracle, on the host meanwhile
    firsts = {}
    for i, (kind, _) in enumerate(stream):
        firsts.setdefault(kind, i)

    def oracle(i):
        kind, a = stream[i]
        if kind == "join":
            return O.join_agg(dim, 0, a[0], fact, fk, a[1], a[2], topk=10, exact=True)
        return O.agg_dict(fact, a[0], a[1], a[2], exact=True)

    try:
        with cf.ThreadPoolExecutor(3) as ex:
            orefs = {kind: ex.submit(oracle, i) for kind, i in firsts.items()}
            print("[case] serve full size: alone", flush=True)
            gpu_ctx.set_slots(1)
            alone = [run(k, a) for k, a in stream]
            print("[case] serve full size: in the mix", flush=True)
            gpu_ctx.set_slots(4)
            try:
                tickets = [submit(k, a) for k, a in stream]
                mixed = [tk.wait() for tk in tickets]
            finally:
                gpu_ctx.set_slots(1)
            for i, (a, m) in enumerate(zip(alone, mixed)):
                same_bits(a, m, f"query {i} ({stream[i][0]})")
            print("[case] serve full size: oracle", flush=True)
            for kind, i in firsts.items():
                ref = orefs[kind].result()
                assert_matches_oracle(alone[i], ref, stream[i][1][2], f"{kind} query {i}")
                if kind != "join":
                    from test_gpu_parity import DENSE_MAX_ULPS, float_ulps
                    ulps = float_ulps(alone[i], ref, stream[i][1][2])
                    print(f"[measure] dense {kind} 60M rows: within {ulps} ulps of the exact", flush=True)
                    assert ulps <= DENSE_MAX_ULPS
    finally:
        ddim.close()
        dfact.close()
