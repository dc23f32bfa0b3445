"""Concurrent and asynchronous queries (vq_ctx_set_slots, vq_submit,
vq_wait): every query of a mixed stream returns the same result when it runs
alone (one slot, synchronous) and when it runs in the mix (several slots,
submitted together, overlapping on the device) — bit for bit, integers and
floats, since every kernel's result is independent of what else runs."""
import numpy as np
import pytest

from vexq_amd import datagen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serving_tables(gpu_ctx):
    dim = datagen.orders(150_000)
    fact = datagen.fact_serving(dim, 600_000)
    ddim, dfact = gpu_ctx.upload(dim), gpu_ctx.upload(fact)
    yield dim, fact, ddim, dfact
    ddim.close()
    dfact.close()


def run_sync(ctx, fact, dim, dfact, ddim, kind, args):
    if kind == "join":
        bp, pp, aggs = args
        return ctx.join_agg(ddim, 0, bp, dfact, fact.index("fk"), pp, aggs, topk=10)
    pred, keys, aggs = args
    return ctx.agg_dict(dfact, pred, keys, aggs)


def submit(ctx, fact, dim, dfact, ddim, kind, args):
    if kind == "join":
        bp, pp, aggs = args
        return ctx.submit_join_agg(ddim, 0, bp, dfact, fact.index("fk"), pp, aggs, topk=10)
    pred, keys, aggs = args
    return ctx.submit_agg_dict(dfact, pred, keys, aggs)


def same(a, b):
    n = a.ngroups
    assert b.ngroups == n
    assert np.array_equal(a.keys[:n], b.keys[:n])
    assert np.array_equal(a.counts[:n], b.counts[:n])
    assert np.array_equal(a.first_row[:n], b.first_row[:n])
    assert np.array_equal(a.ivals[:n], b.ivals[:n])
    assert np.array_equal(a.fvals[:n].view(np.int64), b.fvals[:n].view(np.int64))
    assert np.array_equal(a.valid[:n], b.valid[:n])


@pytest.mark.parametrize("slots,budget", [(4, 0), (8, 0), (4, 1)])
def test_alone_and_in_the_mix_agree(gpu_ctx, serving_tables, slots, budget):
    dim, fact, ddim, dfact = serving_tables
    stream = datagen.serve_stream(fact, dim, 24)
    kinds = {k for k, _ in stream}
    assert kinds == {"scan", "groupby", "join"}
    gpu_ctx.set_slots(1)
    alone = [run_sync(gpu_ctx, fact, dim, dfact, ddim, k, a) for k, a in stream]
    gpu_ctx.set_slots(slots, budget)  # budget 1 byte: one query at a time
    try:
        tickets = [submit(gpu_ctx, fact, dim, dfact, ddim, k, a) for k, a in stream]
        mixed = [t.wait() for t in tickets]
    finally:
        gpu_ctx.set_slots(1)
    for a, b in zip(alone, mixed):
        same(a, b)
    assert any(r.ngroups > 0 for r in alone)


def test_submit_reports_errors(gpu_ctx, serving_tables):
    dim, fact, ddim, dfact = serving_tables
    from vexq_amd import gpu
    pred, keys, aggs = datagen.q_groupby(fact)
    t = gpu_ctx.submit_agg_dict(dfact, pred, [fact.index("price")], aggs)  # float key
    with pytest.raises(gpu.GpuError):
        t.wait()


def test_hash_groupbys_alone_and_in_the_mix(gpu_ctx, serving_tables):
    """Submitted hash GROUP BYs (few keys: per-workgroup partials; many keys:
    partitions), overlapping with each other, give the bits they give alone."""
    from vexq_amd import abi
    X = abi.X
    dim, fact, ddim, dfact = serving_tables
    aggs = [abi.agg_sum(X.col("price") * (1 - X.col("discount")), fact), abi.count_star(fact),
            abi.agg_max("quantity", fact), abi.agg_avg("discount", fact)]
    cases = []
    for i, key in enumerate(["quantity", "fk", "flagA", "quantity", "fk", "flagB"]):
        pred = abi.make_pred(fact, [("shipdate", "<=", int(datagen.DATE_MIN + 400 + 300 * i))])
        cases.append((pred, fact.index(key)))
    gpu_ctx.set_slots(1)
    alone = [gpu_ctx.agg_hash(dfact, p, k, aggs) for p, k in cases]
    assert {r.stats.partitions > 0 for r in alone} == {True, False}  # both ways ran
    gpu_ctx.set_slots(4)
    try:
        mixed = [t.wait() for t in [gpu_ctx.submit_agg_hash(dfact, p, k, aggs) for p, k in cases]]
    finally:
        gpu_ctx.set_slots(1)
    for a, b in zip(alone, mixed):
        same(a, b)


def _outcome(ticket, alone):
    """A submitted query that was canceled, or that finished with the result
    it gives alone."""
    from vexq_amd import abi, gpu
    try:
        got = ticket.wait()
    except gpu.GpuError as e:
        assert e.code == abi.VQ_E_CANCELED, e
        return "canceled"
    same(got, alone)
    return "ok"


def test_cancel_queued_and_running(gpu_ctx, serving_tables):
    """Queries queued behind a many-key hash GROUP BY on one slot are
    canceled one by one (vq_ticket_cancel) and, in a second stream, all at
    once (vq_ctx_cancel): each either finishes canceled or, when it had
    already run, with its result alone; the queries not canceled all return
    their results."""
    from vexq_amd import abi
    dim, fact, ddim, dfact = serving_tables
    stream = datagen.serve_stream(fact, dim, 24, seed=31)
    heavy_aggs = [abi.agg_sum("price", fact), abi.count_star(fact)]
    gpu_ctx.set_slots(1)
    alone = [run_sync(gpu_ctx, fact, dim, dfact, ddim, k, a) for k, a in stream]
    heavy = gpu_ctx.agg_hash(dfact, None, fact.index("fk"), heavy_aggs)
    outcomes = []
    try:
        blocker = gpu_ctx.submit_agg_hash(dfact, None, fact.index("fk"), heavy_aggs)
        tickets = [submit(gpu_ctx, fact, dim, dfact, ddim, k, a) for k, a in stream]
        for t in tickets[1::2]:
            t.cancel()
        same(blocker.wait(), heavy)
        outcomes = [_outcome(t, a) for t, a in zip(tickets, alone)]
        assert all(o == "ok" for o in outcomes[0::2])
        blocker = gpu_ctx.submit_agg_hash(dfact, None, fact.index("fk"), heavy_aggs)
        tickets = [submit(gpu_ctx, fact, dim, dfact, ddim, k, a) for k, a in stream]
        gpu_ctx.cancel_all()
        _outcome(blocker, heavy)
        outcomes += [_outcome(t, a) for t, a in zip(tickets, alone)]
    finally:
        gpu_ctx.set_slots(1)
    assert "canceled" in outcomes
