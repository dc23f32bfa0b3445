"""Seeded synthetic tables for the hot-path workloads.

Every value is a pure function of (seed, stream, row index) through the
splitmix64 finaliser, so a table is identical on every machine and any slice
can be regenerated without the rest.  Shapes follow the workloads the
reference benchmarks on (bench/parallel/bench_test.go:110-180 generates a
6M-row lineitem-like table the same way, from seed 42):

* ``lineitem``  — flagA u8 dictionary code (3 values, weights 1/4,1/4,1/2),
  flagB u8 code (2 values), quantity int64 U{1..50},
  price float64 = round_to_cents(quantity * U{900.00..104950.00} / 100),
  discount float64 = U{0..10}/100, tax float64 = U{0..8}/100,
  shipdate int32 days U{8035..10560} (2,526 days).
* ``orders`` (join build) — key int64 = 1 + 4 i, orderdate int32 days
  U{8035..10440} (2,406 days), priority u8 code (5 values).
* ``lineitem_probe`` (join probe) — fk drawn so each build key gets 1..7 rows,
  shipdate = orderdate(fk) + U{1..121}, quantity/price/discount as lineitem.

Dictionary columns carry global codes in first-occurrence order, the order
storage.DictBuilder assigns (storage/codec.go:55-63).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import math

import numpy as np

from . import abi

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_STREAM = np.uint64(0xD1B54A32D192ED03)

DATE_MIN = 8035            # 1992-01-01 in days since 1970-01-01
SHIP_DAYS = 2526           # lineitem shipdate range
ORDER_DAYS = 2406          # orders orderdate range


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def rand_u64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n pseudo-random uint64 values for rows [offset, offset+n) of a stream."""
    with np.errstate(over="ignore"):
        base = _mix(np.array([seed], dtype=np.uint64) * _GOLDEN
                    + np.array([stream], dtype=np.uint64) * _STREAM)[0]
        i = np.arange(offset, offset + n, dtype=np.uint64) + np.uint64(1)
        return _mix(base + i * _GOLDEN)


def rand_int(seed, stream, n, lo, hi, offset=0) -> np.ndarray:
    """Uniform integers in [lo, hi] (int64)."""
    r = rand_u64(seed, stream, n, offset)
    return (r % np.uint64(hi - lo + 1)).astype(np.int64) + lo


def first_occurrence_codes(values: np.ndarray, names):
    """Map value indices to global codes in first-occurrence order."""
    uniq, first = np.unique(values, return_index=True)
    order = uniq[np.argsort(first, kind="stable")]
    remap = np.zeros(int(values.max()) + 1 if values.size else 1, dtype=np.int64)
    remap[order] = np.arange(order.size)
    codes = remap[values] if values.size else values
    dictionary = [names[v] for v in order]
    return codes, dictionary


@dataclass
class Column:
    name: str
    type: int
    data: np.ndarray
    dict_size: int = 0
    dictionary: list = field(default_factory=list)
    valid: np.ndarray | None = None  # LSB-first validity bitmap (uint8); None: no NULLs


def validity_bitmap(mask: np.ndarray) -> np.ndarray:
    """Row mask (True = holds a value) -> LSB-first bitmap bytes, as the
    reference's NullBitmap (exec/batch.go:113-121), padded to whole words."""
    m = np.asarray(mask, dtype=bool)
    pad = (-m.size) % 64
    b = np.packbits(np.concatenate([m, np.zeros(pad, dtype=bool)]), bitorder="little")
    return np.ascontiguousarray(b)


def with_nulls(col: Column, null_mask: np.ndarray) -> Column:
    """The column with the rows of null_mask set to NULL (their values zeroed,
    as the reference's vectors hold zero under a NULL bit)."""
    data = col.data.copy()
    data[np.asarray(null_mask, dtype=bool)] = 0
    return Column(col.name, col.type, data, col.dict_size, col.dictionary,
                  validity_bitmap(~np.asarray(null_mask, dtype=bool)))


@dataclass
class Table:
    columns: list
    rg_rows: np.ndarray

    @property
    def nrows(self) -> int:
        return int(self.columns[0].data.shape[0]) if self.columns else 0

    def index(self, name: str) -> int:
        for i, c in enumerate(self.columns):
            if c.name == name:
                return i
        raise KeyError(name)

    def col(self, name: str) -> np.ndarray:
        return self.columns[self.index(name)].data

    def take(self, order: np.ndarray) -> "Table":
        def valid(c):
            if c.valid is None:
                return None
            m = np.unpackbits(c.valid, bitorder="little")[: self.nrows].astype(bool)
            return validity_bitmap(m[order])

        cols = [Column(c.name, c.type, np.ascontiguousarray(c.data[order]), c.dict_size,
                       list(c.dictionary), valid(c)) for c in self.columns]
        return Table(cols, row_groups(len(order)))


def row_groups(nrows: int, rg: int = abi.ROW_GROUP_ROWS) -> np.ndarray:
    n_full, rem = divmod(nrows, rg)
    sizes = [rg] * n_full + ([rem] if rem else [])
    return np.array(sizes, dtype=np.int64)


_NP = {abi.VQ_U8: np.uint8, abi.VQ_U16: np.uint16, abi.VQ_DATE: np.int32,
       abi.VQ_I64: np.int64, abi.VQ_F64: np.float64}


def _col(name, typ, data, dictionary=None):
    arr = np.ascontiguousarray(data, dtype=_NP[typ])
    d = list(dictionary or [])
    return Column(name, typ, arr, len(d), d)


def _price(seed, stream_base, quantity, n, offset=0):
    # quantity * uniform(900.00, 104950.00) / 100, rounded to cents
    unit_cents = rand_int(seed, stream_base, n, 90000, 10495000, offset)
    cents = (quantity * unit_cents + 50) // 100
    return cents.astype(np.float64) / 100.0


def lineitem(nrows: int, seed: int = 42, offset: int = 0) -> Table:
    """The dict-coded GROUP BY table (8 columns, 38 referenced bytes/row)."""
    n = nrows
    a = rand_u64(seed, 1, n, offset) % np.uint64(4)
    a = np.minimum(a, np.uint64(2)).astype(np.int64)        # 1/4, 1/4, 1/2
    flag_a, dict_a = first_occurrence_codes(a, ["A", "N", "R"])
    b = (rand_u64(seed, 2, n, offset) % np.uint64(2)).astype(np.int64)
    flag_b, dict_b = first_occurrence_codes(b, ["F", "O"])
    quantity = rand_int(seed, 3, n, 1, 50, offset)
    price = _price(seed, 4, quantity, n, offset)
    discount = rand_int(seed, 5, n, 0, 10, offset).astype(np.float64) / 100.0
    tax = rand_int(seed, 6, n, 0, 8, offset).astype(np.float64) / 100.0
    shipdate = rand_int(seed, 7, n, DATE_MIN, DATE_MIN + SHIP_DAYS - 1, offset)
    cols = [
        _col("flagA", abi.VQ_U8, flag_a, dict_a),
        _col("flagB", abi.VQ_U8, flag_b, dict_b),
        _col("quantity", abi.VQ_I64, quantity),
        _col("price", abi.VQ_F64, price),
        _col("discount", abi.VQ_F64, discount),
        _col("tax", abi.VQ_F64, tax),
        _col("shipdate", abi.VQ_DATE, shipdate),
    ]
    return Table(cols, row_groups(n))


def lineitem_sorted(nrows: int, seed: int = 42) -> Table:
    """lineitem written sorted by shipdate (zone maps then prune windows)."""
    t = lineitem(nrows, seed)
    order = np.argsort(t.col("shipdate"), kind="stable")
    return t.take(order)


def orders(nrows: int, seed: int = 42) -> Table:
    """Join build side: unique keys 1 + 4 i."""
    n = nrows
    key = 1 + 4 * np.arange(n, dtype=np.int64)
    orderdate = rand_int(seed, 11, n, DATE_MIN, DATE_MIN + ORDER_DAYS - 1)
    pr = (rand_u64(seed, 12, n) % np.uint64(5)).astype(np.int64)
    prio, dict_p = first_occurrence_codes(pr, ["1-URGENT", "2-HIGH", "3-MEDIUM",
                                               "4-NOT SPECIFIED", "5-LOW"])
    cols = [
        _col("key", abi.VQ_I64, key),
        _col("orderdate", abi.VQ_DATE, orderdate),
        _col("priority", abi.VQ_U8, prio, dict_p),
    ]
    return Table(cols, row_groups(n))


def probe_counts(nbuild: int, nprobe: int, seed: int = 42) -> np.ndarray:
    """Rows per build key: U{1..7}, adjusted (in key order) to sum to nprobe
    while staying inside [1, 7]."""
    c = rand_int(seed, 13, nbuild, 1, 7)
    diff = int(nprobe - c.sum())
    if diff > 0:
        room = 7 - c
        cum = np.cumsum(room)
        k = int(np.searchsorted(cum, diff))
        if k >= nbuild:
            raise ValueError("nprobe exceeds 7 rows per key")
        add = room.copy()
        add[k + 1:] = 0
        add[k] = diff - (cum[k - 1] if k else 0)
        c = c + add
    elif diff < 0:
        room = c - 1
        cum = np.cumsum(room)
        k = int(np.searchsorted(cum, -diff))
        if k >= nbuild:
            raise ValueError("nprobe below 1 row per key")
        sub = room.copy()
        sub[k + 1:] = 0
        sub[k] = -diff - (cum[k - 1] if k else 0)
        c = c - sub
    return c


def lineitem_probe(build: Table, nrows: int, seed: int = 42) -> Table:
    """Join probe side: each build key appears 1..7 times, shuffled."""
    nb = build.nrows
    counts = probe_counts(nb, nrows, seed)
    brow = np.repeat(np.arange(nb, dtype=np.int64), counts)
    perm = np.argsort(rand_u64(seed, 14, nrows), kind="stable")
    brow = brow[perm]
    fk = build.col("key")[brow]
    shipdate = build.col("orderdate")[brow].astype(np.int64) + rand_int(seed, 15, nrows, 1, 121)
    quantity = rand_int(seed, 16, nrows, 1, 50)
    price = _price(seed, 17, quantity, nrows)
    discount = rand_int(seed, 18, nrows, 0, 10).astype(np.float64) / 100.0
    cols = [
        _col("fk", abi.VQ_I64, fk),
        _col("shipdate", abi.VQ_DATE, shipdate),
        _col("quantity", abi.VQ_I64, quantity),
        _col("price", abi.VQ_F64, price),
        _col("discount", abi.VQ_F64, discount),
    ]
    return Table(cols, row_groups(nrows))


def q12_tables(nb: int, npr: int, seed: int = 16):
    """orders (the build side: key, orderdate, priority codes of the five
    TPC-H priority strings) and a lineitem probe side with Q12's columns:
    ship mode codes (7 modes), commit and receipt dates around the ship
    date."""
    b = orders(nb, seed)
    p = lineitem_probe(b, npr, seed)
    ship = p.col("shipdate").astype(np.int64)
    n = p.nrows
    commit = ship + rand_int(seed, 93, n, -30, 30)
    receipt = ship + rand_int(seed, 94, n, 1, 30)
    mode = rand_int(seed, 95, n, 0, 6).astype(np.uint8)
    p = Table(list(p.columns) + [_col("commitdate", abi.VQ_DATE, commit.astype(np.int32)),
                                 _col("receiptdate", abi.VQ_DATE, receipt.astype(np.int32)),
                                 _col("shipmode", abi.VQ_U8, mode,
                                      ["REG AIR", "AIR", "RAIL", "SHIP", "TRUCK", "MAIL", "FOB"])],
              row_groups(n))
    return b, p


def q_q12(b: Table, p: Table, lo: int = DATE_MIN + 731, hi: int = DATE_MIN + 1096):
    """TPC-H Q12 (planner/parallel_join_bench_test.go:19-30) over q12_tables:
    lineitem rows with shipmode IN ('MAIL', 'SHIP'), commitdate < receiptdate,
    shipdate < commitdate, receiptdate in [lo, hi) (1994-01-01 to
    1995-01-01 by default), joined to orders on the
    order key, GROUP BY shipmode, SUM(CASE WHEN o_orderpriority IN
    ('1-URGENT', '2-HIGH') THEN 1 ELSE 0 END), SUM(CASE WHEN o_orderpriority
    <> '1-URGENT' AND o_orderpriority <> '2-HIGH' THEN 1 ELSE 0 END).
    Returns (probe predicate, group columns, aggregates, the two codes)."""
    X = abi.X
    v = abi.JoinView(p, b)
    pdict = b.columns[b.index("priority")].dictionary
    urgent = [pdict.index(x) for x in ("1-URGENT", "2-HIGH")]
    modes = p.columns[p.index("shipmode")].dictionary
    pp = abi.make_pred(p, [("shipmode", "in", [modes.index("MAIL"), modes.index("SHIP")]),
                           ("commitdate", "<", X.col("receiptdate")),
                           ("shipdate", "<", X.col("commitdate")),
                           ("receiptdate", ">=", lo), ("receiptdate", "<", hi)])
    aggs = [abi.agg_sum(X.case([("b.priority", "in", urgent)], 1, 0), v),
            abi.agg_sum(X.case([("b.priority", "<>", urgent[0]), ("b.priority", "<>", urgent[1])], 1, 0), v)]
    return pp, [p.index("shipmode")], aggs, urgent


SEGMENTS = ["AUTOMOBILE", "BUILDING", "FURNITURE", "HOUSEHOLD", "MACHINERY"]


def q3_tables(nc: int, no: int, nl: int, seed: int = 31):
    """TPC-H Q3's three tables: customer (custkey 1..nc, market segment codes
    of the five segments), orders (the join build side's key, orderdate and
    priority, plus the ordering customer and a ship priority) and a lineitem
    probe side of its orders."""
    c = Table([_col("custkey", abi.VQ_I64, 1 + np.arange(nc, dtype=np.int64)),
               _col("mktsegment", abi.VQ_U8, rand_int(seed, 97, nc, 0, 4).astype(np.uint8), SEGMENTS)],
              row_groups(nc))
    o = orders(no, seed)
    o = Table(list(o.columns) + [_col("custkey", abi.VQ_I64, rand_int(seed, 98, no, 1, nc)),
                                 _col("shippriority", abi.VQ_I64, rand_int(seed, 99, no, 0, 1))],
              o.rg_rows)
    li = lineitem_probe(o, nl, seed)
    return c, o, li


def q_q3(c: Table, o: Table, li: Table, cutoff: int = DATE_MIN + 1169):
    """TPC-H Q3 (planner/parallel_join_bench_test.go:32-44): customers in the
    BUILDING segment, their orders dated before 1995-03-15, lineitems shipped
    after it, GROUP BY l_orderkey (o_orderdate, o_shippriority riding with
    it), SUM(l_extendedprice), ORDER BY revenue DESC LIMIT 10.  Returns the
    arguments of join_agg_nested after the tables: (inner key, inner pred,
    build fk, build key, build pred, probe key, probe pred, pay columns,
    aggregates, top-k)."""
    ip = abi.make_pred(c, [("mktsegment", "=", SEGMENTS.index("BUILDING"))])
    bp = abi.make_pred(o, [("orderdate", "<", cutoff)])
    pp = abi.make_pred(li, [("shipdate", ">", cutoff)])
    aggs = [abi.agg_sum("price", li)]
    return (c.index("custkey"), ip, o.index("custkey"), o.index("key"), bp, li.index("fk"), pp,
            [o.index("orderdate"), o.index("shippriority")], aggs, 10)


def fact_serving(dim: Table, nrows: int, seed: int = 42) -> Table:
    """The served stream's fact table: lineitem's dictionary and measure
    columns plus the join probe side's foreign key (each dimension key 1..7
    times, shuffled) with shipdate = orderdate + U{1..121}."""
    p = lineitem_probe(dim, nrows, seed)
    li = lineitem(nrows, seed)
    cols = [li.columns[li.index("flagA")], li.columns[li.index("flagB")],
            p.columns[p.index("quantity")], p.columns[p.index("price")],
            p.columns[p.index("discount")], li.columns[li.index("tax")],
            p.columns[p.index("shipdate")], p.columns[p.index("fk")]]
    return Table(cols, row_groups(nrows))


def _ranks(seed, stream, n, groups, skew):
    """n draws of a rank in [0, groups): uniform (skew 0) or Zipf(skew)
    truncated to `groups` ranks (inverse CDF over the rank weights)."""
    u = (rand_u64(seed, stream, n) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    if skew <= 0:
        return np.minimum((u * groups).astype(np.int64), groups - 1)
    w = 1.0 / np.arange(1, groups + 1, dtype=np.float64) ** skew
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return np.minimum(np.searchsorted(cdf, u, side="right"), groups - 1).astype(np.int64)


def keyed_table(nrows: int, groups: int, key_type: int = abi.VQ_I64, skew: float = 0.0,
                seed: int = 42) -> Table:
    """A GROUP BY instance of the batched workload: a key column with
    `groups` distinct values (int64 keys spread over the int64 range, or u16
    codes), uniform or Zipf-skewed, and config 0's measures."""
    n = nrows
    rank = _ranks(seed, 30, n, groups, skew)
    if key_type == abi.VQ_U16:
        if groups > 65536:
            raise ValueError("u16 keys hold at most 65536 groups")
        perm = np.argsort(rand_u64(seed, 31, groups), kind="stable")
        key = perm[rank].astype(np.uint16)
    else:
        key = _mix(rank.astype(np.uint64) + np.uint64(seed * 7919)).view(np.int64)
    quantity = rand_int(seed, 32, n, 1, 50)
    price = _price(seed, 33, quantity, n)
    discount = rand_int(seed, 34, n, 0, 10).astype(np.float64) / 100.0
    shipdate = rand_int(seed, 35, n, DATE_MIN, DATE_MIN + SHIP_DAYS - 1)
    cols = [
        _col("key", key_type, key),
        _col("quantity", abi.VQ_I64, quantity),
        _col("price", abi.VQ_F64, price),
        _col("discount", abi.VQ_F64, discount),
        _col("shipdate", abi.VQ_DATE, shipdate),
    ]
    return Table(cols, row_groups(n))


def join_instance(nbuild: int, ratio: float, skew: float = 0.0, seed: int = 42):
    """A join instance of the batched workload: a build side of unique keys
    with an orderdate, and a probe side of ratio x nbuild rows whose foreign
    keys are drawn from the build keys (uniform or Zipf-skewed), with
    shipdate and config 0's measures."""
    b = orders(nbuild, seed)
    npr = max(1, int(nbuild * ratio))
    rank = _ranks(seed, 40, npr, nbuild, skew)
    fk = b.col("key")[rank]
    quantity = rand_int(seed, 41, npr, 1, 50)
    price = _price(seed, 42, quantity, npr)
    discount = rand_int(seed, 43, npr, 0, 10).astype(np.float64) / 100.0
    shipdate = b.col("orderdate")[rank].astype(np.int64) + rand_int(seed, 44, npr, 1, 121)
    p = Table([
        _col("fk", abi.VQ_I64, fk),
        _col("shipdate", abi.VQ_DATE, shipdate),
        _col("quantity", abi.VQ_I64, quantity),
        _col("price", abi.VQ_F64, price),
        _col("discount", abi.VQ_F64, discount),
    ], row_groups(npr))
    return b, p


# ---- the workload queries ---------------------------------------------------

def q_groupby(t: Table):
    """Filtered dict-coded GROUP BY (flagA, flagB) with eight aggregates:
    SUM(quantity), SUM(price), SUM(price*(1-discount)),
    SUM(price*(1-discount)*(1+tax)), AVG(quantity), AVG(price), AVG(discount),
    COUNT(*); predicate shipdate <= max - 90."""
    X = abi.X
    pred = abi.make_pred(t, [("shipdate", "<=", DATE_MIN + SHIP_DAYS - 1 - 90)])
    disc_price = X.col("price") * (1 - X.col("discount"))
    charge = X.col("price") * (1 - X.col("discount")) * (1 + X.col("tax"))
    aggs = [
        abi.agg_sum("quantity", t), abi.agg_sum("price", t), abi.agg_sum(disc_price, t),
        abi.agg_sum(charge, t), abi.agg_avg("quantity", t), abi.agg_avg("price", t),
        abi.agg_avg("discount", t), abi.count_star(t),
    ]
    keys = [t.index("flagA"), t.index("flagB")]
    return pred, keys, aggs


def q_scan_sum(t: Table, seed: int = 42, which: int = 0):
    """Range-predicate scan: shipdate in a random 365-day window, discount
    BETWEEN d-0.01 AND d+0.01 (d = k/100, k in 2..9), quantity < 24;
    SUM(price*discount)."""
    r = rand_u64(seed, 100 + which, 2)
    lo = DATE_MIN + int(r[0] % np.uint64(SHIP_DAYS - 365))
    d = int(2 + r[1] % np.uint64(8)) / 100.0
    X = abi.X
    pred = abi.make_pred(t, [("shipdate", ">=", lo), ("shipdate", "<", lo + 365),
                             ("discount", ">=", d - 0.01), ("discount", "<=", d + 0.01),
                             ("quantity", "<", 24)])
    aggs = [abi.agg_sum(X.col("price") * X.col("discount"), t)]
    return pred, [], aggs


JOIN_CUTOFF = DATE_MIN + 1155      # orderdate < cutoff keeps ~48% of orders


def q_join(build: Table, probe: Table, cutoff: int = JOIN_CUTOFF):
    """orders JOIN lineitem ON key = fk WHERE orderdate < c AND shipdate > c
    GROUP BY key, SUM(price*(1-discount)), ORDER BY 2 DESC LIMIT 10."""
    X = abi.X
    bpred = abi.make_pred(build, [("orderdate", "<", cutoff)])
    ppred = abi.make_pred(probe, [("shipdate", ">", cutoff)])
    aggs = [abi.agg_sum(X.col("price") * (1 - X.col("discount")), probe)]
    return bpred, ppred, aggs


def q_groupby_cut(t: Table, cut: int):
    """q_groupby with the shipdate cutoff `cut` (shipdate <= cut)."""
    pred, keys, aggs = q_groupby(t)
    pred = abi.make_pred(t, [("shipdate", "<=", int(cut))])
    return pred, keys, aggs


def serve_stream(fact: Table, dim: Table, nq: int, seed: int = 7):
    """A served query mix: 50% range scans with random windows (q_scan_sum),
    30% GROUP BYs with random shipdate cutoffs, 20% joins with random date
    cutoffs.  Returns [(kind, args)] with kind in scan/groupby/join."""
    r = rand_u64(seed, 200, 3 * nq)
    out = []
    for i in range(nq):
        u = int(r[3 * i] % np.uint64(100))
        if u < 50:
            out.append(("scan", q_scan_sum(fact, seed, which=1000 + i)))
        elif u < 80:
            cut = DATE_MIN + SHIP_DAYS - 1 - int(r[3 * i + 1] % np.uint64(366))
            out.append(("groupby", q_groupby_cut(fact, cut)))
        else:
            cutoff = DATE_MIN + 900 + int(r[3 * i + 2] % np.uint64(501))
            out.append(("join", q_join(dim, fact, cutoff)))
    return out


def q_keyed(t: Table, cut: int = DATE_MIN + SHIP_DAYS - 1 - 90):
    """GROUP BY key: SUM(price*(1-discount)), SUM(quantity), AVG(price),
    COUNT(*), MIN(price), MAX(quantity) WHERE shipdate <= cut."""
    X = abi.X
    pred = abi.make_pred(t, [("shipdate", "<=", int(cut))])
    aggs = [abi.agg_sum(X.col("price") * (1 - X.col("discount")), t), abi.agg_sum("quantity", t),
            abi.agg_avg("price", t), abi.count_star(t), abi.agg_min("price", t),
            abi.agg_max("quantity", t)]
    return pred, t.index("key"), aggs


def batch_specs(n: int = 256, seed: int = 42, min_rows: int = 10_000, max_rows: int = 4_000_000):
    """The batched workload's instances: half GROUP BYs, half joins; rows
    log-uniform in [min_rows, max_rows] (the probe side's, for a join); GROUP
    BY cardinality log-uniform in [4, 1M] (at most the rows), u16 keys when
    it fits in 16 bits and the draw says so, else int64; join build:probe
    uniform in [1:2, 1:16]; Zipf(1.1) key skew on every other instance."""
    r = rand_u64(seed, 300, 6 * n)
    u = (r >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    specs = []
    for i in range(n):
        rows = int(round(math.exp(math.log(min_rows) + u[6 * i] * (math.log(max_rows) - math.log(min_rows)))))
        skew = 1.1 if i % 2 else 0.0
        if u[6 * i + 1] < 0.5:
            groups = int(round(math.exp(math.log(4) + u[6 * i + 2] * (math.log(1_000_000) - math.log(4)))))
            groups = max(4, min(groups, rows))
            kt = abi.VQ_U16 if (groups <= 65536 and u[6 * i + 3] < 0.5) else abi.VQ_I64
            specs.append({"kind": "hash", "rows": rows, "groups": groups, "key_type": kt,
                          "skew": skew, "seed": 1000 + i})
        else:
            ratio = 2.0 + 14.0 * u[6 * i + 4]
            specs.append({"kind": "join", "rows": rows, "build": max(16, int(rows / ratio)),
                          "ratio": ratio, "skew": skew, "seed": 1000 + i})
    return specs


def q_join_instance(build: Table, probe: Table, cutoff: int = JOIN_CUTOFF + 100):
    """A batched join instance's query: orderdate < c, shipdate > c - 60,
    GROUP BY key, SUM(price*(1-discount)), COUNT(*), MAX(quantity)."""
    X = abi.X
    bpred = abi.make_pred(build, [("orderdate", "<", cutoff)])
    ppred = abi.make_pred(probe, [("shipdate", ">", cutoff - 60)])
    aggs = [abi.agg_sum(X.col("price") * (1 - X.col("discount")), probe), abi.count_star(probe),
            abi.agg_max("quantity", probe)]
    return bpred, ppred, aggs
