"""ctypes mirror of include/vexq_gpu.h plus small builders for the query
description (predicates, the left-fold arithmetic, aggregates).

The builders follow the reference planner's typing rules:
  * an integer literal next to a FLOAT64 operand becomes a float literal
    (coercePair, planner/physical.go:773-818);
  * mixed int/float arithmetic casts the integer column to float
    (buildBinExpr + CastIntToFloatExpr, planner/physical.go:572-589);
  * AccumTypeFor decides the accumulator (exec/aggregate.go:51-72).
"""
from __future__ import annotations

import ctypes as C

# ---- constants (include/vexq_gpu.h) ----------------------------------------
VQ_OK, VQ_E_ARG, VQ_E_DEVICE, VQ_E_BUDGET, VQ_E_UNSUPPORTED, VQ_E_NOMEM, VQ_E_CORRUPT, VQ_E_CANCELED = range(8)
VQ_U8, VQ_U16, VQ_DATE, VQ_I64, VQ_F64 = 1, 2, 3, 4, 5
VQ_STR = 6  # accumulator of MIN/MAX over a dictionary (string) column
VQ_EQ, VQ_NE, VQ_LT, VQ_LE, VQ_GT, VQ_GE = range(6)
VQ_IN = 6        # column IN (set of values in [0, 64)): ival is the mask
VQ_COLCMP = 16   # op | VQ_COLCMP: column OP column ival
VQ_BUILD_COL = 0x10000  # col | VQ_BUILD_COL: a build-side column (vq_join_group)
VQ_F_COL, VQ_F_LIT, VQ_F_LIT_OP_COL, VQ_F_COL_OP_LIT = range(4)
VQ_ADD, VQ_SUB, VQ_MUL, VQ_DIV = range(4)
VQ_AGG_COUNT, VQ_AGG_SUM, VQ_AGG_AVG, VQ_AGG_MIN, VQ_AGG_MAX = range(5)
VQ_MAX_TERMS, VQ_MAX_FACTORS, VQ_MAX_AGGS, VQ_MAX_WHEN = 8, 4, 8, 2
ROW_GROUP_ROWS = 65536

STATUS_NAMES = {
    VQ_OK: "ok", VQ_E_ARG: "invalid argument", VQ_E_DEVICE: "device error",
    VQ_E_BUDGET: "device-memory budget exceeded", VQ_E_UNSUPPORTED: "unsupported shape",
    VQ_E_NOMEM: "host allocation failure", VQ_E_CORRUPT: "corrupt file", VQ_E_CANCELED: "canceled",
}

WIDTH = {VQ_U8: 1, VQ_U16: 2, VQ_DATE: 4, VQ_I64: 8, VQ_F64: 8}


class Zone(C.Structure):
    _fields_ = [("min_bits", C.c_uint64), ("max_bits", C.c_uint64),
                ("null_count", C.c_int64), ("has_minmax", C.c_int32), ("pad_", C.c_int32)]


class ColDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("dict_size", C.c_int32), ("data", C.c_void_p),
                ("valid", C.c_void_p)]


class Term(C.Structure):
    _fields_ = [("col", C.c_int32), ("op", C.c_int32), ("ival", C.c_int64), ("fval", C.c_double)]


class Pred(C.Structure):
    _fields_ = [("nterms", C.c_int32), ("pad_", C.c_int32), ("terms", Term * VQ_MAX_TERMS)]


class Factor(C.Structure):
    _fields_ = [("kind", C.c_int32), ("col", C.c_int32), ("op", C.c_int32),
                ("combine", C.c_int32), ("fval", C.c_double), ("ival", C.c_int64)]


class Expr(C.Structure):
    _fields_ = [("nfactors", C.c_int32), ("is_float", C.c_int32),
                ("f", Factor * VQ_MAX_FACTORS),
                ("nwhen", C.c_int32), ("pad_", C.c_int32), ("when", Term * VQ_MAX_WHEN),
                ("else_f", C.c_double), ("else_i", C.c_int64)]


class Agg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("accum", C.c_int32), ("arg", Expr)]


class AggResult(C.Structure):
    _fields_ = [("capacity", C.c_int64), ("ngroups", C.c_int64),
                ("keys", C.POINTER(C.c_int64)), ("counts", C.POINTER(C.c_int64)),
                ("first_row", C.POINTER(C.c_int64)), ("ivals", C.POINTER(C.c_int64)),
                ("fvals", C.POINTER(C.c_double)), ("valid", C.POINTER(C.c_uint8))]


class Stats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("main_kernel_ms", C.c_double),
                ("rows_scanned", C.c_int64), ("rg_total", C.c_int64), ("rg_scanned", C.c_int64),
                ("bytes_alg", C.c_int64), ("build_rows", C.c_int64), ("build_keys", C.c_int64),
                ("partitions", C.c_int32), ("radix_bits", C.c_int32), ("grid", C.c_int32),
                ("tile_rows", C.c_int32)]


# ---- predicate builder -----------------------------------------------------
_OPS = {"==": VQ_EQ, "=": VQ_EQ, "!=": VQ_NE, "<>": VQ_NE, "<": VQ_LT, "<=": VQ_LE,
        ">": VQ_GT, ">=": VQ_GE}


VQ_Q_AGG_DICT = 1
VQ_Q_JOIN_AGG = 2
VQ_Q_AGG_HASH = 3
VQ_BATCH_TUNE = 1
VQ_BATCH_NO_GRAPH = 2


class Query(C.Structure):
    """vq_query: one query for vq_submit."""
    _fields_ = [("kind", C.c_int32), ("nkeys", C.c_int32), ("table", C.c_void_p),
                ("pred", C.POINTER(Pred)), ("key_cols", C.POINTER(C.c_int32)),
                ("build", C.c_void_p), ("build_pred", C.POINTER(Pred)),
                ("build_key", C.c_int32), ("probe_key", C.c_int32), ("naggs", C.c_int32),
                ("radix_bits", C.c_int32), ("aggs", C.POINTER(Agg)), ("topk", C.c_int32),
                ("pad_", C.c_int32)]


def make_pred(table, terms) -> Pred:
    """terms: iterable of (column name, op string, literal).  BETWEEN is two
    terms (>= lo, <= hi), the reference's own rewrite (exec/expr.go:1319-1336).
    op "in" takes an iterable of values in [0, 64) (dictionary codes); a
    literal X.col(name) compares with another column of the row."""
    terms = list(terms)
    if len(terms) > VQ_MAX_TERMS:
        raise ValueError("too many predicate terms")
    p = Pred()
    p.nterms = len(terms)
    for i, (name, op, lit) in enumerate(terms):
        _fill_term(table, p.terms[i], name, op, lit)
    return p


def _fill_term(table, t, name, op, lit):
    ci = table.index(name)
    t.col = ci
    if op == "in":
        mask = 0
        for v in lit:
            if not 0 <= int(v) < 64:
                raise ValueError("IN values must lie in [0, 64)")
            mask |= 1 << int(v)
        t.op = VQ_IN
        t.ival = mask - (1 << 64) if mask >= 1 << 63 else mask
        return
    if isinstance(lit, X):
        if not lit._single():
            raise TypeError("a column compare takes one column, X.col(name)")
        t.op = _OPS[op] | VQ_COLCMP
        t.ival = table.index(lit.factors[0][1].col)
        return
    t.op = _OPS[op]
    if table.columns[ci].type == VQ_F64:
        t.fval = float(lit)
    else:
        if isinstance(lit, float):
            raise TypeError(f"float literal against integer column {name}")
        t.ival = int(lit)


# ---- arithmetic builder: left fold of up to four factors -------------------
class _Col:
    def __init__(self, name):
        self.name = name


class _F:
    """One factor: column, literal, (lit op col) or (col op lit)."""

    def __init__(self, kind, col=None, op=VQ_ADD, lit=0):
        self.kind, self.col, self.op, self.lit = kind, col, op, lit


class X:
    """A left fold ((f0 c1 f1) c2 f2) ... built with Python operators, e.g.
    X.col("price") * (1 - X.col("discount")) * (1 + X.col("tax"))."""

    def __init__(self, factors, when=None, else_=0):
        self.factors = factors  # list of (combine, _F)
        self.when = when        # CASE WHEN: terms as make_pred takes them
        self.else_ = else_

    @staticmethod
    def col(name):
        return X([(VQ_ADD, _F(VQ_F_COL, col=name))])

    @staticmethod
    def case(when, then, else_=0):
        """CASE WHEN <when, a conjunction of up to VQ_MAX_WHEN predicate
        terms> THEN <then, a left fold> ELSE <literal> END (CaseExpr with one
        WHEN and a literal ELSE, exec/expr.go:1386-1435)."""
        if isinstance(then, (int, float)):
            then = X.lit(then)
        if then.when is not None:
            raise ValueError("nested CASE")
        when = list(when)
        if not 1 <= len(when) <= VQ_MAX_WHEN:
            raise ValueError(f"a WHEN takes 1 to {VQ_MAX_WHEN} terms")
        return X(list(then.factors), when, else_)

    @staticmethod
    def lit(v):
        return X([(VQ_ADD, _F(VQ_F_LIT, lit=v))])

    def _single(self):
        return self.when is None and len(self.factors) == 1 and self.factors[0][1].kind == VQ_F_COL

    def _fold(self, other, op):
        if self.when is not None or getattr(other, "when", None) is not None:
            raise ValueError("a CASE is an aggregate's whole argument")
        if isinstance(other, (int, float)):
            other = X.lit(other)
        if len(other.factors) != 1:
            raise ValueError("right operand must be a single factor")
        return X(self.factors + [(op, other.factors[0][1])])

    def _lit_side(self, lit, op, lit_left):
        if self.when is not None:
            raise ValueError("a CASE is an aggregate's whole argument")
        # (lit op col) / (col op lit): only on a bare column
        if not self._single():
            if lit_left:
                raise ValueError("literal-op-expression needs a bare column")
            return self._fold(lit, op)
        name = self.factors[0][1].col
        kind = VQ_F_LIT_OP_COL if lit_left else VQ_F_COL_OP_LIT
        return X([(VQ_ADD, _F(kind, col=name, op=op, lit=lit))])

    def __mul__(self, o):
        return self._fold(o, VQ_MUL)

    def __truediv__(self, o):
        return self._fold(o, VQ_DIV)

    def __add__(self, o):
        if isinstance(o, (int, float)) and self._single():
            return self._lit_side(o, VQ_ADD, False)
        return self._fold(o, VQ_ADD)

    def __sub__(self, o):
        if isinstance(o, (int, float)) and self._single():
            return self._lit_side(o, VQ_SUB, False)
        return self._fold(o, VQ_SUB)

    def __radd__(self, lit):
        return self._lit_side(lit, VQ_ADD, True)

    def __rsub__(self, lit):
        return self._lit_side(lit, VQ_SUB, True)

    def __rmul__(self, lit):
        return self._lit_side(lit, VQ_MUL, True)

    def __rtruediv__(self, lit):
        return self._lit_side(lit, VQ_DIV, True)

    def to_c(self, table) -> Expr:
        if len(self.factors) > VQ_MAX_FACTORS:
            raise ValueError("too many factors")
        is_float = False
        for _, f in self.factors:
            if f.col is not None and table.columns[table.index(f.col)].type == VQ_F64:
                is_float = True
            if f.kind != VQ_F_COL and isinstance(f.lit, float):
                is_float = True
        if self.when is not None and isinstance(self.else_, float):
            is_float = True
        e = Expr()
        e.nfactors = len(self.factors)
        e.is_float = int(is_float)
        if self.when is not None:
            e.nwhen = len(self.when)
            for k, (name, op, lit) in enumerate(self.when):
                _fill_term(table, e.when[k], name, op, lit)
            if is_float:
                e.else_f = float(self.else_)
            else:
                e.else_i = int(self.else_)
        for k, (comb, f) in enumerate(self.factors):
            cf = e.f[k]
            cf.kind = f.kind
            cf.col = table.index(f.col) if f.col is not None else 0
            cf.op = f.op
            cf.combine = comb
            if f.kind != VQ_F_COL:
                if is_float:
                    cf.fval = float(f.lit)
                else:
                    cf.ival = int(f.lit)
        return e


class JoinView:
    """The column names of a GROUP BY over joined rows (vq_join_group):
    probe columns by their names, build columns as "b.<name>", which resolve
    to col | VQ_BUILD_COL.  Pass it as the table to X.to_c / agg_* /
    make_pred-style builders."""

    def __init__(self, probe, build):
        self.probe, self.build = probe, build
        self.columns = _JoinCols(probe, build)

    def index(self, name: str) -> int:
        if name.startswith("b."):
            return VQ_BUILD_COL | self.build.index(name[2:])
        return self.probe.index(name)


class _JoinCols:
    def __init__(self, probe, build):
        self.probe, self.build = probe, build

    def __getitem__(self, i):
        return self.build.columns[i & ~VQ_BUILD_COL] if i & VQ_BUILD_COL else self.probe.columns[i]


def _agg(kind, x, table) -> Agg:
    a = Agg()
    a.kind = kind
    if x is None:
        a.arg.nfactors = 0
        a.accum = VQ_I64
        return a
    if isinstance(x, str):
        x = X.col(x)
    a.arg = x.to_c(table)
    if kind == VQ_AGG_AVG:
        a.accum = VQ_F64
    elif kind == VQ_AGG_COUNT:
        a.accum = VQ_I64
    else:
        a.accum = VQ_F64 if a.arg.is_float else VQ_I64
    return a


def count_star(table=None):
    return _agg(VQ_AGG_COUNT, None, table)


def agg_count(x, table):
    """COUNT(x): the rows where x holds a value (COUNT(*) where x reads no
    column with NULLs)."""
    return _agg(VQ_AGG_COUNT, x, table)


def agg_sum(x, table):
    return _agg(VQ_AGG_SUM, x, table)


def agg_avg(x, table):
    return _agg(VQ_AGG_AVG, x, table)


def agg_min(x, table):
    return _agg(VQ_AGG_MIN, x, table)


def agg_max(x, table):
    return _agg(VQ_AGG_MAX, x, table)


def _agg_str(kind, name, table) -> Agg:
    a = _agg(kind, name, table)
    a.accum = VQ_STR
    return a


def agg_min_str(name, table):
    """MIN over the strings of dictionary column `name` (Go's bytewise string
    order); the result is the code of the smallest string."""
    return _agg_str(VQ_AGG_MIN, name, table)


def agg_max_str(name, table):
    """MAX over the strings of dictionary column `name`."""
    return _agg_str(VQ_AGG_MAX, name, table)


def dict_rank(dictionary) -> "_np.ndarray":
    """rank[code]: the position of code's string among the dictionary's
    strings in bytewise order (what vq_table_dict_order takes)."""
    order = sorted(range(len(dictionary)), key=lambda c: _as_bytes(dictionary[c]))
    rank = _np.empty(len(dictionary), dtype=_np.uint32)
    rank[order] = _np.arange(len(dictionary), dtype=_np.uint32)
    return rank


def _as_bytes(s):
    return s if isinstance(s, (bytes, bytearray)) else str(s).encode()


def agg_array(aggs):
    arr = (Agg * max(1, len(aggs)))()
    for i, a in enumerate(aggs):
        arr[i] = a
    return arr


# ---- result buffers -----------------------------------------------------------
import numpy as _np  # noqa: E402


class AggOut:
    """Numpy view of a vq_agg_result: one row per group."""

    def __init__(self, capacity: int, nkeys: int, naggs: int):
        cap = max(1, int(capacity))
        self.nkeys, self.naggs = nkeys, naggs
        # the library writes every field of the ngroups rows it returns
        self._keys = _np.empty(cap * max(nkeys, 1), dtype=_np.int64)
        self._counts = _np.empty(cap, dtype=_np.int64)
        self._first = _np.empty(cap, dtype=_np.int64)
        self._ivals = _np.empty(cap * max(naggs, 1), dtype=_np.int64)
        self._fvals = _np.empty(cap * max(naggs, 1), dtype=_np.float64)
        self._valid = _np.empty(cap * max(naggs, 1), dtype=_np.uint8)
        r = AggResult()
        r.capacity = cap
        r.ngroups = 0
        r.keys = self._keys.ctypes.data_as(C.POINTER(C.c_int64))
        r.counts = self._counts.ctypes.data_as(C.POINTER(C.c_int64))
        r.first_row = self._first.ctypes.data_as(C.POINTER(C.c_int64))
        r.ivals = self._ivals.ctypes.data_as(C.POINTER(C.c_int64))
        r.fvals = self._fvals.ctypes.data_as(C.POINTER(C.c_double))
        r.valid = self._valid.ctypes.data_as(C.POINTER(C.c_uint8))
        self.c = r
        self.stats = None

    @property
    def ngroups(self) -> int:
        return int(self.c.ngroups)

    @property
    def keys(self):
        n = self.ngroups
        return self._keys[: n * max(self.nkeys, 1)].reshape(n, max(self.nkeys, 1))[:, : self.nkeys]

    @property
    def counts(self):
        return self._counts[: self.ngroups]

    @property
    def first_row(self):
        return self._first[: self.ngroups]

    @property
    def ivals(self):
        n = self.ngroups
        return self._ivals[: n * max(self.naggs, 1)].reshape(n, max(self.naggs, 1))[:, : self.naggs]

    @property
    def fvals(self):
        n = self.ngroups
        return self._fvals[: n * max(self.naggs, 1)].reshape(n, max(self.naggs, 1))[:, : self.naggs]

    @property
    def valid(self):
        n = self.ngroups
        return self._valid[: n * max(self.naggs, 1)].reshape(n, max(self.naggs, 1))[:, : self.naggs]


def col_descs(table):
    """ColDesc array over a datagen.Table (keeps references alive on it)."""
    arr = (ColDesc * max(1, len(table.columns)))()
    for i, c in enumerate(table.columns):
        arr[i].type = c.type
        arr[i].dict_size = c.dict_size
        arr[i].data = c.data.ctypes.data
        v = getattr(c, "valid", None)  # a LSB-first validity bitmap (uint8), or None
        arr[i].valid = v.ctypes.data if v is not None else None
    return arr


def rg_array(table):
    rg = _np.ascontiguousarray(table.rg_rows, dtype=_np.int64)
    return rg, rg.ctypes.data_as(C.POINTER(C.c_int64))
