"""Host binding of libvexqgpu.so (the gfx950 operators) through its C ABI.

The library must have been built (``python -m vexq_amd.build``); loading it is
mandatory and a missing build raises at import of this module's first use.
There is no CPU fallback here: every call runs the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libvexqgpu.so")
_lib = None


class GpuError(RuntimeError):
    def __init__(self, code: int, msg: str, needed: int = 0):
        super().__init__(f"{abi.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code
        # result rows a call needed when it failed for want of capacity
        self.needed = needed


def lib():
    """The loaded C ABI library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"{_LIB_PATH} is not built: run `python -m vexq_amd.build`")
    L = C.CDLL(_LIB_PATH)
    P = C.POINTER
    L.vq_last_error.restype = C.c_char_p
    L.vq_ctx_create.argtypes = [C.c_int, C.c_uint64, P(C.c_void_p)]
    L.vq_ctx_destroy.argtypes = [C.c_void_p]
    L.vq_ctx_set_kernel_repeats.argtypes = [C.c_void_p, C.c_int32]
    L.vq_ctx_set_slots.argtypes = [C.c_void_p, C.c_int32, C.c_uint64]
    L.vq_submit.argtypes = [C.c_void_p, P(abi.Query), P(abi.AggResult), P(abi.Stats),
                            P(C.c_void_p)]
    L.vq_ticket_done.argtypes = [C.c_void_p]
    L.vq_ticket_cancel.argtypes = [C.c_void_p]
    L.vq_ctx_cancel.argtypes = [C.c_void_p]
    L.vq_agg_dict_multi.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, P(abi.Pred), C.c_int32,
                                    P(abi.Agg), P(abi.AggResult), P(abi.Stats)]
    L.vq_batch_run.argtypes = [C.c_void_p, C.c_int32, P(abi.Query), P(abi.AggResult),
                               P(abi.Stats), C.c_int32]
    L.vq_wait.argtypes = [C.c_void_p]
    L.vq_ctx_bytes_in_use.argtypes = [C.c_void_p]
    L.vq_ctx_bytes_in_use.restype = C.c_uint64
    L.vq_table_upload.argtypes = [C.c_void_p, C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32,
                                  P(C.c_int64), P(abi.Zone), P(C.c_void_p)]
    L.vq_table_upload_vxq.argtypes = [C.c_void_p, C.c_char_p, C.c_uint64, C.c_int32, P(C.c_int32),
                                      P(C.c_void_p), P(abi.Stats)]
    L.vq_table_column.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32)]
    L.vq_table_dict_entry.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_char_p), P(C.c_int64)]
    L.vq_table_dict_order.argtypes = [C.c_void_p, C.c_int32, P(C.c_uint32)]
    L.vq_table_destroy.argtypes = [C.c_void_p]
    L.vq_table_rows.argtypes = [C.c_void_p]
    L.vq_table_rows.restype = C.c_int64
    L.vq_table_zones.argtypes = [C.c_void_p, P(abi.Zone), C.c_int64]
    L.vq_zone_prune.argtypes = [C.c_void_p, P(abi.Pred), P(C.c_uint8)]
    L.vq_filter.argtypes = [C.c_void_p, C.c_void_p, P(abi.Pred), P(C.c_void_p), P(abi.Stats)]
    L.vq_selvec_count.argtypes = [C.c_void_p]
    L.vq_selvec_count.restype = C.c_int64
    L.vq_selvec_copy.argtypes = [C.c_void_p, P(C.c_uint32), C.c_int64]
    L.vq_selvec_destroy.argtypes = [C.c_void_p]
    L.vq_agg_dict.argtypes = [C.c_void_p, C.c_void_p, P(abi.Pred), C.c_int32, P(C.c_int32),
                              C.c_int32, P(abi.Agg), P(abi.AggResult), P(abi.Stats)]
    L.vq_agg_dict_noalloc.argtypes = [C.c_void_p, C.c_void_p, P(abi.Pred), C.c_int32, P(C.c_int32),
                                      C.c_int32, P(abi.Agg), P(abi.AggResult), P(abi.Stats)]
    L.vq_join_agg.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, P(abi.Pred), C.c_void_p,
                              C.c_int32, P(abi.Pred), C.c_int32, P(abi.Agg), C.c_int32,
                              C.c_int32, P(abi.AggResult), P(abi.Stats)]
    L.vq_agg_hash.argtypes = [C.c_void_p, C.c_void_p, P(abi.Pred), C.c_int32, C.c_int32,
                              P(abi.Agg), P(abi.AggResult), P(abi.Stats)]
    L.vq_join_agg_nested.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, P(abi.Pred), C.c_void_p, C.c_int32,
                                     C.c_int32, P(abi.Pred), C.c_void_p, C.c_int32, P(abi.Pred), C.c_int32,
                                     P(C.c_int32), C.c_int32, P(abi.Agg), C.c_int32, C.c_int32,
                                     P(abi.AggResult), P(C.c_int64), P(abi.Stats)]
    L.vq_join_group.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, P(abi.Pred), C.c_void_p,
                                C.c_int32, P(abi.Pred), C.c_int32, P(C.c_int32), C.c_int32,
                                P(abi.Agg), P(abi.AggResult), P(abi.Stats)]
    L.vq_radix_bits_for.argtypes = [C.c_int64]
    L.vq_radix_bits_for.restype = C.c_int32
    L.vq_radix_hash.argtypes = [C.c_int64]
    L.vq_radix_hash.restype = C.c_uint64
    L.vq_join_table_capacity.argtypes = [C.c_int64]
    L.vq_join_table_capacity.restype = C.c_int64
    _lib = L
    return L


def _check(rc: int, out: "abi.AggOut | None" = None):
    """Raises GpuError on a failed call.  With `out`, the result buffer the
    call was given: on failure it is left empty (ngroups 0) and the rows a
    capacity failure needed are carried in the error's `needed`."""
    if rc != abi.VQ_OK:
        needed = 0
        if out is not None:
            needed = int(out.c.ngroups)
            out.c.ngroups = 0
        raise GpuError(rc, lib().vq_last_error().decode(), needed)


def _pp(x):
    return C.byref(x) if x is not None else None


class Context:
    """vq_ctx: one device, one stream, a device-memory budget."""

    def __init__(self, device: int = 0, budget_bytes: int = 0):
        h = C.c_void_p()
        _check(lib().vq_ctx_create(device, budget_bytes, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().vq_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def bytes_in_use(self) -> int:
        return int(lib().vq_ctx_bytes_in_use(self.h))

    def set_kernel_repeats(self, repeats: int) -> None:
        """Timing aid: launch each query's main kernel `repeats` times inside
        its event pair (stats.main_kernel_ms is the mean per launch)."""
        _check(lib().vq_ctx_set_kernel_repeats(self.h, repeats))

    def upload(self, table, zones=None) -> "DeviceTable":
        return DeviceTable(self, table, zones)

    def upload_vxq(self, image: bytes, cols) -> "VxqTable":
        """A table straight from a .vxq file image (vq_table_upload_vxq):
        the file columns `cols` checked and decoded on the device."""
        return VxqTable(self, image, cols)

    def cancel_all(self) -> None:
        """vq_ctx_cancel: every queued or running submitted query."""
        _check(lib().vq_ctx_cancel(self.h))

    def set_slots(self, nslots: int, serve_budget: int = 0) -> None:
        """nslots concurrent query slots; vq_submit admission bound in bytes."""
        _check(lib().vq_ctx_set_slots(self.h, nslots, serve_budget))

    # -- asynchronous queries ------------------------------------------------
    def submit_agg_dict(self, dt: "DeviceTable", pred, key_cols, aggs,
                        capacity: int = 4096) -> "Ticket":
        q = abi.Query()
        q.kind = abi.VQ_Q_AGG_DICT
        q.table = dt.h
        keep = [abi.agg_array(aggs), (C.c_int32 * max(1, len(key_cols)))(*key_cols)]
        if pred is not None:
            keep.append(pred)
            q.pred = C.pointer(pred)
        q.nkeys = len(key_cols)
        q.key_cols = C.cast(keep[1], C.POINTER(C.c_int32))
        q.naggs = len(aggs)
        q.aggs = C.cast(keep[0], C.POINTER(abi.Agg))
        return self._submit(q, abi.AggOut(capacity, len(key_cols), len(aggs)), keep)

    def submit_join_agg(self, build: "DeviceTable", build_key: int, build_pred,
                        probe: "DeviceTable", probe_key: int, probe_pred, aggs,
                        radix_bits: int = -1, topk: int = 0,
                        capacity: int | None = None) -> "Ticket":
        q = abi.Query()
        q.kind = abi.VQ_Q_JOIN_AGG
        q.build = build.h
        q.table = probe.h
        keep = [abi.agg_array(aggs)]
        if build_pred is not None:
            keep.append(build_pred)
            q.build_pred = C.pointer(build_pred)
        if probe_pred is not None:
            keep.append(probe_pred)
            q.pred = C.pointer(probe_pred)
        q.build_key, q.probe_key = build_key, probe_key
        q.naggs = len(aggs)
        q.aggs = C.cast(keep[0], C.POINTER(abi.Agg))
        q.radix_bits, q.topk = radix_bits, topk
        if capacity is None:
            capacity = topk if topk > 0 else max(1, min(build.nrows, probe.nrows))
        return self._submit(q, abi.AggOut(capacity, 1, len(aggs)), keep)

    def submit_agg_hash(self, dt: "DeviceTable", pred, key_col: int, aggs,
                        capacity: int | None = None) -> "Ticket":
        q, out, keep = self.batch_hash(dt, pred, key_col, aggs, capacity)
        return self._submit(q, out, keep)

    # -- batches ------------------------------------------------------------
    @staticmethod
    def batch_hash(dt: "DeviceTable", pred, key_col: int, aggs, capacity: int | None = None):
        """A hash GROUP BY instance for batch_run."""
        q = abi.Query()
        q.kind = abi.VQ_Q_AGG_HASH
        q.table = dt.h
        kc = (C.c_int32 * 1)(key_col)
        keep = [abi.agg_array(aggs), kc]
        if pred is not None:
            keep.append(pred)
            q.pred = C.pointer(pred)
        q.nkeys = 1
        q.key_cols = C.cast(kc, C.POINTER(C.c_int32))
        q.naggs = len(aggs)
        q.aggs = C.cast(keep[0], C.POINTER(abi.Agg))
        return q, abi.AggOut(capacity or max(1, dt.nrows), 1, len(aggs)), keep

    @staticmethod
    def batch_join(build: "DeviceTable", build_key: int, build_pred, probe: "DeviceTable",
                   probe_key: int, probe_pred, aggs, topk: int = 0, capacity: int | None = None):
        """A join + aggregate instance for batch_run."""
        q = abi.Query()
        q.kind = abi.VQ_Q_JOIN_AGG
        q.build, q.table = build.h, probe.h
        keep = [abi.agg_array(aggs)]
        if build_pred is not None:
            keep.append(build_pred)
            q.build_pred = C.pointer(build_pred)
        if probe_pred is not None:
            keep.append(probe_pred)
            q.pred = C.pointer(probe_pred)
        q.build_key, q.probe_key = build_key, probe_key
        q.naggs = len(aggs)
        q.aggs = C.cast(keep[0], C.POINTER(abi.Agg))
        q.radix_bits, q.topk = -1, topk
        cap = capacity or (topk if topk > 0 else max(1, min(build.nrows, probe.nrows)))
        return q, abi.AggOut(cap, 1, len(aggs)), keep

    def batch_run(self, instances, flags: int = 0):
        """Runs [batch_hash / batch_join instances] as one batch; returns
        their results (AggOut, with .stats)."""
        n = len(instances)
        qs = (abi.Query * max(n, 1))(*[q for q, _, _ in instances])
        outs = (abi.AggResult * max(n, 1))(*[o.c for _, o, _ in instances])
        sts = (abi.Stats * max(n, 1))()
        _check(lib().vq_batch_run(self.h, n, qs, outs, sts, flags))
        res = []
        for i, (_, o, _) in enumerate(instances):
            o.c.ngroups = outs[i].ngroups
            o.stats = sts[i]
            res.append(o)
        return res

    def _submit(self, q, out, keep) -> "Ticket":
        st = abi.Stats()
        h = C.c_void_p()
        _check(lib().vq_submit(self.h, C.byref(q), C.byref(out.c), C.byref(st), C.byref(h)))
        out.stats = st
        return Ticket(h, out, keep)

    # -- operators ----------------------------------------------------------
    def filter(self, dt: "DeviceTable", pred) -> np.ndarray:
        sv = C.c_void_p()
        st = abi.Stats()
        _check(lib().vq_filter(self.h, dt.h, _pp(pred), C.byref(sv), C.byref(st)))
        try:
            n = lib().vq_selvec_count(sv)
            rows = np.empty(max(n, 0), dtype=np.uint32)
            _check(lib().vq_selvec_copy(sv, rows.ctypes.data_as(C.POINTER(C.c_uint32)), n))
        finally:
            lib().vq_selvec_destroy(sv)
        self.last_stats = st
        return rows

    def filter_count(self, dt: "DeviceTable", pred) -> int:
        """Filter -> selection vector left in HBM; returns the row count only."""
        sv = C.c_void_p()
        st = abi.Stats()
        _check(lib().vq_filter(self.h, dt.h, _pp(pred), C.byref(sv), C.byref(st)))
        n = lib().vq_selvec_count(sv)
        lib().vq_selvec_destroy(sv)
        self.last_stats = st
        return int(n)

    def agg_dict(self, dt: "DeviceTable", pred, key_cols, aggs, capacity: int | None = None,
                 out: abi.AggOut | None = None) -> abi.AggOut:
        """The dictionary-code GROUP BY.  `aggs`: a list of aggregates or an
        array abi.agg_array made of them (built once, for a query run many
        times); `out`: a result buffer of an earlier call to fill again."""
        if out is None:
            if capacity is None:  # at most one group per code tuple, and per row
                space = 1
                for k in key_cols:  # (a key with NULLs: one group more, the NULL one)
                    c = dt.table.columns[k]
                    space *= max(1, c.dict_size + (getattr(c, "valid", None) is not None))
                capacity = max(1, min(space, dt.nrows))
            out = abi.AggOut(capacity, len(key_cols), len(aggs))
        elif out.nkeys != len(key_cols) or out.naggs != len(aggs):
            raise ValueError("agg_dict: result buffer shaped for another query")
        keys = (C.c_int32 * max(1, len(key_cols)))(*key_cols)
        arr = aggs if isinstance(aggs, C.Array) else abi.agg_array(aggs)
        st = abi.Stats()
        _check(lib().vq_agg_dict(self.h, dt.h, _pp(pred), len(key_cols), keys, len(aggs), arr,
                                 C.byref(out.c), C.byref(st)), out)
        out.stats = st
        return out

    def agg_dict_multi(self, dt: "DeviceTable", preds, aggs) -> list:
        """vq_agg_dict_multi: global aggregates (no GROUP BY) of one shape,
        one per predicate in `preds` (None: every row), in one pass over the
        table.  Returns one AggOut per query."""
        n = len(preds)
        arr = aggs if isinstance(aggs, C.Array) else abi.agg_array(aggs)
        outs = [abi.AggOut(1, 0, len(aggs)) for _ in range(n)]
        parr = (abi.Pred * n)(*[p if p is not None else abi.Pred() for p in preds])
        res = (abi.AggResult * n)(*[o.c for o in outs])
        st = (abi.Stats * n)()
        rc = lib().vq_agg_dict_multi(self.h, dt.h, n, parr, len(aggs), arr, res, st)
        for i, o in enumerate(outs):
            o.c.ngroups = res[i].ngroups
            o.stats = st[i]
        _check(rc)
        return outs

    def agg_dict_noalloc(self, dt: "DeviceTable", pred, key_cols, aggs,
                         capacity: int = 4096) -> abi.AggOut:
        """agg_dict through vq_agg_dict_noalloc: no device allocation (the
        library's fixed per-device scratch array)."""
        keys = (C.c_int32 * max(1, len(key_cols)))(*key_cols)
        arr = abi.agg_array(aggs)
        out = abi.AggOut(capacity, len(key_cols), len(aggs))
        st = abi.Stats()
        _check(lib().vq_agg_dict_noalloc(self.h, dt.h, _pp(pred), len(key_cols), keys, len(aggs), arr,
                                         C.byref(out.c), C.byref(st)), out)
        out.stats = st
        return out

    def agg_hash(self, dt: "DeviceTable", pred, key_col: int, aggs,
                 capacity: int | None = None, out: abi.AggOut | None = None) -> abi.AggOut:
        """GROUP BY one integer column of any cardinality.  `out`: a result
        buffer of an earlier call to fill again (its pages already mapped; the
        previous result in it is overwritten)."""
        arr = abi.agg_array(aggs)
        if out is None:
            if capacity is None:
                capacity = max(1, dt.nrows)
            out = abi.AggOut(capacity, 1, len(aggs))
        elif out.nkeys != 1 or out.naggs != len(aggs):
            raise ValueError("agg_hash: result buffer shaped for another query")
        st = abi.Stats()
        _check(lib().vq_agg_hash(self.h, dt.h, _pp(pred), key_col, len(aggs), arr,
                                 C.byref(out.c), C.byref(st)), out)
        out.stats = st
        return out

    def join_agg(self, build: "DeviceTable", build_key: int, build_pred, probe: "DeviceTable",
                 probe_key: int, probe_pred, aggs, radix_bits: int = -1, topk: int = 0,
                 capacity: int | None = None) -> abi.AggOut:
        arr = abi.agg_array(aggs)
        if capacity is None:
            capacity = topk if topk > 0 else max(1, min(build.nrows, probe.nrows))
        out = abi.AggOut(capacity, 1, len(aggs))
        st = abi.Stats()
        _check(lib().vq_join_agg(self.h, build.h, build_key, _pp(build_pred), probe.h, probe_key,
                                 _pp(probe_pred), len(aggs), arr, radix_bits, topk,
                                 C.byref(out.c), C.byref(st)), out)
        out.stats = st
        return out


    def join_group(self, build: "DeviceTable", build_key: int, build_pred, probe: "DeviceTable",
                   probe_key: int, probe_pred, key_cols, aggs, capacity: int | None = None,
                   out: abi.AggOut | None = None) -> abi.AggOut:
        """vq_join_group: GROUP BY probe columns over the joined rows;
        aggregates over either side (build columns flagged VQ_BUILD_COL,
        abi.JoinView resolves names)."""
        arr = aggs if isinstance(aggs, C.Array) else abi.agg_array(aggs)
        karr = (C.c_int32 * len(key_cols))(*key_cols)
        if out is None:
            out = abi.AggOut(capacity if capacity is not None else max(1, probe.nrows), len(key_cols), len(aggs))
        st = abi.Stats()
        _check(lib().vq_join_group(self.h, build.h, build_key, _pp(build_pred), probe.h, probe_key,
                                   _pp(probe_pred), len(key_cols), karr, len(aggs), arr,
                                   C.byref(out.c), C.byref(st)), out)
        out.stats = st
        return out


    def join_agg_nested(self, inner: "DeviceTable", inner_key: int, inner_pred, build: "DeviceTable",
                        build_fk: int, build_key: int, build_pred, probe: "DeviceTable", probe_key: int,
                        probe_pred, pay_cols, aggs, radix_bits: int = -1, topk: int = 0,
                        capacity: int | None = None):
        """vq_join_agg_nested: (result, payload [ngroups, npay] int64)."""
        import numpy as np
        arr = aggs if isinstance(aggs, C.Array) else abi.agg_array(aggs)
        parr = (C.c_int32 * max(1, len(pay_cols)))(*pay_cols)
        if capacity is None:
            capacity = topk if topk > 0 else max(1, min(build.nrows, probe.nrows))
        out = abi.AggOut(capacity, 1, len(aggs))
        pay = np.zeros(capacity * max(1, len(pay_cols)), dtype=np.int64)
        st = abi.Stats()
        _check(lib().vq_join_agg_nested(self.h, inner.h, inner_key, _pp(inner_pred), build.h, build_fk, build_key,
                                        _pp(build_pred), probe.h, probe_key, _pp(probe_pred), len(pay_cols), parr,
                                        len(aggs), arr, radix_bits, topk, C.byref(out.c),
                                        pay.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(st)), out)
        out.stats = st
        n = out.ngroups
        return out, pay[: n * len(pay_cols)].reshape(n, len(pay_cols))


class Ticket:
    """A submitted query: wait() returns its result (raises on failure)."""

    def __init__(self, h, out, keep):
        self.h, self.out, self._keep = h, out, keep

    def done(self) -> bool:
        return self.h is None or bool(lib().vq_ticket_done(self.h))

    def wait(self) -> abi.AggOut:
        if self.h is not None:
            h, self.h = self.h, None
            _check(lib().vq_wait(h), self.out)
        return self.out

    def cancel(self) -> None:
        """vq_ticket_cancel: wait() then raises GpuError(VQ_E_CANCELED) unless
        the query had finished."""
        if self.h is not None:
            _check(lib().vq_ticket_cancel(self.h))

    def __del__(self):
        # The library writes into this ticket's result arrays and stats until
        # the query ends: a ticket dropped unwaited waits here, so they are
        # not freed under it, and its vq_ticket is released.
        if getattr(self, "h", None) is not None:
            try:
                h, self.h = self.h, None
                lib().vq_wait(h)
            except Exception:
                pass


class DeviceTable:
    """vq_table: columns resident in HBM after one bulk upload."""

    def __init__(self, ctx: Context, table, zones=None):
        self.ctx = ctx
        self.table = table
        self.nrows = table.nrows
        self.nrg = len(table.rg_rows)
        descs = abi.col_descs(table)
        rg, rgp = abi.rg_array(table)
        zp = None
        if zones is not None:
            zp = zones.ctypes.data_as(C.POINTER(abi.Zone)) if hasattr(zones, "ctypes") else zones
        h = C.c_void_p()
        _check(lib().vq_table_upload(ctx.h, table.nrows, len(table.columns), descs, len(rg), rgp,
                                     zp, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().vq_table_destroy(self.h)
            self.h = None

    def dict_order(self, col: int, rank=None) -> None:
        """vq_table_dict_order: the byte order of dictionary column col's
        strings (by default from the host table's dictionary), for MIN/MAX
        over strings."""
        if rank is None:
            rank = abi.dict_rank(self.table.columns[col].dictionary)
        r = np.ascontiguousarray(rank, dtype=np.uint32)
        _check(lib().vq_table_dict_order(self.h, col, r.ctypes.data_as(C.POINTER(C.c_uint32))))

    def zones(self):
        n = self.nrg * len(self.table.columns)
        z = (abi.Zone * max(n, 1))()
        _check(lib().vq_table_zones(self.h, z, n))
        return z

    def prune(self, pred) -> np.ndarray:
        keep = np.zeros(max(self.nrg, 1), dtype=np.uint8)
        _check(lib().vq_zone_prune(self.h, _pp(pred), keep.ctypes.data_as(C.POINTER(C.c_uint8))))
        return keep[: self.nrg]


class _ColInfo:
    def __init__(self, name, typ, dict_size):
        self.name, self.type, self.dict_size = name, typ, dict_size


class _Schema:
    """The column types and dictionary sizes of a device table (what the
    query helpers read from a host table)."""

    def __init__(self, cols, rg_rows):
        self.columns = cols
        self.rg_rows = rg_rows

    def index(self, name):
        for i, c in enumerate(self.columns):
            if c.name == name:
                return i
        raise KeyError(name)


class VxqTable(DeviceTable):
    """vq_table ingested from a .vxq image on the device."""

    def __init__(self, ctx: Context, image: bytes, cols, names=None):
        self.ctx = ctx
        cols = list(cols)
        arr = (C.c_int32 * max(1, len(cols)))(*cols)
        h = C.c_void_p()
        st = abi.Stats()
        _check(lib().vq_table_upload_vxq(ctx.h, image, len(image), len(cols), arr, C.byref(h),
                                         C.byref(st)))
        self.h = h
        self.stats = st
        self.nrows = int(lib().vq_table_rows(h))
        self.nrg = int(st.rg_total)
        info = []
        for i in range(len(cols)):
            t, d = C.c_int32(), C.c_int32()
            _check(lib().vq_table_column(h, i, C.byref(t), C.byref(d)))
            info.append(_ColInfo(names[i] if names else f"c{cols[i]}", t.value, d.value))
        self.table = _Schema(info, None)

    def dictionary(self, col: int):
        """The strings of column col's global codes, in code order."""
        out = []
        for code in range(self.table.columns[col].dict_size):
            p, n = C.c_char_p(), C.c_int64()
            _check(lib().vq_table_dict_entry(self.h, col, code, C.byref(p), C.byref(n)))
            out.append(C.string_at(p, n.value).decode())
        return out
