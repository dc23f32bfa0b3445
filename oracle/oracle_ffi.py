"""ctypes binding of oracle/liboracle.so — the CPU restatement of vexq's
hot-path operators.

TEST INFRASTRUCTURE ONLY: imported by tests/ (the checker) and by bench.py's
cpu_baseline leg (the timed CPU baseline).  The product path never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from vexq_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_FILTER = os.path.join(HERE, "_ref", "libfilter_ref.so")
_lib = None


def build(ref_dir: str | None = None) -> None:
    """make -C oracle (and the reference filter kernel where its source is)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    rd = ref_dir or "/root/reference"
    if os.path.exists(os.path.join(rd, "bench/simd_filter/cpp/filter_kernel.cpp")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF_DIR={rd}"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    P = C.POINTER
    L.or_radix_hash.argtypes = [C.c_int64]
    L.or_radix_hash.restype = C.c_uint64
    L.or_radix_bits_for.argtypes = [C.c_int64]
    L.or_radix_bits_for.restype = C.c_int32
    L.or_join_table_capacity.argtypes = [C.c_int64]
    L.or_join_table_capacity.restype = C.c_int64
    L.or_xorshift_i64.argtypes = [C.c_int64, C.c_uint64, P(C.c_int64)]
    L.or_exact_sum.argtypes = [P(C.c_double), C.c_int64]
    L.or_exact_sum.restype = C.c_double
    L.or_set_dicts.argtypes = [C.c_int32, C.c_void_p, C.c_void_p]
    L.or_zone_maps.argtypes = [C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32, P(C.c_int64),
                               P(abi.Zone)]
    L.or_zone_prune.argtypes = [C.c_int32, P(abi.ColDesc), C.c_int32, P(abi.Zone), P(abi.Pred),
                                P(C.c_uint8)]
    L.or_filter.argtypes = [C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32, P(C.c_int64),
                            P(abi.Zone), P(abi.Pred), P(C.c_uint32)]
    L.or_filter.restype = C.c_int64
    L.or_agg_dict.argtypes = [C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32, P(C.c_int64),
                              P(abi.Zone), P(abi.Pred), C.c_int32, P(C.c_int32), C.c_int32,
                              P(abi.Agg), C.c_int32, P(abi.AggResult)]
    L.or_agg_hash.argtypes = [C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32, P(C.c_int64),
                              P(abi.Zone), P(abi.Pred), C.c_int32, C.c_int32, P(abi.Agg),
                              C.c_int32, P(abi.AggResult)]
    L.or_filter_parallel.argtypes = [C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32, P(C.c_int64),
                                     P(abi.Zone), P(abi.Pred), C.c_int32, P(C.c_uint32)]
    L.or_filter_parallel.restype = C.c_int64
    L.or_agg_hash_parallel.argtypes = [C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32,
                                       P(C.c_int64), P(abi.Zone), P(abi.Pred), C.c_int32,
                                       C.c_int32, P(abi.Agg), C.c_int32, P(abi.AggResult)]
    L.or_agg_dict_parallel.argtypes = [C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32,
                                       P(C.c_int64), P(abi.Zone), P(abi.Pred), C.c_int32,
                                       P(C.c_int32), C.c_int32, P(abi.Agg), C.c_int32,
                                       P(abi.AggResult)]
    join_args = [C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32, P(abi.Pred), C.c_int64,
                 C.c_int32, P(abi.ColDesc), C.c_int32, P(abi.Pred), C.c_int32, P(abi.Agg),
                 C.c_int32]
    L.or_join_agg.argtypes = join_args + [C.c_int32, P(abi.AggResult)]
    L.or_join_agg_nested.argtypes = ([C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32, P(abi.Pred),
                                      C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32, C.c_int32, P(abi.Pred),
                                      C.c_int64, C.c_int32, P(abi.ColDesc), C.c_int32, P(abi.Pred),
                                      C.c_int32, P(C.c_int32), C.c_int32, P(abi.Agg), C.c_int32, C.c_int32,
                                      P(abi.AggResult), P(C.c_int64)])
    L.or_join_group.argtypes = join_args[:10] + [C.c_int32, P(C.c_int32), C.c_int32, P(abi.Agg), C.c_int32,
                                                 P(abi.AggResult)]
    L.or_join_agg_parallel.argtypes = join_args + [C.c_int32, P(abi.AggResult)]
    _lib = L
    return L


def _pp(x):
    return C.byref(x) if x is not None else None


def radix_hash(key: int) -> int:
    return int(lib().or_radix_hash(key))


def radix_bits_for(est: int) -> int:
    return int(lib().or_radix_bits_for(est))


def join_table_capacity(keys: int) -> int:
    return int(lib().or_join_table_capacity(keys))


def exact_sum(x) -> float:
    a = np.ascontiguousarray(x, dtype=np.float64)
    return float(lib().or_exact_sum(a.ctypes.data_as(C.POINTER(C.c_double)), a.size))


def zone_maps(table) -> "C.Array":
    descs = abi.col_descs(table)
    rg, rgp = abi.rg_array(table)
    n = len(rg) * len(table.columns)
    z = (abi.Zone * max(n, 1))()
    lib().or_zone_maps(table.nrows, len(table.columns), descs, len(rg), rgp, z)
    return z


def zone_prune(table, zones, pred) -> np.ndarray:
    descs = abi.col_descs(table)
    keep = np.zeros(max(len(table.rg_rows), 1), dtype=np.uint8)
    lib().or_zone_prune(len(table.columns), descs, len(table.rg_rows), zones, _pp(pred),
                        keep.ctypes.data_as(C.POINTER(C.c_uint8)))
    return keep[: len(table.rg_rows)]


def filter_rows(table, pred, zones=None, nthreads: int = 0) -> np.ndarray:
    """Serial selection vector, or (nthreads > 0) the morsel-parallel CPU
    baseline of the same scan (the same rows)."""
    zones = zones if zones is not None else zone_maps(table)
    descs = abi.col_descs(table)
    rg, rgp = abi.rg_array(table)
    out = np.empty(max(table.nrows, 1), dtype=np.uint32)
    if nthreads:
        n = lib().or_filter_parallel(table.nrows, len(table.columns), descs, len(rg), rgp, zones,
                                     _pp(pred), nthreads, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    else:
        n = lib().or_filter(table.nrows, len(table.columns), descs, len(rg), rgp, zones, _pp(pred),
                            out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out[:n].copy()


class _Dicts:
    """The table's dictionary strings handed to the oracle for the length of
    one call (MIN/MAX over strings compares them)."""

    def __init__(self, table, aggs):
        self.on = any(a.accum == abi.VQ_STR for a in aggs)
        if not self.on:
            return
        n = len(table.columns)
        self.strs = (C.c_void_p * n)()
        self.lens = (C.c_void_p * n)()
        self.keep = []
        for i, c in enumerate(table.columns):
            if not c.dictionary:
                continue
            bs = [abi._as_bytes(x) for x in c.dictionary]
            sa = (C.c_char_p * len(bs))(*bs)
            la = (C.c_int64 * len(bs))(*[len(b) for b in bs])
            self.keep += [bs, sa, la]
            self.strs[i] = C.cast(sa, C.c_void_p)
            self.lens[i] = C.cast(la, C.c_void_p)

    def __enter__(self):
        if self.on:
            lib().or_set_dicts(len(self.strs), C.cast(self.strs, C.c_void_p), C.cast(self.lens, C.c_void_p))
        return self

    def __exit__(self, *a):
        if self.on:
            lib().or_set_dicts(0, None, None)


def agg_dict(table, pred, key_cols, aggs, exact: bool = True, capacity: int = 1 << 16,
             zones=None, nthreads: int = 0) -> abi.AggOut:
    """Serial restatement (nthreads == 0) or the morsel-parallel CPU baseline."""
    zones = zones if zones is not None else zone_maps(table)
    descs = abi.col_descs(table)
    rg, rgp = abi.rg_array(table)
    keys = (C.c_int32 * max(1, len(key_cols)))(*key_cols)
    arr = abi.agg_array(aggs)
    out = abi.AggOut(capacity, len(key_cols), len(aggs))
    with _Dicts(table, aggs):
        if nthreads:
            rc = lib().or_agg_dict_parallel(table.nrows, len(table.columns), descs, len(rg), rgp, zones,
                                            _pp(pred), len(key_cols), keys, len(aggs), arr, nthreads,
                                            C.byref(out.c))
        else:
            rc = lib().or_agg_dict(table.nrows, len(table.columns), descs, len(rg), rgp, zones,
                                   _pp(pred), len(key_cols), keys, len(aggs), arr, int(exact),
                                   C.byref(out.c))
    if rc:
        raise RuntimeError(f"oracle agg_dict failed: {rc}")
    return out


def agg_hash(table, pred, key_col, aggs, exact: bool = True, capacity: int | None = None,
             zones=None, nthreads: int = 0) -> abi.AggOut:
    """Serial GROUP BY one integer column (HashAggregate's general key path),
    or (nthreads > 0) the morsel-parallel CPU baseline of the same query."""
    zones = zones if zones is not None else zone_maps(table)
    descs = abi.col_descs(table)
    rg, rgp = abi.rg_array(table)
    arr = abi.agg_array(aggs)
    out = abi.AggOut(capacity or max(1, table.nrows), 1, len(aggs))
    if nthreads:
        rc = lib().or_agg_hash_parallel(table.nrows, len(table.columns), descs, len(rg), rgp, zones,
                                        _pp(pred), key_col, len(aggs), arr, nthreads, C.byref(out.c))
    else:
        rc = lib().or_agg_hash(table.nrows, len(table.columns), descs, len(rg), rgp, zones,
                               _pp(pred), key_col, len(aggs), arr, int(exact), C.byref(out.c))
    if rc:
        raise RuntimeError(f"oracle agg_hash failed: {rc}")
    return out


def join_agg(build, build_key, build_pred, probe, probe_key, probe_pred, aggs, topk=0,
             exact: bool = True, capacity: int | None = None, nthreads: int = 0) -> abi.AggOut:
    bd = abi.col_descs(build)
    pd = abi.col_descs(probe)
    arr = abi.agg_array(aggs)
    if capacity is None:
        capacity = topk if topk > 0 else max(1, min(build.nrows, probe.nrows))
    out = abi.AggOut(capacity, 1, len(aggs))
    args = (build.nrows, len(build.columns), bd, build_key, _pp(build_pred), probe.nrows,
            len(probe.columns), pd, probe_key, _pp(probe_pred), len(aggs), arr, topk)
    if nthreads:
        rc = lib().or_join_agg_parallel(*args, nthreads, C.byref(out.c))
    else:
        rc = lib().or_join_agg(*args, int(exact), C.byref(out.c))
    if rc:
        raise RuntimeError(f"oracle join_agg failed: {rc}")
    return out


def join_group(build, build_key, build_pred, probe, probe_key, probe_pred, key_cols, aggs,
               exact: bool = True, capacity: int | None = None) -> abi.AggOut:
    """or_join_group: GROUP BY probe columns over the joined rows, aggregates
    over either side (build columns flagged VQ_BUILD_COL; abi.JoinView builds
    them by name)."""
    bd = abi.col_descs(build)
    pd = abi.col_descs(probe)
    arr = abi.agg_array(aggs)
    karr = (C.c_int32 * len(key_cols))(*key_cols)
    if capacity is None:
        capacity = max(1, probe.nrows)
    out = abi.AggOut(capacity, len(key_cols), len(aggs))
    rc = lib().or_join_group(build.nrows, len(build.columns), bd, build_key, _pp(build_pred), probe.nrows,
                             len(probe.columns), pd, probe_key, _pp(probe_pred), len(key_cols), karr,
                             len(aggs), arr, int(exact), C.byref(out.c))
    if rc:
        raise RuntimeError(f"oracle join_group failed: {rc}")
    return out


def join_agg_nested(inner, inner_key, inner_pred, build, build_fk, build_key, build_pred, probe, probe_key,
                    probe_pred, pay_cols, aggs, topk=0, exact: bool = True, capacity: int | None = None):
    """or_join_agg_nested: (result, payload [ngroups, npay] int64)."""
    idesc, bd, pd = abi.col_descs(inner), abi.col_descs(build), abi.col_descs(probe)
    arr = abi.agg_array(aggs)
    parr = (C.c_int32 * max(1, len(pay_cols)))(*pay_cols)
    if capacity is None:
        capacity = topk if topk > 0 else max(1, probe.nrows)
    out = abi.AggOut(capacity, 1, len(aggs))
    pay = np.zeros(capacity * max(1, len(pay_cols)), dtype=np.int64)
    rc = lib().or_join_agg_nested(inner.nrows, len(inner.columns), idesc, inner_key, _pp(inner_pred), build.nrows,
                                  len(build.columns), bd, build_fk, build_key, _pp(build_pred), probe.nrows,
                                  len(probe.columns), pd, probe_key, _pp(probe_pred), len(pay_cols), parr, len(aggs),
                                  arr, topk, int(exact), C.byref(out.c),
                                  pay.ctypes.data_as(C.POINTER(C.c_int64)))
    if rc:
        raise RuntimeError(f"oracle join_agg_nested failed: {rc}")
    n = out.ngroups
    return out, pay[: n * len(pay_cols)].reshape(n, len(pay_cols))


# ---- the reference's own AVX2 filter kernel (built from its source) ---------
_ref = None


def ref_filter_lib():
    """bench/simd_filter/cpp/filter_kernel.cpp compiled by `make -C oracle ref`,
    or None where the checkout was not present at build time."""
    global _ref
    if _ref is None and os.path.exists(REF_FILTER):
        L = C.CDLL(REF_FILTER)
        for name in ("_Z18filter_gt_i64_avx2PKlmlPj", "_Z20filter_gt_i64_scalarPKlmlPj"):
            f = getattr(L, name)
            f.argtypes = [C.POINTER(C.c_int64), C.c_size_t, C.c_int64, C.POINTER(C.c_uint32)]
            f.restype = C.c_size_t
        for name in ("_Z18filter_gt_i32_avx2PKimiPj", "_Z20filter_gt_i32_scalarPKimiPj"):
            f = getattr(L, name)
            f.argtypes = [C.POINTER(C.c_int32), C.c_size_t, C.c_int32, C.POINTER(C.c_uint32)]
            f.restype = C.c_size_t
        _ref = L
    return _ref


def ref_filter_gt_i64(data: np.ndarray, threshold: int, avx2: bool = True) -> np.ndarray:
    L = ref_filter_lib()
    d = np.ascontiguousarray(data, dtype=np.int64)
    out = np.empty(max(d.size, 1), dtype=np.uint32)
    f = L._Z18filter_gt_i64_avx2PKlmlPj if avx2 else L._Z20filter_gt_i64_scalarPKlmlPj
    n = f(d.ctypes.data_as(C.POINTER(C.c_int64)), d.size, threshold,
          out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out[:n].copy()


def xorshift_i64(n: int, seed: int = 42) -> np.ndarray:
    """The reference benchmark's data (bench/simd_filter/cpp/main.cpp:28-45):
    xorshift64* from state 42, values (x * 0x2545F4914F6CDD1D) % 50 + 1."""
    out = np.empty(max(n, 1), dtype=np.int64)
    lib().or_xorshift_i64(n, seed, out.ctypes.data_as(C.POINTER(C.c_int64)))
    return out[:n]
