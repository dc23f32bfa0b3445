"""The C ABI library: it loads, exports every symbol include/vexq_gpu.h
declares, its structs have the layout the C compiler gives them, and its
host-only functions agree with the oracle.  No GPU needed."""
import ctypes as C
import os
import re
import subprocess

import pytest

from oracle import oracle_ffi as O
from vexq_amd import abi, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vexq_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vq_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for want in ("vq_ctx_create", "vq_table_upload", "vq_filter", "vq_agg_dict", "vq_join_agg",
                 "vq_zone_prune", "vq_last_error"):
        assert want in names


def test_library_exports_every_declared_symbol():
    lib = gpu.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", gpu._LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (vq_\w+)", out))
    assert set(declared_functions()) <= exported


def test_struct_layout_matches_c(tmp_path):
    """Compile a probe against the header with the host C compiler and compare
    every struct's size and field offsets with the ctypes mirror."""
    structs = {"vq_zone": abi.Zone, "vq_col_desc": abi.ColDesc, "vq_term": abi.Term,
               "vq_pred": abi.Pred, "vq_factor": abi.Factor, "vq_expr": abi.Expr,
               "vq_agg": abi.Agg, "vq_agg_result": abi.AggResult, "vq_stats": abi.Stats,
               "vq_query": abi.Query}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def test_host_functions_match_oracle():
    lib = gpu.lib()
    for est in (0, 65535, 65536, 131073, 300_000, 1 << 40):
        assert lib.vq_radix_bits_for(est) == O.radix_bits_for(est)
    for k in (0, 1, -1, 12345, -(2**63), 2**63 - 1):
        assert lib.vq_radix_hash(k) == O.radix_hash(k)
    for keys in (-1, 0, 1, 7, 8, 100, 700_000):
        assert lib.vq_join_table_capacity(keys) == O.join_table_capacity(keys)


def test_last_error_starts_empty_and_null_args_fail_cleanly():
    lib = gpu.lib()
    assert lib.vq_last_error() in (b"",) or isinstance(lib.vq_last_error(), bytes)
    # NULL out-pointer: rejected before any device call
    assert lib.vq_ctx_create(0, 0, None) == abi.VQ_E_ARG
    assert b"NULL" in lib.vq_last_error()
    assert lib.vq_agg_dict(None, None, None, 0, None, 0, None, None, None) == abi.VQ_E_ARG


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(gpu, "_lib", None)
    monkeypatch.setattr(gpu, "_LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="not built"):
        gpu.lib()
