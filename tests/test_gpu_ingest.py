"""Device-side ingestion (vq_table_upload_vxq): a .vxq image written as the
reference's storage.Writer writes it, checked and decoded on the GPU, gives
the same table as the upload of decoded columns -- the same query results
bit for bit, the same zone maps and pruning -- with dictionary codes remapped
to global codes in first-seen order; corrupt blocks, dictionaries and
footers are refused with VQ_E_CORRUPT naming the place."""
import numpy as np
import pytest

from oracle import oracle_ffi as O
from oracle import vxq
from vexq_amd import abi, datagen, gpu

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    n = a.ngroups
    assert b.ngroups == n
    for f in ("keys", "counts", "first_row", "ivals", "valid"):
        np.testing.assert_array_equal(getattr(a, f)[:n], getattr(b, f)[:n])
    np.testing.assert_array_equal(a.fvals[:n].view(np.int64), b.fvals[:n].view(np.int64))


def _first_seen(t):
    """t with every dictionary column recoded so its codes follow first
    occurrence in row order (what a scan of the file produces)."""
    cols = []
    for c in t.columns:
        if c.dict_size:
            codes, names = datagen.first_occurrence_codes(c.data.astype(np.int64), list(c.dictionary))
            c = datagen.Column(c.name, c.type, codes.astype(c.data.dtype), len(names), names)
        cols.append(c)
    return datagen.Table(cols, t.rg_rows)


def _odd_groups(n):
    sizes, i, pat = [], 0, [65536, 1003, 4097, 65536, 1, 777]
    while sum(sizes) + pat[i % len(pat)] <= n:
        sizes.append(pat[i % len(pat)])
        i += 1
    if n - sum(sizes):
        sizes.append(n - sum(sizes))
    return np.array(sizes, dtype=np.int64)


@pytest.mark.parametrize("layout", ["64k", "odd"])
def test_ingested_table_answers_as_the_decoded_upload(gpu_ctx, layout):
    n = 700_001
    t = datagen.lineitem_sorted(n, seed=31)
    t = _first_seen(datagen.Table(t.columns, datagen.row_groups(n) if layout == "64k" else _odd_groups(n)))
    print(f"[case] ingest layout={layout}", flush=True)
    img = vxq.table_to_vxq(t)
    vt = gpu_ctx.upload_vxq(img, list(range(len(t.columns))))
    dt = gpu_ctx.upload(t)
    try:
        assert vt.nrows == n and vt.nrg == len(t.rg_rows)
        for i, c in enumerate(t.columns):
            info = vt.table.columns[i]
            assert info.type == c.type and info.dict_size == c.dict_size
            if c.dict_size:
                assert vt.dictionary(i) == list(c.dictionary)
        # zone maps: the footer's raw bits equal the ones computed at upload
        zv, zd = vt.zones(), dt.zones()
        for a, b in zip(zv, zd):
            assert (a.min_bits, a.max_bits, a.has_minmax, a.null_count) == \
                   (b.min_bits, b.max_bits, b.has_minmax, b.null_count)
        pred, keys, aggs = datagen.q_groupby(t)
        same_bits(gpu_ctx.agg_dict(vt, pred, keys, aggs), gpu_ctx.agg_dict(dt, pred, keys, aggs))
        for which in range(2):
            sp, sk, sa = datagen.q_scan_sum(t, which=which)
            np.testing.assert_array_equal(vt.prune(sp), dt.prune(sp))
            same_bits(gpu_ctx.agg_dict(vt, sp, sk, sa), gpu_ctx.agg_dict(dt, sp, sk, sa))
            np.testing.assert_array_equal(gpu_ctx.filter(vt, sp), gpu_ctx.filter(dt, sp))
        ha = [abi.count_star(t), abi.agg_sum("price", t)]
        same_bits(gpu_ctx.agg_hash(vt, pred, t.index("quantity"), ha),
                  gpu_ctx.agg_hash(dt, pred, t.index("quantity"), ha))
        st = vt.stats
        print(f"[measure] ingest {n} rows: {st.bytes_alg} B in {st.device_ms:.3f} ms on the device",
              flush=True)
    finally:
        vt.close()
        dt.close()


def test_per_row_group_dictionaries_remap_to_first_seen_codes(gpu_ctx):
    """The int-key remap KAT (exec/agg_intkey_test.go:221-310) through the
    device: two row groups with differently ordered local dictionaries, the
    GROUP BY over the global codes."""
    import json
    import os
    kat = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kats.json")))["intkey_remap"]
    strs = [kat["dictionary"][f] for f in kat["flag"]]
    img = vxq.write_vxq([("flag", vxq.T_STRING), ("value", vxq.T_INT64)],
                        [strs, np.array(kat["value"], np.int64)], kat["row_groups"])
    vt = gpu_ctx.upload_vxq(img, [0, 1], )
    try:
        assert vt.dictionary(0) == kat["dictionary"]
        t = datagen.Table([datagen.Column("flag", abi.VQ_U8, np.zeros(8, np.uint8), 3, kat["dictionary"]),
                           datagen.Column("value", abi.VQ_I64, np.zeros(8, np.int64))],
                          np.array(kat["row_groups"], np.int64))
        got = gpu_ctx.agg_dict(vt, None, [0], [abi.agg_sum("value", t), abi.count_star(t)])
        names = vt.dictionary(0)
        res = {names[int(k)]: [int(got.ivals[i, 0]), int(got.ivals[i, 1])] for i, k in enumerate(got.keys[:, 0])}
        assert res == kat["expected"]
        assert [names[int(k)] for k in got.keys[:, 0]] == ["R", "A", "N"]  # first-occurrence order
    finally:
        vt.close()


def _expect_corrupt(ctx, img, cols, pattern):
    with pytest.raises(gpu.GpuError) as e:
        ctx.upload_vxq(img, cols)
    assert e.value.code == abi.VQ_E_CORRUPT, str(e.value)
    assert pattern in str(e.value), str(e.value)


def test_corruption_is_refused(gpu_ctx):
    print("[case] ingest corruption", flush=True)
    t = datagen.lineitem(200_000, seed=33)
    img = bytearray(vxq.table_to_vxq(t))
    ctx = gpu_ctx
    before = ctx.bytes_in_use
    # a byte of the first block (TestCRCCorruption, storage_test.go:232-271)
    bad = bytearray(img)
    bad[100] ^= 0xFF
    _expect_corrupt(ctx, bytes(bad), list(range(len(t.columns))), "crc block rg=0 col=flagA")
    # a value in row group 2 of the price column
    _, rgs = vxq.read_footer(bytes(img))
    pc = t.index("price")
    off = rgs[2][2][pc]["sec"] + 128 + 8 * 17
    bad = bytearray(img)
    bad[off] ^= 0x01
    _expect_corrupt(ctx, bytes(bad), list(range(len(t.columns))), "crc block rg=2 col=price")
    # ... which a query leaving that column out never reads
    vt = ctx.upload_vxq(bytes(bad), [0, 1])
    vt.close()
    # a dictionary blob
    fa = rgs[1][2][0]
    bad = bytearray(img)
    bad[fa["sec"] + fa["doff"] + 9] ^= 0x20
    _expect_corrupt(ctx, bytes(bad), [0], "dictionary checksum")
    # the footer (TestFooterCorruption, storage_test.go:273-289)
    _expect_corrupt(ctx, bytes(img[:-1]), [0], "magic")
    bad = bytearray(img)
    bad[-20] ^= 0x01
    _expect_corrupt(ctx, bytes(bad), [0], "footer")
    assert ctx.bytes_in_use == before


def _blocks(img, vs_of):
    """(row group, column, offset, CRC'd length) of every block in the image,
    from its footer: [128-byte bitmap][rows x vs][CRC] back to back from the
    start of each column section (storage/rowgroup.go:101-151)."""
    _, rgs = vxq.read_footer(img)
    out = []
    for g, (_, rows, cols) in enumerate(rgs):
        for ci, m in enumerate(cols):
            pos = m["sec"]
            for b0 in range(0, rows, 1024):
                n = 128 + min(1024, rows - b0) * vs_of[ci]
                out.append((g, ci, pos, n))
                pos += n + 4
    return out


def test_every_block_edge_is_under_the_checksum(gpu_ctx):
    """A flipped bit at each edge of a block -- its first bitmap byte, its
    first and last value bytes, each byte of its stored CRC -- is refused and
    named, for blocks starting at every offset mod 16 and of full and partial
    length; the untouched image loads."""
    print("[case] ingest block edges", flush=True)
    rng = np.random.default_rng(35)
    sizes = [1024 * 3 + 5, 999, 2048, 1, 3001, 17, 1030, 2500]
    n = sum(sizes)
    names = ["N", "R", "A", "xyz", "q"]
    strs = [names[i] for i in rng.integers(0, len(names), n)]
    fields = [("s", vxq.T_STRING), ("v", vxq.T_INT64), ("d", vxq.T_DATE), ("f", vxq.T_FLOAT64)]
    cols = [strs, rng.integers(-2**62, 2**62, n), rng.integers(0, 20000, n).astype(np.int32),
            rng.standard_normal(n)]
    img = vxq.write_vxq(fields, cols, sizes)
    vt = gpu_ctx.upload_vxq(img, [0, 1, 2, 3])
    vt.close()
    fnames = [f[0] for f in fields]
    blocks = _blocks(img, [4, 8, 4, 8])
    picked, seen = [], set()
    for b in blocks:  # one block per (offset mod 16, full or partial)
        key = (b[2] % 16, b[3] == 128 + 1024 * (8 if b[1] in (1, 3) else 4))
        if key not in seen:
            seen.add(key)
            picked.append(b)
    assert len({b[2] % 16 for b in picked}) == 16
    before = gpu_ctx.bytes_in_use
    for g, ci, pos, ln in picked:
        for at in (0, 127, 128, ln - 1, ln, ln + 1, ln + 2, ln + 3):
            bad = bytearray(img)
            bad[pos + at] ^= 1 << (at % 8)
            _expect_corrupt(gpu_ctx, bytes(bad), [0, 1, 2, 3], f"crc block rg={g} col={fnames[ci]} off={pos}")
    assert gpu_ctx.bytes_in_use == before


def test_nulls_and_bool_are_unsupported(gpu_ctx):
    print("[case] ingest nulls", flush=True)
    n = 5000
    v = np.arange(n, dtype=np.int64)
    ok = np.ones(n, bool)
    ok[17] = False
    img = vxq.write_vxq([("v", vxq.T_INT64), ("w", vxq.T_INT64)], [v, v], [n], valid=[ok, None])
    with pytest.raises(gpu.GpuError) as e:
        gpu_ctx.upload_vxq(img, [0])
    assert e.value.code == abi.VQ_E_UNSUPPORTED and "NULL" in str(e.value)
    vt = gpu_ctx.upload_vxq(img, [1])  # the column without NULLs is fine
    try:
        t = datagen.Table([datagen.Column("w", abi.VQ_I64, v)], np.array([n], np.int64))
        got = gpu_ctx.agg_dict(vt, None, [], [abi.agg_sum("w", t)])
        assert int(got.ivals[0, 0]) == int(v.sum())
    finally:
        vt.close()
