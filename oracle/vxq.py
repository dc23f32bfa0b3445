"""The .vxq columnar file format restated in Python: a writer that produces
files the reference's storage.Reader would read, and a CPU reader/decoder.

TEST INFRASTRUCTURE ONLY: the tests use it to make .vxq images and to decode
them on the CPU as the check of the device-side ingestion.  The product path
never uses it.

Format (storage/format.go:3-66, storage/writer.go:37-511,
storage/reader.go:295-408, storage/rowgroup.go:95-189, storage/codec.go:9-247,
internal/encoding/crc.go:9-40), all integers little-endian:

  "VXQ1"
  per row group, per column, a section:
    blocks of up to 1024 rows (BlockRows), each
      [128 B null bitmap, 1 = valid, LSB first][rows x value][4 B CRC-32 IEEE
       of the bitmap and the values]                    (writer.go:230-337)
    STRING columns: values are uint32 codes into the row group's own
    dictionary (first occurrence = code 0, writer.go:371-414), whose blob
    follows the blocks:
      [4 B n][n x 4 B offset][4 B data length][data][4 B CRC]  (codec.go:80-104)
  footer body:                                            (writer.go:418-462)
    [4 B ncols] per field [2 B name length][name][1 B type][1 B encoding][1 B flags]
    [4 B nrg] per row group [8 B file offset][4 B rows] per column
      [8 B section offset][8 B section length][8 B null count][8 B sum]
      [8 B min][8 B max][1 B has min/max][4 B dict offset][4 B dict length]
  trailer: [4 B CRC of the body][8 B body length]["VXQ1"]

The CRC is zlib's crc32 (CRC-32/ISO-HDLC, polynomial 0xEDB88320 reflected,
init and final xor 0xFFFFFFFF), which is Go's crc32.IEEE table.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

MAGIC = b"VXQ1"
BLOCK_ROWS = 1024
ROW_GROUP_ROWS = 65536
NULL_BITMAP_BYTES = 128
# storage/format.go:20-26
T_INT64, T_FLOAT64, T_BOOL, T_STRING, T_DATE = 0x01, 0x02, 0x03, 0x04, 0x05
ENC_PLAIN, ENC_RLE, ENC_DICT = 0, 1, 2
VALUE_SIZE = {T_INT64: 8, T_FLOAT64: 8, T_DATE: 4, T_STRING: 4}


class VxqError(Exception):
    """A checksum or structure that does not verify (storage.ErrChecksum,
    ErrCorruptFooter, ErrBadMagic)."""


def crc32(b) -> int:
    return zlib.crc32(bytes(b)) & 0xFFFFFFFF


def _bitmap(valid: np.ndarray) -> bytes:
    """The 128-byte null bitmap of one block (FullBitmap / subBitmap,
    codec.go:28-40, writer.go:497-508): bit i set when row i is valid."""
    bits = np.zeros(BLOCK_ROWS, dtype=np.uint8)
    bits[: len(valid)] = valid.astype(np.uint8)
    return np.packbits(bits, bitorder="little").tobytes()


def _dict_blob(strings) -> bytes:
    """DictBuilder.Marshal (codec.go:80-104)."""
    data = b"".join(s.encode() for s in strings)
    offs, o = [], 0
    for s in strings:
        offs.append(o)
        o += len(s.encode())
    b = struct.pack("<I", len(strings)) + b"".join(struct.pack("<I", x) for x in offs)
    b += struct.pack("<I", len(data)) + data
    return b + struct.pack("<I", crc32(b))


def _local_codes(col, start, rows, ok):
    """One row group's dictionary and codes of a STRING column (the writer's
    DictBuilder, codec.go:42-77: codes in first-occurrence order, NULL rows
    0xFFFFFFFF).  col: a list of str, or (global codes, names) -- then the
    strings are names[code]."""
    codes = np.full(rows, 0xFFFFFFFF, dtype=np.uint32)
    if isinstance(col, tuple):
        g = np.asarray(col[0][start:start + rows], dtype=np.int64)
        gv = g[ok]
        if gv.size == 0:
            return [], codes
        uniq, first = np.unique(gv, return_index=True)
        order = uniq[np.argsort(first, kind="stable")]
        local = np.zeros(int(order.max()) + 1, dtype=np.int64)
        local[order] = np.arange(order.size)
        codes[ok] = local[gv].astype(np.uint32)
        return [col[1][int(x)] for x in order], codes
    index, keys = {}, []
    for i in range(rows):
        if not ok[i]:
            continue
        s = col[start + i]
        c = index.get(s)
        if c is None:
            c = index[s] = len(keys)
            keys.append(s)
        codes[i] = c
    return keys, codes


def _zone(typ, vals, valid):
    """The section's ZoneMap as the writer computes it (writer.go:205-414):
    (null count, sum bits, min bits, max bits, has min/max)."""
    nulls = int((~valid).sum())
    v = vals[valid]
    if typ == T_STRING:
        return nulls, 0, 0, 0, 0  # set by the caller from the dictionary
    if len(v) == 0:
        return nulls, 0, 0, 0, 0
    if typ == T_INT64:
        s = int(np.sum(v.astype(np.int64), dtype=np.int64))
        return nulls, s, int(v.min()) & (2**64 - 1), int(v.max()) & (2**64 - 1), 1
    if typ == T_FLOAT64:
        acc = 0.0
        for x in v.tolist():  # row order, as the writer sums
            acc += x
        sb = struct.unpack("<q", struct.pack("<d", acc))[0]
        mn = struct.unpack("<Q", struct.pack("<d", float(v.min())))[0]
        mx = struct.unpack("<Q", struct.pack("<d", float(v.max())))[0]
        return nulls, sb, mn, mx, 1
    # DATE: raw bits zero-extended from uint32
    return nulls, 0, int(v.min()) & 0xFFFFFFFF, int(v.max()) & 0xFFFFFFFF, 1


def write_vxq(fields, columns, rg_rows, valid=None) -> bytes:
    """A .vxq image.  fields: [(name, type)]; columns: per field a numpy array
    (int64 / float64 / int32 days) or, for T_STRING, a list of str or a pair
    (codes, names); rg_rows: rows per row group; valid: optional per field
    boolean arrays (False = NULL)."""
    n = sum(int(r) for r in rg_rows)
    out = bytearray(MAGIC)
    rgs = []
    start = 0
    for rows in rg_rows:
        rows = int(rows)
        rg_off = len(out)
        cols = []
        for fi, (name, typ) in enumerate(fields):
            vals = columns[fi][start:start + rows]
            ok = np.ones(rows, dtype=bool) if valid is None or valid[fi] is None else \
                np.asarray(valid[fi][start:start + rows], dtype=bool)
            sec = len(out)
            dict_off = dict_len = 0
            if typ == T_STRING:
                keys, codes = _local_codes(columns[fi], start, rows, ok)
                payload_vals = codes
            elif typ == T_INT64:
                payload_vals = np.asarray(vals, dtype="<i8")
            elif typ == T_FLOAT64:
                payload_vals = np.asarray(vals, dtype="<f8")
            elif typ == T_DATE:
                payload_vals = np.asarray(vals, dtype="<i4")
            else:
                raise ValueError("BOOL columns are not written here")
            for b0 in range(0, rows, BLOCK_ROWS):
                b1 = min(rows, b0 + BLOCK_ROWS)
                blk = _bitmap(ok[b0:b1]) + payload_vals[b0:b1].astype(payload_vals.dtype.newbyteorder("<")).tobytes()
                out += blk + struct.pack("<I", crc32(blk))
            if typ == T_STRING:
                blob = _dict_blob(keys)
                dict_off = len(out) - sec
                out += blob
                dict_len = len(blob)
                z = (int((~ok).sum()), 0, 0, max(0, len(keys) - 1), 1 if keys else 0)
            else:
                z = _zone(typ, np.asarray(payload_vals), ok)
            cols.append((sec, len(out) - sec, z, dict_off, dict_len))
        rgs.append((rg_off, rows, cols))
        start += rows
    assert start == n
    body = bytearray(struct.pack("<I", len(fields)))
    for name, typ in fields:
        nb = name.encode()
        body += struct.pack("<H", len(nb)) + nb + bytes([typ, ENC_DICT if typ == T_STRING else ENC_PLAIN, 0])
    body += struct.pack("<I", len(rgs))
    for rg_off, rows, cols in rgs:
        body += struct.pack("<qI", rg_off, rows)
        for sec, seclen, (nulls, s, mn, mx, has), doff, dlen in cols:
            body += struct.pack("<qqqqQQBII", sec, seclen, nulls, s, mn, mx, has, doff, dlen)
    out += body + struct.pack("<I", crc32(body)) + struct.pack("<q", len(body)) + MAGIC
    return bytes(out)


def read_footer(img: bytes):
    """readFooter + parseFooter (reader.go:295-408): (fields, row groups),
    row group = (file offset, rows, [column meta dict])."""
    if len(img) < 16 + 4 or img[:4] != MAGIC or img[-4:] != MAGIC:
        raise VxqError("bad magic")
    crc, blen = struct.unpack("<Iq", img[-16:-4])
    if blen < 0 or blen > len(img) - 16 - 4:
        raise VxqError("corrupt footer")
    body = img[len(img) - 16 - blen: len(img) - 16]
    if crc32(body) != crc:
        raise VxqError("footer checksum")
    p = 0
    (ncols,) = struct.unpack_from("<I", body, p)
    p += 4
    fields = []
    for _ in range(ncols):
        (nl,) = struct.unpack_from("<H", body, p)
        p += 2
        name = body[p:p + nl].decode()
        p += nl
        fields.append((name, body[p], body[p + 1], body[p + 2]))
        p += 3
    (nrg,) = struct.unpack_from("<I", body, p)
    p += 4
    rgs = []
    for _ in range(nrg):
        off, rows = struct.unpack_from("<qI", body, p)
        p += 12
        cols = []
        for _ in range(ncols):
            sec, seclen, nulls, s, mn, mx, has, doff, dlen = struct.unpack_from("<qqqqQQBII", body, p)
            p += 57
            cols.append(dict(sec=sec, seclen=seclen, nulls=nulls, sum=s, min=mn, max=mx, has=has,
                             doff=doff, dlen=dlen))
        rgs.append((off, rows, cols))
    return fields, rgs


def _unmarshal_dict(blob: bytes):
    """UnmarshalDictionary (codec.go:216-247)."""
    if len(blob) < 4 or crc32(blob[:-4]) != struct.unpack("<I", blob[-4:])[0]:
        raise VxqError("dictionary checksum")
    (n,) = struct.unpack_from("<I", blob, 0)
    offs = list(struct.unpack_from(f"<{n}I", blob, 4))
    (tl,) = struct.unpack_from("<I", blob, 4 + 4 * n)
    data = blob[8 + 4 * n: 8 + 4 * n + tl]
    return [data[offs[i]:(offs[i + 1] if i + 1 < n else tl)].decode() for i in range(n)]


def decode(img: bytes, cols):
    """The columns `cols` of an image, decoded as TableScan reads them
    (NextBlock with its per-block CRC check, rowgroup.go:101-151) and with
    every STRING column's per-row-group codes remapped to global codes in
    first-seen order (getOrBuildRemap, exec/agg_intkey.go:118-134).
    Returns ([(type, values, valid, global dictionary or None)] per column,
    rows per row group)."""
    fields, rgs = read_footer(img)
    out = []
    for c in cols:
        typ = fields[c][1]
        vs = VALUE_SIZE.get(typ)
        if vs is None:
            raise VxqError("BOOL column")
        vals, valid, gdict, gindex = [], [], [], {}
        for _, rows, cm in rgs:
            m = cm[c]
            pos = m["sec"]
            remap = None
            if typ == T_STRING:
                blob = img[m["sec"] + m["doff"]: m["sec"] + m["doff"] + m["dlen"]]
                local = _unmarshal_dict(blob)
                remap = np.empty(max(1, len(local)), dtype=np.int64)
                for i, s in enumerate(local):
                    if s not in gindex:
                        gindex[s] = len(gdict)
                        gdict.append(s)
                    remap[i] = gindex[s]
            for b0 in range(0, rows, BLOCK_ROWS):
                r = min(BLOCK_ROWS, rows - b0)
                blen = NULL_BITMAP_BYTES + r * vs
                blk = img[pos:pos + blen]
                if crc32(blk) != struct.unpack_from("<I", img, pos + blen)[0]:
                    raise VxqError(f"crc block off={pos}")
                bits = np.unpackbits(np.frombuffer(blk[:NULL_BITMAP_BYTES], np.uint8),
                                     bitorder="little")[:r].astype(bool)
                raw = blk[NULL_BITMAP_BYTES:]
                if typ == T_STRING:
                    codes = np.frombuffer(raw, "<u4").astype(np.int64)
                    v = np.where(bits, remap[np.where(bits, codes, 0)], 0)
                elif typ == T_INT64:
                    v = np.frombuffer(raw, "<i8")
                elif typ == T_FLOAT64:
                    v = np.frombuffer(raw, "<f8")
                else:
                    v = np.frombuffer(raw, "<i4")
                vals.append(v)
                valid.append(bits)
                pos += blen + 4
        out.append((typ, np.concatenate(vals) if vals else np.zeros(0), np.concatenate(valid) if valid else
                    np.zeros(0, bool)))
        fields_dict = gdict if typ == T_STRING else None
        out[-1] = out[-1] + (fields_dict,)
    return out, [rows for _, rows, _ in rgs]


def table_to_vxq(t, rg_rows=None) -> bytes:
    """A datagen table as a .vxq image: dictionary-code columns as STRING
    columns of their dictionary's strings, int64 as INT64, float64 as
    FLOAT64, dates as DATE; row groups as the table's (or rg_rows)."""
    from vexq_amd import abi
    fields, cols = [], []
    for c in t.columns:
        if c.dict_size > 0:
            fields.append((c.name, T_STRING))
            cols.append((c.data, list(c.dictionary)))
        elif c.type == abi.VQ_F64:
            fields.append((c.name, T_FLOAT64))
            cols.append(c.data)
        elif c.type == abi.VQ_DATE:
            fields.append((c.name, T_DATE))
            cols.append(c.data)
        else:
            fields.append((c.name, T_INT64))
            cols.append(c.data.astype(np.int64))
    return write_vxq(fields, cols, t.rg_rows if rg_rows is None else rg_rows)
