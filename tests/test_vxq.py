"""The .vxq restatement (oracle/vxq.py) pinned against what the reference's
own storage tests and comments fix, and against zlib's CRC-32 -- it is the
check of the device-side ingestion.  CPU only."""
import json
import os
import struct

import numpy as np
import pytest

from oracle import vxq

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "kats.json")))


def test_crc_is_crc32_ieee():
    # the CRC-32/ISO-HDLC check value (hash/crc32's IEEE table,
    # internal/encoding/crc.go:9-16)
    assert vxq.crc32(b"123456789") == 0xCBF43926
    assert vxq.crc32(b"") == 0


def test_section_sizes_match_the_readers_arithmetic():
    """storage/reader.go:14-46: a full row group of 8-byte values is
    64 x (128 B bitmap + 8 KiB values + 4 B CRC) = 520.25 KiB, of 4-byte
    values 264.25 KiB."""
    n = vxq.ROW_GROUP_ROWS
    img = vxq.write_vxq([("a", vxq.T_INT64), ("d", vxq.T_DATE)],
                        [np.arange(n, dtype=np.int64), np.zeros(n, np.int32)], [n])
    _, rgs = vxq.read_footer(img)
    assert rgs[0][2][0]["seclen"] == int(520.25 * 1024)
    assert rgs[0][2][1]["seclen"] == int(264.25 * 1024)
    assert rgs[0][2][0]["sec"] == 4  # right after the magic


def test_round_trip_1m_rows():
    """TestRoundTrip1MRows (storage/storage_test.go:161-230): ids 0..N-1 and
    i * 0.5 in 65,536-row groups read back block by block."""
    N = 1_000_000
    ids = np.arange(N, dtype=np.int64)
    vals = ids * 0.5
    rg = [vxq.ROW_GROUP_ROWS] * (N // vxq.ROW_GROUP_ROWS) + [N % vxq.ROW_GROUP_ROWS]
    img = vxq.write_vxq([("id", vxq.T_INT64), ("val", vxq.T_FLOAT64)], [ids, vals], rg)
    cols, rgs = vxq.decode(img, [0, 1])
    assert rgs == rg
    np.testing.assert_array_equal(cols[0][1], ids)
    np.testing.assert_array_equal(cols[1][1], vals)
    assert cols[0][2].all() and cols[1][2].all()


def test_crc_corruption_is_detected():
    """TestCRCCorruption (storage_test.go:232-271): a byte flipped at offset
    100 of a 100-row file fails the first block's checksum."""
    img = bytearray(vxq.write_vxq([("x", vxq.T_INT64)], [np.zeros(100, np.int64)], [100]))
    img[100] ^= 0xFF
    with pytest.raises(vxq.VxqError, match="crc block"):
        vxq.decode(bytes(img), [0])


def test_truncated_footer_is_detected():
    """TestFooterCorruption (storage_test.go:273-289)."""
    img = vxq.write_vxq([("x", vxq.T_INT64)], [np.zeros(0, np.int64)], [])
    with pytest.raises(vxq.VxqError):
        vxq.read_footer(img[:-1])


def test_string_dictionary_round_trip():
    """TestStringDictionary (storage_test.go:394-467): 4 row groups of
    'A'+i%50 strings; each dictionary stays small, the strings read back."""
    N = vxq.ROW_GROUP_ROWS * 4
    strs = [chr(ord("A") + i % 50) for i in range(N)]
    img = vxq.write_vxq([("s", vxq.T_STRING)], [strs], [vxq.ROW_GROUP_ROWS] * 4)
    _, rgs = vxq.read_footer(img)
    assert all(r[2][0]["dlen"] <= 5 * 1024 for r in rgs)
    (typ, codes, ok, gdict), = vxq.decode(img, [0])[0]
    assert [gdict[c] for c in codes[:200]] == strs[:200]
    assert [gdict[c] for c in codes] == strs


def test_global_codes_follow_first_occurrence_across_row_groups():
    """The int-key remap KAT (exec/agg_intkey_test.go:221-310): two row
    groups whose local dictionaries order R, A, N differently; global codes
    in first-seen order R=0, A=1, N=2, and the grouped sums."""
    kat = KATS["intkey_remap"]
    strs = [kat["dictionary"][f] for f in kat["flag"]]
    img = vxq.write_vxq([("flag", vxq.T_STRING), ("value", vxq.T_INT64)],
                        [strs, np.array(kat["value"], np.int64)], kat["row_groups"])
    _, rgs = vxq.read_footer(img)
    d0 = vxq._unmarshal_dict(img[rgs[0][2][0]["sec"] + rgs[0][2][0]["doff"]:][:rgs[0][2][0]["dlen"]])
    d1 = vxq._unmarshal_dict(img[rgs[1][2][0]["sec"] + rgs[1][2][0]["doff"]:][:rgs[1][2][0]["dlen"]])
    assert d0 != d1  # the row groups' local codes differ
    (_, codes, _, gdict), (_, vals, _, _) = vxq.decode(img, [0, 1])[0]
    assert gdict == kat["dictionary"]
    np.testing.assert_array_equal(codes, kat["flag"])
    for s, (total, cnt) in kat["expected"].items():
        m = codes == gdict.index(s)
        assert int(vals[m].sum()) == total and int(m.sum()) == cnt


def test_footer_zone_maps_are_the_writers():
    """The footer's zone maps carry raw bits (storage/schema.go:26-35): int64
    min/max as uint64, dates zero-extended from uint32, float64 bits, the
    int64 sum wrapping, the float64 sum the row-order fold."""
    v = np.array([5, -3, 7], np.int64)
    d = np.array([-2, 9, 4], np.int32)
    f = np.array([0.5, 0.25, 2.0])
    img = vxq.write_vxq([("v", vxq.T_INT64), ("d", vxq.T_DATE), ("f", vxq.T_FLOAT64)], [v, d, f], [3])
    _, rgs = vxq.read_footer(img)
    cv, cd, cf = rgs[0][2]
    assert (cv["min"], cv["max"], cv["sum"]) == ((-3) & (2**64 - 1), 7, 9)
    assert (cd["min"], cd["max"]) == ((-2) & 0xFFFFFFFF, 9)
    assert cf["min"] == struct.unpack("<Q", struct.pack("<d", 0.25))[0]
    assert cf["sum"] == struct.unpack("<q", struct.pack("<d", 2.75))[0]
