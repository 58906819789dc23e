"""Byte layouts against the reference's own code (tests/golden/layouts.json,
gen_layout_vectors.py): projection tuples as kern_form_minimal_tuple
(xpu_common.cu:1096-1209) writes them, and a KDS_FORMAT_HASH inner buffer as
KDS_HASH_FIRST_ITEM / KDS_HASH_NEXT_ITEM (xpu_common.h:1424-1451) read it."""
import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from pgstrom import abi, kds, hashing  # noqa: E402
from pgstrom.kds import Column  # noqa: E402
from helpers import project_all_query, run_oracle_projection  # noqa: E402
import layoutcases as LC  # noqa: E402

with open(os.path.join(HERE, "golden", "layouts.json")) as _f:
    GOLDEN = json.load(_f)


def tuple_chunk(rows=None):
    rows = LC.tuple_rows() if rows is None else rows
    cols = []
    for j, (name, t) in enumerate(LC.TUPLE_COLDEFS):
        vals = [r[j] for r in rows]
        nulls = np.array([v is None for v in vals])
        fill = b"" if t == "text" else 0
        cols.append(Column(name, t, [fill if v is None else v for v in vals], nulls))
    return cols


def dst_tuples(dst):
    """The kern_tupitem bytes of a kds_dst, in row-index order."""
    raw = bytes(dst)
    hd, _ = kds.kds_colmeta(memoryview(raw))
    head_sz = abi.kds_head_length(hd.nr_colmeta)
    ridx = np.frombuffer(raw, dtype=np.uint64, count=hd.nitems, offset=head_sz)
    out = []
    for o in ridx:
        off = hd.length - int(o)
        t_len = struct.unpack_from("<I", raw, off)[0]
        out.append(raw[off:off + t_len])
    return out


def external_tuple_chunk():
    """Heap pages (KDS_FORMAT_BLOCK) of the rows whose texts are TOAST
    pointers or compressed datums, which only a heap page can hold."""
    cols = tuple_chunk(LC.tuple_rows_external())
    return cols, kds.build_block_kds(cols, seed=8)


def test_oracle_projection_external_tuples_match_reference():
    """A compressed or external text is copied into the projected tuple as it
    is, and a TOAST pointer sets HEAP_HASEXTERNAL (xpu_text_datum_write,
    xpu_textlib.cu:312-335; kern_form_minimal_tuple)."""
    cols, src = external_tuple_chunk()
    dst, res = run_oracle_projection(project_all_query(cols), src, 1 << 20)
    assert [t.hex() for t in dst_tuples(dst)] == GOLDEN["tuples_external"]


def test_builder_tuples_match_reference():
    """The host builder of ROW / HASH inputs (kds.form_tupitem) writes the
    reference's bytes, TOAST pointers and compressed datums included."""
    for rowid, (row, exp) in enumerate(zip(LC.tuple_rows(), GOLDEN["tuples"])):
        assert kds.form_tupitem(LC.TUPLE_COLDEFS, row, rowid).hex() == exp, rowid
    for rowid, (row, exp) in enumerate(zip(LC.tuple_rows_external(), GOLDEN["tuples_external"])):
        assert kds.form_tupitem(LC.TUPLE_COLDEFS, row, rowid).hex() == exp, rowid


def test_oracle_projection_tuples_match_reference():
    cols = tuple_chunk()
    blob = project_all_query(cols)
    dst, res = run_oracle_projection(blob, kds.build_arrow_kds(cols), 1 << 20)
    got = dst_tuples(dst)
    assert [t.hex() for t in got] == GOLDEN["tuples"]


def walk(buf, h):
    """KDS_HASH_FIRST_ITEM / KDS_HASH_NEXT_ITEM restated: row ids on the chain
    of hash h whose item hash equals h."""
    raw = bytes(buf)
    hd, _ = kds.kds_colmeta(memoryview(raw))
    body = abi.kds_head_length(hd.nr_colmeta)
    off = struct.unpack_from("<Q", raw, body + 8 * (h % hd.hash_nslots))[0]
    out = []
    while off not in (0, 0xffffffffffffffff):
        item = hd.length - off
        nxt, t_len, th = struct.unpack_from("<QII", raw, item)
        if th == h:
            out.append(struct.unpack_from("<I", raw, item + 8 + t_len - 4)[0])
        off = nxt
    return out


def test_hash_layout_matches_reference():
    coldefs, rows, hashes = LC.hash_inner()
    buf = kds.build_hash_kds(coldefs, rows, hashes)
    g = GOLDEN["hash"]
    # the recorded walk is the reference's reading of exactly these bytes
    assert hashlib.sha256(bytes(buf)).hexdigest() == g["sha256"]
    for k, exp in zip(g["probe_keys"], g["walk"]):
        h = hashing.hash_any(struct.pack("<i", k))
        assert walk(buf, h) == exp, k
        assert sorted(exp) == [i for i, r in enumerate(rows) if r[0] == k]
