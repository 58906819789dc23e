"""Write tests/golden/layouts.json: byte layouts read or written by the
reference's own code (oracle/_ref/libpgstrom_ref.so, `make -C oracle ref`):

  tuples   kern_form_minimal_tuple (xpu_common.cu:1096-1209) over rows of
           mixed types with NULL patterns and texts of several lengths, into
           a KDS_FORMAT_ROW kds_dst of those columns: the kern_tupitem bytes
  tuples_external  the same over rows whose texts are TOAST pointers or
           inline compressed datums (read by xpu_text_datum_heap_read as
           length -1, written as they are)
  hash     a KDS_FORMAT_HASH inner buffer built by this repository's builder
           (pgstrom.kds.build_hash_kds, keys with duplicates), walked with the
           reference's KDS_HASH_FIRST_ITEM / KDS_HASH_NEXT_ITEM
           (xpu_common.h:1424-1451) for the hash of every key 0..47: the row
           ids on each chain whose hash matches, in chain order; with a
           digest of the buffer's bytes

Run in the build container: make -C oracle ref && python3 tests/golden/gen_layout_vectors.py
"""
import ctypes as C
import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "pg-strom_amd"))
sys.path.insert(0, HERE)

from pgstrom import abi, kds, hashing, opcodes  # noqa: E402
import layoutcases as LC  # noqa: E402


def main():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libpgstrom_ref.so"))
    vp = C.c_void_p
    L.ref_form_tuple.restype = C.c_int
    L.ref_form_tuple.argtypes = [vp, C.c_int, vp, vp, vp, C.c_uint32, vp, C.c_int]
    L.ref_hash_walk.restype = C.c_int64
    L.ref_hash_walk.argtypes = [vp, C.c_int, vp, vp, vp, C.c_int64]

    head = bytes(kds.kds_header(abi.KDS_FORMAT_ROW, LC.TUPLE_COLDEFS, 1 << 20))
    hbuf = C.create_string_buffer(head, len(head))
    types = (C.c_int32 * len(LC.TUPLE_COLDEFS))(*[opcodes.TYPES[t] for _, t in LC.TUPLE_COLDEFS])

    def form(rows):
        return [form_one(rowid, row) for rowid, row in enumerate(rows)]

    def form_one(rowid, row):
        data = bytearray()
        offs = []
        for (_, t), v in zip(LC.TUPLE_COLDEFS, row):
            if v is None:
                offs.append(-1)
                continue
            data += b"\0" * ((-len(data)) % 8)
            offs.append(len(data))
            data += kds.encode_datum(t, v)
        data += b"\0" * 16
        dbuf = C.create_string_buffer(bytes(data), len(data))
        ob = (C.c_int64 * len(offs))(*offs)
        out = C.create_string_buffer(4096)
        n = L.ref_form_tuple(hbuf, len(row), types, dbuf, ob, rowid, out, 4096)
        assert n > 0
        return out.raw[:n].hex()

    tuples = form(LC.tuple_rows())
    tuples_ext = form(LC.tuple_rows_external())

    coldefs, rows, hashes = LC.hash_inner()
    buf = kds.build_hash_kds(coldefs, rows, hashes)
    probe = [hashing.hash_any(struct.pack("<i", k)) for k in range(48)]
    hb = C.create_string_buffer(bytes(buf), len(buf))
    ph = (C.c_uint32 * len(probe))(*probe)
    counts = (C.c_uint32 * len(probe))()
    rowids = (C.c_uint32 * 4096)()
    total = L.ref_hash_walk(hb, len(probe), ph, counts, rowids, 4096)
    assert total >= 0
    walk, pos = [], 0
    for c in counts:
        walk.append(list(rowids[pos:pos + c]))
        pos += c
    with open(os.path.join(HERE, "layouts.json"), "w") as f:
        json.dump({"tuples": tuples, "tuples_external": tuples_ext,
                   "hash": {"sha256": hashlib.sha256(bytes(buf)).hexdigest(),
                            "probe_keys": list(range(48)), "walk": walk}}, f, indent=0)
        f.write("\n")
    print("%d tuples, %d chains (%d items)" % (len(tuples), len(walk), total))


if __name__ == "__main__":
    main()
