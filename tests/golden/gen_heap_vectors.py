"""Write tests/golden/heap_deform.json: the reference's own heap-tuple deform
(ExecLoadVarsOuterHeap, xpu_common.cu:1587-1620 with kern_extract_heap_tuple
:194-257 and __kern_extract_common_tuple :60-130, built into
oracle/_ref/libpgstrom_ref.so by `make -C oracle ref`) on every LP_NORMAL
tuple of two small KDS_FORMAT_BLOCK chunks made by pgstrom.kds.build_block_kds:

  mixed     datagen.heap_table(300): int2, int4, bool, float8, text with short
            and 4-byte varlena headers, int8, float4, text; 10% NULLs;
            invisible rows (LP_UNUSED), LP_DEAD and LP_REDIRECT items
  external  int4, text holding TOAST pointers in some rows, int8, text

The fixture holds each chunk's sha256 (so the test knows it rebuilt the same
bytes) and, per tuple in page order, every attribute: null, the integer value,
the float's bit pattern, the text, or "external" (the reference keeps the
datum packed: length -1).  Run in the build container."""
import ctypes
import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "pg-strom_amd"))

from pgstrom import abi, datagen, kds  # noqa: E402
from pgstrom.kds import Column, ExternalDatum  # noqa: E402

TYPE_CODE = {"bool": 1, "int1": 2, "int2": 3, "int4": 4, "int8": 5, "float4": 7,
             "float8": 8, "text": 11}


def mixed_chunk():
    cols = datagen.heap_table(300, fragment=7)
    inv = np.random.default_rng(11).random(300) < 0.1
    return cols, kds.build_block_kds(cols, invisible=inv, extra_items=True, seed=3,
                                     empty_pages=(1,))


def external_chunk():
    r = np.random.default_rng(12)
    n = 120
    a = r.integers(0, 1000, size=n, dtype=np.int32)
    g = [ExternalDatum(18) if i % 7 == 3 else (ExternalDatum(1) if i % 11 == 5 else
                                               b"x" * int(r.integers(0, 200)))
         for i in range(n)]
    c = r.integers(-10**9, 10**9, size=n, dtype=np.int64)
    d = [bytes("row%03d" % i, "ascii") for i in range(n)]
    gn = r.random(n) < 0.1
    cols = [Column("a", "int4", a), Column("g", "text", g, gn), Column("c", "int8", c),
            Column("d", "text", d)]
    return cols, kds.build_block_kds(cols, seed=4)


def chunk_tuples(buf):
    """(page offset, item offset) of every LP_NORMAL item, in scan order."""
    hd = abi.kern_data_store_head.from_buffer_copy(bytes(buf[:ctypes.sizeof(abi.kern_data_store_head)]))
    out = []
    for b in range(hd.nitems):
        pg = hd.block_offset + kds.BLCKSZ * b
        pd_lower = struct.unpack_from("<H", buf, pg + 12)[0]
        nlp = 0 if pd_lower <= 24 else (pd_lower - 24) // 4
        for k in range(nlp):
            lp = struct.unpack_from("<I", buf, pg + 24 + 4 * k)[0]
            if (lp >> 15) & 3 == kds.LP_NORMAL:
                out.append(pg + (lp & 0x7fff))
    return out


def deform_all(lib, cols, buf):
    n = len(cols)
    resnos = (ctypes.c_int16 * n)(*range(1, n + 1))
    types = (ctypes.c_int32 * n)(*[TYPE_CODE[c.typname] for c in cols])
    vals = (ctypes.c_int64 * n)()
    offs = (ctypes.c_int64 * n)()
    lens = (ctypes.c_int32 * n)()
    nulls = (ctypes.c_uint8 * n)()
    raw = bytes(buf)
    cbuf = ctypes.create_string_buffer(raw, len(raw))
    base = ctypes.addressof(cbuf)
    rows = []
    for off in chunk_tuples(raw):
        rc = lib.ref_heap_deform(ctypes.c_void_p(base), ctypes.c_void_p(base + off), n,
                                 resnos, types, vals, offs, lens, nulls)
        assert rc == 0
        row = []
        for j, c in enumerate(cols):
            if nulls[j]:
                row.append(None)
            elif c.typname == "text":
                if lens[j] < 0:
                    row.append("external")
                else:
                    p = off + offs[j]
                    row.append(raw[p:p + lens[j]].decode("ascii"))
            else:
                row.append(int(vals[j]))
        rows.append(row)
    return rows


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libpgstrom_ref.so"))
    lib.ref_heap_deform.restype = ctypes.c_int
    out = {}
    for name, make in (("mixed", mixed_chunk), ("external", external_chunk)):
        cols, buf = make()
        out[name] = {
            "sha256": hashlib.sha256(bytes(buf)).hexdigest(),
            "columns": [[c.name, c.typname] for c in cols],
            "rows": deform_all(lib, cols, buf),
        }
    with open(os.path.join(HERE, "heap_deform.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()
