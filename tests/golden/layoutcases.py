"""Inputs of tests/golden/layouts.json (shared by gen_layout_vectors.py and
the tests that replay it)."""
import numpy as np

from pgstrom import hashing
from pgstrom.kds import CompressedDatum, ExternalDatum

TUPLE_COLDEFS = [("a", "int2"), ("b", "int4"), ("c", "text"), ("d", "float8"), ("e", "bool"),
                 ("f", "int8"), ("g", "text"), ("h", "float4")]


def tuple_rows():
    """Rows of TUPLE_COLDEFS: no NULL, all NULL, single NULLs at each
    attribute, and seeded mixes; texts of 0..300 bytes."""
    r = np.random.default_rng(424242)
    base = [-7, 123456, b"abc", -2.5, True, -(2 ** 40), b"", 0.75]
    rows = [list(base), [None] * 8]
    for j in range(8):
        row = list(base)
        row[j] = None
        rows.append(row)
    for i in range(14):
        row = [int(r.integers(-32768, 32767)), int(r.integers(-2**31, 2**31 - 1)),
               bytes(r.integers(65, 91, size=int(r.integers(0, 300)), dtype=np.uint8)),
               float(r.normal()), bool(r.integers(0, 2)), int(r.integers(-2**63, 2**63 - 1)),
               bytes(r.integers(97, 123, size=int(r.integers(0, 130)), dtype=np.uint8)),
               float(np.float32(r.normal()))]
        for j in range(8):
            if r.random() < 0.3:
                row[j] = None
        rows.append(row)
    return rows


def tuple_rows_external():
    """Rows of TUPLE_COLDEFS whose texts include TOAST pointers (on-disk and
    indirect tags) and inline compressed datums, which a tuple carries as
    they are (xpu_text_datum_write with length -1)."""
    base = [-7, 123456, b"abc", -2.5, True, -(2 ** 40), b"", 0.75]
    kinds = [ExternalDatum(18), ExternalDatum(1), CompressedDatum(5000),
             CompressedDatum(70000, b"\x01" * 37)]
    rows = []
    for i, x in enumerate(kinds):
        for col in (2, 6):
            row = list(base)
            row[col] = x
            if i % 2:
                row[0] = None
            rows.append(row)
    row = list(base)
    row[2], row[6] = kinds[0], kinds[2]
    rows.append(row)
    return rows


def hash_inner():
    """(coldefs, rows, hashes) of a KDS_FORMAT_HASH inner relation keyed by
    pk with duplicate keys (pk = i mod 37 for 90 rows)."""
    coldefs = [("pk", "int4"), ("category", "int2"), ("txt", "text")]
    rows, hashes = [], []
    for i in range(90):
        pk = i % 37
        rows.append((pk, i % 25, b"t%03d" % i))
        hashes.append(int(hashing.hash_keys_np([(np.array([pk], dtype=np.int32), 4)])[0]))
    return coldefs, rows, hashes
