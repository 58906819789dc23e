"""Arrow IPC files for the chunk producer's tests (written with pyarrow),
and the same rows as kds.Column lists for the test-side Arrow builder."""
import decimal

import numpy as np

from pgstrom.kds import Column, DecimalValues

COLDEFS = [("a", "int4"), ("b", "float8"), ("c", "int8"), ("t", "text"), ("lt", "text"),
           ("f", "bool"), ("d", "numeric"), ("fb", "bytea"), ("s", "int2")]


def _rows(n, seed):
    r = np.random.default_rng(seed)
    out = {
        "a": [None if i % 11 == 3 else int(x) for i, x in enumerate(r.integers(-10**6, 10**6, n))],
        "b": [float(x) for x in r.normal(size=n)],
        "c": [int(x) for x in r.integers(-2**62, 2**62, n)],
        "t": [None if i % 7 == 1 else "t%d" % int(x) for i, x in enumerate(r.integers(0, 10**5, n))],
        "lt": ["large-%d" % i for i in range(n)],
        "f": [bool(x) for x in r.integers(0, 2, n)],
        "d": [None if i % 13 == 5 else decimal.Decimal(int(x)).scaleb(-3)
              for i, x in enumerate(r.integers(-10**12, 10**12, n))],
        "fb": [bytes([i % 256, 1, 2, 3]) for i in range(n)],
        "s": [int(x) for x in r.integers(-30000, 30000, n)],
        "unused": [float(i) for i in range(n)],
    }
    return out


def write_file(path, sizes=(1000, 0, 4097), seed=3, compression=None):
    """An Arrow IPC file with one record batch per size; returns the
    per-batch rows (dict of column lists)."""
    import pyarrow as pa
    import pyarrow.ipc as ipc
    schema = pa.schema([("a", pa.int32()), ("b", pa.float64()), ("c", pa.int64()),
                        ("t", pa.utf8()), ("lt", pa.large_utf8()), ("f", pa.bool_()),
                        ("d", pa.decimal128(20, 3)), ("fb", pa.binary(4)), ("unused", pa.float64()),
                        ("s", pa.int16())])
    batches = []
    opts = ipc.IpcWriteOptions(compression=compression) if compression else None
    with ipc.new_file(path, schema, options=opts) as w:
        for k, n in enumerate(sizes):
            rows = _rows(n, seed + k)
            w.write_batch(pa.record_batch([pa.array(rows[f.name], type=f.type) for f in schema],
                                          schema=schema))
            batches.append(rows)
    return batches


def columns(rows):
    """The rows as kds.Column objects in COLDEFS order."""
    n = len(rows["a"])
    cols = []
    for name, t in COLDEFS:
        vals = rows[name]
        nulls = np.array([v is None for v in vals], dtype=bool)
        if t == "numeric":
            ints = [0 if v is None else int(v.scaleb(3)) for v in vals]
            cols.append(Column(name, t, DecimalValues(ints, 3, 20), nulls))
        elif t == "text":
            cols.append(Column(name, t, [b"" if v is None else v.encode() for v in vals], nulls))
        elif t == "bytea":
            cols.append(Column(name, t, [b"" if v is None else v for v in vals], nulls))
        else:
            cols.append(Column(name, t, [0 if v is None else v for v in vals], nulls))
    assert all(len(c) == n for c in cols)
    return cols
