"""Native build of an inner relation's KDS_FORMAT_HASH buffer
(xpuhip_inner_hash_setup: get_tuple_hashvalue, gpu_join.c:1548-1584, and
__innerPreloadSetupHashBuffer, gpu_join.c:2128-2164) against the
reference-pinned layout and hash functions.  Host code only: no GPU."""
import ctypes as C
import hashlib
import json
import os
import struct
import sys

import numpy as np
import pytest

from pgstrom import abi, hashing, kds, xpuhip

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import layoutcases as LC  # noqa: E402

GOLDEN = json.load(open(os.path.join(HERE, "golden", "layouts.json")))
M32 = 0xffffffff


def _key_hash(typname, v):
    """The device type's hash of one key (codegen.c:544-660), from the
    hash_any pinned by the reference's vectors; 0 for NULL."""
    if v is None:
        return 0
    if typname == "bool":
        return hashing.hash_any(bytes([1 if v else 0]))
    if typname in ("text", "bytea"):
        return hashing.hash_any(v)
    if typname == "bpchar":
        return hashing.hash_any(v.rstrip(b" "))
    fmt = {"int2": "<h", "int4": "<i", "int8": "<q", "float4": "<f", "float8": "<d",
           "date": "<i"}[typname]
    return hashing.hash_any(struct.pack(fmt, v))


def _row_hash(coldefs, row, key_attnums):
    h = M32
    for a in key_attnums:
        x = _key_hash(coldefs[a - 1][1], row[a - 1])
        h = (((h >> 3) | (h << 29)) & M32) ^ x
    return h ^ M32


def test_matches_reference_layout():
    """The golden case's rows through the native builder give the bytes the
    reference's KDS_HASH_FIRST/NEXT_ITEM walk was recorded on."""
    coldefs, rows, hashes = LC.hash_inner()
    mt = [kds.form_minimal_tuple(coldefs, r) for r in rows]
    buf = xpuhip.inner_hash_setup(coldefs, mt, [1])
    assert hashlib.sha256(bytes(buf)).hexdigest() == GOLDEN["hash"]["sha256"]


def _random_inner(seed, n):
    r = np.random.default_rng(seed)
    coldefs = [("a", "int4"), ("t", "text"), ("b", "int8"), ("f", "float8"),
               ("c", "bpchar"), ("s", "int2"), ("ok", "bool"), ("d", "date")]
    rows = []
    for i in range(n):
        def maybe(v):
            return None if r.random() < 0.1 else v
        rows.append((maybe(int(r.integers(-50, 50))),
                     maybe(bytes(r.integers(97, 100, size=int(r.integers(0, 300)), dtype=np.uint8))),
                     maybe(int(r.integers(-2**62, 2**62))),
                     maybe(float(r.choice([0.0, -0.0, 1.5, float(r.normal())]))),
                     maybe(b"ab" + b" " * int(r.integers(0, 4))),
                     maybe(int(r.integers(-300, 300))),
                     maybe(bool(r.integers(0, 2))),
                     maybe(int(r.integers(-10**5, 10**5)))))
    return coldefs, rows


@pytest.mark.parametrize("keys", [[1], [2], [3, 4], [5], [6, 7, 8], [8, 2, 1]])
def test_keys_hash_and_layout(keys):
    """Every supported key type, NULL keys, several keys: the buffer equals
    the layout of the serial loop over the hashes the type functions give."""
    coldefs, rows = _random_inner(3, 700)
    hashes = [_row_hash(coldefs, row, keys) for row in rows]
    exp = kds.build_hash_kds(coldefs, rows, hashes)
    got = xpuhip.inner_hash_setup(coldefs, [kds.form_minimal_tuple(coldefs, r) for r in rows], keys)
    assert np.array_equal(got, exp)


def test_threads_do_not_change_the_bytes():
    coldefs = [("k", "int4"), ("v", "int8"), ("t", "text")]
    r = np.random.default_rng(9)
    keys = r.integers(0, 3000, size=40000)
    rows = [(int(k), int(i), b"x%d" % i) for i, k in enumerate(keys)]
    mt = [kds.form_minimal_tuple(coldefs, row) for row in rows]
    one = xpuhip.inner_hash_setup(coldefs, mt, [1], nthreads=1)
    many = xpuhip.inner_hash_setup(coldefs, mt, [1], nthreads=8)
    assert np.array_equal(one, many)
    hashes = hashing.hash_keys_np([(keys.astype(np.int32), 4)])
    assert np.array_equal(one, kds.build_hash_kds(coldefs, rows, [int(h) for h in hashes]))


def test_appending_batches():
    """Two calls on one buffer (base nitems/usage) equal one call."""
    coldefs = [("k", "int4"), ("t", "text")]
    rows = [(i % 11, b"r%d" % i) for i in range(300)]
    mt = [kds.form_minimal_tuple(coldefs, row) for row in rows]
    whole = xpuhip.inner_hash_setup(coldefs, mt, [1], nslots=20000)
    buf = np.zeros_like(whole)
    head = kds.kds_header(abi.KDS_FORMAT_HASH, coldefs, hash_nslots=20000)
    buf[:len(head)] = np.frombuffer(bytes(head), dtype=np.uint8)
    abi.kern_data_store_head.from_buffer(buf).length = len(buf)
    keys = np.array([1], dtype=np.int16)
    for part in (mt[:120], mt[120:]):
        bufs = [np.frombuffer(t, dtype=np.uint8) for t in part]
        ptrs = (C.c_void_p * len(part))(*[b.ctypes.data for b in bufs])
        err = abi.kern_errorbuf()
        assert xpuhip.lib().xpuhip_inner_hash_setup(buf.ctypes.data, ptrs, len(part),
                                                    keys.ctypes.data, 1, 2, C.byref(err)) == 0
    assert np.array_equal(buf, whole)


def test_errors():
    coldefs = [("k", "int4"), ("t", "text"), ("n", "numeric")]
    mt = [kds.form_minimal_tuple(coldefs, (1, b"x", None))]
    with pytest.raises(xpuhip.XpuError, match="out of range"):
        xpuhip.inner_hash_setup(coldefs, mt, [4])
    # a compressed text key: the backend detoasts before hashing
    ext = kds.form_minimal_tuple(coldefs, (1, kds.CompressedDatum(5000), None))
    with pytest.raises(xpuhip.XpuError, match="compressed or external"):
        xpuhip.inner_hash_setup(coldefs, [ext], [2])
    # too small a buffer
    buf = xpuhip.inner_hash_setup(coldefs, mt, [1])
    small = buf.copy()
    hd = abi.kern_data_store_head.from_buffer(small)
    hd.nitems = 0
    hd.usage = 0
    hd.length = len(small) - 8
    bufs = [np.frombuffer(t, dtype=np.uint8) for t in mt]
    ptrs = (C.c_void_p * 1)(bufs[0].ctypes.data)
    keys = np.array([1], dtype=np.int16)
    err = abi.kern_errorbuf()
    assert xpuhip.lib().xpuhip_inner_hash_setup(small.ctypes.data, ptrs, 1, keys.ctypes.data, 1,
                                                1, C.byref(err)) == -1
    assert b"too small" in bytes(err.message)


def test_numeric_key_hash_is_the_device_types():
    """numeric join keys hash by value (devtype_numeric_hash: the parsed
    scale and 128-bit value), as the reference's HashValue over the same
    datum does: every item's hash equals the oracle's HashValue, and
    1.5 / 1.50 / a short-header 1.5 hash alike."""
    import oracle
    from pgstrom import codegen as cg, opcodes
    from pgstrom.kds import numeric_image
    imgs = [numeric_image(v) for v in ("1.5", "-2", "0", "1e20", "0.001", "NaN", "Infinity")] + \
        [numeric_image("1.50"), numeric_image("1.5", short_varlena=True),
         numeric_image("1.5", long_form=True)]
    coldefs = [("n", "numeric"), ("i", "int4")]
    mt = [kds.form_minimal_tuple(coldefs, (im, j)) for j, im in enumerate(imgs)]
    buf = xpuhip.inner_hash_setup(coldefs, mt, [1])
    b = bytearray()
    cg.emit_expr(b, cg.HashValue(cg.Var(0, "numeric")))
    raw = memoryview(bytes(buf))
    hd, cms = kds.kds_colmeta(raw)
    ridx = np.frombuffer(raw, dtype=np.uint64, count=hd.nitems,
                         offset=abi.kds_head_length(hd.nr_colmeta) + 8 * hd.hash_nslots)
    for j, im in enumerate(imgs):
        _, _, h = kds.decode_tupitem(cms, raw, hd.length - int(ridx[j]))
        exp = oracle.eval_expr(bytes(b), [opcodes.TYPES["numeric"]], [im], opcodes.TYPES["int4"])
        assert h == struct.unpack("<I", exp)[0], (j, im.hex())
    hs = [kds.decode_tupitem(cms, raw, hd.length - int(ridx[j]))[2] for j in (0, 7, 8, 9)]
    assert len(set(hs)) == 1
