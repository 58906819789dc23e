"""PostgreSQL's hash_any (Bob Jenkins' lookup3 with PG's seed) on the host, as
the reference backend uses it to hash inner-relation join keys while building
KDS_FORMAT_HASH buffers (gpu_join.c:1548-1584 with the type hashes of
codegen.c:543-602; the device side is xpu_common.cu:2920-3041).  A vectorized
numpy form covers fixed-width keys of 1..8 bytes."""
import struct

import numpy as np

M32 = 0xffffffff


def _rot(x, k):
    return ((x << k) | (x >> (32 - k))) & M32


def _final(a, b, c):
    c ^= b; c = (c - _rot(b, 14)) & M32
    a ^= c; a = (a - _rot(c, 11)) & M32
    b ^= a; b = (b - _rot(a, 25)) & M32
    c ^= b; c = (c - _rot(b, 16)) & M32
    a ^= c; a = (a - _rot(c, 4)) & M32
    b ^= a; b = (b - _rot(a, 14)) & M32
    c ^= b; c = (c - _rot(b, 24)) & M32
    return a, b, c


def _mix(a, b, c):
    a = (a - c) & M32; a ^= _rot(c, 4);  c = (c + b) & M32
    b = (b - a) & M32; b ^= _rot(a, 6);  a = (a + c) & M32
    c = (c - b) & M32; c ^= _rot(b, 8);  b = (b + a) & M32
    a = (a - c) & M32; a ^= _rot(c, 16); c = (c + b) & M32
    b = (b - a) & M32; b ^= _rot(a, 19); a = (a + c) & M32
    c = (c - b) & M32; c ^= _rot(b, 4);  b = (b + a) & M32
    return a, b, c


def hash_any(data):
    k = bytes(data)
    n = len(k)
    a = b = c = (0x9e3779b9 + n + 3923095) & M32
    pos = 0
    while n - pos >= 12:
        x, y, z = struct.unpack_from("<III", k, pos)
        a = (a + x) & M32
        b = (b + y) & M32
        c = (c + z) & M32
        a, b, c = _mix(a, b, c)
        pos += 12
    rest = k[pos:]
    r = len(rest)
    if r >= 9:
        c = (c + (int.from_bytes(rest[8:r], "little") << 8)) & M32
    if r >= 5:
        b = (b + int.from_bytes(rest[4:min(r, 8)], "little")) & M32
    if r >= 1:
        a = (a + int.from_bytes(rest[0:min(r, 4)], "little")) & M32
    a, b, c = _final(a, b, c)
    return c


def _vrot(x, k):
    return ((x << np.uint32(k)) | (x >> np.uint32(32 - k))).astype(np.uint32)


def hash_fixed_np(values, width):
    """hash_any of the little-endian `width`-byte images of integer/float
    values (width <= 8), vectorized."""
    v = np.ascontiguousarray(values)
    raw = v.view(np.uint8).reshape(len(v), v.dtype.itemsize)[:, :width]
    pad = np.zeros((len(v), 8), dtype=np.uint8)
    pad[:, :width] = raw
    words = pad.view("<u4").reshape(len(v), 2)
    with np.errstate(over="ignore"):
        seed = np.uint32((0x9e3779b9 + width + 3923095) & M32)
        a = (words[:, 0] + seed).astype(np.uint32)
        b = (words[:, 1] + seed).astype(np.uint32)
        c = np.full(len(v), seed, dtype=np.uint32)
        c ^= b; c -= _vrot(b, 14)
        a ^= c; a -= _vrot(c, 11)
        b ^= a; b -= _vrot(a, 25)
        c ^= b; c -= _vrot(b, 16)
        a ^= c; a -= _vrot(c, 4)
        b ^= a; b -= _vrot(a, 14)
        c ^= b; c -= _vrot(b, 24)
    return c


def hash_merge_np(prev, nxt):
    """pg_hash_merge (xpu_common.h:3472-3476), vectorized."""
    prev = prev.astype(np.uint32)
    return (((prev >> np.uint32(3)) | (prev << np.uint32(29))) ^ nxt).astype(np.uint32)


def hash_keys_np(cols_widths, nulls=None):
    """HashValue of a list of fixed-width key columns: h = ~0, merged with the
    hash of each non-NULL key, then complemented (xpu_common.cu:1363-1398);
    the host side merges 0 for a NULL key (gpu_join.c:1574-1580)."""
    n = len(cols_widths[0][0])
    h = np.full(n, M32, dtype=np.uint32)
    for k, (vals, width) in enumerate(cols_widths):
        hk = hash_fixed_np(vals, width)
        if nulls is not None and nulls[k] is not None:
            hk = np.where(nulls[k], np.uint32(0), hk)
        h = hash_merge_np(h, hk)
    return (h ^ np.uint32(M32)).astype(np.uint32)
