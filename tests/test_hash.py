"""hash_any parity: the host mirror (pgstrom.hashing) and the oracle against
the reference's own pg_hash_any output (tests/golden/pg_hash_any.json, written
by tests/golden/gen_hash_vectors.py from oracle/_ref)."""
import json
import os
import struct

import numpy as np

import oracle
from pgstrom import hashing

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VECTORS = json.load(open(os.path.join(ROOT, "tests", "golden", "pg_hash_any.json")))


def test_vectors_present():
    assert len(VECTORS) >= 90


def test_host_hash_any():
    for v in VECTORS:
        assert hashing.hash_any(bytes.fromhex(v["hex"])) == v["hash"], v


def test_oracle_hash_any():
    for v in VECTORS:
        assert oracle.hash_any(bytes.fromhex(v["hex"])) == v["hash"], v


def test_vectorized_fixed_width():
    fixed = [v for v in VECTORS if len(bytes.fromhex(v["hex"])) in (1, 2, 4, 8)]
    for v in fixed:
        b = bytes.fromhex(v["hex"])
        arr = np.frombuffer(b, dtype={1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[len(b)])
        assert int(hashing.hash_fixed_np(arr, len(b))[0]) == v["hash"]


def test_hash_merge_composition():
    """HashValue of one int4 key = ~merge(~0, hash_any(key))."""
    keys = np.array([1, 2, 500000, -7], dtype=np.int32)
    h = hashing.hash_keys_np([(keys, 4)])
    for k, hv in zip(keys, h):
        x = hashing.hash_any(struct.pack("<i", int(k)))
        merged = (((0xffffffff >> 3) | (0xffffffff << 29)) & 0xffffffff) ^ x
        assert int(hv) == merged ^ 0xffffffff
