"""Heap-page (KDS_FORMAT_BLOCK) and row (KDS_FORMAT_ROW) sources on the CPU:
the oracle's tuple walk against the reference's own deform (golden vectors
from oracle/_ref), against the Arrow path on the same rows, and the MVCC rule
of row chunks against its closed form."""
import hashlib
import json
import os
import struct
import sys

import numpy as np
import pytest

from pgstrom import abi, codegen as cg, datagen, kds, workloads as W
from pgstrom.kds import Column
from helpers import project_all_query, rows_of
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import gen_heap_vectors as G  # noqa: E402

GOLDEN = json.load(open(os.path.join(HERE, "golden", "heap_deform.json")))


def oracle_rows(blob, src, dst_length=16 << 20):
    o = oracle.OracleSession(blob)
    rc, dst, res = o.exec(src, dst_length=dst_length, nthreads=2)
    o.close()
    return rc, dst, res


def as_golden(typname, v):
    """An oracle-decoded value in the fixture's encoding."""
    if v is None:
        return None
    if typname == "text":
        return v.decode("ascii")
    if typname == "float8":
        return struct.unpack("<q", struct.pack("<d", v))[0]
    if typname == "float4":
        return struct.unpack("<I", struct.pack("<f", v))[0]
    if typname == "bool":
        return 1 if v else 0
    return int(v)


def test_golden_mixed_chunk_bytes():
    _, buf = G.mixed_chunk()
    assert hashlib.sha256(bytes(buf)).hexdigest() == GOLDEN["mixed"]["sha256"]


def test_oracle_deform_matches_reference_mixed():
    """Every attribute of every LP_NORMAL tuple equals what the reference's
    ExecLoadVarsOuterHeap loads (NULLs, alignment, short and 4-byte varlena
    headers, float bit patterns)."""
    cols, buf = G.mixed_chunk()
    rc, dst, res = oracle_rows(project_all_query(cols), buf)
    assert rc == 0
    got = [[as_golden(c.typname, v) for c, v in zip(cols, r)] for r in rows_of(dst)]
    assert got == GOLDEN["mixed"]["rows"]
    assert res.nitems_raw == len(GOLDEN["mixed"]["rows"])


def test_oracle_walks_past_external_datums():
    """Attributes after a TOAST pointer load as the reference loads them; the
    pointer itself stays packed (length -1, xpu_text_datum_heap_read) and a
    projection copies it as it is (xpu_text_datum_write)."""
    cols, buf = G.external_chunk()
    assert hashlib.sha256(bytes(buf)).hexdigest() == GOLDEN["external"]["sha256"]
    rc, dst, _ = oracle_rows(project_all_query(cols, names=["a", "c", "d"]), buf)
    assert rc == 0
    got = [[as_golden(t, v) for t, v in zip(("int4", "int8", "text"), r)] for r in rows_of(dst)]
    exp = [[r[0], r[2], r[3]] for r in GOLDEN["external"]["rows"]]
    assert got == exp
    assert any(r[1] == "external" for r in GOLDEN["external"]["rows"])
    rc, dst, _ = oracle_rows(project_all_query(cols), buf)
    assert rc == 0
    raw = bytes(dst)
    g = cols[1]
    for i, v in enumerate(g.values):
        if isinstance(v, kds.ExternalDatum) and not g.nulls[i]:
            assert v.encode() in raw


@pytest.mark.parametrize("case", ["empty", "one", "plain", "holes", "sparse"])
def test_oracle_block_equals_arrow(case):
    """A BLOCK chunk scans exactly like the Arrow chunk of its visible rows:
    the same kds_dst bytes (both keep row order)."""
    n = {"empty": 0, "one": 1}.get(case, 3000)
    cols = datagen.heap_table(n, fragment=1)
    rng = np.random.default_rng(5)
    inv = rng.random(n) < (0.2 if case in ("holes", "sparse") else 0.0)
    blk = kds.build_block_kds(cols, invisible=inv, extra_items=case != "plain",
                              max_per_page=7 if case == "sparse" else None,
                              empty_pages=(0, 2) if case == "sparse" else ())
    arr = kds.build_arrow_kds(datagen.take_rows(cols, ~inv))
    blob = project_all_query(cols, quals=lambda p, v: cg.Func("int4lt", v["a"],
                                                              cg.Const("int4", 700000)))
    rc1, d1, r1 = oracle_rows(blob, blk)
    rc2, d2, r2 = oracle_rows(blob, arr)
    assert rc1 == 0 and rc2 == 0
    assert np.array_equal(d1, d2)
    assert r1.nitems_in == r2.nitems_in


def test_oracle_block_groupby_equals_arrow():
    cols = datagen.preagg_table(20000, fragment=2)
    blk = kds.build_block_kds(cols, extra_items=True)
    arr = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    outs = []
    for src in (blk, arr):
        o = oracle.OracleSession(blob)
        rc, _, _ = o.exec(src, nthreads=1)
        assert rc == 0
        outs.append(o.final())
        o.close()
    actions = [a for a, _, _ in W.preagg_tlist(None, cols[0], None, None, None)]
    assert W.decode_groups(outs[0], actions) == W.decode_groups(outs[1], actions)


def visible(xmin, xmax, xids):
    """__gpuscan_check_row_visibility (cuda_gpuscan.cu:614-642) in closed
    form: xmin frozen or in the transaction, and xmax unset, or neither
    frozen nor in the transaction."""
    if xmin == 0:
        return False
    if xmin != kds.FROZEN_XID and xmin not in xids:
        return False
    if xmax == 0:
        return True
    return xmax != kds.FROZEN_XID and xmax not in xids


def test_visibility_rule_known_answers():
    xids = [700, 701]
    cases = [((2, 0), True), ((0, 0), False), ((700, 0), True), ((999, 0), False),
             ((2, 2), False), ((2, 701), False), ((2, 999), True), ((701, 999), True),
             ((701, 700), False), ((999, 999), False)]
    for (xmin, xmax), exp in cases:
        assert visible(xmin, xmax, xids) == exp, (xmin, xmax)


def test_oracle_row_chunk_visibility():
    """A ROW chunk with embedded xmin/xmax scans exactly like the Arrow chunk
    of its visible rows; empty row-index entries load nothing."""
    n = 2500
    cols = datagen.heap_table(n, fragment=3)
    rng = np.random.default_rng(9)
    xids = [700, 701, 705]
    pool_min = [2, 2, 2, 700, 705, 0, 999]
    pool_max = [0, 0, 0, 2, 701, 999, 1234]
    xact = [(pool_min[rng.integers(len(pool_min))], pool_max[rng.integers(len(pool_max))])
            for _ in range(n)]
    holes = set(int(i) for i in rng.choice(n, 40, replace=False))
    keep = np.array([visible(a, b, xids) and i not in holes for i, (a, b) in enumerate(xact)])
    row = kds.build_row_kds(cols, xact=xact, holes=holes)
    arr = kds.build_arrow_kds(datagen.take_rows(cols, keep))
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    vs = [plan.scan_var(c.name) for c in cols]
    plan.set_projection(vs, [(c.name, c.typname) for c in cols])
    blob = plan.build(xids=xids)
    rc1, d1, _ = oracle_rows(blob, row)
    rc2, d2, _ = oracle_rows(blob, arr)
    assert rc1 == 0 and rc2 == 0
    assert 0 < keep.sum() < n
    assert np.array_equal(d1, d2)


def test_block_builder_page_invariants():
    """Every page: line pointers below pd_upper, items inside the page and
    MAXALIGN'ed, tuple headers consistent with the columns."""
    cols = datagen.heap_table(4000, fragment=4)
    buf = kds.build_block_kds(cols, extra_items=True)
    hd = abi.kern_data_store_head.from_buffer_copy(bytes(buf[:256]))
    raw = bytes(buf)
    seen = 0
    for b in range(hd.nitems):
        pg = hd.block_offset + kds.BLCKSZ * b
        pd_lower, pd_upper = struct.unpack_from("<HH", raw, pg + 12)
        assert 24 <= pd_lower <= pd_upper <= kds.BLCKSZ
        for k in range((pd_lower - 24) // 4):
            lp = struct.unpack_from("<I", raw, pg + 24 + 4 * k)[0]
            off, flags, ln = lp & 0x7fff, (lp >> 15) & 3, lp >> 17
            if flags in (kds.LP_NORMAL, kds.LP_DEAD):
                assert off >= pd_upper and off % 8 == 0 and off + ln <= kds.BLCKSZ
            if flags == kds.LP_NORMAL:
                natts = struct.unpack_from("<H", raw, pg + off + 18)[0] & 0x7ff
                assert natts == len(cols)
                seen += 1
    assert seen == 4000
