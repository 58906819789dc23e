"""Heap-page (KDS_FORMAT_BLOCK) and row (KDS_FORMAT_ROW) sources on the GPU:
the chunk is deformed into column tiles on the device and run through the
same kernels as an Arrow chunk.  Every result is compared with the oracle
over the same chunk (kds_dst bytes for scans, group states for group-bys)
and, at full size, with the GPU's own Arrow path over the visible rows."""
import numpy as np
import pytest

from pgstrom import abi, codegen as cg, datagen, kds, workloads as W, xpuhip
import oracle
from helpers import (skip_interpreter_join_groupby, assert_groups_equal, project_all_query, run_gpu_groupby,
                     run_gpu_projection, run_oracle_groupby, run_oracle_projection)

pytestmark = pytest.mark.gpu

C1_ACTIONS = [a for a, _, _ in W.preagg_tlist(None, type("K", (), {"typname": "int4"})(),
                                              None, None, None)]


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 5000, 200_000])
def test_block_scan_bytes(gpu, n, jit):
    """C0's scan over heap pages: kds_dst bytes equal the oracle's."""
    cols = datagen.scan_table(n, fragment=11)
    inv = np.random.default_rng(n).random(n) < 0.1
    src = kds.build_block_kds(cols, invisible=inv, extra_items=True)
    blob = W.scan_query(cols)
    got = run_gpu_projection(blob, src, 32 << 20, tuning=dict(jit=jit))
    exp, res = run_oracle_projection(blob, src, 32 << 20)
    assert len(got) == 1
    assert np.array_equal(got[0], exp)


@pytest.mark.parametrize("layout", ["mixed", "sparse"])
def test_block_project_all_types(gpu, jit, layout):
    """Every attribute type and alignment: int2/int4/bool/float8/text (short
    and 4-byte headers)/int8/float4/text with NULLs; invisible rows, dead and
    redirect items, empty pages."""
    n = 20_000
    cols = datagen.heap_table(n, fragment=12)
    inv = np.random.default_rng(3).random(n) < 0.15
    src = kds.build_block_kds(cols, invisible=inv, extra_items=True,
                              max_per_page=5 if layout == "sparse" else None,
                              empty_pages=(0, 3, 4) if layout == "sparse" else ())
    blob = project_all_query(cols, quals=lambda p, v: cg.Func(
        "int4lt", v["a"], cg.Const("int4", 800_000)))
    got = run_gpu_projection(blob, src, 64 << 20, tuning=dict(jit=jit))
    exp, _ = run_oracle_projection(blob, src, 64 << 20)
    assert np.array_equal(got[0], exp)


def test_block_scan_resume(gpu, jit):
    """A kds_dst too small for a heap chunk: resumes number the tile rows the
    same way on every call; the parts concatenate to one big run."""
    cols = datagen.scan_table(150_000, fragment=13)
    src = kds.build_block_kds(cols, extra_items=True)
    blob = W.scan_query(cols)
    parts = run_gpu_projection(blob, src, 1 << 20, tuning=dict(jit=jit))
    assert len(parts) > 1
    exp, _ = run_oracle_projection(blob, src, 64 << 20)
    rows = [r for p in parts for r in kds.parse_row_kds(p)]
    exp_rows = kds.parse_row_kds(exp)
    assert [r[0] for r in rows] == [r[0] for r in exp_rows]


def test_block_groupby(gpu, jit):
    """C1's group-by over heap pages of several chunks."""
    chunks = []
    for f in range(3):
        cols = datagen.preagg_table(60_000, fragment=20 + f)
        chunks.append(kds.build_block_kds(cols, extra_items=True, seed=f))
    blob = W.preagg_query(cols)
    got = run_gpu_groupby(blob, chunks, tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, chunks)
    assert_groups_equal(got, exp, C1_ACTIONS)


def test_block_star_join(gpu, jit):
    """C2's star join with the fact table on heap pages."""
    skip_interpreter_join_groupby(jit)
    fact, dims = datagen.star_tables(100_000, dim_sizes=(20_000, 2_000, 2_556))
    src = kds.build_block_kds(fact)
    km = W.star_inner(dims)
    blob = W.star_query(fact)
    got = run_gpu_groupby(blob, [src], kmrels=km, tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, [src], kmrels=km)
    assert_groups_equal(got, exp, [abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF,
                                   abi.KAGG_ACTION__VREF])


def test_row_chunk_visibility(gpu, jit):
    """A KDS_FORMAT_ROW chunk with embedded xmin/xmax and empty row-index
    entries: the visible rows scan exactly as the oracle scans them."""
    n = 30_000
    cols = datagen.heap_table(n, fragment=14)
    rng = np.random.default_rng(21)
    xids = [700, 701, 705]
    pool_min = [2, 2, 2, 700, 705, 0, 999]
    pool_max = [0, 0, 0, 2, 701, 999, 1234]
    xact = [(pool_min[rng.integers(7)], pool_max[rng.integers(7)]) for _ in range(n)]
    holes = set(int(i) for i in rng.choice(n, 300, replace=False))
    src = kds.build_row_kds(cols, xact=xact, holes=holes)
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    vs = [plan.scan_var(c.name) for c in cols]
    plan.set_projection(vs, [(c.name, c.typname) for c in cols])
    blob = plan.build(xids=xids)
    got = run_gpu_projection(blob, src, 64 << 20, tuning=dict(jit=jit))
    exp, res = run_oracle_projection(blob, src, 64 << 20)
    assert res.nitems_in > 0
    assert np.array_equal(got[0], exp)


def test_block_device_entry(gpu):
    """xpuhip_task_exec_device_tuples on a chunk already in device memory
    gives the host entry point's rows; the task stats report the heap
    passes."""
    cols = datagen.scan_table(300_000, fragment=15)
    src = kds.build_block_kds(cols)
    blob = W.scan_query(cols)
    host = run_gpu_projection(blob, src, 64 << 20)[0]
    d_src = xpuhip.DeviceBuffer(len(src), host=src)
    d_dst = xpuhip.DeviceBuffer(64 << 20)
    s = xpuhip.Session(blob)
    rc, res, _ = s.exec_device_tuples(d_src, src[:4096], d_dst, 64 << 20)
    st = xpuhip.last_task_stats()
    s.close()
    assert rc == 0
    assert st["heap_rows"] == 300_000 and st["heap_ms"] > 0
    dev = d_dst.to_host()
    assert kds.parse_row_kds(dev) == kds.parse_row_kds(host)


def test_block_external_datum(gpu, jit):
    """A TOAST pointer in a text column: the columns after it load as the
    reference loads them, and a projection of the column copies the pointer
    as it is (kds_dst bytes equal the oracle's); an operator that needs its
    bytes fails the task with the reference's fallback error when the
    session has no CPU fallback."""
    import os
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import gen_heap_vectors as G
    cols, src = G.external_chunk()
    for names in (["a", "c", "d"], None):
        blob = project_all_query(cols, names=names)
        got = run_gpu_projection(blob, src, 8 << 20, tuning=dict(jit=jit))
        exp, _ = run_oracle_projection(blob, src, 8 << 20)
        assert len(got) == 1 and np.array_equal(got[0], exp)
    blob = project_all_query(cols, quals=lambda plan, v: cg.Func("text_ne", v["g"],
                                                                 cg.Const("text", b"q")))
    s = xpuhip.Session(blob)
    s.set_tuning(jit=jit)
    with pytest.raises(xpuhip.XpuError) as ei:
        s.exec(src, dst_length=8 << 20)
    s.close()
    assert ei.value.code == abi.ERRCODE_SUSPEND_FALLBACK
    assert "compressed or external" in str(ei.value)


def test_block_corrupt_item_is_an_error(gpu):
    """A line pointer reaching past its page fails the task with an error
    (the passes bound every access by the page and the item)."""
    cols = datagen.scan_table(2000, fragment=16)
    src = kds.build_block_kds(cols)
    hd = abi.kern_data_store_head.from_buffer(src)
    lp_at = hd.block_offset + kds.BLCKSZ * 1 + 24 + 4 * 3
    lp = int(src[lp_at:lp_at + 4].view(np.uint32)[0])
    src[lp_at:lp_at + 4] = np.array([(lp & 0x1ffff) | (8000 << 17)], dtype=np.uint32).view(np.uint8)
    s = xpuhip.Session(W.scan_query(cols))
    with pytest.raises(xpuhip.XpuError) as ei:
        s.exec(src, dst_length=8 << 20)
    s.close()
    assert "out of its chunk" in str(ei.value)


def test_c0_heap_full_size_equals_arrow(gpu):
    """C0 at 1M rows on heap pages: the rows equal those of the Arrow chunk
    of the same table on the GPU (a size-independent property)."""
    cols = datagen.scan_table(1_000_000)
    blk = kds.build_block_kds(cols)
    arr = kds.build_arrow_kds(cols)
    blob = W.scan_query(cols)
    a = run_gpu_projection(blob, blk, 64 << 20)
    b = run_gpu_projection(blob, arr, 64 << 20)
    assert len(a) == 1 and len(b) == 1
    assert np.array_equal(a[0], b[0])
