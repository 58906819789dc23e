"""Session open validation without a GPU (xpuhip_session_check)."""
import struct

import numpy as np
import pytest

from pgstrom import abi, codegen as cg, datagen, kds, workloads as W, xpuhip


def test_workload_sessions_compile():
    for blob in (W.scan_query(datagen.scan_table(10)),
                 W.preagg_query(datagen.preagg_table(10)),
                 W.star_query(datagen.star_tables(10, dim_sizes=(10, 10, 10))[0])):
        rc, err = xpuhip.session_check(blob)
        assert rc == 0, err.describe()


def test_corrupted_magic_rejected():
    blob = bytearray(W.scan_query(datagen.scan_table(10)))
    s = abi.kern_session_info_head.from_buffer_copy(bytes(blob[:ctypes_size()]))
    off = s.xpucode_scan_quals
    ln = struct.unpack_from("<I", blob, off)[0]
    blob[off + ln - 1] ^= 0xff
    rc, err = xpuhip.session_check(bytes(blob))
    assert rc == abi.ERRCODE_DEVICE_ERROR
    assert b"magic" in err.message


def ctypes_size():
    import ctypes
    return ctypes.sizeof(abi.kern_session_info_head)


def test_unsupported_function_rejected():
    """An opcode this build does not execute fails at session open, not in a kernel."""
    blob = bytearray(W.scan_query(datagen.scan_table(10)))
    s = abi.kern_session_info_head.from_buffer_copy(bytes(blob[:ctypes_size()]))
    off = s.xpucode_scan_quals
    # first argument of the AND is int4lt: retag it with an unknown opcode and fix its magic
    arg = off + 40
    ln = struct.unpack_from("<I", blob, arg)[0]
    exptype = struct.unpack_from("<i", blob, arg + 4)[0]
    struct.pack_into("<i", blob, arg + 20, 9000)
    magic = (abi.KERN_EXPRESSION_MAGIC ^ (exptype << 6) ^ (9000 << 14)) & 0xffffffff
    struct.pack_into("<I", blob, arg + ln - 4, magic)
    rc, err = xpuhip.session_check(bytes(blob))
    assert rc == abi.ERRCODE_DEVICE_ERROR
    assert b"not supported" in err.message


def test_register_pressure_uses_large_file():
    """Expressions that need more than the 16-register file still compile."""
    cols = datagen.scan_table(10)
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    a = plan.scan_var("a")
    e = a
    for i in range(12):
        e = cg.Func("int4pl", e, cg.Func("int4mul", a, cg.Const("int4", i)))
    plan.scan_quals = cg.Func("int4gt", e, cg.Const("int4", 0))
    plan.set_projection([a], [("a", "int4")])
    rc, err = xpuhip.session_check(plan.build())
    assert rc == 0, err.describe()


def test_explain_short_circuit_guard():
    """The division inside an AND is compiled under the AND's guard register,
    so a row where the first argument is false never raises."""
    cols = [datagen.scan_table(10)[0]]
    plan = cg.Plan([("a", "int4")])
    a = plan.scan_var("a")
    plan.scan_quals = cg.And(cg.Func("int4ne", a, cg.Const("int4", 0)),
                             cg.Func("int4gt", cg.Func("int4div", cg.Const("int4", 10), a),
                                     cg.Const("int4", 2)))
    plan.set_projection([a], [("a", "int4")])
    text = xpuhip.explain(plan.build())
    lines = text.splitlines()
    guard = [l for l in lines if "G_NOTFALSE" in l][0]
    gdst = int(guard.split("d=")[1].split()[0])
    div = [l for l in lines if "DIV_I" in l][0]
    assert int(div.split("g=")[1].split()[0]) == gdst, text


@pytest.mark.parametrize("which", ["c1", "fragment", "nogroup"])
def test_jit_prepare_compiles(which):
    """The kernel specialized for a scan + group-by plan compiles with hiprtc
    for gfx950 (no GPU needed)."""
    cols = datagen.preagg_table(1000)
    if which == "c1":
        blob = W.preagg_query(cols)
    elif which == "fragment":
        a = kds.Column("a", "int4", np.arange(1000, dtype=np.int32))
        blob = W.fragment_query([a] + cols, k=500, m=16)
    else:
        plan = cg.Plan([(c.name, c.typname) for c in cols])
        plan.set_groupby([(abi.KAGG_ACTION__NROWS_ANY, None, ("n", "int8")),
                          (abi.KAGG_ACTION__PSUM_FP, plan.scan_var("w"), ("f", "bytea"))])
        blob = plan.build(final_length=4 << 20)
    xpuhip.jit_prepare(blob)


@pytest.mark.parametrize("which", ["c0", "c2", "c2_left"])
def test_jit_prepare_pipeline_compiles(which):
    """Projection and join plans specialise too (the pipeline kernel)."""
    if which == "c0":
        blob = W.scan_query(datagen.scan_table(10))
    else:
        fact, _ = datagen.star_tables(10, dim_sizes=(10, 10, 10))
        blob = W.star_query(fact, left_outer=(which == "c2_left"))
    xpuhip.jit_prepare(blob)


def test_pinned_hash_session_compiles():
    """A projection into a pinned inner buffer: the projection's HashValue is
    compiled after the projected values, into the register the tuple's hash
    comes from; the blob carries the buffer's KDS_FORMAT_HASH header."""
    _, dims = datagen.star_tables(10, dim_sizes=(100, 10, 10))
    blob = W.dim_scan_pinned(dims[0], 0)
    rc, err = xpuhip.session_check(blob)
    assert rc == 0, err.describe()
    s = abi.kern_session_info_head.from_buffer_copy(bytes(blob[:ctypes_size()]))
    assert s.xpu_task_flags & abi.DEVTASK__PINNED_HASH_RESULTS
    hd = abi.kern_data_store_head.from_buffer_copy(bytes(blob[s.groupby_kds_final:
                                                              s.groupby_kds_final + ctypes_kds()]))
    assert hd.format == abi.KDS_FORMAT_HASH and hd.hash_nslots == 137
    text = xpuhip.explain(blob)
    final = text[text.index("[final"):]
    assert "HASH_FIN" in final and "MOV        d=1" in final


def test_pinned_hash_session_needs_its_hash():
    """DEVTASK__PINNED_HASH_RESULTS without the projection's hash is refused."""
    _, dims = datagen.star_tables(10, dim_sizes=(100, 10, 10))
    blob = bytearray(W.dim_scan_pinned(dims[0], 0))
    s = abi.kern_session_info_head.from_buffer_copy(bytes(blob[:ctypes_size()]))
    struct.pack_into("<I", blob, s.xpucode_projection + 40, 0)      # u.proj.hash = 0
    rc, err = xpuhip.session_check(bytes(blob))
    assert rc == abi.ERRCODE_DEVICE_ERROR
    assert b"projection hash" in err.message


def ctypes_kds():
    import ctypes
    return ctypes.sizeof(abi.kern_data_store_head)


def test_pinned_multirels_chunks():
    """kern_multirels chunks naming pinned buffers: pinned_buffer set,
    buffer_id the inner plan's query_plan_id, no KDS in the buffer."""
    import ctypes
    km = W.star_inner_pinned(ndims=3)
    for d in range(3):
        ch = abi.kern_multirels_chunk.from_buffer_copy(
            bytes(km[abi.KMRELS_CHUNKS_OFFSET + d * ctypes.sizeof(abi.kern_multirels_chunk):
                     abi.KMRELS_CHUNKS_OFFSET + (d + 1) * ctypes.sizeof(abi.kern_multirels_chunk)]))
        assert ch.pinned_buffer and ch.kds_offset == 0
        assert ch.buffer_id == W.PINNED_PLAN_ID + d
