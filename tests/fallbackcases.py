"""Inputs and closed-form expectations of the CPU-fallback tests (shared by
test_fallback.py on the oracle and test_gpu_fallback.py on the GPU).

The chunk is heap pages whose text column g holds TOAST pointers (on-disk and
indirect tags), inline compressed datums, NULLs and plain texts.  A query
that needs g's bytes sends the rows whose g is compressed or external to the
CPU (HandleErrorIfCpuFallback, xpu_common.cu:1898-1955): each becomes a
kern_fallbackitem of depth 0, l_state 0, not matched, holding the scan
slot's tuple (a, g, c, d) with g as it is; every other row runs on the GPU."""
import numpy as np

from pgstrom import codegen as cg, kds
from pgstrom.kds import Column, CompressedDatum, ExternalDatum

COLDEFS = [("a", "int4"), ("g", "text"), ("c", "int8"), ("d", "text")]


def rows(n=600, seed=21):
    r = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k = i % 13
        if k == 3:
            g = ExternalDatum(18)
        elif k == 7:
            g = ExternalDatum(1)
        elif k == 9:
            g = CompressedDatum(3000 + i)
        elif k == 11:
            g = None
        else:
            g = bytes(r.integers(97, 123, size=int(r.integers(0, 40)), dtype=np.uint8))
        out.append((int(r.integers(0, 1000)), g, int(r.integers(-10**9, 10**9)),
                    b"row%05d" % i))
    return out


def chunk(rs):
    cols = []
    for j, (name, t) in enumerate(COLDEFS):
        vals = [row[j] for row in rs]
        nulls = np.array([v is None for v in vals])
        fill = b"" if t == "text" else 0
        cols.append(Column(name, t, [fill if v is None else v for v in vals], nulls))
    return cols, kds.build_block_kds(cols, seed=5)


def is_raw(g):
    return isinstance(g, (ExternalDatum, CompressedDatum))


def plan_of():
    plan = cg.Plan(COLDEFS)
    return plan, {n: plan.scan_var(n) for n, _ in COLDEFS}


def q_textlen(fallback=True, encoding="UTF8"):
    """SELECT a, c, d WHERE length(g) > 5"""
    plan, v = plan_of()
    plan.scan_quals = cg.Func("int4gt", cg.Func("textlen", v["g"]), cg.Const("int4", 5))
    plan.set_projection([v["a"], v["c"], v["d"]], [("a", "int4"), ("c", "int8"), ("d", "text")])
    return plan.build(fallback=fallback, encoding=encoding)


def q_texteq(fallback=True):
    """SELECT a, g, d WHERE g <> 'zz' (the projection copies g as it is; the
    comparison needs its bytes)"""
    plan, v = plan_of()
    plan.scan_quals = cg.Func("text_ne", v["g"], cg.Const("text", b"zz"))
    plan.set_projection([v["a"], v["g"], v["d"]], [("a", "int4"), ("g", "text"), ("d", "text")])
    return plan.build(fallback=fallback)


def q_groupby_g(fallback=True):
    """SELECT g, count(*), sum(c) GROUP BY g"""
    from pgstrom import abi
    plan, v = plan_of()
    plan.set_groupby([(abi.KAGG_ACTION__NROWS_ANY, None, ("n", "int8")),
                      (abi.KAGG_ACTION__PSUM_INT64, v["c"], ("s", "bytea")),
                      (abi.KAGG_ACTION__VREF, v["g"], ("g", "text"))])
    return plan.build(final_nrows=1000, fallback=fallback)


def expected_fallback(rs):
    """(depth, l_state, matched, tuple bytes without the row id) of the rows
    that fall back, in input order; the tuple is the scan slot's minimal
    tuple (kern_form_minimal_tuple, as kds.form_tupitem writes it: pinned by
    the reference's bytes in test_layouts.py)."""
    return [(0, 0, 0, kds.form_tupitem(COLDEFS, list(row), 0)[:-4])
            for row in rs if is_raw(row[1])]


def raw_textlen(g):
    """pgfn_textlen of a compressed or external datum under a single-byte
    encoding (xpu_textlib.cu:676-703): the TOAST pointer's va_extinfo, or the
    word the reference copies into toast_compress_header's tcinfo from
    VARDATA (the one after the compression info)."""
    b = g.encode()
    off = 6 if b[0] == 0x01 else 8
    w = b[off:off + 4] + b"\0" * max(0, off + 4 - len(b))
    return int.from_bytes(w[:4], "little") & 0x3fffffff
