"""GpuScan parity on the GPU: kds_dst bytes of the HIP path equal the
oracle's, row for row (both keep input order)."""
import numpy as np
import pytest

from pgstrom import abi, codegen as cg, datagen, kds, workloads as W
from helpers import run_gpu_projection, run_oracle_projection

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 255, 256, 257, 10000, 1_000_000])
def test_c0_scan_bytes(gpu, n, jit):
    cols = datagen.scan_table(n)
    src = kds.build_arrow_kds(cols)
    blob = W.scan_query(cols)
    dst_len = 64 << 20
    got = run_gpu_projection(blob, src, dst_len, tuning=dict(jit=jit))
    exp, res = run_oracle_projection(blob, src, dst_len)
    assert len(got) == 1
    assert np.array_equal(got[0], exp)


def test_c0_scan_param(gpu):
    cols = datagen.scan_table(50000, fragment=3)
    src = kds.build_arrow_kds(cols)
    blob = W.scan_query(cols, k=123456, use_param=True)
    got = run_gpu_projection(blob, src, 16 << 20)
    exp, _ = run_oracle_projection(blob, src, 16 << 20)
    assert np.array_equal(got[0], exp)


def test_c0_scan_resume(gpu, jit):
    """A kds_dst too small for the chunk: 'd' returns, the caller ships the
    partial buffer and resumes; the concatenation equals one big run."""
    cols = datagen.scan_table(200_000)
    src = kds.build_arrow_kds(cols)
    blob = W.scan_query(cols)
    small = 1 << 20
    parts = run_gpu_projection(blob, src, small, tuning=dict(jit=jit))
    assert len(parts) > 1
    rows = []
    for p in parts:
        rows += [r[0] for r in kds.parse_row_kds(p)]
    exp, _ = run_oracle_projection(blob, src, 64 << 20)
    assert rows == [r[0] for r in kds.parse_row_kds(exp)]


def _expr_plan(cols, quals, targets, coldefs):
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    v = {c.name: plan.scan_var(c.name) for c in cols}
    plan.scan_quals = quals(v, plan)
    plan.set_projection(targets(v), coldefs)
    return plan.build()


EXPRS = [
    # three-valued OR / NOT / IS NULL over NULL-bearing columns
    (lambda v, p: cg.Or(cg.Func("int4gt", v["a"], cg.Const("int4", 900000)),
                        cg.Not(cg.Func("float8lt", v["b"], cg.Const("float8", 1.0)))),
     lambda v: [v["a"], v["b"]], [("a", "int4"), ("b", "float8")]),
    (lambda v, p: cg.NullTest(v["d"], isnull=True),
     lambda v: [v["c"]], [("c", "int8")]),
    # arithmetic and casts in quals and targets
    (lambda v, p: cg.Func("int8lt", cg.Func("int48pl", v["a"], v["c"]), cg.Const("int8", 0)),
     lambda v: [cg.Func("int4_to_float8", v["a"]), cg.Func("float8mul", v["b"], v["b"])],
     [("af", "float8"), ("bb", "float8")]),
    # text comparison against a constant
    (lambda v, p: cg.Func("text_lt", v["d"], cg.Const("text", b"M")),
     lambda v: [v["d"], v["a"]], [("d", "text"), ("a", "int4")]),
    # short-circuit guard: a division that would fail is never reached
    (lambda v, p: cg.And(cg.Func("int4ne", v["a"], cg.Const("int4", 0)),
                         cg.Func("int4gt", cg.Func("int4div", cg.Const("int4", 1000000), v["a"]),
                                 cg.Const("int4", 2))),
     lambda v: [v["a"]], [("a", "int4")]),
    # coalesce and boolean tests
    (lambda v, p: cg.BoolTest("IsNotFalse", cg.Func("float8ge", v["b"], cg.Const("float8", 0.5))),
     lambda v: [cg.Coalesce("int8", v["c"], cg.Const("int8", -1))], [("c", "int8")]),
]


@pytest.mark.parametrize("i", range(len(EXPRS)))
def test_scan_expressions(gpu, i, jit):
    quals, targets, coldefs = EXPRS[i]
    cols = datagen.scan_table(30000, null_frac=0.2)
    cols[0].values[:50] = 0          # a = 0 rows for the division guard
    src = kds.build_arrow_kds(cols)
    blob = _expr_plan(cols, quals, targets, coldefs)
    got = run_gpu_projection(blob, src, 16 << 20, tuning=dict(jit=jit))
    exp, _ = run_oracle_projection(blob, src, 16 << 20)
    assert np.array_equal(got[0], exp)


def test_scan_error_raised(gpu, jit):
    """int4 overflow in a target raises the reference's error text."""
    from pgstrom import xpuhip
    cols = datagen.scan_table(1000, null_frac=0.0)
    cols[0].values[7] = 2**31 - 1
    src = kds.build_arrow_kds(cols)
    blob = _expr_plan(cols, lambda v, p: cg.Const("bool", True),
                      lambda v: [cg.Func("int4pl", v["a"], cg.Const("int4", 1))], [("a", "int4")])
    with pytest.raises(xpuhip.XpuError) as ei:
        run_gpu_projection(blob, src, 4 << 20, tuning=dict(jit=jit))
    assert ei.value.code == abi.ERRCODE_DEVICE_ERROR
    assert b"int4pl: value out of range" in ei.value.errbuf.message


@pytest.mark.parametrize("layout", ["utf8", "large", "fixed"])
def test_arrow_text_layouts(gpu, jit, layout):
    """Text columns as Utf8, LargeUtf8 (64-bit offsets) and FixedSizeBinary:
    kds_dst bytes equal the oracle's (xpu_textlib.cu:19-54)."""
    n = 50_000
    r = datagen.rng(7, 2)
    a = kds.Column("a", "int4", r.integers(0, 100, size=n, dtype=np.int32))
    nulls = r.random(n) < 0.1
    t = kds.Column("t", "text", datagen.random_text(r, n, 16, 16 if layout == "fixed" else 60,
                                                    nulls=nulls), nulls)
    cols = [a, t]
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    va, vt = plan.scan_var("a"), plan.scan_var("t")
    plan.scan_quals = cg.Func("int4lt", va, cg.Const("int4", 60))
    plan.set_projection([vt, va], [("t", "text"), ("a", "int4")])
    blob = plan.build()
    src = kds.build_arrow_kds(cols, text_layout=layout)
    got = run_gpu_projection(blob, src, 16 << 20, tuning=dict(jit=jit))
    exp, _ = run_oracle_projection(blob, src, 16 << 20)
    assert np.array_equal(got[0], exp)


def test_unmappable_arrow_type_is_an_error(gpu, jit):
    from pgstrom import xpuhip
    cols = [kds.Column("t", "text", [b"x", b"yy"] * 100)]
    src = kds.build_arrow_kds(cols)
    cm = abi.kern_colmeta.from_buffer(src, abi.KDS_COLMETA_OFFSET)
    cm.attopts.tag = abi.ArrowType__Int
    plan = cg.Plan([("t", "text")])
    vt = plan.scan_var("t")
    plan.set_projection([vt], [("t", "text")])
    s = xpuhip.Session(plan.build())
    s.set_tuning(jit=jit)
    with pytest.raises(xpuhip.XpuError) as ei:
        s.exec(src, dst_length=1 << 20)
    s.close()
    assert "not a mappable Arrow data type" in str(ei.value)
