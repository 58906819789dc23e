"""The expression constructs on the GPU against the oracle over a 40K-row
chunk (both kernels): scalar op ANY / ALL with NULL elements and NULL
scalars, hashed IN lists with and without NULL elements, CASE maps with a
NULL result and no ELSE, searched CASE, DISTINCT FROM, LEAST / GREATEST,
LIKE and substr, in the projection and in the scan quals; kds_dst bytes
must be equal."""
import numpy as np
import pytest

from pgstrom import codegen as cg, datagen, kds
from pgstrom.kds import Column
from helpers import run_gpu_projection, run_oracle_projection

pytestmark = pytest.mark.gpu


def _table(n=40_000):
    r = datagen.rng(11, 3)
    x = r.integers(-4, 12, size=n, dtype=np.int32)
    y = r.integers(-3, 4, size=n, dtype=np.int32)
    t = datagen.random_text(r, n, 0, 6)
    return [Column("x", "int4", x, r.random(n) < 0.1), Column("y", "int4", y, r.random(n) < 0.1),
            Column("t", "text", t, r.random(n) < 0.1)]


def _targets(v):
    x, y, t = v["x"], v["y"], v["t"]
    C = lambda typ, val: cg.Const(typ, val)
    return [
        cg.ScalarArrayOp(False, "int4lt", x, 3, "int4", [1, 3, 7]),
        cg.ScalarArrayOp(False, "int4ne", x, 3, "int4", [7, None, 1]),
        cg.ScalarArrayOp(True, "int4eq", x, 3, "int4", [5, None, -2]),
        cg.ScalarArrayOp(True, "int4ge", x, 3, "int4", []),
        cg.HashedArrayOp("bool", "int4", x, [(v_, True) for v_ in (1, 7, 9, 11)] + [(None, True)],
                         in_list=True),
        cg.HashedArrayOp("bool", "int4", x, [(v_, True) for v_ in (1, 7, 9, 11)], in_list=True),
        cg.HashedArrayOp("int4", "int4", x, [(1, 10), (3, 30), (5, None), (7, 70)]),
        cg.HashedArrayOp("text", "int4", x, [(0, b"zero"), (1, b"one"), (2, b""), (3, b"x")],
                         b"other"),
        cg.CaseWhen("int4", [(cg.Func("int4gt", x, y), cg.Func("int4mi", x, y)),
                             (cg.NullTest(y, True), C("int4", -1))],
                    cg.Func("int4div", x, y)),
        cg.DistinctFrom(cg.Func("int4eq", x, y)),
        cg.MinMax("least", "int4", x, y, C("int4", 2)),
        cg.MinMax("greatest", "text", t, C("text", b"m")),
        cg.Func("textlike", t, C("text", b"%a_")),
        cg.Func("texticnlike", t, C("text", b"A%")),
        cg.Func("substr", t, x, cg.Func("int4abs", y)),
        cg.Func("textlen", t),
    ]


def _plan(cols, quals=None, group=0):
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    v = {c.name: plan.scan_var(c.name) for c in cols}
    plan.temp_slot("int4", 0)       # element slot of the ANY / ALL arrays (slot 3)
    ts = _targets(v)
    ts = ts[:8] if group == 0 else ts[8:]
    if quals is not None:
        plan.scan_quals = quals(v)
    plan.set_projection(ts, [("r%d" % i, e.typname) for i, e in enumerate(ts)])
    return plan.build()


@pytest.mark.parametrize("group", [0, 1])
def test_gpu_constructs_projection(gpu, jit, group):
    cols = _table()
    # CASE's x / y runs only where y is not NULL and x <= y: y = 0 there must
    # still raise, so keep y non-zero on those rows
    y = cols[1].values
    y[y == 0] = 1
    src = kds.build_arrow_kds(cols)
    blob = _plan(cols, group=group)
    got = run_gpu_projection(blob, src, 32 << 20, tuning=dict(jit=jit))
    exp, _ = run_oracle_projection(blob, src, 32 << 20)
    assert len(got) == 1 and np.array_equal(got[0], exp)


def test_gpu_constructs_quals(gpu, jit):
    """The constructs as scan quals: x IN (...), CASE, LIKE, NOT (x < ALL)."""
    cols = _table()
    y = cols[1].values
    y[y == 0] = 1
    src = kds.build_arrow_kds(cols)

    def quals(v):
        x, y, t = v["x"], v["y"], v["t"]
        return cg.And(
            cg.HashedArrayOp("bool", "int4", x, [(v_, True) for v_ in (0, 1, 2, 3, 5, 8, 9)],
                             in_list=True),
            cg.Not(cg.ScalarArrayOp(False, "int4lt", y, 3, "int4", [-1, 0])),
            cg.Or(cg.Func("textlike", t, cg.Const("text", b"%b%")),
                  cg.CaseWhen("bool", [(cg.Func("int4gt", x, cg.Const("int4", 4)),
                                        cg.Const("bool", True))], cg.Const("bool", False))))
    blob = _plan(cols, quals, group=1)
    got = run_gpu_projection(blob, src, 32 << 20, tuning=dict(jit=jit))
    exp, res = run_oracle_projection(blob, src, 32 << 20)
    assert res.nitems_out > 100
    assert len(got) == 1 and np.array_equal(got[0], exp)
