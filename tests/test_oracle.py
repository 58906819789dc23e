"""The oracle pinned against closed forms computed independently with numpy
on the same seeded data (it is the checker of the GPU path)."""
import collections
import math

import numpy as np
import pytest

import oracle
from pgstrom import abi, codegen as cg, datagen, kds, workloads as W


def _scan_expected(cols, k):
    a, b, c, d = cols
    an = a.nulls if a.nulls is not None else np.zeros(len(a), bool)
    bn = b.nulls if b.nulls is not None else np.zeros(len(b), bool)
    sel = np.nonzero(~an & (a.values < k) & ~bn & (b.values > 0.0))[0]
    return [((None if c.isnull(i) else int(c.values[i])),
             (None if d.isnull(i) else bytes(d.values[i]))) for i in sel]


@pytest.mark.parametrize("n", [0, 1, 1000, 100_000])
def test_oracle_scan_closed_form(n):
    cols = datagen.scan_table(n)
    src = kds.build_arrow_kds(cols)
    o = oracle.OracleSession(W.scan_query(cols))
    rc, dst, res = o.exec(src, dst_length=32 << 20, nthreads=3)
    assert rc == 0
    rows = kds.parse_row_kds(dst)
    assert [r[0] for r in rows] == _scan_expected(cols, 500000)
    assert [r[1] for r in rows] == list(range(len(rows)))


def test_oracle_scan_nan_and_negzero():
    """float8gt sorts NaN above everything (xpu_basetype.cu:475-476); -0.0 is
    not > 0.0."""
    n = 6
    cols = [kds.Column("a", "int4", np.zeros(n, np.int32)),
            kds.Column("b", "float8", np.array([np.nan, -0.0, 0.0, 1e-300, -1.0, np.inf])),
            kds.Column("c", "int8", np.arange(n)),
            kds.Column("d", "text", [b"x"] * n)]
    o = oracle.OracleSession(W.scan_query(cols))
    rc, dst, _ = o.exec(kds.build_arrow_kds(cols), dst_length=1 << 20)
    assert [r[0][0] for r in kds.parse_row_kds(dst)] == [0, 3, 5]


def test_oracle_preagg_closed_form():
    cols = datagen.preagg_table(200_000)
    o = oracle.OracleSession(W.preagg_query(cols))
    o.exec(kds.build_arrow_kds(cols), nthreads=4)
    acts = [a for a, _, _ in W.preagg_tlist(None, cols[0], None, None, None)]
    got = W.decode_groups(o.final(), acts)
    key, v, w, x = cols
    uk, inv = np.unique(key.values, return_inverse=True)
    assert len(got) == len(uk)
    for i, kk in enumerate(uk):
        sel = inv == i
        st = got[(int(kk),)]
        n = int(sel.sum())
        assert st[0] == n
        assert st[1] == (n, int(v.values[sel].astype(object).sum()))
        assert st[2][0] == n
        assert math.isclose(st[2][1], math.fsum(w.values[sel]), rel_tol=1e-12)
        assert st[3] == (1, int(x.values[sel].min()))
        assert st[4] == (1, int(x.values[sel].max()))


def test_oracle_star_join_closed_form():
    fact, dims = datagen.star_tables(100_000, dim_sizes=(5000, 700, 256))
    km = W.star_inner(dims)
    o = oracle.OracleSession(W.star_query(fact))
    rc, _, res = o.exec(kds.build_arrow_kds(fact), kmrels=km, nthreads=2)
    assert rc == 0
    got = W.decode_groups(o.final(), [abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF,
                                      abi.KAGG_ACTION__VREF])
    fk1, fk2, fk3, q, p = fact
    m = (fk1.values % 10 == 0) & (fk2.values % 10 == 1) & (fk3.values % 10 == 2)
    c1 = dims[0][1].values[fk1.values[m] - 1]
    c3 = dims[2][1].values[fk3.values[m] - 1]
    amt = q.values[m].astype(object) * p.values[m].astype(object)
    exp = collections.defaultdict(lambda: [0, 0])
    for a_, b_, s_ in zip(c1, c3, amt):
        exp[(int(a_), int(b_))][0] += 1
        exp[(int(a_), int(b_))][1] += int(s_)
    assert {k: v[0] for k, v in got.items()} == {k: tuple(v) for k, v in exp.items()}


def test_oracle_error_division_by_zero_guarded():
    """An AND stops at its first false argument, so a division by zero in a
    later argument raises nothing for that row (xpu_common.cu:472-503)."""
    cols = [kds.Column("a", "int4", np.array([0, 2, 5], np.int32))]
    plan = cg.Plan([("a", "int4")])
    a = plan.scan_var("a")
    plan.scan_quals = cg.And(cg.Func("int4ne", a, cg.Const("int4", 0)),
                             cg.Func("int4gt", cg.Func("int4div", cg.Const("int4", 10), a),
                                     cg.Const("int4", 2)))
    plan.set_projection([a], [("a", "int4")])
    o = oracle.OracleSession(plan.build())
    rc, dst, _ = o.exec(kds.build_arrow_kds(cols), dst_length=1 << 20)
    assert rc == 0
    assert [r[0][0] for r in kds.parse_row_kds(dst)] == [2]


def test_oracle_error_planted():
    """A planted mistake (dividing by a zero row without a guard) is an error."""
    cols = [kds.Column("a", "int4", np.array([0, 2], np.int32))]
    plan = cg.Plan([("a", "int4")])
    a = plan.scan_var("a")
    plan.scan_quals = cg.Func("int4gt", cg.Func("int4div", cg.Const("int4", 10), a),
                              cg.Const("int4", 2))
    plan.set_projection([a], [("a", "int4")])
    o = oracle.OracleSession(plan.build())
    rc, _, _ = o.exec(kds.build_arrow_kds(cols), dst_length=1 << 20)
    assert rc == abi.ERRCODE_DEVICE_ERROR
    assert b"division by zero" in o.err.message


def test_window_headers_match_sliced_chunks():
    """Window headers over one columnar chunk describe the same rows as
    chunks built from the sliced columns (checked through the oracle)."""
    n = 5000
    cols = datagen.preagg_table(n)
    cols[2] = kds.Column("w", "float8", cols[2].values, nulls=np.arange(n) % 7 == 0)
    base = kds.build_arrow_kds(cols)
    windows = [(0, 10), (64, 1000), (4096, 904), (8, 0)]
    stride = 4096
    hdrs = kds.window_headers(base, windows, len(windows) * stride, stride)
    buf = np.concatenate([hdrs, base])
    blob = W.preagg_query(cols)
    acts = [a for a, _, _ in W.preagg_tlist(None, cols[0], None, None, None)]
    for i, (r0, m) in enumerate(windows):
        o = oracle.OracleSession(blob)
        o.exec(buf[i * stride:], nthreads=2)
        got = W.decode_groups(o.final(), acts)
        o.close()
        sl = [kds.Column(c.name, c.typname, c.values[r0:r0 + m],
                         nulls=None if c.nulls is None else c.nulls[r0:r0 + m]) for c in cols]
        o = oracle.OracleSession(blob)
        o.exec(kds.build_arrow_kds(sl), nthreads=2)
        exp = W.decode_groups(o.final(), acts)
        o.close()
        assert got == exp


@pytest.mark.parametrize("layout", ["large", "fixed"])
def test_oracle_arrow_text_layouts(layout):
    """Text columns as LargeUtf8 (64-bit offsets) and FixedSizeBinary read the
    same values as Utf8 (xpu_textlib.cu:19-54): identical kds_dst bytes."""
    from pgstrom import codegen as cg
    n = 3000
    r = datagen.rng(7, 1)
    a = kds.Column("a", "int4", r.integers(0, 100, size=n, dtype=np.int32))
    nulls = r.random(n) < 0.1
    t = kds.Column("t", "text", datagen.random_text(r, n, 12, 12 if layout == "fixed" else 40,
                                                    nulls=nulls), nulls)
    cols = [a, t]
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    va, vt = plan.scan_var("a"), plan.scan_var("t")
    plan.scan_quals = cg.Func("int4lt", va, cg.Const("int4", 60))
    plan.set_projection([vt, va], [("t", "text"), ("a", "int4")])
    blob = plan.build()
    outs = []
    for lay in ("utf8", layout):
        o = oracle.OracleSession(blob)
        rc, dst, _ = o.exec(kds.build_arrow_kds(cols, text_layout=lay), dst_length=4 << 20)
        o.close()
        assert rc == 0
        outs.append(dst)
    assert np.array_equal(outs[0], outs[1])


def test_oracle_unmappable_arrow_type():
    """A text column whose Arrow type is not a varlena layout is an error."""
    from pgstrom import codegen as cg
    import ctypes
    cols = [kds.Column("t", "text", [b"x", b"yy"])]
    src = kds.build_arrow_kds(cols)
    cm = abi.kern_colmeta.from_buffer(src, abi.KDS_COLMETA_OFFSET)
    cm.attopts.tag = abi.ArrowType__Int
    plan = cg.Plan([("t", "text")])
    vt = plan.scan_var("t")
    plan.set_projection([vt], [("t", "text")])
    o = oracle.OracleSession(plan.build())
    rc, _, _ = o.exec(src, dst_length=1 << 20)
    o.close()
    assert rc == abi.ERRCODE_DEVICE_ERROR
