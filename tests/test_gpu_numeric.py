"""numeric on the GPU against the oracle: projections of numeric columns
(copied as they are) and expressions (parsed and written in the reference's
encoding), numeric sums and averages (PSUM_NUMERIC / PAVG_NUMERIC) at the
typmod's scale over heap pages and Arrow Decimal128 chunks, numeric grouping
keys, and a hash join on a numeric key whose inner relation the library
builds natively.  Both kernels (interpreter and plan-specialised) run."""
import decimal
import random

import numpy as np
import pytest

from pgstrom import abi, codegen as cg, kds, xpuhip
from pgstrom.kds import Column, DecimalValues, numeric_image
from helpers import assert_groups_equal, run_gpu_projection, run_oracle_groupby, \
    run_oracle_projection
import oracle

pytestmark = pytest.mark.gpu
D = decimal.Decimal


def _decimals(n, seed, digits=14, scales=(-2, 9)):
    r = random.Random(seed)
    out = []
    for _ in range(n):
        mag = r.randint(0, digits)
        out.append(D(r.randint(-10**mag, 10**mag)).scaleb(-r.randint(*scales)))
    return out


def _heap_chunk(keys, vals, seed=1):
    """int4 k, numeric n (on-page form; None = NULL), int8 m"""
    n = len(keys)
    imgs = [numeric_image("0") if v is None else
            (v if isinstance(v, bytes) else numeric_image(v, short_varlena=True)) for v in vals]
    cols = [Column("k", "int4", np.asarray(keys, dtype=np.int32)),
            Column("n", "numeric", imgs, np.array([v is None for v in vals])),
            Column("m", "int8", np.arange(n, dtype=np.int64) * 3 - 7)]
    return cols, kds.build_block_kds(cols, seed=seed)


PROJECTIONS = {
    # the column itself (copied), a sum with a constant, a negation
    "copy_add_neg": lambda n, m: ([n, cg.Func("numeric_add", n, cg.Const("numeric", "1.5")),
                                   cg.Func("numeric_uminus", n)],
                                  [("n", "numeric"), ("a", "numeric"), ("c", "numeric")]),
    # a product with an integer column, abs, a float cast, an integer cast
    "mul_abs_casts": lambda n, m: ([cg.Func("numeric_mul", n, cg.Func("int8_to_numeric", m)),
                                    cg.Func("numeric_abs", n), cg.Func("numeric_to_float8", n),
                                    cg.Func("int8_to_numeric", m)],
                                   [("b", "numeric"), ("d", "numeric"), ("f", "float8"),
                                    ("g", "numeric")]),
}


@pytest.mark.parametrize("which", sorted(PROJECTIONS))
def test_projection(gpu, jit, which):
    """WHERE n > -1000 projecting numeric expressions: the column is copied
    byte for byte, the expressions are written as the reference writes
    them.  (A numeric value takes three registers: seven numeric targets in
    one plan exceed the 32-register file, which the session open refuses.)"""
    vals = _decimals(3000, 3) + [D("NaN"), D("Infinity"), D("-Infinity"), None]
    keys = list(range(len(vals)))
    cols, src = _heap_chunk(keys, vals)
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    n, m = plan.scan_var("n"), plan.scan_var("m")
    plan.scan_quals = cg.Func("numeric_gt", n, cg.Const("numeric", "-1000"))
    targets, coldefs = PROJECTIONS[which](n, m)
    plan.set_projection(targets, coldefs)
    blob = plan.build()
    got = run_gpu_projection(blob, src, 8 << 20, tuning=dict(jit=jit))
    exp, _ = run_oracle_projection(blob, src, 8 << 20)
    assert len(got) == 1 and np.array_equal(got[0], exp)
    rows = [r[0] for r in kds.parse_row_kds(got[0])]
    assert len(rows) > 1000
    if which == "copy_add_neg":
        want = {numeric_image(v, short_varlena=True) for v in vals if v is not None}
        assert all(r[0] in want for r in rows)


def _sum_plan(coldefs, typmod):
    plan = cg.Plan(coldefs)
    k, nv = plan.scan_var("k"), plan.scan_var("n")
    plan.set_groupby([(abi.KAGG_ACTION__NROWS_ANY, None, ("c", "int8")),
                      (abi.KAGG_ACTION__PSUM_NUMERIC, nv, ("s", "bytea", typmod)),
                      (abi.KAGG_ACTION__PAVG_NUMERIC, nv, ("a", "bytea", typmod)),
                      (abi.KAGG_ACTION__VREF, k, ("k", "int4"))])
    return plan.build(final_nrows=200)


ACTS = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_NUMERIC,
        abi.KAGG_ACTION__PAVG_NUMERIC, abi.KAGG_ACTION__VREF]


@pytest.mark.parametrize("typmod", [-1, ((20 << 16) | 4) + 4])
def test_numeric_sum_heap(gpu, jit, typmod):
    vals = _decimals(20000, 7)
    for i in range(0, len(vals), 131):
        vals[i] = None
    vals[10], vals[11], vals[12] = D("NaN"), D("Infinity"), D("-Infinity")
    keys = [i % 37 for i in range(len(vals))]
    cols, src = _heap_chunk(keys, vals)
    blob = _sum_plan([(c.name, c.typname) for c in cols], typmod)
    s = xpuhip.Session(blob)
    s.set_tuning(jit=jit)
    for _ in range(2):
        assert s.exec(src)[0] == 0
    got = s.final()
    s.close()
    exp = run_oracle_groupby(blob, [src, src])
    groups = assert_groups_equal(got, exp, ACTS)
    # the closed form: finite values truncated to the sum's scale
    scale = 15 if typmod < 0 else 4
    for (kv,), st in groups.items():
        fin = [v for v, kk in zip(vals, keys) if kk == kv and v is not None and v.is_finite()]
        total = 2 * sum(int(v.scaleb(scale).to_integral_value(rounding=decimal.ROUND_DOWN))
                        for v in fin)
        assert st[1][1] == 2 * len(fin) and st[1][2] == total


def test_numeric_sum_arrow_decimal(gpu, jit):
    """An Arrow Decimal128 column at scale 3 (xpu_numeric_datum_arrow_read)."""
    r = random.Random(5)
    n = 50000
    ints = [r.randint(-10**15, 10**15) for _ in range(n)]
    nulls = np.array([i % 101 == 0 for i in range(n)])
    keys = np.array([i % 53 for i in range(n)], dtype=np.int32)
    cols = [Column("k", "int4", keys), Column("n", "numeric", DecimalValues(ints, 3), nulls)]
    src = kds.build_arrow_kds(cols)
    blob = _sum_plan([("k", "int4"), ("n", "numeric")], ((30 << 16) | 2) + 4)
    s = xpuhip.Session(blob)
    s.set_tuning(jit=jit)
    assert s.exec(src)[0] == 0
    got = s.final()
    s.close()
    exp = run_oracle_groupby(blob, [src])
    groups = assert_groups_equal(got, exp, ACTS)
    for (kv,), st in groups.items():
        # scale 3 to scale 2: one digit cut toward zero
        want = sum(int(D(v) / 10) for v, kk, nl in zip(ints, keys, nulls) if kk == kv and not nl)
        assert st[1][2] == want


def test_numeric_grouping_key(gpu, jit):
    """GROUP BY n: equal values group together whatever their image (1.5 and
    1.50 are one group); the key tuple keeps the image it was first seen
    with, so the test gives every value one image."""
    base = [D("1.5"), D("-2"), D("0"), D("1e20"), D("0.001"), D("NaN"), D("123456789.123456789")]
    vals = [base[i % len(base)] for i in range(6000)] + [None] * 50
    keys = list(range(len(vals)))
    cols, src = _heap_chunk(keys, vals)
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    nv, m = plan.scan_var("n"), plan.scan_var("m")
    plan.set_groupby([(abi.KAGG_ACTION__NROWS_ANY, None, ("c", "int8")),
                      (abi.KAGG_ACTION__PSUM_INT64, m, ("s", "bytea")),
                      (abi.KAGG_ACTION__VREF, nv, ("n", "numeric"))])
    blob = plan.build(final_nrows=50)
    s = xpuhip.Session(blob)
    s.set_tuning(jit=jit)
    assert s.exec(src)[0] == 0
    got = s.final()
    s.close()
    exp = run_oracle_groupby(blob, [src])
    groups = assert_groups_equal(got, exp, [abi.KAGG_ACTION__NROWS_ANY,
                                            abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF])
    assert len(groups) == len(base) + 1


def test_numeric_join_key(gpu, jit):
    """fact JOIN dim ON fact.n = dim.n: the inner relation hashed by the
    library's host build (numeric keys by value), probed on the device."""
    dim_vals = [D(i).scaleb(-2) for i in range(-300, 300, 3)]
    coldefs = [("dn", "numeric"), ("tag", "int4")]
    mt = [kds.form_minimal_tuple(coldefs, (numeric_image(v), i % 9))
          for i, v in enumerate(dim_vals)]
    inner = xpuhip.inner_hash_setup(coldefs, mt, [1])
    km = kds.build_multirels([(inner, False)])
    r = random.Random(11)
    vals = [D(r.randint(-400, 400)).scaleb(-2) for _ in range(8000)]
    cols, src = _heap_chunk(list(range(len(vals))), vals)
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    nv, m = plan.scan_var("n"), plan.scan_var("m")
    d = plan.add_inner(coldefs)
    dn, tag = plan.inner_var(d, "dn"), plan.inner_var(d, "tag")
    plan.add_join(d, [nv], [cg.Func("numeric_eq", nv, dn)])
    plan.set_groupby([(abi.KAGG_ACTION__NROWS_ANY, None, ("c", "int8")),
                      (abi.KAGG_ACTION__PSUM_INT64, m, ("s", "bytea")),
                      (abi.KAGG_ACTION__VREF, tag, ("tag", "int4"))])
    blob = plan.build(final_nrows=20)
    s = xpuhip.Session(blob)
    s.set_tuning(jit=jit)
    s.attach_inner(km)
    assert s.exec(src)[0] == 0
    got = s.final()
    s.close()
    exp = run_oracle_groupby(blob, [src], kmrels=km)
    groups = assert_groups_equal(got, exp, [abi.KAGG_ACTION__NROWS_ANY,
                                            abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF])
    dims = set(dim_vals)
    assert sum(st[0] for st in groups.values()) == sum(v in dims for v in vals)
