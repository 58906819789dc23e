"""numeric on the oracle, pinned by what the mathematics fixes (the
reference's own vectors pin every function case by case in
test_expr_golden.py): images round-trip to their decimal value, finite sums
and products are exact, integer casts round half away from zero, exact
quotients are exact, and numeric sums keep the scale of the argument's
typmod, truncating finer digits, with NaN and the infinities flagged.  A
plausible slip (a digit group in the wrong order, a lost sign, a scale off by
one, a rescale that rounds) fails one of these."""
import decimal
import struct

import numpy as np
import pytest

from pgstrom import abi, codegen as cg, kds, opcodes, workloads as W
from pgstrom.kds import Column, numeric_image, numeric_value
import oracle

D = decimal.Decimal
T = opcodes.TYPES


def _blob(expr):
    b = bytearray()
    cg.emit_expr(b, expr)
    return bytes(b)


def _eval(expr, slot_types, images):
    return oracle.eval_expr(_blob(expr), [T[t] for t in slot_types], images, T[expr.typname])


def _random_decimals(n, seed, digits=18, scale_range=(-6, 12)):
    import random
    r = random.Random(seed)
    out = []
    for _ in range(n):
        mag = r.randint(0, digits)
        v = r.randint(-10**mag, 10**mag)
        out.append(D(v).scaleb(-r.randint(scale_range[0], scale_range[1] - 1)))
    return out


@pytest.mark.parametrize("form", ["short", "long", "1b"])
def test_images_round_trip(form):
    """-(-x) parses x and writes it back in the reference's encoding: the
    same value, whatever header x came with."""
    N = lambda i: cg.Var(i, "numeric")
    e = cg.Func("numeric_uminus", cg.Func("numeric_uminus", N(0)))
    for v in _random_decimals(400, 3, digits=30):
        img = numeric_image(v, short_varlena=(form == "1b"), long_form=(form == "long"))
        got = _eval(e, ["numeric"], [img])
        assert numeric_value(got) == v, (v, got.hex())
        # NumericLong with a 4-byte header, display scale a multiple of 4
        assert got[0] & 3 == 0 and struct.unpack_from("<H", got, 4)[0] & 0x8000 == 0
        assert (struct.unpack_from("<H", got, 4)[0] & 0x3FFF) % 4 == 0


@pytest.mark.parametrize("fn,op", [("numeric_add", lambda a, b: a + b),
                                   ("numeric_sub", lambda a, b: a - b),
                                   ("numeric_mul", lambda a, b: a * b)])
def test_finite_arithmetic_is_exact(fn, op):
    N = lambda i: cg.Var(i, "numeric")
    xs = _random_decimals(60, 5, digits=15, scale_range=(-3, 8))
    ys = _random_decimals(60, 6, digits=15, scale_range=(-3, 8))
    with decimal.localcontext() as ctx:
        ctx.prec = 100
        for x, y in zip(xs, ys):
            got = _eval(cg.Func(fn, N(0), N(1)), ["numeric", "numeric"],
                        [numeric_image(x), numeric_image(y)])
            assert numeric_value(got) == op(x, y), (fn, x, y)


def test_exact_division_and_mod_as_division():
    """Quotients with a finite decimal expansion are exact; numeric_mod
    runs the reference's division (its pgfn_numeric_mod calls
    __decimal_div)."""
    N = lambda i: cg.Var(i, "numeric")
    for x, y, q in [("1", "4", "0.25"), ("-7.5", "2.5", "-3"), ("1e10", "-8", "-1250000000"),
                    ("0.001", "0.2", "0.005"), ("123.456", "1000", "0.123456")]:
        for fn in ("numeric_div", "numeric_mod"):
            got = _eval(cg.Func(fn, N(0), N(1)), ["numeric", "numeric"],
                        [numeric_image(x), numeric_image(y)])
            assert numeric_value(got) == D(q), (fn, x, y, got)
    got = _eval(cg.Func("numeric_div", N(0), N(1)), ["numeric", "numeric"],
                [numeric_image("1"), numeric_image("0")])
    assert got == ("E", abi.ERRCODE_DEVICE_ERROR, "division by zero")


def test_integer_casts_round_half_away_from_zero():
    N = lambda i: cg.Var(i, "numeric")
    for v in _random_decimals(300, 9, digits=12, scale_range=(0, 6)) + \
            [D("2.5"), D("-2.5"), D("0.5"), D("-0.5"), D("1.49999")]:
        exp = int(v.to_integral_value(rounding=decimal.ROUND_HALF_UP))
        got = _eval(cg.Func("numeric_to_int8", N(0)), ["numeric"], [numeric_image(v)])
        assert struct.unpack("<q", got)[0] == exp, v


def _sum_chunk(values):
    n = len(values)
    keys = np.arange(n, dtype=np.int32) % 7
    cols = [Column("k", "int4", keys), Column("n", "numeric",
                                              [numeric_image("0") if v is None else numeric_image(v)
                                               for v in values],
                                              np.array([v is None for v in values]))]
    return cols, kds.build_block_kds(cols, seed=2), keys


@pytest.mark.parametrize("typmod_scale", [None, 0, 2, 6])
def test_numeric_sum_keeps_the_typmod_scale(typmod_scale):
    """sum(n) / avg(n) states: the count of finite values and their sum at
    the typmod's scale (15 without a typmod), each value's finer digits
    truncated toward zero (__normalize_numeric_int128); NaN and the
    infinities set their flags and are not counted."""
    vals = _random_decimals(900, 11, digits=14, scale_range=(-2, 9))
    for i in range(0, 900, 97):
        vals[i] = None
    vals[5], vals[13], vals[27] = D("NaN"), D("Infinity"), D("-Infinity")
    cols, src, keys = _sum_chunk(vals)
    plan = cg.Plan([("k", "int4"), ("n", "numeric")])
    k, nv = plan.scan_var("k"), plan.scan_var("n")
    typmod = -1 if typmod_scale is None else ((30 << 16) | typmod_scale) + 4
    plan.set_groupby([(abi.KAGG_ACTION__PSUM_NUMERIC, nv, ("s", "bytea", typmod)),
                      (abi.KAGG_ACTION__PAVG_NUMERIC, nv, ("a", "bytea", typmod)),
                      (abi.KAGG_ACTION__VREF, k, ("k", "int4"))])
    o = oracle.OracleSession(plan.build(final_nrows=100))
    assert o.exec(src, nthreads=2)[0] == 0
    groups = W.decode_groups(o.final(), [abi.KAGG_ACTION__PSUM_NUMERIC,
                                         abi.KAGG_ACTION__PAVG_NUMERIC, abi.KAGG_ACTION__VREF])
    o.close()
    scale = 15 if typmod_scale is None else typmod_scale
    assert len(groups) == 7
    for (kv,), states in groups.items():
        mine = [v for v, kk in zip(vals, keys) if kk == kv and v is not None]
        finite = [v for v in mine if v.is_finite()]
        flags = 0
        for v in mine:
            if v.is_nan():
                flags |= abi.__PAGG_NUMERIC_ATTRS__NAN
            elif v.is_infinite():
                flags |= abi.__PAGG_NUMERIC_ATTRS__PINF if v > 0 else abi.__PAGG_NUMERIC_ATTRS__NINF
        exp_sum = sum(int((v.scaleb(scale)).to_integral_value(rounding=decimal.ROUND_DOWN))
                      for v in finite)
        for attrs, n, s in states:
            assert attrs == scale | flags
            assert n == len(finite)
            assert s == exp_sum


def test_float_to_numeric_exact_on_short_dyadic_values():
    """float8 -> numeric (__decimal_from_float8): the significand is
    halved to binary exponent 0 while factors of 10 keep it in range, so a
    value with few significant bits and at most 15 decimal digits comes out
    exactly; longer ones are cut to 15 digits (rounding of the halvings
    aside, which the reference's own vectors pin)."""
    F = cg.Var(0, "float8")
    for d in (1.0, -2.5, 0.75, 123456.0, 2.0 ** 40, -0.125, 1e14, 3.0, 0.0009765625):
        got = numeric_value(_eval(cg.Func("float8_to_numeric", F), ["float8"],
                                  [struct.pack("<d", d)]))
        exact = D(d)
        q = exact.scaleb(-exact.adjusted()).quantize(D("1.00000000000000"), rounding=decimal.ROUND_DOWN)
        assert got == q.scaleb(exact.adjusted()), (d, got)
