"""Inputs of the device-function golden vectors (shared by the generator,
gen_expr_vectors.py, and the tests that replay the vectors): per type a
table of edge-case values as PostgreSQL datum images, and the expression
templates over variable slots that exercise the constructs beyond single
functions (SQL three-valued AND / OR / NOT, NullTest, BoolTest, COALESCE,
HashValue, short-circuit guards)."""
import struct

from pgstrom.kds import numeric_image as _num

NAN = float("nan")
INF = float("inf")


def _text(s):
    b = s.encode() if isinstance(s, str) else bytes(s)
    return struct.pack("<I", (len(b) + 4) << 2) + b


VALUES = {
    "bool": [struct.pack("<?", v) for v in (False, True)],
    "int1": [struct.pack("<b", v) for v in (0, 1, -1, 5, -7, 11, 127, -128, 100, -100, 64)],
    "int2": [struct.pack("<h", v) for v in (0, 1, -1, 7, -7, 32767, -32768, 181, 182, 300,
                                            -300, 255, 128)],
    "int4": [struct.pack("<i", v) for v in (0, 1, -1, 7, -7, 2**31 - 1, -2**31, 46340, 46341,
                                            123456, -654321, 32767, 128, 500000)],
    "int8": [struct.pack("<q", v) for v in (0, 1, -1, 7, -7, 2**63 - 1, -2**63, 3037000499,
                                            3037000500, 10**12, -10**12, 2**31, 2**40 + 3)],
    "float4": [struct.pack("<f", v) for v in (0.0, -0.0, 1.0, -1.0, 2.5, 0.1, 3.0e38, -3.0e38,
                                              1e-38, INF, -INF, NAN, 1.5e10, 127.5, -32768.5,
                                              2147483520.0, 9.2233715e18)],
    "float8": [struct.pack("<d", v) for v in (0.0, -0.0, 1.0, -1.0, 2.5, 3.5, -2.5, 0.1, 1e308,
                                              -1e308, 5e-324, INF, -INF, NAN, 2147483647.5,
                                              -2147483648.5, 2147483647.4, 9.3e18,
                                              -9.2233720368547758e18, 1e39, 127.5, 3.4028235e38)],
    "text": [_text(s) for s in ("", "a", "abc", "abd", "ab", "ABC", "abc ", "b", "été",
                                "zzzzzzzzzzzzzzzzzzzzzzzzzzzzzz", "a%c", "_b")],
    # numeric datums as PostgreSQL stores them: NumericShort under a 4-byte or
    # a 1-byte varlena header, NumericLong, NaN and the infinities; then
    # values at the edges of the integer casts, and datums the reference
    # rejects (more than 39 digits, a body shorter than its header, an
    # unknown special header)
    "numeric": [_num(v) for v in ("0", "1", "-1", "1.5", "-2.25", "0.001", "123.456", "1e20",
                                  "-1e20", "99999999999999999999.99", "1e-10", "NaN",
                                  "Infinity", "-Infinity")] +
               [_num("3.14", short_varlena=True),
                _num("12345678901234567890123456789012345678"),
                _num("-42", short_varlena=True), _num("5.5", long_form=True), _num("0.10"),
                _num("2.5"), _num("-2.5"), _num("-0.5"), _num("0.4"), _num("32767.5"),
                _num("2147483647.4"), _num("9223372036854775807"), _num("9223372036854775808"),
                _num("1e-30"), _num("10000"), _num("-7e300", long_form=True),
                _num("123456789012345678901234567890123456789012345"),
                struct.pack("<I", 5 << 2) + b"\x80",
                struct.pack("<IH", 6 << 2, 0xE000)],
    # bpchar values as a char(n) column stores them: padded with spaces
    "bpchar": [_text(s) for s in ("", "a   ", "abc", "abc  ", "abd ", "ab", "ABC", " a", "été ",
                                  "b")],
}

# the session's database encoding of the golden vectors (text functions
# count characters by it)
ENCODING = "UTF8"

# LIKE patterns (the second argument of the LIKE family)
PATTERNS = [_text(s) for s in ("%", "", "a%", "%c", "a_c", "%b%", "abc", "ABC", "a%c%",
                               "_", "__", "%_", "\\%", "a\\", "%\\", "a\\_c", "é%", "_t_",
                               "%%c", "b%", "%d")]
LIKE_TEXTS = [_text(s) for s in ("", "a", "abc", "abd", "ab", "ABC", "aXc", "a%c", "a_c", "%",
                                 "été", "xbx", "abcabc", "a\\")]

# substr / substring arguments: (text index, start, count)
SUBSTR_STARTS = [-3, -1, 0, 1, 2, 3, 5, 100]
SUBSTR_COUNTS = [0, 1, 2, 3, 100, -1]


def numeric_canon(img):
    """A numeric result image as the golden comparison reads it.  The
    reference's __decimal_to_varlena keeps at most 10 base-10000 digits in a
    stack array (PG_MAX_DATA, xpu_numeric.h:47-49, 275-372); a value that
    needs 11 (a 38-digit integer at scale 4) is written one slot below it,
    over the sign flag, whose value is then the stack's.  The digits, scale
    and weight of such an image are compared, not its sign bit."""
    if img is None or isinstance(img, (tuple, list)) or len(img) < 8 or img[0] & 1:
        return img
    head = struct.unpack_from("<H", img, 4)[0]
    if head & 0x8000 == 0 and (len(img) - 8) // 2 > 10:
        return img[:4] + struct.pack("<H", head & ~0x4000) + img[6:]
    return img


def func_arg_sets(kind, types):
    """Explicit argument tuples for function kinds whose inputs are not a
    plain cross product: (value lists per argument)."""
    if kind == "like":
        return [VALUES[types[0]] if types[0] == "bpchar" else LIKE_TEXTS, PATTERNS]
    if kind == "substr":
        texts = [VALUES["text"][i] for i in (0, 2, 8, 9)]
        sets = [texts, [struct.pack("<i", v) for v in SUBSTR_STARTS]]
        if len(types) == 3:
            sets.append([struct.pack("<i", v) for v in SUBSTR_COUNTS])
        return sets
    return None

# binary functions take the cross product of the first BINARY_N values of
# each argument type (the full tables for unary functions)
BINARY_N = {"bool": 2, "int1": 9, "int2": 10, "int4": 11, "int8": 11, "float4": 14,
            "float8": 16, "text": 12, "bpchar": 10, "numeric": 16}


_INT_FMT = {"int1": "<b", "int2": "<h", "int4": "<i", "int8": "<q"}


def host_traps(name, types, imgs):
    """True for the one input the reference's device code cannot run on a
    host CPU: TYPEmod(MIN, -1) computes x % y directly (xpu_basetype.cu:
    748-769), which traps on x86.  PostgreSQL's int4mod defines it as 0,
    which this build and the oracle return (pinned by the closed form
    x % -1 = 0 in the tests instead)."""
    if not name.endswith("mod") or None in imgs or types[0] not in _INT_FMT:
        return False
    fx, fy = _INT_FMT[types[0]], _INT_FMT[types[1]]
    x = struct.unpack(fx, imgs[0])[0]
    y = struct.unpack(fy, imgs[1])[0]
    bits = 8 * struct.calcsize(fx)
    return y == -1 and x == -(1 << (bits - 1))


def arg_tuples(types, kind=None):
    """Index tuples into the argument value lists (func_arg_lists)."""
    sets = func_arg_sets(kind, types)
    if sets is not None:
        out = [[]]
        for vals in sets:
            out = [o + [i] for o in out for i in range(len(vals))]
        return [tuple(o) for o in out] + [tuple(None if j == k else 0 for j in range(len(types)))
                                          for k in range(len(types))]
    if len(types) == 1:
        return [(i,) for i in range(len(VALUES[types[0]]))] + [(None,)]
    a, b = types
    out = [(i, j) for i in range(BINARY_N[a]) for j in range(BINARY_N[b])]
    return out + [(None, 0), (0, None), (None, None)]


def templates(cg):
    """(name, expression over slots, slot types, compare_nulls) of the
    composite constructs; each is evaluated over every combination of its
    slots' values (bool: false / true / NULL; int4: a few values and NULL)."""
    B = lambda i: cg.Var(i, "bool")
    I = lambda i: cg.Var(i, "int4")
    T = lambda i: cg.Var(i, "text")
    N = lambda i: cg.Var(i, "numeric")
    out = [
        ("and2", cg.And(B(0), B(1)), ["bool", "bool"], 0),
        ("and3", cg.And(B(0), B(1), B(2)), ["bool", "bool", "bool"], 0),
        ("or2", cg.Or(B(0), B(1)), ["bool", "bool"], 0),
        ("or3", cg.Or(B(0), B(1), B(2)), ["bool", "bool", "bool"], 0),
        ("not", cg.Not(B(0)), ["bool"], 0),
        ("not_and_or", cg.Not(cg.Or(cg.And(B(0), B(1)), B(2))), ["bool", "bool", "bool"], 0),
        ("isnull_int4", cg.NullTest(I(0), True), ["int4"], 0),
        ("isnotnull_text", cg.NullTest(T(0), False), ["text"], 0),
        ("coalesce3", cg.Coalesce("int4", I(0), I(1), I(2)), ["int4", "int4", "int4"], 0),
        ("coalesce_text", cg.Coalesce("text", T(0), T(1)), ["text", "text"], 0),
        ("hash_int4_text", cg.HashValue(I(0), T(1)), ["int4", "text"], 0),
        ("eq_nulls", cg.Func("int4eq", I(0), I(1)), ["int4", "int4"], 1),
        ("ne_nulls", cg.Func("int4ne", I(0), I(1)), ["int4", "int4"], 1),
        ("text_eq_nulls", cg.Func("text_eq", T(0), T(1)), ["text", "text"], 1),
        # AND stops at the first false: the division is evaluated only for
        # rows whose first argument is not false
        ("and_guard", cg.And(B(0), cg.Func("int4eq", cg.Func("int4div", I(1), I(2)), I(1))),
         ["bool", "int4", "int4"], 0),
        ("or_guard", cg.Or(B(0), cg.Func("int4eq", cg.Func("int4div", I(1), I(2)), I(1))),
         ["bool", "int4", "int4"], 0),
        ("nested_arith", cg.Func("int8mul", cg.Func("int4_to_int8", cg.Func("int4pl", I(0), I(1))),
                                 cg.Func("int4_to_int8", I(2))), ["int4", "int4", "int4"], 0),
    ]
    for kind in cg.BoolTest.KINDS:
        out.append(("booltest_" + kind, cg.BoolTest(kind, B(0)), ["bool"], 0))
    C = lambda t, v: cg.Const(t, v)
    out += [
        ("distinct_int4", cg.DistinctFrom(cg.Func("int4eq", I(0), I(1))), ["int4", "int4"], 0),
        ("distinct_text", cg.DistinctFrom(cg.Func("text_eq", T(0), T(1))), ["text", "text"], 0),
        ("least3", cg.MinMax("least", "int4", I(0), I(1), I(2)), ["int4", "int4", "int4"], 0),
        ("greatest3", cg.MinMax("greatest", "int4", I(0), I(1), I(2)),
         ["int4", "int4", "int4"], 0),
        ("least_text", cg.MinMax("least", "text", T(0), T(1)), ["text", "text"], 0),
        ("greatest_text", cg.MinMax("greatest", "text", T(0), T(1), C("text", b"abc")),
         ["text", "text"], 0),
        ("case_searched", cg.CaseWhen("int4", [(cg.Func("int4gt", I(0), C("int4", 0)), I(1)),
                                               (cg.Func("int4lt", I(0), C("int4", 0)), I(2))],
                                      C("int4", -99)), ["int4", "int4", "int4"], 0),
        ("case_no_else", cg.CaseWhen("text", [(B(0), T(1)), (cg.Not(B(0)), C("text", b"no"))]),
         ["bool", "text"], 0),
        # later WHEN / THEN arms are evaluated only where no earlier arm matched
        ("case_guard", cg.CaseWhen("int4", [(cg.Func("int4eq", I(1), C("int4", 0)), C("int4", 0)),
                                            (B(0), cg.Func("int4div", I(2), I(1)))],
                                   cg.Func("int4div", I(1), I(2))), ["bool", "int4", "int4"], 0),
        ("case_key", cg.CaseWhen("int4", [(cg.Func("int4eq", I(3), C("int4", 1)), C("int4", 10)),
                                          (cg.Func("int4eq", I(3), C("int4", 7)), I(1))],
                                 I(2), key=(3, cg.Func("int4pl", I(0), C("int4", 0)))),
         ["int4", "int4", "int4", "int4"], 0),
        ("any_int4", cg.ScalarArrayOp(True, "int4eq", I(0), 1, "int4", [1, 7, 2**31 - 1]),
         ["int4", "int4"], 0),
        ("any_int4_nulls", cg.ScalarArrayOp(True, "int4eq", I(0), 1, "int4", [7, None, 1]),
         ["int4", "int4"], 0),
        ("any_lt", cg.ScalarArrayOp(True, "int4lt", I(0), 1, "int4", [-5, 0]), ["int4", "int4"], 0),
        ("any_empty", cg.ScalarArrayOp(True, "int4eq", I(0), 1, "int4", []), ["int4", "int4"], 0),
        ("any_text", cg.ScalarArrayOp(True, "text_eq", T(0), 1, "text", [b"abd", b"", b"zz"]),
         ["text", "text"], 0),
        ("in_hashed", cg.HashedArrayOp("bool", "int4", I(0), [(v, True) for v in (1, 7, 9, 11, 0)],
                                       in_list=True), ["int4"], 0),
        ("in_hashed_text", cg.HashedArrayOp("bool", "text", T(0),
                                            [(v, True) for v in (b"abc", b"x", b"y", b"", b"q")],
                                            in_list=True), ["text"], 0),
        ("case_hashed", cg.HashedArrayOp("int4", "int4", I(0),
                                         [(1, 100), (7, 700), (2**31 - 1, -1), (-2, 20)], -5),
         ["int4"], 0),
        ("case_hashed_text", cg.HashedArrayOp("text", "int4", I(0),
                                              [(0, b"zero"), (1, b"one"), (7, b"seven"),
                                               (-2, b"")], b"other"), ["int4"], 0),
    ]
    out += [
        ("least_numeric", cg.MinMax("least", "numeric", N(0), N(1), N(2)),
         ["numeric", "numeric", "numeric"], 0),
        ("greatest_numeric", cg.MinMax("greatest", "numeric", N(0), N(1)),
         ["numeric", "numeric"], 0),
        ("coalesce_numeric", cg.Coalesce("numeric", N(0), N(1)), ["numeric", "numeric"], 0),
        ("case_numeric", cg.CaseWhen("numeric",
                                     [(cg.Func("numeric_gt", N(0), C("numeric", "0")), N(1))],
                                     C("numeric", "-1.25")), ["numeric", "numeric"], 0),
        ("hash_numeric", cg.HashValue(N(0), I(1)), ["numeric", "int4"], 0),
        ("numeric_eq_nulls", cg.Func("numeric_eq", N(0), N(1)), ["numeric", "numeric"], 1),
        ("distinct_numeric", cg.DistinctFrom(cg.Func("numeric_eq", N(0), N(1))),
         ["numeric", "numeric"], 0),
        # (the inner product always has a finite operand: the reference
        # leaves the value of NaN * NaN unset, xpu_numeric.h:527-597, and a
        # later operator would read it)
        ("numeric_expr", cg.Func("numeric_add", cg.Func("numeric_mul", N(0),
                                                        cg.Func("int4_to_numeric", I(2))), N(1)),
         ["numeric", "numeric", "int4"], 0),
        ("numeric_to_int_guard", cg.And(B(0), cg.Func("int4gt", cg.Func("numeric_to_int4", N(1)),
                                                      C("int4", 0))), ["bool", "numeric"], 0),
    ]
    return out


TEMPLATE_VALUES = {
    "bool": [struct.pack("<?", False), struct.pack("<?", True), None],
    "int4": [struct.pack("<i", v) for v in (0, 1, -2, 7, 2**31 - 1)] + [None],
    "text": [_text(s) for s in ("", "abc", "abd")] + [None],
    "numeric": [_num("0"), _num("1.5"), _num("-2.25"), _num("NaN"),
                _num("3.14", short_varlena=True), _num("1e20"), None],
}


def func_arg_lists(kind, types):
    """The value list each argument's index refers to."""
    sets = func_arg_sets(kind, types)
    if sets is not None:
        return sets
    return [VALUES[t] for t in types]


def template_traps(name, imgs):
    """Inputs of a template the reference's host build cannot evaluate:
    HashedArrayOp with a non-bool result dereferences a NULL address where
    the result is NULL (no matching key and no ELSE, or a NULL THEN value;
    pgfn_HashedArrayOp, xpu_common.cu:1084, reads it with a heap reader that
    does not check for NULL, xpu_basetype.cu:150-159).  This build returns
    NULL there, as PostgreSQL does."""
    return False
