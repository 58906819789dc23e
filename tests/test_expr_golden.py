"""The oracle's device functions against the reference's own: every vector of
tests/golden/expr_vectors.json (each device function this build executes, on
edge-case inputs, plus SQL three-valued logic, NullTest, BoolTest, COALESCE,
HashValue composition and short-circuit guards), evaluated by the reference's
pgfn_* code compiled from its sources (gen_expr_vectors.py), replayed through
the oracle: equal datum images, equal NULLs, equal error codes and messages."""
import json
import os
import struct
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from pgstrom import codegen as cg, opcodes  # noqa: E402
import exprcases as X  # noqa: E402
import oracle  # noqa: E402

with open(os.path.join(HERE, "golden", "expr_vectors.json")) as _f:
    GOLDEN = json.load(_f)


def _blob(expr):
    b = bytearray()
    cg.emit_expr(b, expr)
    return bytes(b)


def _norm(out):
    """Golden result as the oracle returns it."""
    if out is None:
        return None
    if isinstance(out, list):
        return ("E", out[1], out[2])
    return bytes.fromhex(out)


def _canon(typname, img):
    """Datum images compare by value: a text's header length form aside."""
    if typname == "numeric":
        return X.numeric_canon(img)
    if img is None or isinstance(img, tuple) or typname not in ("text", "bytea", "bpchar"):
        return img
    if img[0] & 1:
        return img[1:]
    return img[4:(struct.unpack_from("<I", img)[0] >> 2)]


def test_golden_covers_every_function():
    assert set(GOLDEN["funcs"]) == set(opcodes.FUNCS)


@pytest.mark.parametrize("name", sorted(GOLDEN["funcs"]))
def test_oracle_function_matches_reference(name):
    g = GOLDEN["funcs"][name]
    args, rtype = g["args"], g["result"]
    assert opcodes.FUNCS[name][0] == g["opcode"]
    blob = _blob(cg.Func(name, *[cg.Var(i, t) for i, t in enumerate(args)]))
    codes = [opcodes.TYPES[t] for t in args]
    lists = X.func_arg_lists(opcodes.FUNCS[name][1], list(args))
    bad = []
    for case in g["cases"]:
        idx, exp = case[:-1], _norm(case[-1])
        imgs = [None if i is None else lists[j][i] for j, i in enumerate(idx)]
        got = oracle.eval_expr(blob, codes, imgs, opcodes.TYPES[rtype], encoding=X.ENCODING)
        if _canon(rtype, got) != _canon(rtype, exp):
            bad.append((idx, got, exp))
    assert not bad, "%d of %d cases differ, first: %r" % (len(bad), len(g["cases"]), bad[:3])


@pytest.mark.parametrize("name", sorted(GOLDEN["templates"]))
def test_oracle_construct_matches_reference(name):
    g = GOLDEN["templates"][name]
    expr = {n: (e, s, c) for n, e, s, c in X.templates(cg)}[name][0]
    blob = _blob(expr)
    slots, rtype = g["slots"], g["result"]
    codes = [opcodes.TYPES[t] for t in slots]
    bad = []
    for case in g["cases"]:
        idx, exp = case[:-1], _norm(case[-1])
        imgs = [X.TEMPLATE_VALUES[t][i] for t, i in zip(slots, idx)]
        got = oracle.eval_expr(blob, codes, imgs, opcodes.TYPES[rtype], g["compare_nulls"],
                               encoding=X.ENCODING)
        if _canon(rtype, got) != _canon(rtype, exp):
            bad.append((idx, got, exp))
    assert not bad, "%d of %d cases differ, first: %r" % (len(bad), len(g["cases"]), bad[:3])


@pytest.mark.parametrize("t", ["int1", "int2", "int4", "int8"])
def test_oracle_mod_minus_one(t):
    """TYPEmod(MIN, -1): the one input the reference's host build cannot run
    (x % y traps on x86); PostgreSQL defines it as 0, and so does the oracle
    (x % -1 = 0 for every x)."""
    fmt = X._INT_FMT[t]
    bits = 8 * struct.calcsize(fmt)
    blob = _blob(cg.Func(t + "mod", cg.Var(0, t), cg.Var(1, t)))
    for x in (-(1 << (bits - 1)), -1, 0, 5, (1 << (bits - 1)) - 1):
        got = oracle.eval_expr(blob, [opcodes.TYPES[t]] * 2,
                               [struct.pack(fmt, x), struct.pack(fmt, -1)], opcodes.TYPES[t])
        assert got == struct.pack(fmt, 0)


@pytest.mark.parametrize("name", sorted(GOLDEN["funcs"]))
def test_library_compiles_every_golden_function(name):
    """The HIP library's session compiler accepts a projection of each
    function the vectors cover (the GPU replays them in
    test_gpu_expr_golden.py)."""
    from pgstrom import xpuhip
    args = GOLDEN["funcs"][name]["args"]
    plan = cg.Plan([("c%d" % i, t) for i, t in enumerate(args)])
    vs = [plan.scan_var("c%d" % i) for i in range(len(args))]
    e = cg.Func(name, *vs)
    plan.set_projection([e], [("r", e.typname)])
    rc, err = xpuhip.session_check(plan.build())
    assert rc == 0, err.describe()


@pytest.mark.parametrize("name", sorted(GOLDEN["templates"]))
def test_library_compiles_every_golden_construct(name):
    """The session compiler accepts a projection of each construct (CASE,
    DISTINCT FROM, LEAST/GREATEST, ANY/ALL, hashed IN lists ...)."""
    from pgstrom import xpuhip
    from test_gpu_expr_golden import _plan_blob
    g = GOLDEN["templates"][name]
    expr = {n: e for n, e, _, _ in X.templates(cg)}[name]
    rc, err = xpuhip.session_check(_plan_blob(expr, g["slots"]))
    assert rc == 0, err.describe()
