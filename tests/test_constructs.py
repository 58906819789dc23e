"""Expression constructs where the reference's device code departs from
PostgreSQL, pinned here by PostgreSQL's definitions on the oracle (the
golden vectors of the reference pin everything else):

  - scalar op ALL (array): false at the first false comparison, else NULL
    if a comparison was NULL, else true (ExecEvalScalarArrayOp); the
    reference's pgfn_ScalarArrayOp returns true where a comparison is false
    (xpu_common.cu:848-880 overwrites the false result after its loop);
  - scalar = ANY (hashed constant list) with a NULL element: NULL when no
    element matches (ExecEvalHashedScalarArrayOp's has_nulls); the
    reference's pgfn_HashedArrayOp gives false;
  - x % -1 = 0 for the minimum integer (PostgreSQL's int4mod), which the
    reference's host build cannot evaluate (x86 traps)."""
import itertools
import os
import struct
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from pgstrom import codegen as cg, opcodes  # noqa: E402
import oracle  # noqa: E402

I = lambda i: cg.Var(i, "int4")
P4 = lambda v: None if v is None else struct.pack("<i", v)


def _blob(e):
    b = bytearray()
    cg.emit_expr(b, e)
    return bytes(b)


def _bool(img):
    return None if img is None else bool(img[0])


def and3(vals):
    if any(v is False for v in vals):
        return False
    if any(v is None for v in vals):
        return None
    return True


def or3(vals):
    if any(v is True for v in vals):
        return True
    if any(v is None for v in vals):
        return None
    return False


OPS = {"int4lt": lambda x, y: x < y, "int4ne": lambda x, y: x != y, "int4eq": lambda x, y: x == y,
       "int4ge": lambda x, y: x >= y}
ARRAYS = [[1, 3, 7], [7, None, 1], [None], [], [5, 5, None, -2]]
SCALARS = [-3, 1, 3, 5, 7, 100, None]


@pytest.mark.parametrize("use_any", [False, True])
@pytest.mark.parametrize("fname", sorted(OPS))
def test_oracle_saop_is_postgresql(use_any, fname):
    for arr in ARRAYS:
        blob = _blob(cg.ScalarArrayOp(use_any, fname, I(0), 1, "int4", arr))
        for x in SCALARS:
            got = _bool(oracle.eval_expr(blob, [opcodes.TYPES["int4"]] * 2, [P4(x), None],
                                         opcodes.TYPES["bool"]))
            cmps = [None if (x is None or e is None) else OPS[fname](x, e) for e in arr]
            exp = (or3 if use_any else and3)(cmps)
            assert got == exp, (fname, use_any, arr, x)


def test_oracle_hashed_in_list_with_null():
    items = [(v, True) for v in (1, 7, 9, 11)] + [(None, True)]
    blob = _blob(cg.HashedArrayOp("bool", "int4", I(0), items, in_list=True))
    for x, exp in ((7, True), (2, None), (None, None), (11, True)):
        got = _bool(oracle.eval_expr(blob, [opcodes.TYPES["int4"]], [P4(x)], opcodes.TYPES["bool"]))
        assert got == exp, x
    # without the NULL element a miss is false
    blob = _blob(cg.HashedArrayOp("bool", "int4", I(0), items[:-1], in_list=True))
    assert _bool(oracle.eval_expr(blob, [opcodes.TYPES["int4"]], [P4(2)],
                                  opcodes.TYPES["bool"])) is False


def test_oracle_case_map_null_then_and_no_else():
    """CASE x WHEN 5 THEN NULL ... and a CASE map without ELSE give NULL
    (the reference's device code reads a NULL address there)."""
    blob = _blob(cg.HashedArrayOp("int4", "int4", I(0), [(1, 10), (3, 30), (5, None), (7, 70)]))
    t = opcodes.TYPES["int4"]
    for x, exp in ((3, P4(30)), (5, None), (4, None), (None, None)):
        assert oracle.eval_expr(blob, [t], [P4(x)], t) == exp, x
