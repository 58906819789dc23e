"""The GPU's device functions against the reference's own: every vector of
tests/golden/expr_vectors.json (computed by the reference's pgfn_* code,
gen_expr_vectors.py) replayed on the GPU through the C ABI.  Per function,
one scan + projection session projects the function of the chunk's columns:
the rows whose reference result is a value or NULL run as one chunk and
every projected datum must equal the reference's image; each row on which
the reference raised an error runs alone and must fail with the reference's
error code and message.  The composite constructs (three-valued AND / OR /
NOT, NullTest, BoolTest, COALESCE, HashValue, short-circuit guards) run the
same way over their slot values."""
import json
import os
import struct
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from pgstrom import abi, codegen as cg, kds, opcodes, xpuhip  # noqa: E402
from pgstrom.kds import Column  # noqa: E402
import exprcases as X  # noqa: E402

pytestmark = pytest.mark.gpu

with open(os.path.join(HERE, "golden", "expr_vectors.json")) as _f:
    GOLDEN = json.load(_f)

FMT = {"bool": "<?", "int1": "<b", "int2": "<h", "int4": "<i", "int8": "<q",
       "float4": "<f", "float8": "<d"}


def _value(t, img):
    if t == "numeric":
        return bytes(img)
    if t in ("text", "bytea", "bpchar"):
        if img[0] & 1:
            return bytes(img[1:])
        return bytes(img[4:(struct.unpack_from("<I", img)[0] >> 2)])
    return struct.unpack(FMT[t], img)[0]


def _column(name, t, imgs):
    nulls = np.array([i is None for i in imgs], dtype=bool)
    if t == "numeric":
        vals = [kds.numeric_image("0") if i is None else bytes(i) for i in imgs]
    elif t in ("text", "bytea", "bpchar"):
        vals = [b"" if i is None else _value(t, i) for i in imgs]
    else:
        vals = [0 if i is None else _value(t, i) for i in imgs]
    return Column(name, t, vals, nulls)


def _plan_blob(expr, slot_types):
    """A scan + projection of expr over columns c0.. of slot_types; slots
    beyond the columns the expression reads (CASE keys, array elements)
    become temporaries."""
    plan = cg.Plan([("c%d" % i, t) for i, t in enumerate(slot_types)])
    used = _slots_read(expr)
    vars_ = {}
    for i, t in enumerate(slot_types):
        if i in used:
            vars_[i] = plan.scan_var("c%d" % i)
    temps = _slots_written(expr)
    for i in sorted(temps):
        while len(plan.slots) <= i:
            plan.temp_slot(temps[i], 0)
    e = _rebind(expr, vars_)
    plan.set_projection([e], [("r", e.typname)])
    return plan.build()


def _session(expr, slot_types, jit=None):
    s = xpuhip.Session(_plan_blob(expr, slot_types))
    if jit is not None:
        s.set_tuning(jit=jit)
    return s


def _walk(e, fn):
    fn(e)
    for attr in ("args", "whens"):
        for a in getattr(e, attr, []) or []:
            for x in (a if isinstance(a, tuple) else (a,)):
                _walk(x, fn)
    for attr in ("arg", "cmp", "scalar", "default"):
        x = getattr(e, attr, None)
        if isinstance(x, cg.Expr):
            _walk(x, fn)
    if getattr(e, "key", None):
        _walk(e.key[1], fn)


def _slots_written(expr):
    """Slots an expression defines itself: CASE keys and array elements."""
    out = {}

    def fn(e):
        if isinstance(e, cg.CaseWhen) and e.key:
            out[e.key[0]] = e.key[1].typname
        if isinstance(e, cg.ScalarArrayOp):
            out[e.elem_slot] = e.elem_typname
    _walk(expr, fn)
    return out


def _slots_read(expr):
    out = set()
    _walk(expr, lambda e: out.add(e.slot) if isinstance(e, cg.Var) else None)
    return out - set(_slots_written(expr))


def _rebind(e, vars_):
    """The expression with its column-slot Vars replaced by the plan's
    column Vars (temporary slots keep their numbers)."""
    if isinstance(e, cg.Var):
        return vars_.get(e.slot, e)
    if hasattr(e, "args"):
        e.args = [_rebind(a, vars_) for a in e.args]
    if hasattr(e, "whens"):
        e.whens = [(_rebind(c, vars_), _rebind(r, vars_)) for c, r in e.whens]
    for attr in ("arg", "cmp", "scalar", "default"):
        x = getattr(e, attr, None)
        if isinstance(x, cg.Expr):
            setattr(e, attr, _rebind(x, vars_))
    if getattr(e, "key", None):
        e.key = (e.key[0], _rebind(e.key[1], vars_))
    return e


def _chunk(cols):
    """Arrow columns; a chunk with numeric columns is heap pages instead,
    which keep each datum's varlena image as the test gives it (an Arrow
    numeric is a Decimal128)."""
    if any(c.typname == "numeric" for c in cols):
        return kds.build_block_kds(cols)
    return kds.build_arrow_kds(cols)


def _run(sess, cols, n):
    src = _chunk(cols)
    rc, dst, res, _ = sess.exec(src, dst_length=max(1 << 16, 256 * n + (1 << 14)))
    assert rc == 0
    return [r[0][0] for r in kds.parse_row_kds(dst)]


def _image(t, v):
    if v is None:
        return None
    if t == "numeric":
        return X.numeric_canon(v)
    if t in ("text", "bytea", "bpchar"):
        return v
    return struct.pack(FMT[t], v)


def _expected(t, out):
    if out is None:
        return None
    img = bytes.fromhex(out)
    if t == "numeric":
        return X.numeric_canon(img)
    return _value(t, img) if t in ("text", "bytea", "bpchar") else img


def _replay(expr_fn, slot_types, rtype, cases, pick, jit=None):
    """cases: [idx..., out]; pick(argument number, index) -> datum image or None."""
    ok = [c for c in cases if not isinstance(c[-1], list)]
    err = [c for c in cases if isinstance(c[-1], list)]
    bad = []
    if ok:
        cols = [_column("c%d" % i, t, [pick(i, c[i]) for c in ok])
                for i, t in enumerate(slot_types)]
        s = _session(expr_fn(), slot_types, jit)
        got = _run(s, cols, len(ok))
        s.close()
        assert len(got) == len(ok)
        for c, g in zip(ok, got):
            if _image(rtype, g) != _expected(rtype, c[-1]):
                bad.append((c[:-1], g, c[-1]))
    s = _session(expr_fn(), slot_types, jit) if err else None
    for c in err:
        cols = [_column("c%d" % i, t, [pick(i, c[i])]) for i, t in enumerate(slot_types)]
        src = _chunk(cols)
        with pytest.raises(xpuhip.XpuError) as e:
            s.exec(src, dst_length=1 << 16)
        got = (e.value.errbuf.errcode, e.value.errbuf.message.decode())
        if got != (c[-1][1], c[-1][2]):
            bad.append((c[:-1], got, c[-1]))
    if s:
        s.close()
    return bad


def _check_function(name, jit=None):
    g = GOLDEN["funcs"][name]
    args = g["args"]
    expr_fn = lambda: cg.Func(name, *[cg.Var(i, t) for i, t in enumerate(args)])
    lists = X.func_arg_lists(opcodes.FUNCS[name][1], list(args))
    bad = _replay(expr_fn, args, g["result"], g["cases"],
                  lambda j, i: None if i is None else lists[j][i], jit)
    assert not bad, "%d of %d cases differ, first: %r" % (len(bad), len(g["cases"]), bad[:3])


@pytest.mark.parametrize("name", sorted(GOLDEN["funcs"]))
def test_gpu_function_matches_reference(gpu, name):
    """Every function, on the interpreter (small chunks run it)."""
    _check_function(name)


# functions with data-dependent control flow over texts and numerics, and
# the float operators with their NaN and overflow paths
JIT_TYPES = ("text", "bpchar", "float4", "float8", "numeric")
JIT_FUNCS = sorted(n for n, f in GOLDEN["funcs"].items()
                   if any(a in JIT_TYPES for a in f["args"]) or f["result"] in JIT_TYPES)


@pytest.mark.parametrize("name", JIT_FUNCS)
def test_gpu_function_matches_reference_specialised(gpu, name):
    """The same on the kernel specialised for the plan (hiprtc), where the
    compiler sees constant operands: text, numeric and float functions."""
    _check_function(name, xpuhip.JIT_ON)


@pytest.mark.parametrize("name", sorted(GOLDEN["templates"]))
def test_gpu_construct_matches_reference(gpu, jit, name):
    g = GOLDEN["templates"][name]
    if g["compare_nulls"]:
        pytest.skip("kmode_compare_nulls applies to group-key comparison only, "
                    "covered by the group-by tests")
    expr_fn = lambda: {n: e for n, e, _, _ in X.templates(cg)}[name]
    slots = g["slots"]
    bad = _replay(expr_fn, slots, g["result"], g["cases"],
                  lambda j, i: X.TEMPLATE_VALUES[slots[j]][i], jit)
    assert not bad, "%d of %d cases differ, first: %r" % (len(bad), len(g["cases"]), bad[:3])
