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

from pgstrom import abi, codegen as cg, kds, xpuhip  # noqa: E402
from pgstrom.kds import Column  # noqa: E402
import exprcases as X  # noqa: E402

pytestmark = pytest.mark.gpu

with open(os.path.join(HERE, "golden", "expr_vectors.json")) as _f:
    GOLDEN = json.load(_f)

FMT = {"bool": "<?", "int1": "<b", "int2": "<h", "int4": "<i", "int8": "<q",
       "float4": "<f", "float8": "<d"}


def _value(t, img):
    if t in ("text", "bytea"):
        if img[0] & 1:
            return bytes(img[1:])
        return bytes(img[4:(struct.unpack_from("<I", img)[0] >> 2)])
    return struct.unpack(FMT[t], img)[0]


def _column(name, t, imgs):
    nulls = np.array([i is None for i in imgs], dtype=bool)
    if t in ("text", "bytea"):
        vals = [b"" if i is None else _value(t, i) for i in imgs]
    else:
        vals = [0 if i is None else _value(t, i) for i in imgs]
    return Column(name, t, vals, nulls)


def _session(expr, slot_types):
    plan = cg.Plan([("c%d" % i, t) for i, t in enumerate(slot_types)])
    vars_ = [plan.scan_var("c%d" % i) for i in range(len(slot_types))]
    e = _rebind(expr, vars_)
    plan.set_projection([e], [("r", e.typname)])
    return xpuhip.Session(plan.build())


def _rebind(e, vars_):
    """The expression with its slot Vars replaced by the plan's column Vars."""
    if isinstance(e, cg.Var):
        return vars_[e.slot]
    for attr in ("args",):
        if hasattr(e, attr):
            setattr(e, attr, [_rebind(a, vars_) for a in getattr(e, attr)])
    if hasattr(e, "arg"):
        e.arg = _rebind(e.arg, vars_)
    return e


def _run(sess, cols, n):
    src = kds.build_arrow_kds(cols)
    rc, dst, res, _ = sess.exec(src, dst_length=max(1 << 16, 256 * n + (1 << 14)))
    assert rc == 0
    return [r[0][0] for r in kds.parse_row_kds(dst)]


def _image(t, v):
    if v is None:
        return None
    if t in ("text", "bytea"):
        return v
    return struct.pack(FMT[t], v)


def _expected(t, out):
    if out is None:
        return None
    img = bytes.fromhex(out)
    return _value(t, img) if t in ("text", "bytea") else img


def _replay(expr_fn, slot_types, rtype, cases, pick):
    """cases: [idx..., out]; pick(type, index) -> datum image or None."""
    ok = [c for c in cases if not isinstance(c[-1], list)]
    err = [c for c in cases if isinstance(c[-1], list)]
    bad = []
    if ok:
        cols = [_column("c%d" % i, t, [pick(t, c[i]) for c in ok])
                for i, t in enumerate(slot_types)]
        s = _session(expr_fn(), slot_types)
        got = _run(s, cols, len(ok))
        s.close()
        assert len(got) == len(ok)
        for c, g in zip(ok, got):
            if _image(rtype, g) != _expected(rtype, c[-1]):
                bad.append((c[:-1], g, c[-1]))
    s = _session(expr_fn(), slot_types) if err else None
    for c in err:
        cols = [_column("c%d" % i, t, [pick(t, c[i])]) for i, t in enumerate(slot_types)]
        src = kds.build_arrow_kds(cols)
        with pytest.raises(xpuhip.XpuError) as e:
            s.exec(src, dst_length=1 << 16)
        got = (e.value.errbuf.errcode, e.value.errbuf.message.decode())
        if got != (c[-1][1], c[-1][2]):
            bad.append((c[:-1], got, c[-1]))
    if s:
        s.close()
    return bad


@pytest.mark.parametrize("name", sorted(GOLDEN["funcs"]))
def test_gpu_function_matches_reference(gpu, name):
    g = GOLDEN["funcs"][name]
    args = g["args"]
    expr_fn = lambda: cg.Func(name, *[cg.Var(i, t) for i, t in enumerate(args)])
    bad = _replay(expr_fn, args, g["result"], g["cases"],
                  lambda t, i: None if i is None else X.VALUES[t][i])
    assert not bad, "%d of %d cases differ, first: %r" % (len(bad), len(g["cases"]), bad[:3])


@pytest.mark.parametrize("name", sorted(GOLDEN["templates"]))
def test_gpu_construct_matches_reference(gpu, name):
    g = GOLDEN["templates"][name]
    if g["compare_nulls"]:
        pytest.skip("kmode_compare_nulls applies to group-key comparison only, "
                    "covered by the group-by tests")
    expr_fn = lambda: {n: e for n, e, _, _ in X.templates(cg)}[name]
    bad = _replay(expr_fn, g["slots"], g["result"], g["cases"],
                  lambda t, i: X.TEMPLATE_VALUES[t][i])
    assert not bad, "%d of %d cases differ, first: %r" % (len(bad), len(g["cases"]), bad[:3])
