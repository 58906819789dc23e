"""CPU fallback on the GPU against the oracle: rows whose depth-0 work needs
the bytes of a compressed or external text go to the session's
KDS_FORMAT_FALLBACK buffer (HandleErrorIfCpuFallback, xpu_common.cu:
1898-1955), the other rows give the oracle's kds_dst bytes or groups, and
the fallback items equal the oracle's (and the closed form of
fallbackcases.py) up to their order, which device atomics decide."""
import numpy as np
import pytest

from pgstrom import abi, codegen as cg, hashing, kds, xpuhip
from helpers import assert_groups_equal, run_oracle_groupby
import fallbackcases as F
import oracle

pytestmark = pytest.mark.gpu


def _items(fb):
    return [] if fb is None else sorted(it[:4] for it in kds.parse_fallback_kds(fb))


def _gpu_projection(blob, src, dst_length, tuning):
    s = xpuhip.Session(blob)
    s.set_tuning(**tuning)
    outs, pos = [], 0
    for _ in range(1000):
        rc, dst, res, pos = s.exec(src, dst_length=dst_length, resume_pos=pos)
        outs.append(dst)
        if rc == 0:
            break
        assert rc == abi.ERRCODE_SUSPEND_NO_SPACE and pos > 0
    fb = s.fallback()
    s.close()
    return outs, fb


def _oracle_projection(blob, src):
    o = oracle.OracleSession(blob)
    rc, dst, _ = o.exec(src, dst_length=8 << 20, nthreads=2)
    assert rc == 0, o.err.describe()
    fb = o.fallback()
    o.close()
    return dst, fb


@pytest.mark.parametrize("query", ["textlen", "texteq"])
def test_projection_fallback(gpu, jit, query):
    rs = F.rows(3000)
    cols, src = F.chunk(rs)
    blob = F.q_textlen() if query == "textlen" else F.q_texteq()
    got, gfb = _gpu_projection(blob, src, 8 << 20, dict(jit=jit))
    exp, efb = _oracle_projection(blob, src)
    assert len(got) == 1 and np.array_equal(got[0], exp)
    assert _items(gfb) == _items(efb) == sorted(F.expected_fallback(rs))


def test_projection_fallback_with_resumes(gpu, jit):
    """kds_dst fills several times ('d'): every fallback row is written once."""
    rs = F.rows(3000)
    cols, src = F.chunk(rs)
    blob = F.q_texteq()
    got, gfb = _gpu_projection(blob, src, 24 << 10, dict(jit=jit))
    assert len(got) > 2
    rows = [r for d in got for r in kds.parse_row_kds(d)]
    exp, efb = _oracle_projection(blob, src)
    assert [v for v, _ in rows] == [v for v, _ in kds.parse_row_kds(exp)]
    assert _items(gfb) == _items(efb)


def test_groupby_fallback(gpu, jit):
    rs = F.rows(3000)
    cols, src = F.chunk(rs)
    blob = F.q_groupby_g()
    s = xpuhip.Session(blob)
    s.set_tuning(jit=jit)
    for _ in range(2):
        rc, _, _, _ = s.exec(src)
        assert rc == 0
    gfb = s.fallback()
    got = s.final()
    s.close()
    o = oracle.OracleSession(blob)
    for _ in range(2):
        assert o.exec(src, nthreads=2)[0] == 0
    efb = o.fallback()
    exp = o.final()
    o.close()
    assert_groups_equal(got, exp, [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64,
                                   abi.KAGG_ACTION__VREF])
    assert _items(gfb) == _items(efb) == sorted(F.expected_fallback(rs) * 2)


def test_join_scan_quals_fallback(gpu, jit):
    """A join whose scan quals need the text: the fact rows that fall back
    carry their depth-0 variables (the inner attributes of the scan slot are
    NULL), in the screen pass of the multi-pass join and in the pipeline."""
    rs = F.rows(4000)
    cols, src = F.chunk(rs)
    plan = cg.Plan(F.COLDEFS)
    v = {n: plan.scan_var(n) for n, _ in F.COLDEFS}
    d = plan.add_inner([("pk", "int4"), ("cat", "int2")])
    pk = plan.inner_var(d, "pk")
    cat = plan.inner_var(d, "cat")
    plan.add_join(d, [v["a"]], [cg.Func("int4eq", v["a"], pk)])
    plan.scan_quals = cg.Func("text_ne", v["g"], cg.Const("text", b"zz"))
    plan.set_groupby([(abi.KAGG_ACTION__NROWS_ANY, None, ("n", "int8")),
                      (abi.KAGG_ACTION__PSUM_INT64, v["c"], ("s", "bytea")),
                      (abi.KAGG_ACTION__VREF, cat, ("cat", "int2"))])
    blob = plan.build(final_nrows=100, fallback=True)
    keys = np.arange(1000, dtype=np.int32)
    hashes = [int(h) for h in hashing.hash_keys_np([(keys, 4)])]
    inner = kds.build_hash_kds([("pk", "int4"), ("cat", "int2")],
                               [(int(k), int(k) % 7) for k in keys], hashes)
    km = kds.build_multirels([(inner, False)])
    s = xpuhip.Session(blob)
    s.set_tuning(jit=jit)
    s.attach_inner(km)
    assert s.exec(src)[0] == 0
    gfb = s.fallback()
    got = s.final()
    s.close()
    exp = run_oracle_groupby(blob, [src], kmrels=km)
    o = oracle.OracleSession(blob)
    o.exec(src, kmrels=km, nthreads=2)
    efb = o.fallback()
    o.close()
    assert_groups_equal(got, exp, [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64,
                                   abi.KAGG_ACTION__VREF])
    assert _items(gfb) == _items(efb)
    assert len(_items(gfb)) == len(F.expected_fallback(rs))


def test_fallback_disabled_is_an_error(gpu, jit):
    rs = F.rows(500)
    cols, src = F.chunk(rs)
    s = xpuhip.Session(F.q_textlen(fallback=False))
    s.set_tuning(jit=jit)
    with pytest.raises(xpuhip.XpuError) as ei:
        s.exec(src, dst_length=1 << 20)
    s.close()
    assert ei.value.code == abi.ERRCODE_SUSPEND_FALLBACK
