"""Record batches of an Arrow IPC file, produced by the library's chunk
producer, through the GPU path against the oracle: a scan with projection
of every column type, and a group-by summing the Decimal128 column (both
kernels)."""
import numpy as np
import pytest

from pgstrom import abi, codegen as cg, kds, xpuhip
from helpers import assert_groups_equal, run_gpu_projection, run_oracle_groupby, \
    run_oracle_projection
import arrowfiles as AF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def arrow_file(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("arrow") / "t.arrow")
    batches = AF.write_file(path, sizes=(1000, 0, 300000))
    return xpuhip.ArrowFile(open(path, "rb").read()), batches


def test_scan_projection(gpu, jit, arrow_file):
    f, batches = arrow_file
    plan = cg.Plan(AF.COLDEFS)
    v = {n: plan.scan_var(n) for n, _ in AF.COLDEFS}
    plan.scan_quals = cg.Func("int4gt", v["a"], cg.Const("int4", 0))
    plan.set_projection([v[n] for n, _ in AF.COLDEFS], AF.COLDEFS)
    blob = plan.build()
    for i in range(len(batches)):
        chunk = f.batch(i, AF.COLDEFS)
        got = run_gpu_projection(blob, chunk, 64 << 20, tuning=dict(jit=jit))
        exp, _ = run_oracle_projection(blob, chunk, 64 << 20)
        assert len(got) == 1 and np.array_equal(got[0], exp)


def test_decimal_sum(gpu, jit, arrow_file):
    f, batches = arrow_file
    plan = cg.Plan(AF.COLDEFS)
    s, d = plan.scan_var("s"), plan.scan_var("d")
    key = cg.Func("int2mod", s, cg.Const("int2", 97))
    acts = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_NUMERIC, abi.KAGG_ACTION__VREF]
    plan.set_groupby([(acts[0], None, ("n", "int8")),
                      (acts[1], d, ("sd", "bytea", ((20 << 16) | 3) + 4)),
                      (acts[2], key, ("k", "int2"))])
    blob = plan.build(final_nrows=200)
    names = [n for n, _ in AF.COLDEFS]
    chunks = [f.batch(i, AF.COLDEFS, [n if n in ("s", "d") else None for n in names])
              for i in range(len(batches))]
    sess = xpuhip.Session(blob)
    sess.set_tuning(jit=jit)
    for c in chunks:
        assert sess.exec(c)[0] == 0
    got = sess.final()
    sess.close()
    exp = run_oracle_groupby(blob, chunks)
    groups = assert_groups_equal(got, exp, acts)
    total = sum(int(x.scaleb(3)) for rows in batches for x in rows["d"] if x is not None)
    assert sum(st[1][2] for st in groups.values()) == total
