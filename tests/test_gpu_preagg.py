"""GpuPreAgg parity on the GPU: the group set and every integer state equal
the oracle's; float8 sums within 1e-12 relative."""
import numpy as np
import pytest

from pgstrom import abi, codegen as cg, datagen, kds, workloads as W
from helpers import assert_groups_equal, run_gpu_groupby, run_oracle_groupby

pytestmark = pytest.mark.gpu

C1_ACTIONS = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__PAVG_FP,
              abi.KAGG_ACTION__PMIN_INT32, abi.KAGG_ACTION__PMAX_INT32, abi.KAGG_ACTION__VREF]


@pytest.mark.parametrize("n", [1, 1000, 1_000_000])
def test_c1_preagg(gpu, n):
    cols = datagen.preagg_table(n)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    got = run_gpu_groupby(blob, [src])
    exp = run_oracle_groupby(blob, [src])
    groups = assert_groups_equal(got, exp, C1_ACTIONS)
    assert sum(g[0] for g in groups.values()) == n


def test_c1_preagg_multi_chunk(gpu):
    """kds_final accumulates over several XpuTaskExec chunks of one session."""
    chunks = [kds.build_arrow_kds(datagen.preagg_table(300_000, fragment=i)) for i in range(3)]
    cols = datagen.preagg_table(10)
    blob = W.preagg_query(cols)
    got = run_gpu_groupby(blob, chunks)
    exp = run_oracle_groupby(blob, chunks)
    assert_groups_equal(got, exp, C1_ACTIONS)


@pytest.mark.parametrize("tuning", [dict(lds_nents=64), dict(lds_nents=512),
                                    dict(rows_per_thread=1, block_size=64),
                                    dict(rows_per_thread=16, block_size=512)])
def test_c1_preagg_tunings(gpu, tuning):
    """LDS table overflow (64 entries) and launch shapes give the same groups."""
    cols = datagen.preagg_table(400_000, fragment=7)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    got = run_gpu_groupby(blob, [src], tuning=tuning)
    exp = run_oracle_groupby(blob, [src])
    assert_groups_equal(got, exp, C1_ACTIONS)


@pytest.mark.parametrize("m", [16, 1000, 100000])
def test_fragment_groupby_mod(gpu, m):
    """Batched-fragment query: WHERE a < k GROUP BY key % m (params)."""
    r = datagen.rng(3, 5)
    n = 300_000
    cols = datagen.preagg_table(n, fragment=5)
    a = kds.Column("a", "int4", r.integers(0, 1_000_000, size=n, dtype=np.int32))
    cols = [a] + cols
    src = kds.build_arrow_kds(cols)
    blob = W.fragment_query(cols, k=400000, m=m)
    got = run_gpu_groupby(blob, [src])
    exp = run_oracle_groupby(blob, [src])
    assert_groups_equal(got, exp, C1_ACTIONS)


def test_preagg_nogroup(gpu):
    """No GROUP BY: a single group tuple in a ROW-format kds_final."""
    cols = datagen.preagg_table(500_000)
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    v = plan.scan_var("v")
    w = plan.scan_var("w")
    x = plan.scan_var("x")
    plan.scan_quals = cg.Func("int4lt", plan.scan_var("key"), cg.Const("int4", 50))
    plan.set_groupby([(abi.KAGG_ACTION__NROWS_ANY, None, ("n", "int8")),
                      (abi.KAGG_ACTION__PSUM_INT64, v, ("s", "bytea")),
                      (abi.KAGG_ACTION__PSUM_FP, w, ("f", "bytea")),
                      (abi.KAGG_ACTION__PMIN_INT32, x, ("mn", "bytea")),
                      (abi.KAGG_ACTION__PMAX_INT32, x, ("mx", "bytea"))])
    blob = plan.build(final_length=4 << 20)
    src = kds.build_arrow_kds(cols)
    acts = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__PSUM_FP,
            abi.KAGG_ACTION__PMIN_INT32, abi.KAGG_ACTION__PMAX_INT32]
    got = run_gpu_groupby(blob, [src])
    exp = run_oracle_groupby(blob, [src])
    groups = assert_groups_equal(got, exp, acts)
    sel = cols[0].values < 50
    assert groups[()][0] == int(sel.sum())
    assert groups[()][1] == (int(sel.sum()), int(cols[1].values[sel].astype(object).sum()))


def test_lds_table_too_large_is_refused(gpu):
    """An LDS table larger than a block's LDS is refused before any launch."""
    from pgstrom import xpuhip
    cols = datagen.preagg_table(1000)
    s = xpuhip.Session(W.preagg_query(cols))
    s.set_tuning(lds_nents=1 << 16)
    with pytest.raises(xpuhip.XpuError) as ei:
        s.exec(kds.build_arrow_kds(cols))
    assert b"exceeds the LDS" in ei.value.errbuf.message
    s.close()
