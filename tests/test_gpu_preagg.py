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
def test_c1_preagg(gpu, n, jit):
    cols = datagen.preagg_table(n)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    got = run_gpu_groupby(blob, [src], tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, [src])
    groups = assert_groups_equal(got, exp, C1_ACTIONS)
    assert sum(g[0] for g in groups.values()) == n


def test_c1_preagg_multi_chunk(gpu, jit):
    """kds_final accumulates over several XpuTaskExec chunks of one session."""
    chunks = [kds.build_arrow_kds(datagen.preagg_table(300_000, fragment=i)) for i in range(3)]
    cols = datagen.preagg_table(10)
    blob = W.preagg_query(cols)
    got = run_gpu_groupby(blob, chunks, tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, chunks)
    assert_groups_equal(got, exp, C1_ACTIONS)


@pytest.mark.parametrize("tuning", [dict(lds_nents=64), dict(lds_nents=512), dict(lds_nents=7),
                                    dict(rows_per_thread=1, block_size=64),
                                    dict(rows_per_thread=2, block_size=192),
                                    dict(rows_per_thread=4, block_size=512, blocks_per_cu=2,
                                         lds_nents=1500),
                                    dict(rows_per_thread=4, block_size=512),
                                    dict(rows_per_thread=4, block_size=1024),
                                    dict(rows_per_thread=2, block_size=1024, lds_nents=64)])
def test_c1_preagg_tunings(gpu, tuning, jit):
    """LDS table overflow (64 and 7 entries) and launch shapes (rows per lane,
    partial waves per block, 1024-thread blocks, which the interpreter runs
    as 512) give the same groups."""
    cols = datagen.preagg_table(400_000, fragment=7)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    got = run_gpu_groupby(blob, [src], tuning=dict(tuning, jit=jit))
    exp = run_oracle_groupby(blob, [src])
    assert_groups_equal(got, exp, C1_ACTIONS)


@pytest.mark.parametrize("m", [16, 1000, 100000])
def test_fragment_groupby_mod(gpu, m, jit):
    """Batched-fragment query: WHERE a < k GROUP BY key % m (params)."""
    r = datagen.rng(3, 5)
    n = 300_000
    cols = datagen.preagg_table(n, fragment=5)
    a = kds.Column("a", "int4", r.integers(0, 1_000_000, size=n, dtype=np.int32))
    cols = [a] + cols
    src = kds.build_arrow_kds(cols)
    blob = W.fragment_query(cols, k=400000, m=m)
    got = run_gpu_groupby(blob, [src], tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, [src])
    assert_groups_equal(got, exp, C1_ACTIONS)


def test_preagg_nogroup(gpu, jit):
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
    got = run_gpu_groupby(blob, [src], tuning=dict(jit=jit))
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


def test_preagg_null_arguments(gpu, jit):
    """Rows with a NULL aggregate argument bypass the LDS table (the row
    counter there stands for every count) and still update every state the
    reference updates: count(*) counts them, sum/avg/min/max skip them."""
    r = datagen.rng(4, 0)
    n = 400_000
    cols = datagen.preagg_table(n, fragment=11)
    key, v, w, x = cols
    cols = [key,
            kds.Column("v", "int8", v.values, nulls=r.random(n) < 0.05),
            kds.Column("w", "float8", w.values, nulls=r.random(n) < 0.05),
            kds.Column("x", "int4", x.values, nulls=r.random(n) < 0.3)]
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    got = run_gpu_groupby(blob, [src], tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, [src])
    groups = assert_groups_equal(got, exp, C1_ACTIONS)
    assert sum(g[0] for g in groups.values()) == n


def test_preagg_null_keys(gpu, jit):
    """NULL grouping keys form their own group (packed key NULL bit)."""
    r = datagen.rng(4, 1)
    n = 200_000
    cols = datagen.preagg_table(n, fragment=12)
    cols[0] = kds.Column("key", "int4", cols[0].values, nulls=r.random(n) < 0.1)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    got = run_gpu_groupby(blob, [src], tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, [src])
    groups = assert_groups_equal(got, exp, C1_ACTIONS)
    assert (None,) in groups


def test_preagg_rejected_rows_raise_nothing(gpu, jit):
    """WHERE x <> 0 GROUP BY key: sum(v / x).  Rows with x = 0 fail the scan
    quals, so their division by zero is never evaluated (the scan quals and
    the aggregate arguments run in one pass over the row)."""
    n = 300_000
    cols = datagen.preagg_table(n, fragment=13)
    key, v, w, x = cols
    xv = x.values.copy()
    xv[::7] = 0
    cols = [key, v, w, kds.Column("x", "int4", xv)]
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    k = plan.scan_var("key")
    vv = plan.scan_var("v")
    xx = plan.scan_var("x")
    plan.scan_quals = cg.Func("int4ne", xx, cg.Const("int4", 0))
    q = cg.Func("int84div", vv, xx)
    plan.set_groupby([(abi.KAGG_ACTION__NROWS_ANY, None, ("n", "int8")),
                      (abi.KAGG_ACTION__PSUM_INT64, q, ("s", "bytea")),
                      (abi.KAGG_ACTION__VREF, k, ("key", "int4"))])
    blob = plan.build(final_nrows=10000)
    src = kds.build_arrow_kds(cols)
    acts = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF]
    got = run_gpu_groupby(blob, [src], tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, [src])
    groups = assert_groups_equal(got, exp, acts)
    assert sum(g[0] for g in groups.values()) == n - len(xv[::7])


def test_preagg_error_in_argument(gpu, jit):
    """A division by zero in an aggregate argument of a qualifying row is
    reported with the reference's message."""
    from pgstrom import xpuhip
    n = 100_000
    cols = datagen.preagg_table(n, fragment=14)
    key, v, w, x = cols
    xv = x.values.copy()
    xv[12345] = 0
    cols = [key, v, w, kds.Column("x", "int4", xv)]
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    k = plan.scan_var("key")
    vv = plan.scan_var("v")
    xx = plan.scan_var("x")
    q = cg.Func("int84div", vv, xx)
    plan.set_groupby([(abi.KAGG_ACTION__PSUM_INT64, q, ("s", "bytea")),
                      (abi.KAGG_ACTION__VREF, k, ("key", "int4"))])
    blob = plan.build(final_nrows=10000)
    s = xpuhip.Session(blob)
    s.set_tuning(jit=jit)
    with pytest.raises(xpuhip.XpuError) as ei:
        s.exec(kds.build_arrow_kds(cols))
    assert b"division by zero" in ei.value.errbuf.message
    s.close()


def test_spill_overflow(gpu, jit):
    """A 7-entry LDS table sends almost every row to the spill buffer, which
    holds a quarter of the rows: the rest merge into kds_final directly.
    Both paths together give the oracle's groups."""
    from pgstrom import xpuhip
    n = 5_000_000
    cols = datagen.preagg_table(n, fragment=15)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    s = xpuhip.Session(blob)
    s.set_tuning(lds_nents=7, jit=jit)
    rc, _, _, _ = s.exec(src)
    assert rc == 0
    st = xpuhip.last_task_stats()
    assert st["spill_rows"] > 0 and st["lds_miss"] > 0
    got = s.final()
    s.close()
    exp = run_oracle_groupby(blob, [src])
    groups = assert_groups_equal(got, exp, C1_ACTIONS)
    assert sum(g[0] for g in groups.values()) == n


def test_session_autotune(gpu):
    """Autotuning on the caller's chunk resets kds_final; the tuned session
    gives the oracle's groups."""
    from pgstrom import xpuhip
    n = 1_000_000
    cols = datagen.preagg_table(n, fragment=16)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    d = xpuhip.DeviceBuffer(len(src), src)
    s = xpuhip.Session(blob)
    best = s.autotune(d, n)
    assert best["rows_per_thread"] in (1, 2, 4)
    rc, _, _ = s.exec_device(d, n)
    assert rc == 0
    got = s.final()
    s.close()
    d.free()
    exp = run_oracle_groupby(blob, [src])
    assert_groups_equal(got, exp, C1_ACTIONS)


@pytest.mark.parametrize("lds_nents", [0, 64])
def test_preagg_int128_sums(gpu, jit, lds_nents):
    """sum(int8) over full-range values: group sums leave the int64 range, so
    the 128-bit state (LDS halves, spill records, partial-state dumps and the
    kds_final carry) must be exact; a 64-entry LDS table forces spills."""
    r = datagen.rng(9, 1)
    n = 400_000
    key = r.integers(0, 300, size=n, dtype=np.int32)
    v = r.integers(-2**63, 2**63 - 1, size=n, dtype=np.int64, endpoint=True)
    v[::7] = 2**63 - 1
    v[3::11] = -2**63
    cols = [kds.Column("key", "int4", key), kds.Column("v", "int8", v)]
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    kv = plan.scan_var("key")
    plan.set_groupby([(abi.KAGG_ACTION__NROWS_ANY, None, ("n", "int8")),
                      (abi.KAGG_ACTION__PSUM_INT64, plan.scan_var("v"), ("s", "bytea")),
                      (abi.KAGG_ACTION__VREF, kv, ("key", "int4"))])
    blob = plan.build(final_nrows=1000)
    src = kds.build_arrow_kds(cols)
    tuning = dict(jit=jit)
    if lds_nents:
        tuning["lds_nents"] = lds_nents
    got = run_gpu_groupby(blob, [src, src], tuning=tuning)
    exp = run_oracle_groupby(blob, [src, src])
    acts = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF]
    groups = assert_groups_equal(got, exp, acts)
    vs = v.astype(object)
    for k in (0, 17, 299):
        sel = key == k
        assert groups[(k,)][1] == (2 * int(sel.sum()), 2 * int(vs[sel].sum()))
    assert any(abs(g[1][1]) > 2**63 for g in groups.values())
