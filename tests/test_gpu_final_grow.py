"""Group-by kds_final growth: a session whose kds_final has room for far
fewer groups than the data holds completes every task, as the reference's
service grows the buffer and resumes the kernel (gpu_service.c:3868-3884,
__expandGpuQueryGroupByBuffer :2418-2496).  Groups that find no room go to
the session's overflow table during the launch; kds_final then grows and
takes them.  The groups equal the oracle's, whose kds_final is large enough
from the start."""
import numpy as np
import pytest

from pgstrom import abi, codegen as cg, datagen, kds, workloads as W, xpuhip
from helpers import assert_groups_equal, run_oracle_groupby

pytestmark = pytest.mark.gpu

C1_ACTIONS = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__PAVG_FP,
              abi.KAGG_ACTION__PMIN_INT32, abi.KAGG_ACTION__PMAX_INT32, abi.KAGG_ACTION__VREF]
# room for the head, 20,200 hash slots and about 7,000 group tuples
SMALL_FINAL = 2 << 20


def _table(n, ngroups, seed, key_type="int4", null_frac=0.0):
    r = np.random.default_rng(seed)
    keys = r.integers(0, ngroups, size=n)
    if key_type == "text":
        key = kds.Column("key", "text", [b"group-%07d" % k for k in keys])
    else:
        key = kds.Column("key", key_type, keys.astype(kds.NP_DTYPES[key_type]))
    v = r.integers(0, 1_000_000, size=n, dtype=np.int64)
    w = r.random(n) * 1000.0
    x = r.integers(-2**31, 2**31 - 1, size=n, dtype=np.int32, endpoint=True)
    vn = (r.random(n) < null_frac) if null_frac else None
    return [key, kds.Column("v", "int8", v, nulls=vn), kds.Column("w", "float8", w),
            kds.Column("x", "int4", x)]


def _query(cols, final_length):
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    key = plan.scan_var("key")
    plan.set_groupby(W.preagg_tlist(plan, key, plan.scan_var("v"), plan.scan_var("w"),
                                    plan.scan_var("x")))
    return plan.build(final_nrows=100, final_length=final_length)


def _run(blob, chunks, tuning):
    s = xpuhip.Session(blob)
    s.set_tuning(**tuning)
    grows = 0
    for src in chunks:
        rc, _, _, _ = s.exec(src)
        assert rc == 0
        grows += xpuhip.last_task_stats()["final_grows"]
    fin = s.final()
    s.close()
    return fin, grows


@pytest.mark.parametrize("n,ngroups", [(400_000, 50_000), (1_200_000, 600_000)])
def test_final_grows_within_a_task(gpu, jit, n, ngroups):
    cols = _table(n, ngroups, seed=ngroups)
    src = kds.build_arrow_kds(cols)
    got, grows = _run(_query(cols, SMALL_FINAL), [src], dict(jit=jit))
    assert grows >= 1
    exp = run_oracle_groupby(_query(cols, 1 << 30), [src])
    groups = assert_groups_equal(got, exp, C1_ACTIONS)
    assert sum(g[0] for g in groups.values()) == n


def test_final_grows_over_chunks(gpu, jit):
    """Each chunk brings new groups: kds_final grows during the first task
    (by at least 1 GiB, as the reference grows it) and keeps the groups of
    the earlier tasks while the later ones add theirs."""
    chunks = []
    for i in range(4):
        cols = _table(150_000, 20_000, seed=100 + i)
        cols[0] = kds.Column("key", "int4", cols[0].values + np.int32(15_000 * i))
        chunks.append(kds.build_arrow_kds(cols))
    got, grows = _run(_query(cols, SMALL_FINAL), chunks, dict(jit=jit))
    assert grows >= 1
    exp = run_oracle_groupby(_query(cols, 1 << 30), chunks)
    assert_groups_equal(got, exp, C1_ACTIONS)


def test_final_grows_with_null_arguments(gpu, jit):
    """Rows with a NULL argument reach kds_final row by row (not through the
    LDS table); their groups overflow as well."""
    cols = _table(300_000, 40_000, seed=7, null_frac=0.3)
    src = kds.build_arrow_kds(cols)
    got, grows = _run(_query(cols, SMALL_FINAL), [src], dict(jit=jit))
    assert grows >= 1
    exp = run_oracle_groupby(_query(cols, 1 << 30), [src])
    assert_groups_equal(got, exp, C1_ACTIONS)


def test_final_grows_text_keys(gpu, jit):
    """Text grouping keys: every row takes the find-or-insert in kds_final;
    overflow groups carry their key bytes into the merge."""
    cols = _table(200_000, 30_000, seed=9, key_type="text")
    src = kds.build_arrow_kds(cols)
    got, grows = _run(_query(cols, SMALL_FINAL), [src], dict(jit=jit))
    assert grows >= 1
    exp = run_oracle_groupby(_query(cols, 1 << 30), [src])
    assert_groups_equal(got, exp, C1_ACTIONS)


def test_final_fits_does_not_grow(gpu):
    cols = _table(200_000, 1_000, seed=3)
    src = kds.build_arrow_kds(cols)
    got, grows = _run(_query(cols, SMALL_FINAL), [src], dict(jit=xpuhip.JIT_ON))
    assert grows == 0
    exp = run_oracle_groupby(_query(cols, SMALL_FINAL), [src])
    assert_groups_equal(got, exp, C1_ACTIONS)


@pytest.mark.parametrize("final_length", [SMALL_FINAL, 1 << 30])
def test_one_session_from_several_threads(gpu, final_length):
    """Four threads drive one group-by session at once, as the service's
    worker threads do (each call takes its own stream and scratch; launches
    share kds_final through device atomics and hold the session's
    final-buffer lock shared, growth holds it exclusively): the groups equal
    the oracle's over all chunks."""
    import threading
    chunks = []
    for i in range(8):
        cols = _table(120_000, 30_000, seed=500 + i)
        chunks.append(kds.build_arrow_kds(cols))
    blob = _query(cols, final_length)
    s = xpuhip.Session(blob)
    s.set_tuning(jit=xpuhip.JIT_ON)
    errors = []

    def work(mine):
        try:
            for src in mine:
                rc, _, _, _ = s.exec(src)
                assert rc == 0
        except Exception as e:      # surfaced after join
            errors.append(e)

    threads = [threading.Thread(target=work, args=(chunks[i::4],)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    got = s.final()
    s.close()
    exp = run_oracle_groupby(_query(cols, 1 << 30), chunks)
    groups = assert_groups_equal(got, exp, C1_ACTIONS)
    assert sum(g[0] for g in groups.values()) == 8 * 120_000
