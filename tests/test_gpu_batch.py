"""Batched launches: many scan + group-by fragments of one plan shape (each
with its own chunk, parameter values and kds_final) in one persistent
launch give, fragment by fragment, the groups of running each alone."""
import numpy as np
import pytest

from pgstrom import abi, datagen, kds, workloads as W, xpuhip
from helpers import assert_groups_equal, run_oracle_groupby

pytestmark = pytest.mark.gpu

ACTIONS = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__PAVG_FP,
           abi.KAGG_ACTION__PMIN_INT32, abi.KAGG_ACTION__PMAX_INT32, abi.KAGG_ACTION__VREF]


def fragment(i, n):
    r = datagen.rng(5, i)
    cols = datagen.preagg_table(n, seed_workload=5, fragment=i)
    a = kds.Column("a", "int4", r.integers(0, 1_000_000, size=n, dtype=np.int32))
    k = int(r.integers(100_000, 900_001))
    m = int(round(np.exp(r.uniform(np.log(16), np.log(100_000)))))
    cols = [a] + cols
    return cols, W.fragment_query(cols, k=k, m=m), kds.build_arrow_kds(cols)


SIZES = [1, 100, 5000, 70_000, 300_000, 1_200_000, 64, 2_000_000, 40_000, 700_000]


@pytest.fixture(scope="module")
def frags(gpu):
    out = []
    for i, n in enumerate(SIZES):
        cols, blob, src = fragment(i, n)
        out.append((n, blob, src, xpuhip.DeviceBuffer(len(src), src)))
    yield out
    for _, _, _, d in out:
        d.free()


def run_batch(frags, tuning, graph, repeat=1):
    sessions = []
    for n, blob, src, d in frags:
        s = xpuhip.Session(blob)
        s.set_tuning(**tuning)
        sessions.append(s)
    tasks = [(s, d, n) for s, (n, _, _, d) in zip(sessions, frags)]
    res = None
    for _ in range(repeat):
        for s in sessions:
            s.reset()
        res = xpuhip.batch_exec(tasks, graph=graph)
    finals = [s.final() for s in sessions]
    for s in sessions:
        s.close()
    return finals, res


@pytest.mark.parametrize("graph", [False, True], ids=["stream", "graph"])
def test_batch_matches_oracle(frags, jit, graph):
    finals, res = run_batch(frags, dict(jit=jit), graph, repeat=2 if graph else 1)
    for (n, blob, src, _), fin, r in zip(frags, finals, res):
        exp = run_oracle_groupby(blob, [src])
        assert_groups_equal(fin, exp, ACTIONS)
        assert r.nitems_raw == n


def test_batch_small_table_spills(frags, jit):
    """A 64-entry LDS table: most rows of every fragment go through the
    per-fragment spill regions and the batched spill pass."""
    finals, _ = run_batch(frags, dict(jit=jit, lds_nents=64), True)
    st = xpuhip.last_task_stats()
    assert st["spill_rows"] > 0
    for (n, blob, src, _), fin in zip(frags, finals):
        exp = run_oracle_groupby(blob, [src])
        assert_groups_equal(fin, exp, ACTIONS)


def test_batch_rejects_mixed_shapes(frags):
    n, blob, src, d = frags[3]
    s1 = xpuhip.Session(blob)
    s2 = xpuhip.Session(W.preagg_query(datagen.preagg_table(10)))
    with pytest.raises(xpuhip.XpuError) as ei:
        xpuhip.batch_exec([(s1, d, n), (s2, d, n)])
    assert b"one plan shape" in ei.value.errbuf.message
    s1.close()
    s2.close()


def test_batch_autotune_then_exec(frags):
    """The autotuner tries launch shapes over the batch, leaves every
    kds_final empty, and the tuned batch gives the oracle's groups."""
    sessions = [xpuhip.Session(blob) for _, blob, _, _ in frags]
    tasks = [(s, d, n) for s, (n, _, _, d) in zip(sessions, frags)]
    best = xpuhip.batch_autotune(tasks)
    assert best["block_size"] in (256, 512)
    xpuhip.batch_exec(tasks, graph=True)
    for s, (n, blob, src, _) in zip(sessions, frags):
        exp = run_oracle_groupby(blob, [src])
        assert_groups_equal(s.final(), exp, ACTIONS)
        s.close()
