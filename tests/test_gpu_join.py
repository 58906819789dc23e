"""GpuJoin parity on the GPU: hash-join probe of host-built inner hash tables,
feeding a group-by or a projection."""
import numpy as np
import pytest

from pgstrom import abi, codegen as cg, datagen, hashing, kds, workloads as W, xpuhip
import oracle
from helpers import (skip_interpreter_join_groupby, assert_groups_equal, run_gpu_groupby, run_gpu_projection,
                     run_oracle_groupby, run_oracle_projection)

pytestmark = pytest.mark.gpu

C2_ACTIONS = [abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF, abi.KAGG_ACTION__VREF]


FILTERS = [xpuhip.JOIN_FILTER_AUTO, xpuhip.JOIN_FILTER_OFF]
FILTER_IDS = ["filter", "nofilter"]
PASSES = [xpuhip.JOIN_PASSES_ON, xpuhip.JOIN_PASSES_OFF]
PASSES_IDS = ["passes", "fused"]


@pytest.mark.parametrize("join_passes", PASSES, ids=PASSES_IDS)
@pytest.mark.parametrize("join_filter", FILTERS, ids=FILTER_IDS)
@pytest.mark.parametrize("left_outer", [False, True])
def test_c2_star_join_preagg(gpu, left_outer, jit, join_filter, join_passes):
    skip_interpreter_join_groupby(jit)
    fact, dims = datagen.star_tables(300_000, dim_sizes=(20_000, 2_000, 2_556))
    src = kds.build_arrow_kds(fact)
    km = W.star_inner(dims)
    blob = W.star_query(fact, left_outer=left_outer)
    got = run_gpu_groupby(blob, [src], kmrels=km,
                          tuning=dict(jit=jit, join_filter=join_filter, join_passes=join_passes))
    exp = run_oracle_groupby(blob, [src], kmrels=km)
    assert_groups_equal(got, exp, C2_ACTIONS)


@pytest.mark.parametrize("join_passes", PASSES, ids=PASSES_IDS)
@pytest.mark.parametrize("join_filter", FILTERS, ids=FILTER_IDS)
def test_c2_star_join_stats(gpu, jit, join_filter, join_passes):
    """Rows in, rows out of every depth and groups out: the same counts as
    the oracle's with and without the inner-relation filters (the filters
    skip bucket walks that cannot match and change nothing else)."""
    skip_interpreter_join_groupby(jit)
    fact, dims = datagen.star_tables(200_000, dim_sizes=(20_000, 2_000, 2_556))
    src = kds.build_arrow_kds(fact)
    km = W.star_inner(dims)
    blob = W.star_query(fact)
    s = xpuhip.Session(blob)
    s.set_tuning(jit=jit, join_filter=join_filter, join_passes=join_passes)
    s.attach_inner(km)
    rc, _, res, _ = s.exec(src)
    s.close()
    assert rc == 0
    o = oracle.OracleSession(blob)
    orc, _, ores = o.exec(src, kmrels=km, nthreads=4)
    o.close()
    assert orc == 0
    assert res.nitems_in == ores.nitems_in
    assert res.nitems_out == ores.nitems_out
    for d in range(3):
        assert res.stats[d].nitems_out == ores.stats[d].nitems_out, d
    assert 0 < res.stats[2].nitems_out < res.stats[0].nitems_out


@pytest.mark.parametrize("join_passes", PASSES, ids=PASSES_IDS)
@pytest.mark.parametrize("join_filter", FILTERS, ids=FILTER_IDS)
def test_join_groupby_multi_match(gpu, jit, join_filter, join_passes):
    """Group-by over a join whose inner relation holds duplicate keys, with
    scan quals rejecting a third of the rows: every (row, match) pair is
    aggregated once."""
    skip_interpreter_join_groupby(jit)
    r = datagen.rng(11)
    n = 150_000
    outer = [kds.Column("k", "int4", r.integers(0, 5000, size=n, dtype=np.int32)),
             kds.Column("v", "int8", r.integers(0, 10**9, size=n, dtype=np.int64))]
    keys = r.integers(0, 8000, size=3000, dtype=np.int32)
    rows = [(int(k), int(i % 7), bytes("t%d" % i, "ascii")) for i, k in enumerate(keys)]
    hashes = hashing.hash_keys_np([(keys, 4)])
    inner = kds.build_hash_kds([("k", "int4"), ("g", "int2"), ("t", "text")], rows,
                               [int(h) for h in hashes])
    km = kds.build_multirels([(inner, False)])
    plan = cg.Plan([("k", "int4"), ("v", "int8")])
    k = plan.scan_var("k")
    v = plan.scan_var("v")
    plan.scan_quals = cg.Func("int8ne", cg.Func("int8mod", v, plan.param("int8", 3)),
                              plan.param("int8", 0))
    d = plan.add_inner([("k", "int4"), ("g", "int2"), ("t", "text")])
    ik = plan.inner_var(d, "k")
    g = plan.inner_var(d, "g")
    plan.add_join(d, [k], [cg.Func("int4eq", k, ik)])
    plan.set_groupby([
        (abi.KAGG_ACTION__NROWS_ANY, None, ("nrows", "int8")),
        (abi.KAGG_ACTION__PSUM_INT64, v, ("psum_v", "bytea")),
        (abi.KAGG_ACTION__VREF, g, ("g", "int2")),
    ])
    blob = plan.build(final_nrows=16)
    src = kds.build_arrow_kds(outer)
    acts = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF]
    got = run_gpu_groupby(blob, [src], kmrels=km,
                          tuning=dict(jit=jit, join_filter=join_filter, join_passes=join_passes))
    exp = run_oracle_groupby(blob, [src], kmrels=km)
    groups = assert_groups_equal(got, exp, acts)
    assert len(groups) == 7


def test_join_projection_multi_match(gpu, jit):
    """Inner relation with duplicate keys: every match is emitted, in the
    oracle's (row-major) order."""
    r = datagen.rng(9)
    n = 20000
    outer = [kds.Column("k", "int4", r.integers(0, 500, size=n, dtype=np.int32)),
             kds.Column("v", "int8", r.integers(0, 10**9, size=n, dtype=np.int64))]
    keys = r.integers(0, 600, size=900, dtype=np.int32)
    rows = [(int(k), int(i), bytes("t%d" % i, "ascii")) for i, k in enumerate(keys)]
    hashes = hashing.hash_keys_np([(keys, 4)])
    inner = kds.build_hash_kds([("k", "int4"), ("id", "int2"), ("t", "text")], rows,
                               [int(h) for h in hashes])
    km = kds.build_multirels([(inner, False)])
    plan = cg.Plan([("k", "int4"), ("v", "int8")])
    k = plan.scan_var("k")
    v = plan.scan_var("v")
    d = plan.add_inner([("k", "int4"), ("id", "int2"), ("t", "text")])
    ik = plan.inner_var(d, "k")
    iid = plan.inner_var(d, "id")
    it = plan.inner_var(d, "t")
    plan.add_join(d, [k], [cg.Func("int4eq", k, ik)],
                  other_quals=[cg.Func("int2lt", iid, cg.Const("int2", 800))])
    plan.set_projection([v, iid, it], [("v", "int8"), ("id", "int2"), ("t", "text")])
    blob = plan.build()
    src = kds.build_arrow_kds(outer)
    got = run_gpu_projection(blob, src, 32 << 20, kmrels=km, tuning=dict(jit=jit))
    exp, res = run_oracle_projection(blob, src, 32 << 20, kmrels=km)
    assert len(got) == 1
    assert np.array_equal(got[0], exp)
    assert res.nitems_out > n


@pytest.mark.parametrize("join_passes", PASSES, ids=PASSES_IDS)
def test_join_groupby_heavy_duplicates(gpu, join_passes):
    """One join key held by 600 inner items (a chain longer than the exact
    multiplicity count covers): every outer row with that key matches all
    of them, and the passes size their records for it."""
    r = datagen.rng(12)
    n = 20_000
    outer = [kds.Column("k", "int4", r.integers(0, 10, size=n, dtype=np.int32)),
             kds.Column("v", "int8", r.integers(0, 1000, size=n, dtype=np.int64))]
    keys = np.concatenate([np.full(600, 3, dtype=np.int32),
                           r.integers(100, 200, size=200, dtype=np.int32)])
    rows = [(int(k), int(i % 5), bytes("t%d" % i, "ascii")) for i, k in enumerate(keys)]
    hashes = hashing.hash_keys_np([(keys, 4)])
    inner = kds.build_hash_kds([("k", "int4"), ("g", "int2"), ("t", "text")], rows,
                               [int(h) for h in hashes])
    km = kds.build_multirels([(inner, False)])
    plan = cg.Plan([("k", "int4"), ("v", "int8")])
    k = plan.scan_var("k")
    v = plan.scan_var("v")
    d = plan.add_inner([("k", "int4"), ("g", "int2"), ("t", "text")])
    ik = plan.inner_var(d, "k")
    g = plan.inner_var(d, "g")
    plan.add_join(d, [k], [cg.Func("int4eq", k, ik)])
    plan.set_groupby([
        (abi.KAGG_ACTION__NROWS_ANY, None, ("nrows", "int8")),
        (abi.KAGG_ACTION__PSUM_INT64, v, ("psum_v", "bytea")),
        (abi.KAGG_ACTION__VREF, g, ("g", "int2")),
    ])
    blob = plan.build(final_nrows=16)
    src = kds.build_arrow_kds(outer)
    acts = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF]
    got = run_gpu_groupby(blob, [src], kmrels=km, tuning=dict(jit=xpuhip.JIT_ON, join_passes=join_passes))
    exp = run_oracle_groupby(blob, [src], kmrels=km)
    groups = assert_groups_equal(got, exp, acts)
    assert sum(g[0] for g in groups.values()) == 600 * int(np.sum(outer[0].values == 3))
