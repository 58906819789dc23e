"""GpuJoin parity on the GPU: hash-join probe of host-built inner hash tables,
feeding a group-by or a projection."""
import numpy as np
import pytest

from pgstrom import abi, codegen as cg, datagen, kds, workloads as W
from helpers import (assert_groups_equal, run_gpu_groupby, run_gpu_projection,
                     run_oracle_groupby, run_oracle_projection)

pytestmark = pytest.mark.gpu

C2_ACTIONS = [abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF, abi.KAGG_ACTION__VREF]


@pytest.mark.parametrize("left_outer", [False, True])
def test_c2_star_join_preagg(gpu, left_outer, jit):
    fact, dims = datagen.star_tables(300_000, dim_sizes=(20_000, 2_000, 2_556))
    src = kds.build_arrow_kds(fact)
    km = W.star_inner(dims)
    blob = W.star_query(fact, left_outer=left_outer)
    got = run_gpu_groupby(blob, [src], kmrels=km, tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, [src], kmrels=km)
    assert_groups_equal(got, exp, C2_ACTIONS)


def test_join_projection_multi_match(gpu, jit):
    """Inner relation with duplicate keys: every match is emitted, in the
    oracle's (row-major) order."""
    r = datagen.rng(9)
    n = 20000
    outer = [kds.Column("k", "int4", r.integers(0, 500, size=n, dtype=np.int32)),
             kds.Column("v", "int8", r.integers(0, 10**9, size=n, dtype=np.int64))]
    keys = r.integers(0, 600, size=900, dtype=np.int32)
    rows = [(int(k), int(i), bytes("t%d" % i, "ascii")) for i, k in enumerate(keys)]
    from pgstrom import hashing
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
