"""GPU-Sort of the group-by results on the GPU: the final buffer's row index
sorted by the session's SortKeys on the device (key words + LSD radix sort)
gives the oracle's order (cuda_gpusort.cu:273-342), with LIMIT keeping the
leading rows; every sort ends in the group key, so the order is total and
the row sequences must agree exactly."""
import numpy as np
import pytest

from pgstrom import abi, datagen, kds, xpuhip
from pgstrom.kds import Column
from helpers import assert_groups_equal, run_oracle_groupby
import sortcases as S

pytestmark = pytest.mark.gpu

C1_ACTIONS = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__PAVG_FP,
              abi.KAGG_ACTION__PMIN_INT32, abi.KAGG_ACTION__PMAX_INT32, abi.KAGG_ACTION__VREF]


def _gpu_sorted(blob, chunks, tuning=None):
    s = xpuhip.Session(blob)
    if tuning:
        s.set_tuning(**tuning)
    for src in chunks:
        rc, _, _, _ = s.exec(src)
        assert rc == 0
    fin = s.final()
    st = xpuhip.last_task_stats()
    again = s.final()
    s.close()
    return fin, st, again


@pytest.mark.parametrize("case", range(len(S.KNOWN)))
def test_gpu_sort_known_orders(gpu, case):
    keys, limit, expected = S.KNOWN[case]
    cols = S.tiny_table()
    fin, st, again = _gpu_sorted(S.sorted_query(cols, keys, limit), [kds.build_arrow_kds(cols)])
    assert S.group_keys_in_order(fin) == expected
    assert S.group_keys_in_order(again) == expected
    assert st["sort_items"] == 5


SORTS = [
    ([S.sum_key(False), S.key_vref()], 0),
    ([S.sum_key(False), S.key_vref()], 100),
    ([S.avg_key(True), S.key_vref()], 0),
    ([S.min_key(True, True), S.max_key(False), S.key_vref(False)], 0),
    ([S.count_vref(False), S.key_vref()], 25),
    ([S.key_vref(False)], 0),
]


@pytest.mark.parametrize("which", range(len(SORTS)))
def test_gpu_sort_matches_oracle(gpu, jit, which):
    """C1's table (400K rows, Zipf keys): the sorted groups, in order."""
    keys, limit = SORTS[which]
    cols = datagen.preagg_table(400_000, fragment=21)
    src = kds.build_arrow_kds(cols)
    blob = S.sorted_query(cols, keys, limit)
    fin, st, _ = _gpu_sorted(blob, [src], dict(jit=jit))
    exp = run_oracle_groupby(blob, [src])
    got_order = S.group_keys_in_order(fin)
    assert got_order == S.group_keys_in_order(exp)
    assert len(got_order) == (limit or len(got_order))
    assert st["sort_items"] > limit and st["sort_passes"] >= 1
    assert_groups_equal(fin, exp, C1_ACTIONS)
    if limit:
        # the kept rows' tuples are packed (kern_buffer_simple_limit)
        hd, _ = kds.kds_colmeta(memoryview(bytes(fin)))
        assert hd.length == abi.kds_head_length(hd.nr_colmeta) + 8 * hd.nitems + hd.usage
        assert hd.usage <= 512 * limit
        # each packed tuple's row id is its sorted position (KERN_TUPITEM_SET_ROWID,
        # cuda_gpusort.cu:1446)
        rows = kds.parse_row_kds(fin)
        assert [rowid for _, rowid in rows] == list(range(len(rows)))


def test_gpu_sort_million_groups(gpu):
    """About a million groups: sum(v) DESC, key; the first rows and the whole
    order against the oracle."""
    r = np.random.default_rng(77)
    n = 3_000_000
    key = r.integers(0, 1_000_000, size=n, dtype=np.int32)
    cols = [Column("key", "int4", key),
            Column("v", "int8", r.integers(-1000, 1_000_000, size=n, dtype=np.int64)),
            Column("w", "float8", r.random(n) * 1000.0),
            Column("x", "int4", r.integers(-2**31, 2**31 - 1, size=n, dtype=np.int32))]
    src = kds.build_arrow_kds(cols)
    blob = S.sorted_query(cols, [S.sum_key(False), S.key_vref()], 0,
                          final_nrows=1_000_000, final_length=1 << 30)
    fin, st, _ = _gpu_sorted(blob, [src])
    ngroups = len(np.unique(key))
    assert st["sort_items"] == ngroups
    exp = run_oracle_groupby(blob, [src])
    assert S.group_keys_in_order(fin) == S.group_keys_in_order(exp)


@pytest.mark.parametrize("typed", range(len(S.TYPED)))
@pytest.mark.parametrize("asc, nulls_first", [(True, False), (False, True), (False, False)])
def test_gpu_sort_key_types(gpu, typed, asc, nulls_first):
    """Group keys of int2 / int8 / float8 / float4 / bool (negative values,
    extremes, -0.0 beside 0.0), ASC and DESC, NULLs first and last: the
    distinct keys in order, as the oracle."""
    typname, dtype, values = S.TYPED[typed]
    cols = S.typed_table(typname, dtype, values)
    src = kds.build_arrow_kds(cols)
    blob = S.sorted_query(cols, [(abi.KSORT_KEY_KIND__VREF, S.KEY, typname, asc, nulls_first)])
    fin, _, _ = _gpu_sorted(blob, [src])
    got = [None if g is None else float(g) for g in S.group_keys_in_order(fin)]
    exp = [None if e is None else float(e) for e in S.typed_expected(dtype, values, asc, nulls_first)]
    assert got == exp
    assert S.group_keys_in_order(run_oracle_groupby(blob, [src])) == S.group_keys_in_order(fin)


@pytest.mark.parametrize("nulls_first", [False, True])
def test_gpu_sort_null_states(gpu, nulls_first):
    """Groups whose min(x) saw only NULL arguments carry an invalid state,
    which sorts as NULL (cuda_gpusort.cu:66-93)."""
    n = 40
    key = np.arange(n, dtype=np.int32) % 8
    xnull = (key % 3 == 0)
    cols = [Column("key", "int4", key), Column("v", "int8", np.arange(n, dtype=np.int64)),
            Column("w", "float8", np.ones(n)),
            Column("x", "int4", (np.arange(n, dtype=np.int32) * 7) % 11 - 5, xnull)]
    src = kds.build_arrow_kds(cols)
    blob = S.sorted_query(cols, [S.min_key(True, nulls_first), S.key_vref()])
    fin, _, _ = _gpu_sorted(blob, [src])
    exp = run_oracle_groupby(blob, [src])
    order = S.group_keys_in_order(fin)
    assert order == S.group_keys_in_order(exp)
    null_groups = sorted(int(k) for k in set(key[xnull]))
    assert (order[:len(null_groups)] if nulls_first else order[-len(null_groups):]) == null_groups


@pytest.mark.parametrize("func", sorted(S.WINDOW_FUNCS))
@pytest.mark.parametrize("nrows, nkeys, limit", [(20000, 3000, 3), (2_000_000, 200_000, 5)])
def test_gpu_window_rank_matches_oracle(gpu, func, nrows, nkeys, limit):
    """rank() < limit over (PARTITION BY count ORDER BY sum DESC), the group
    key last: the kept groups in the oracle's order with their states, row
    ids renumbered; 200K groups span many scan tiles and long partitions."""
    cols = S.window_table(nrows, nkeys)
    blob = S.window_query(cols, func, limit, final_nrows=nkeys + 10)
    src = kds.build_arrow_kds(cols)
    fin, st, again = _gpu_sorted(blob, [src])
    exp = run_oracle_groupby(blob, [src])
    keys = S.group_keys_in_order(fin)
    assert keys == S.group_keys_in_order(exp)
    assert S.group_keys_in_order(again) == keys
    assert 0 < len(keys) < nkeys
    assert [rid for _, rid in kds.parse_row_kds(fin)] == list(range(len(keys)))
    assert_groups_equal(fin, exp, S.WINDOW_ACTIONS)
    assert st["sort_items"] > len(keys)


def test_gpu_window_rank_single_group(gpu):
    """One group: kept (its rank is 1) and re-packed, as the reference's
    finalize re-packs every kept row."""
    cols = S.window_table(50, 1)
    cols[0] = Column("key", "int4", cols[0].values)
    blob = S.window_query(cols, "rank", 2)
    src = kds.build_arrow_kds(cols)
    fin, _, _ = _gpu_sorted(blob, [src])
    exp = run_oracle_groupby(blob, [src])
    assert S.group_keys_in_order(fin) == S.group_keys_in_order(exp) == [0]
    assert np.array_equal(np.frombuffer(bytes(fin), np.uint8)[:16],
                          np.frombuffer(bytes(exp), np.uint8)[:16])
