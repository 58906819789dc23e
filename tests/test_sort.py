"""GPU-Sort on the CPU: the oracle's restatement of the reference's
comparator (cuda_gpusort.cu:15-343) and LIMIT (gpu_service.c:5099-5197)
against hand-computed orders, and the compiler's checks of SortKeys."""
import pytest

from pgstrom import abi, kds, workloads as W, xpuhip
from helpers import run_oracle_groupby
import sortcases as S


@pytest.mark.parametrize("case", range(len(S.KNOWN)))
def test_oracle_sort_known_orders(case):
    keys, limit, expected = S.KNOWN[case]
    cols = S.tiny_table()
    blob = S.sorted_query(cols, keys, limit)
    fin = run_oracle_groupby(blob, [kds.build_arrow_kds(cols)])
    assert S.group_keys_in_order(fin) == expected


def test_oracle_sort_limit_keeps_leading_rows():
    """LIMIT n keeps the first n rows of the full order."""
    cols = S.tiny_table()
    full = S.group_keys_in_order(run_oracle_groupby(
        S.sorted_query(cols, [S.avg_key(False), S.key_vref()]), [kds.build_arrow_kds(cols)]))
    for n in range(1, 6):
        got = S.group_keys_in_order(run_oracle_groupby(
            S.sorted_query(cols, [S.avg_key(False), S.key_vref()], limit=n),
            [kds.build_arrow_kds(cols)]))
        assert got == full[:n]


def test_oracle_sort_limit_packs_rows():
    """LIMIT packs the kept tuples (kern_buffer_simple_limit,
    cuda_gpusort.cu:1401-1454): the buffer holds only them, and each row id
    is its sorted position (:1446)."""
    cols = S.tiny_table()
    fin = run_oracle_groupby(S.sorted_query(cols, [S.avg_key(False), S.key_vref()], limit=3),
                             [kds.build_arrow_kds(cols)])
    hd, _ = kds.kds_colmeta(memoryview(bytes(fin)))
    assert hd.nitems == 3 and len(fin) == hd.length
    assert hd.length == abi.kds_head_length(hd.nr_colmeta) + 8 * 3 + hd.usage
    assert [rowid for _, rowid in kds.parse_row_kds(fin)] == [0, 1, 2]


def test_sort_numeric_key_needs_int8_sum():
    """A numeric sort key is accepted only over a sum(int8) state (weight
    0), whose raw int128 sums order correctly."""
    cols = S.tiny_table()
    blob = S.sorted_query(cols, [(abi.KSORT_KEY_KIND__PSUM_NUMERIC, S.PAVG, None, True, False)])
    err = _compile_error(blob)
    assert err is not None and "numeric key must be a sum(int8) state" in err


def _compile_error(blob):
    rc, err = xpuhip.session_check(blob)
    return None if rc == 0 else err.message.decode()


@pytest.mark.parametrize("bad, message", [
    ("window_limit", "window rank with a LIMIT"),
    ("window_keys", "window rank key counts out of range"),
    ("anum", "key attribute out of range"),
    ("type", "does not match the final buffer column"),
    ("kind", "kind is not supported"),
])
def test_sort_keys_compile_checks(bad, message):
    cols = S.tiny_table()
    if bad.startswith("window"):
        from pgstrom import codegen as cg, workloads as W
        plan = cg.Plan([(c.name, c.typname) for c in cols])
        key, v, w, x = (plan.scan_var(n) for n in ("key", "v", "w", "x"))
        plan.set_groupby(W.preagg_tlist(plan, key, v, w, x))
        if bad == "window_limit":
            blob = plan.build(sort_keys=S.WINDOW_KEYS, sort_limit=5,
                              window_rank=(abi.KSORT_WINDOW_FUNC__RANK, 3, 1, 1))
        else:
            blob = plan.build(sort_keys=[S.key_vref()],
                              window_rank=(abi.KSORT_WINDOW_FUNC__RANK, 3, 0, 1))
    elif bad == "anum":
        blob = S.sorted_query(cols, [(abi.KSORT_KEY_KIND__VREF, 9, "int4", True, False)])
    elif bad == "type":
        blob = S.sorted_query(cols, [(abi.KSORT_KEY_KIND__VREF, S.KEY, "int8", True, False)])
    else:
        blob = S.sorted_query(cols, [(abi.KSORT_KEY_KIND__PVARIANCE_SAMP, S.PAVG, None, True,
                                      False)])
    err = _compile_error(blob)
    assert err is not None and message in err


def test_sort_keys_compile():
    cols = S.tiny_table()
    for keys, limit, _ in S.KNOWN:
        assert _compile_error(S.sorted_query(cols, keys, limit)) is None


@pytest.mark.parametrize("typed", range(len(S.TYPED)))
@pytest.mark.parametrize("asc, nulls_first", [(True, False), (False, True), (False, False)])
def test_oracle_sort_key_types(typed, asc, nulls_first):
    """Group keys of int2 / int8 / float8 / float4 / bool, ASC and
    DESC, NULLs first and last: the distinct keys in order."""
    typname, dtype, values = S.TYPED[typed]
    cols = S.typed_table(typname, dtype, values)
    blob = S.sorted_query(cols, [(abi.KSORT_KEY_KIND__VREF, S.KEY, typname, asc, nulls_first)])
    got = S.group_keys_in_order(run_oracle_groupby(blob, [kds.build_arrow_kds(cols)]))
    exp = S.typed_expected(dtype, values, asc, nulls_first)
    assert [None if g is None else float(g) for g in got] == \
        [None if e is None else float(e) for e in exp]


def _oracle_final(blob, cols):
    return run_oracle_groupby(blob, [kds.build_arrow_kds(cols)])


@pytest.mark.parametrize("func", sorted(S.WINDOW_FUNCS))
@pytest.mark.parametrize("limit", [2, 4])
def test_oracle_window_rank_matches_sql(func, limit):
    """rank() < limit over (PARTITION BY count ORDER BY sum DESC): the
    oracle's kept rows, in order, equal SQL's definitions applied to the
    groups; ties in sum make the three functions keep different rows."""
    cols = S.window_table(20000, 3000)
    groups = W.decode_groups(_oracle_final(S.sorted_query(cols, [S.key_vref()]), cols),
                             S.WINDOW_ACTIONS)
    fin = _oracle_final(S.window_query(cols, func, limit), cols)
    got = S.group_keys_in_order(fin)
    assert got == S.window_expected(groups, func, limit)
    assert 0 < len(got) < len(groups)
    # packed as kern_windowrank_finalize packs: row ids are the new positions
    assert [rid for _, rid in kds.parse_row_kds(fin)] == list(range(len(got)))


def test_window_rank_functions_differ():
    """The test table separates the three functions (ties in sum exist)."""
    cols = S.window_table(20000, 3000)
    groups = W.decode_groups(_oracle_final(S.sorted_query(cols, [S.key_vref()]), cols),
                             S.WINDOW_ACTIONS)
    kept = {f: S.window_expected(groups, f, 3) for f in S.WINDOW_FUNCS}
    assert len(kept["row_number"]) < len(kept["rank"]) < len(kept["dense_rank"])


def test_oracle_window_rank_known():
    """tiny_table: partition count 1 = {5}; count 2 by sum desc = 1 (10),
    2 (8), NULL (3), 3 (-10); rank() < 3 keeps 5, 1, 2."""
    cols = S.tiny_table()
    for func in S.WINDOW_FUNCS:
        fin = _oracle_final(S.window_query(cols, func, 3), cols)
        assert S.group_keys_in_order(fin) == [5, 1, 2]
