"""GPU-Sort cases shared by the CPU (oracle) and GPU tests: the group-by of
count(*), sum(v), avg(w), min(x), max(x) GROUP BY key with sort keys over its
final buffer (kds_final attributes: 1 count, 2 sum(v) numeric state, 3 avg(w)
state, 4 min(x), 5 max(x), 6 key)."""
import numpy as np

from pgstrom import abi, codegen as cg, kds, workloads as W
from pgstrom.kds import Column

A = abi
NROWS, PSUM, PAVG, PMIN, PMAX, KEY = 1, 2, 3, 4, 5, 6


def key_vref(asc=True, nulls_first=False):
    return (A.KSORT_KEY_KIND__VREF, KEY, "int4", asc, nulls_first)


def count_vref(asc=True, nulls_first=False):
    return (A.KSORT_KEY_KIND__VREF, NROWS, "int8", asc, nulls_first)


def sum_key(asc=True, nulls_first=False):
    return (A.KSORT_KEY_KIND__PSUM_NUMERIC, PSUM, None, asc, nulls_first)


def avg_key(asc=True, nulls_first=False):
    return (A.KSORT_KEY_KIND__PAVG_FP64, PAVG, None, asc, nulls_first)


def min_key(asc=True, nulls_first=False):
    return (A.KSORT_KEY_KIND__PMINMAX_INT64, PMIN, None, asc, nulls_first)


def max_key(asc=True, nulls_first=False):
    return (A.KSORT_KEY_KIND__PMINMAX_INT64, PMAX, None, asc, nulls_first)


def sorted_query(cols, sort_keys, limit=0, final_nrows=10000, final_length=1 << 30):
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    key, v, w, x = (plan.scan_var(n) for n in ("key", "v", "w", "x"))
    plan.set_groupby(W.preagg_tlist(plan, key, v, w, x))
    return plan.build(final_nrows=final_nrows, final_length=final_length, sort_keys=sort_keys,
                      sort_limit=limit)


def tiny_table():
    """Nine rows, five groups (one of them the NULL key):
         key   count  sum(v)  avg(w)  min(x)  max(x)
         1     2      10      4.0     -3      8
         2     2      8       2.0     2       4
         3     2      -10     3.0     1       7
         5     1      100     0.5     0       0
         NULL  2      3       6.0     -5      9"""
    key = np.array([3, 1, 2, 0, 1, 3, 2, 5, 0], dtype=np.int32)
    knull = np.array([0, 0, 0, 1, 0, 0, 0, 0, 1], dtype=bool)
    v = np.array([10, 5, 7, 1, 5, -20, 1, 100, 2], dtype=np.int64)
    w = np.array([1.0, 2.0, 3.0, 4.0, 6.0, 5.0, 1.0, 0.5, 8.0])
    x = np.array([7, -3, 4, 9, 8, 1, 2, 0, -5], dtype=np.int32)
    return [Column("key", "int4", key, knull), Column("v", "int8", v),
            Column("w", "float8", w), Column("x", "int4", x)]


# (sort keys, limit, the group keys in the expected order)
KNOWN = [
    ([key_vref(True, False)], 0, [1, 2, 3, 5, None]),
    ([key_vref(False, True)], 0, [None, 5, 3, 2, 1]),
    ([key_vref(True, True)], 0, [None, 1, 2, 3, 5]),
    ([key_vref(False, False)], 0, [5, 3, 2, 1, None]),
    ([sum_key(False), key_vref()], 0, [5, 1, 2, None, 3]),
    ([sum_key(True), key_vref()], 0, [3, None, 2, 1, 5]),
    ([avg_key(True), key_vref()], 0, [5, 2, 3, 1, None]),
    ([min_key(True), key_vref()], 0, [None, 1, 5, 3, 2]),
    ([max_key(False), key_vref()], 0, [None, 1, 3, 2, 5]),
    ([count_vref(False), key_vref(True, False)], 0, [1, 2, 3, None, 5]),
    ([count_vref(False), key_vref(False, True)], 0, [None, 3, 2, 1, 5]),
    ([sum_key(False), key_vref()], 2, [5, 1]),
    ([count_vref(True), key_vref()], 3, [5, 1, 2]),
]


def group_keys_in_order(kds_final):
    """The group key of each row of a written-back final buffer, in row-index
    order."""
    return [vals[KEY - 1] for vals, _ in kds.parse_row_kds(kds_final)]


# group keys of other types: (typname, numpy dtype, values) with the ascending
# order of the distinct non-NULL values written out; row 0's key is NULL
TYPED = [
    ("int2", np.int16, [7, -300, 12, 7, 32767, -32768, 0]),
    ("int8", np.int64, [2**40, -2**62, 5, -1, 2**40, 0, -(2**40)]),
    ("float8", np.float64, [1.5, -2.25, 1e300, -1e-300, 0.0, -7.0, 1.5]),
    ("float4", np.float32, [0.25, -0.5, 3.0e38, -3.0e38, 0.25, 1.0, -0.0]),
    ("bool", np.bool_, [True, False, True, False, True, False, True]),
]


def typed_table(typname, dtype, values):
    """One row per value plus a NULL-key row in front; v = position."""
    n = len(values) + 1
    key = np.zeros(n, dtype=dtype)
    key[1:] = np.array(values, dtype=dtype)
    knull = np.zeros(n, dtype=bool)
    knull[0] = True
    return [Column("key", typname, key, knull), Column("v", "int8", np.arange(n, dtype=np.int64)),
            Column("w", "float8", np.ones(n)), Column("x", "int4", np.arange(n, dtype=np.int32))]


def typed_expected(dtype, values, asc, nulls_first):
    """Distinct keys (as stored in the column's type) in the comparator's
    order (-0.0 and 0.0 compare equal)."""
    distinct = []
    for v in np.array(values, dtype=dtype).tolist():
        if not any(v == d for d in distinct):
            distinct.append(v)
    distinct.sort(reverse=not asc)
    return [None] + distinct if nulls_first else distinct + [None]


# ---- window rank: rank() over (PARTITION BY count(*) ORDER BY sum(v) DESC)
# with the group key as a trailing sort key, so the sorted order is total
WINDOW_KEYS = [count_vref(True), sum_key(False), key_vref(True, False)]
WINDOW_ACTIONS = [A.KAGG_ACTION__NROWS_ANY, A.KAGG_ACTION__PSUM_INT64, A.KAGG_ACTION__PAVG_FP,
                  A.KAGG_ACTION__PMIN_INT32, A.KAGG_ACTION__PMAX_INT32, A.KAGG_ACTION__VREF]
WINDOW_FUNCS = {"row_number": A.KSORT_WINDOW_FUNC__ROW_NUMBER, "rank": A.KSORT_WINDOW_FUNC__RANK,
                "dense_rank": A.KSORT_WINDOW_FUNC__DENSE_RANK}


def window_table(n, nkeys, seed=4):
    """key int4 uniform over nkeys values (about one row in 500 NULL), v int8
    in 0..3 so that sums tie within a partition, w float8, x int4."""
    r = np.random.default_rng(seed)
    key = r.integers(0, nkeys, size=n, dtype=np.int32)
    knull = r.random(n) < 0.002
    v = r.integers(0, 4, size=n, dtype=np.int64)
    w = r.random(n)
    x = r.integers(-1000, 1000, size=n, dtype=np.int32)
    return [Column("key", "int4", key, knull), Column("v", "int8", v),
            Column("w", "float8", w), Column("x", "int4", x)]


def window_query(cols, func, limit, final_nrows=10000):
    """The group-by with GPU-Sort by WINDOW_KEYS and the window rank
    (func, limit) over partition key 1 (count) and order key 2 (sum)."""
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    key, v, w, x = (plan.scan_var(n) for n in ("key", "v", "w", "x"))
    plan.set_groupby(W.preagg_tlist(plan, key, v, w, x))
    return plan.build(final_nrows=final_nrows, final_length=1 << 30, sort_keys=WINDOW_KEYS,
                      window_rank=(WINDOW_FUNCS[func], limit, 1, 1))


def window_expected(groups, func, limit):
    """SQL's row_number / rank / dense_rank over the groups ({(key,): states}
    of decode_groups), written from their definitions: the group keys kept
    by `rank() < limit`, in the sort's order (count asc, sum desc, key asc
    NULLS LAST)."""
    rows = []
    for (k,), st in groups.items():
        cnt, (_, total) = st[0], st[1]
        rows.append((cnt, total, k))
    rows.sort(key=lambda t: (t[0], -t[1], t[2] is None, t[2] if t[2] is not None else 0))
    out = []
    i = 0
    while i < len(rows):
        j = i
        while j < len(rows) and rows[j][0] == rows[i][0]:
            j += 1
        part = rows[i:j]
        for pos, row in enumerate(part):
            if func == "row_number":
                rank = pos + 1
            elif func == "rank":
                rank = 1 + sum(1 for q in part if q[1] > row[1])
            else:
                rank = len({q[1] for q in part if q[1] >= row[1]})
            if rank < limit:
                out.append(row[2])
        i = j
    return out
