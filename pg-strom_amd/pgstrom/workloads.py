"""The scan / pre-aggregation / star-join workloads as PG-Strom would plan
them: each builder returns the session blob (as the reference backend builds
it) plus the input chunk(s), so tests, smoke and bench drive the library
through exactly the XpuCommand payloads of the reference.

  scan_query      WHERE a < K AND b > 0.0, projecting (c, d)
  preagg_query    GROUP BY key: count(*), sum(v), avg(w), min(x), max(x)
  star_query      fact JOIN dim1, dim2, dim3 (dimension filters ~10%),
                  GROUP BY (dim1.category, dim3.category), sum(quantity*price)
  fragment_query  WHERE a < K GROUP BY key % m with the pre-agg aggregates
"""
import time

import numpy as np

from . import abi, codegen as cg, datagen, hashing, kds

A = abi


def scan_query(cols, k=500000, use_param=False):
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    a = plan.scan_var("a")
    b = plan.scan_var("b")
    c = plan.scan_var("c")
    d = plan.scan_var("d")
    kexpr = plan.param("int4", k) if use_param else cg.Const("int4", k)
    plan.scan_quals = cg.And(cg.Func("int4lt", a, kexpr),
                             cg.Func("float8gt", b, cg.Const("float8", 0.0)))
    plan.set_projection([c, d], [("c", "int8"), ("d", "text")])
    return plan.build()


def preagg_tlist(plan, key_expr, v, w, x):
    """count(*), sum(v), avg(w), min(x), max(x) GROUP BY key, in the order the
    reference planner lays out kds_final: partial aggregates first, then the
    grouping key (gpu_preagg.c:92-501 maps each aggregate to its action)."""
    return [
        (A.KAGG_ACTION__NROWS_ANY, None, ("nrows", "int8")),
        (A.KAGG_ACTION__PSUM_INT64, v, ("psum_v", "bytea")),
        (A.KAGG_ACTION__PAVG_FP, w, ("pavg_w", "bytea")),
        (A.KAGG_ACTION__PMIN_INT32, x, ("pmin_x", "bytea")),
        (A.KAGG_ACTION__PMAX_INT32, x, ("pmax_x", "bytea")),
        (A.KAGG_ACTION__VREF, key_expr, ("key", key_expr.typname)),
    ]


def preagg_query(cols, final_nrows=10000):
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    key = plan.scan_var("key")
    v = plan.scan_var("v")
    w = plan.scan_var("w")
    x = plan.scan_var("x")
    plan.set_groupby(preagg_tlist(plan, key, v, w, x))
    return plan.build(final_nrows=final_nrows)


def fragment_query(cols, k, m, final_nrows=None):
    """Batched fragment: columns a, key, v, w, x; WHERE a < k GROUP BY key % m."""
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    a = plan.scan_var("a")
    key = plan.scan_var("key")
    v = plan.scan_var("v")
    w = plan.scan_var("w")
    x = plan.scan_var("x")
    plan.scan_quals = cg.Func("int4lt", a, plan.param("int4", int(k)))
    kmod = cg.Func("int4mod", key, plan.param("int4", int(m)))
    plan.set_groupby(preagg_tlist(plan, kmod, v, w, x))
    nrows = final_nrows or max(int(m), 100)
    # kds_final sized for the groups (at most m, and at most the 10,000 keys
    # of the generator): many fragment sessions stay resident at once
    return plan.build(final_nrows=nrows, final_length=(4 << 20) + 400 * min(nrows, 20000))


def dim_filter_mod(dim_index):
    """Dimension filters of the star join: pk % 10 = dim_index (~10%)."""
    return 10, dim_index


def star_inner(dims, timings=None):
    """KDS_FORMAT_HASH inner buffers of the three dimensions after their
    filters (pk % 10 = i), keyed by pk, and the kern_multirels holding them,
    built as the backend's inner preload does (gpu_join.c:2098-2164): the
    filtered rows as MinimalTuples, hashed and laid out by the library's
    native xpuhip_inner_hash_setup (its seconds go to timings["hash_build_s"])."""
    from . import xpuhip
    inners = []
    coldefs = [("pk", "int4"), ("category", "int2"), ("txt", "text")]
    spent = 0.0
    for i, dim in enumerate(dims):
        pk, cat, txt = dim
        mod, rem = dim_filter_mod(i)
        sel = np.nonzero(pk.values % mod == rem)[0]
        mtups = [kds.form_minimal_tuple(coldefs, (int(pk.values[j]), int(cat.values[j]),
                                                  bytes(txt.values[j]))) for j in sel]
        t0 = time.perf_counter()
        inners.append((xpuhip.inner_hash_setup(coldefs, mtups, [1]), False))
        spent += time.perf_counter() - t0
    if timings is not None:
        timings["hash_build_s"] = spent
    return kds.build_multirels(inners)


PINNED_PLAN_ID = 0x51a0


def dim_scan_pinned(dim_cols, dim_index, pinned_length=64 << 20):
    """The inner side of the star join as the GPU builds it: a scan of one
    dimension with its filter (pk % 10 = i) projecting (pk, category, txt)
    into a pinned inner buffer keyed by HashValue(pk)
    (DEVTASK__PINNED_HASH_RESULTS).  Its query_plan_id, PINNED_PLAN_ID + i,
    is the buffer_id the join's kern_multirels names."""
    plan = cg.Plan([(c.name, c.typname) for c in dim_cols])
    plan.query_plan_id = PINNED_PLAN_ID + dim_index
    pk = plan.scan_var("pk")
    cat = plan.scan_var("category")
    txt = plan.scan_var("txt")
    mod, rem = dim_filter_mod(dim_index)
    plan.scan_quals = cg.Func("int4eq", cg.Func("int4mod", pk, cg.Const("int4", mod)),
                              cg.Const("int4", rem))
    plan.set_projection([pk, cat, txt], [("pk", "int4"), ("category", "int2"), ("txt", "text")])
    return plan.build(pinned_hash=[pk], pinned_length=pinned_length)


def star_inner_pinned(ndims=3, left_outer=False):
    """kern_multirels whose chunks name the pinned buffers of dim_scan_pinned."""
    return kds.build_multirels([(kds.PinnedInner(PINNED_PLAN_ID + i), left_outer)
                                for i in range(ndims)])


def star_query(fact_cols, ndims=3, left_outer=False):
    plan = cg.Plan([(c.name, c.typname) for c in fact_cols])
    fks = [plan.scan_var("fk%d" % (i + 1)) for i in range(ndims)]
    q = plan.scan_var("quantity")
    p = plan.scan_var("price")
    cats = {}
    for i in range(ndims):
        d = plan.add_inner([("pk", "int4"), ("category", "int2"), ("txt", "text")])
        pk = plan.inner_var(d, "pk")
        if i in (0, ndims - 1):
            cats[i] = plan.inner_var(d, "category")
        plan.add_join(d, [fks[i]], [cg.Func("int4eq", fks[i], pk)], left_outer=left_outer)
    cat1 = cats[0]
    cat3 = cats[ndims - 1]
    amount = cg.Func("int8mul", q, p)
    plan.set_groupby([
        (A.KAGG_ACTION__PSUM_INT64, amount, ("psum_amount", "bytea")),
        (A.KAGG_ACTION__VREF, cat1, ("cat1", "int2")),
        (A.KAGG_ACTION__VREF, cat3, ("cat3", "int2")),
    ])
    return plan.build(final_nrows=625)


# ----------------------------------------------------------------------
# decoding of kds_final group tuples into comparable results
# ----------------------------------------------------------------------
def decode_state(action, raw):
    """Partial state payload (varlena header stripped) -> python tuple
    (layouts: xpu_common.h:2400-2471)."""
    import struct
    if action in (A.KAGG_ACTION__NROWS_ANY, A.KAGG_ACTION__NROWS_COND):
        return raw
    b = bytes(raw)
    if action in (A.KAGG_ACTION__PMIN_INT32, A.KAGG_ACTION__PMIN_INT64,
                  A.KAGG_ACTION__PMAX_INT32, A.KAGG_ACTION__PMAX_INT64):
        attrs, v = struct.unpack_from("<Iq", b)
        return (attrs, v)
    if action in (A.KAGG_ACTION__PMIN_FP64, A.KAGG_ACTION__PMAX_FP64):
        attrs, v = struct.unpack_from("<Id", b)
        return (attrs, v)
    if action in (A.KAGG_ACTION__PSUM_INT, A.KAGG_ACTION__PAVG_INT):
        _, n, s = struct.unpack_from("<Iqq", b)
        return (n, s)
    if action in (A.KAGG_ACTION__PSUM_INT64, A.KAGG_ACTION__PAVG_INT64):
        _, n, lo, hi = struct.unpack_from("<IqQq", b)
        return (n, (hi << 64) | lo)
    if action in (A.KAGG_ACTION__PSUM_FP, A.KAGG_ACTION__PAVG_FP):
        _, n, s = struct.unpack_from("<Iqd", b)
        return (n, s)
    if action in (A.KAGG_ACTION__PSUM_NUMERIC, A.KAGG_ACTION__PAVG_NUMERIC):
        attrs, n, lo, hi = struct.unpack_from("<IqQq", b)
        return (attrs, n, (hi << 64) | lo)
    if action == A.KAGG_ACTION__STDDEV:
        _, n, sx, sxx = struct.unpack_from("<Iqdd", b)
        return (n, sx, sxx)
    raise ValueError(action)


def decode_groups(kds_final, actions):
    """{key tuple: [state per non-key action]} of a kds_final written back as
    KDS_FORMAT_ROW."""
    out = {}
    for vals, _rowid in kds.parse_row_kds(kds_final):
        key = tuple(v for a, v in zip(actions, vals) if a == A.KAGG_ACTION__VREF)
        states = [decode_state(a, v) for a, v in zip(actions, vals) if a != A.KAGG_ACTION__VREF]
        if key in out:
            raise ValueError("duplicate group %r" % (key,))
        out[key] = states
    return out
