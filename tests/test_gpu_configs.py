"""The configurations at their real sizes (or a path they take only at
size) against the oracle:

  - C1, 100M rows resident in HBM (Zipf(1.1) over 10,000 keys): every
    group's count, sum(v), min(x), max(x) exact and avg(w) within 1e-12
    relative, the hottest groups (about 15M rows each) included;
  - C2's multi-slice path: a join record budget small enough that the
    20M-row fact table runs in at least 3 slices of records, each slice
    through the screen and per-depth passes, against the oracle;
  - C4, the served mix: every request of a Poisson-timed mix of scans,
    group-bys and star joins over 1M-4M-row windows of resident tables,
    served with coalescing and overlap, against the oracle's result for
    that window."""
import os

import numpy as np
import pytest

from pgstrom import abi, datagen, kds, serving, workloads as W, xpuhip
from pgstrom.kds import Column
from helpers import assert_groups_equal, run_oracle_groupby, run_oracle_projection

pytestmark = pytest.mark.gpu

C1_ACTIONS = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__PAVG_FP,
              abi.KAGG_ACTION__PMIN_INT32, abi.KAGG_ACTION__PMAX_INT32, abi.KAGG_ACTION__VREF]
C2_ACTIONS = [abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF, abi.KAGG_ACTION__VREF]
NTHREADS = min(16, os.cpu_count() or 1)


def test_c1_100m_rows(gpu):
    n = 100_000_000
    print("C1: generating %d rows" % n, flush=True)
    cols = datagen.preagg_table(n)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    d_src = xpuhip.DeviceBuffer(len(src), src)
    s = xpuhip.Session(blob)
    rc, res, _ = s.exec_device(d_src, n)
    assert rc == 0 and res.nitems_in == n
    got = s.final()
    s.close()
    d_src.free()
    print("C1: oracle", flush=True)
    exp = run_oracle_groupby(blob, [src], nthreads=NTHREADS)
    groups = assert_groups_equal(got, exp, C1_ACTIONS, rtol=1e-12)
    counts = sorted((g[0] for g in groups.values()), reverse=True)
    assert sum(counts) == n and len(groups) == 10_000
    assert counts[0] > 10_000_000       # the hottest Zipf groups are checked too


def test_c2_join_multi_slice(gpu, monkeypatch):
    n = 20_000_000
    fact, dims = datagen.star_tables(n)
    src = kds.build_arrow_kds(fact)
    km = W.star_inner(dims)
    blob = W.star_query(fact)
    # 20 B of records per row (levels 1-3, unique keys): 80 MB = 4M-row slices
    monkeypatch.setenv("XPUHIP_JOIN_REC_BUDGET", str(80 << 20))
    d_src = xpuhip.DeviceBuffer(len(src), src)
    s = xpuhip.Session(blob)
    s.set_tuning(jit=xpuhip.JIT_ON, join_passes=xpuhip.JOIN_PASSES_ON)
    s.attach_inner(km)
    rc, res, _ = s.exec_device(d_src, n)
    st = xpuhip.last_task_stats()
    assert rc == 0
    got = s.final()
    s.close()
    d_src.free()
    assert st["join_slices"] >= 3, st
    exp = run_oracle_groupby(blob, [src], kmrels=km, nthreads=NTHREADS)
    assert_groups_equal(got, exp, C2_ACTIONS)


def _window_cols(cols, r0, n):
    out = []
    for c in cols:
        vals = c.values[r0:r0 + n]
        nulls = None if c.nulls is None else c.nulls[r0:r0 + n]
        out.append(Column(c.name, c.typname, vals, nulls))
    return out


def test_c4_served_mix_against_oracle(gpu):
    res = serving.Resident(scan_rows=4_000_000, preagg_rows=4_000_000, fact_rows=4_000_000,
                           dim_sizes=(2_000_000, 200_000, 2_556), nwin=64)
    try:
        reqs = serving.make_requests(res, 24, 1_000_000, 4_000_000, seed=11)
        arrivals = list(np.cumsum(np.random.default_rng(5).exponential(0.004, len(reqs))))
        backends = serving.Backends(res, 8)
        dsts = [xpuhip.DeviceBuffer(serving.dst_bytes(q["n"])) if q["kind"] == "scan" else None
                for q in reqs]
        srv = xpuhip.Server(hbm_budget=200 << 30, slots=8, max_batch=16)
        st, _ = serving.run(srv, backends, reqs, arrivals, dsts, keep_results=True)
        srv.close()
        backends.close()
        assert all(s.rc == 0 for s in st), [s.err for s in st if s.rc]
        blobs = {k: res.blob(k) for k in serving.KINDS}
        tables = {"scan": res.scan_cols, "preagg": res.preagg_cols, "join": res.fact}
        for i, (q, s) in enumerate(zip(reqs, st)):
            print("C4 request %d: %s, %d rows" % (i, q["kind"], q["n"]), flush=True)
            src = kds.build_arrow_kds(_window_cols(tables[q["kind"]], q["r0"], q["n"]))
            if q["kind"] == "scan":
                exp, _ = run_oracle_projection(blobs["scan"], src, serving.dst_bytes(q["n"]),
                                               nthreads=NTHREADS)
                assert kds.parse_row_kds(dsts[i].to_host()) == kds.parse_row_kds(exp), i
            else:
                exp = run_oracle_groupby(blobs[q["kind"]], [src],
                                         kmrels=res.kmrels if q["kind"] == "join" else None,
                                         nthreads=NTHREADS)
                assert_groups_equal(s.result, exp,
                                    C1_ACTIONS if q["kind"] == "preagg" else C2_ACTIONS)
        for d in dsts:
            if d is not None:
                d.free()
    finally:
        res.free()
