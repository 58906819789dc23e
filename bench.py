"""Benchmark of the MI355X xPU path on its headline workload.

Default (N=1): GpuPreAgg over a 100M-row columnar table resident in HBM
(key int4 Zipf(1.1) over 10,000 values, v int8, w float8, x int4;
GROUP BY key computing count(*), sum(v), avg(w), min(x), max(x)).  One step is
one XpuTaskExec of the whole table through the session (the full hot path:
column loads, group-by into kds_final).  Secondary lines: the 1M-row scan and
the 200M-row star join + group-by, measured the same way.

Prints one JSON line on stdout.  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pg-strom_amd"))

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12   # MI355X HBM3E bytes/s


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def time_steps(run_step, warmup, steps):
    """Runs warmup + steps tasks; returns wall seconds per step, the mean
    kernel time of all passes (HIP events on the task's stream) and the
    per-step stats of the library."""
    from pgstrom import xpuhip
    for _ in range(warmup):
        run_step()
    xpuhip.lib().xpuhip_device_synchronize()
    kms, stats = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        run_step()
        kms.append(xpuhip.last_kernel_ms())
        stats.append(xpuhip.last_task_stats())
    xpuhip.lib().xpuhip_device_synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, float(np.mean(kms)), stats


def bench_preagg(n, warmup, steps):
    from pgstrom import datagen, kds, workloads as W, xpuhip
    log("generating %d-row pre-aggregation table" % n)
    cols = datagen.preagg_table(n)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    d_src = xpuhip.DeviceBuffer(len(src), src)
    del src
    s = xpuhip.Session(blob)
    log("autotuning")
    s.autotune(d_src, n)

    def step():
        rc, res, _ = s.exec_device(d_src, n)
        assert rc == 0

    wall, kms, stats = time_steps(step, warmup, steps)
    fin = s.final()
    groups = W.decode_groups(fin, [a for a, _, _ in W.preagg_tlist(None, cols[0], None, None, None)])
    total = sum(g[0] for g in groups.values())
    if total != (warmup + steps) * n:
        raise RuntimeError("count(*) over all groups is %d, expected %d" % (total, (warmup + steps) * n))
    tuning = s.get_tuning()
    s.close()
    d_src.free()
    algo_bytes = n * (4 + 8 + 8 + 4)
    return dict(rows=n, wall_ms=wall * 1e3, kernel_ms=kms, algo_bytes=algo_bytes,
                groups=len(groups), tuning=tuning, cols=cols,
                pass1_ms=float(np.mean([st["pass1_ms"] for st in stats])),
                spill_ms=float(np.mean([st["spill_ms"] for st in stats])),
                spill_frac=stats[-1]["spill_rows"] / n, jit=stats[-1]["jit"],
                lds_nents=stats[-1]["lds_nents"])


def bench_scan(n, warmup, steps):
    from pgstrom import datagen, kds, workloads as W, xpuhip
    log("generating %d-row scan table" % n)
    cols = datagen.scan_table(n)
    src = kds.build_arrow_kds(cols)
    blob = W.scan_query(cols)
    d_src = xpuhip.DeviceBuffer(len(src), src)
    dst_len = 64 << 20
    d_dst = xpuhip.DeviceBuffer(dst_len)
    s = xpuhip.Session(blob)
    out = {}

    def step():
        rc, res, _ = s.exec_device(d_src, n, d_dst, dst_len)
        assert rc == 0
        out["nout"] = res.nitems_out

    wall, kms, _ = time_steps(step, warmup, steps)
    s.close()
    hd = d_dst.to_host(256)
    usage = int(np.frombuffer(hd[8:16].tobytes(), dtype=np.uint64)[0])
    d_src.free()
    d_dst.free()
    # referenced columns (a, b, c, d offsets) + validity maps + text bytes read, output written
    nout = out["nout"]
    text_bytes = int(cols[3].values.offsets[-1])
    algo = n * (4 + 8 + 8 + 4) + 4 * (n // 8) + text_bytes + usage + 8 * nout
    return dict(rows=n, wall_ms=wall * 1e3, kernel_ms=kms, algo_bytes=algo, nout=nout)


def bench_heap(n, warmup, steps):
    """C0's scan over heap pages: the 1M-row table as a KDS_FORMAT_BLOCK chunk
    of 8 KB pages in device memory (xpuhip_task_exec_device_tuples)."""
    from pgstrom import abi, datagen, kds, workloads as W, xpuhip
    log("generating %d-row scan table on heap pages" % n)
    cols = datagen.scan_table(n)
    src = kds.build_block_kds(cols)
    hd = abi.kern_data_store_head.from_buffer_copy(bytes(src[:256]))
    npages = hd.nitems
    blob = W.scan_query(cols)
    d_src = xpuhip.DeviceBuffer(len(src), src)
    head = np.array(src[:abi.kds_head_length(hd.nr_colmeta)])
    dst_len = 64 << 20
    d_dst = xpuhip.DeviceBuffer(dst_len)
    s = xpuhip.Session(blob)
    out = {}

    def step():
        rc, res, _ = s.exec_device_tuples(d_src, head, d_dst, dst_len)
        assert rc == 0
        out["nout"] = res.nitems_out

    wall, kms, stats = time_steps(step, warmup, steps)
    s.close()
    usage = int(np.frombuffer(d_dst.to_host(256)[8:16].tobytes(), dtype=np.uint64)[0])
    d_src.free()
    d_dst.free()
    heap_ms = float(np.mean([st["heap_ms"] for st in stats]))
    nout = out["nout"]
    page_bytes = npages * kds.BLCKSZ
    # deform pass: every page read once; tiles of the four referenced columns
    # written (a 4 B, b 8 B, c 8 B, d an 8 B (offset, length) pair, plus a
    # validity bit each)
    tile_bytes = n * (4 + 8 + 8 + 8) + 4 * (n // 8)
    return dict(rows=n, pages=npages, wall_ms=wall * 1e3, kernel_ms=kms, heap_ms=heap_ms,
                nout=nout, page_bytes=page_bytes, tile_bytes=tile_bytes,
                algo_bytes=page_bytes + usage + 8 * nout)


def heap_line(hp):
    total_ms = hp["heap_ms"] + hp["kernel_ms"]
    return dict(workload="GpuScan 1M rows on 8 KB heap pages (KDS_FORMAT_BLOCK, %d pages): "
                         "WHERE a < 500000 AND b > 0.0 projecting (c, d)" % hp["pages"],
                rows_per_s=hp["rows"] / (total_ms / 1e3), device_ms=total_ms,
                heap_passes_ms=hp["heap_ms"], pipeline_ms=hp["kernel_ms"], wall_ms=hp["wall_ms"],
                hbm_gbs=hp["algo_bytes"] / (total_ms / 1e3) / 1e9,
                frac=hp["algo_bytes"] / (total_ms / 1e3) / HBM_PEAK,
                heap_passes_gbs=(hp["page_bytes"] + hp["tile_bytes"]) / (hp["heap_ms"] / 1e3) / 1e9,
                rows_out=hp["nout"])


def bench_star(n, warmup, steps):
    from pgstrom import datagen, kds, workloads as W, xpuhip
    log("generating %d-row star-join fact table and dimensions" % n)
    fact, dims = datagen.star_tables(n)
    tm = {}
    km = W.star_inner(dims, tm)
    host_build_s = tm["hash_build_s"]
    src = kds.build_arrow_kds(fact)
    blob = W.star_query(fact)
    d_src = xpuhip.DeviceBuffer(len(src), src)
    del src
    s = xpuhip.Session(blob)
    s.attach_inner(km)

    def step():
        rc, res, _ = s.exec_device(d_src, n)
        assert rc == 0

    wall, kms, _ = time_steps(step, warmup, steps)
    s.close()
    d_src.free()
    algo = n * (4 * 3 + 8 + 8)
    return dict(rows=n, wall_ms=wall * 1e3, kernel_ms=kms, algo_bytes=algo,
                host_build_s=host_build_s)


def sort_query(cols, sort_keys, limit=0):
    from pgstrom import codegen as cg, workloads as W
    plan = cg.Plan([(c.name, c.typname) for c in cols])
    key, v, w, x = (plan.scan_var(nm) for nm in ("key", "v", "w", "x"))
    plan.set_groupby(W.preagg_tlist(plan, key, v, w, x))
    return plan.build(final_nrows=1_000_000, final_length=1 << 30, sort_keys=sort_keys,
                      sort_limit=limit)


def bench_sort(n, warmup, steps):
    """GPU-Sort downstream of the group-by: n rows with an int4 key uniform
    over 1M values (about a million groups) aggregated as C1's query, then
    ORDER BY sum(v) DESC, key over the final buffer.  A step is one
    xpuhip_task_final: key pass, radix passes, permute and the write-back;
    sort_ms is the device time of the sort passes alone (HIP events).  The
    oracle's qsort of the same final buffer is timed beside it (the final
    with the sort keys minus the final without them)."""
    from pgstrom import abi, kds, xpuhip
    from pgstrom.kds import Column
    r = np.random.default_rng(77)
    cols = [Column("key", "int4", r.integers(0, 1_000_000, size=n, dtype=np.int32)),
            Column("v", "int8", r.integers(-1000, 1_000_000, size=n, dtype=np.int64)),
            Column("w", "float8", r.random(n) * 1000.0),
            Column("x", "int4", r.integers(-2**31, 2**31 - 1, size=n, dtype=np.int32))]
    src = kds.build_arrow_kds(cols)
    keys = [(abi.KSORT_KEY_KIND__PSUM_NUMERIC, 2, None, False, False),
            (abi.KSORT_KEY_KIND__VREF, 6, "int4", True, False)]
    blob = sort_query(cols, keys)
    s = xpuhip.Session(blob)
    rc, _, _, _ = s.exec(src)
    assert rc == 0
    sort_ms, walls = [], []
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        fin = s.final()
        dt = time.perf_counter() - t0
        st = xpuhip.last_task_stats()
        if i >= warmup:
            sort_ms.append(st["sort_ms"])
            walls.append(dt)
    s.close()
    items, passes = st["sort_items"], st["sort_passes"]
    nwords = 4      # numeric sum: class + two halves; int4 key: one word
    # algorithmic bytes: key pass (row index in, words and item numbers out),
    # per radix pass the digit word read by the histogram plus the words and
    # item numbers read and written by the scatter, the permute's reads and
    # writes
    algo = items * ((8 + 8 * nwords + 4) + passes * (8 + 2 * (8 * nwords + 4)) + (4 + 8 + 8))
    ms = float(np.mean(sort_ms))
    # the oracle's qsort of the same groups
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cpu = {}
    for name, b in (("sorted", blob), ("plain", sort_query(cols, None))):
        o = oracle.OracleSession(b)
        rc, _, _ = o.exec(src, nthreads=min(os.cpu_count() or 1, 16))
        assert rc == 0
        t0 = time.perf_counter()
        o.final()
        cpu[name] = time.perf_counter() - t0
        o.close()
    cpu_sort_s = max(cpu["sorted"] - cpu["plain"], 1e-9)
    return dict(workload="GPU-Sort of %d groups (C1's aggregates over %d rows, key uniform "
                         "over 1M): ORDER BY sum(v) DESC, key" % (items, n),
                sort_ms=ms, final_wall_ms=float(np.mean(walls)) * 1e3, items=items,
                radix_passes=passes, items_per_s=items / (ms / 1e3),
                achieved_gbs=algo / (ms / 1e3) / 1e9, hbm_frac=algo / (ms / 1e3) / HBM_PEAK,
                algo_bytes=algo,
                cpu_baseline={"value": items / cpu_sort_s, "unit": "items/s", "cores": 1,
                              "kind": "port", "sample": "the same %d groups, oracle qsort "
                              "(final with sort keys minus final without)" % items})


def bench_pinned(warmup, steps):
    """C2's three dimensions (2M / 200K / 2,556 rows, resident in HBM as Arrow
    chunks) built on the GPU: one scan session per dimension with its filter
    projects (pk, category, txt) into a pinned inner buffer, and the star
    join's session attaches the three buffers by buffer_id (a re-chained copy
    each, plus the Bloom filters).  A step is that whole build."""
    from pgstrom import datagen, kds, workloads as W, xpuhip
    fact, dims = datagen.star_tables(10)
    chunks = []
    for d in dims:
        a = kds.build_arrow_kds(d)
        chunks.append((xpuhip.DeviceBuffer(len(a), a), len(d[0])))
    blobs = [W.dim_scan_pinned(d, i) for i, d in enumerate(dims)]
    star_blob = W.star_query(fact)
    km = W.star_inner_pinned()
    out = {}

    def step():
        ss = [xpuhip.Session(b) for b in blobs]
        kept = 0
        for s, (dbuf, n) in zip(ss, chunks):
            s.set_tuning(jit=xpuhip.JIT_ON)
            rc, res, _ = s.exec_device(dbuf, n)
            assert rc == 0
            kept += res.nitems_out
        j = xpuhip.Session(star_blob)
        j.attach_inner(km)
        j.close()
        for s in ss:
            s.close()
        out["kept"] = kept

    wall, _, _ = time_steps(step, warmup, steps)
    for dbuf, _ in chunks:
        dbuf.free()
    rows = sum(n for _, n in chunks)
    return dict(workload="C2's dimensions (2M / 200K / 2,556 rows, pk % 10 filters) built on the "
                         "GPU into pinned inner buffers and attached to the star join",
                build_ms=wall * 1e3, dim_rows=rows, rows_kept=out["kept"],
                rows_per_s=rows / wall)


def scan_line(sc):
    return dict(workload="GpuScan 1M rows: WHERE a < 500000 AND b > 0.0 projecting (c, d)",
                rows_per_s=sc["rows"] / (sc["kernel_ms"] / 1e3), kernel_ms=sc["kernel_ms"],
                wall_ms=sc["wall_ms"], hbm_gbs=sc["algo_bytes"] / (sc["kernel_ms"] / 1e3) / 1e9,
                frac=sc["algo_bytes"] / (sc["kernel_ms"] / 1e3) / HBM_PEAK, rows_out=sc["nout"])


def join_line(st):
    return dict(workload="star GpuJoin + GpuPreAgg 200M fact rows, dims 2M/200K/2556 (10% filters), "
                         "GROUP BY (dim1.category, dim3.category) sum(quantity*price)",
                rows_per_s=st["rows"] / (st["kernel_ms"] / 1e3), kernel_ms=st["kernel_ms"],
                wall_ms=st["wall_ms"], hbm_gbs=st["algo_bytes"] / (st["kernel_ms"] / 1e3) / 1e9,
                frac=st["algo_bytes"] / (st["kernel_ms"] / 1e3) / HBM_PEAK,
                host_inner_build_s=st["host_build_s"])


class _DevPtr:
    def __init__(self, ptr):
        self.ptr = ptr


def bench_batch(nfrag, base_rows, warmup, steps):
    """Batched form: nfrag scan + pre-agg fragments (own predicate constant
    a < U[1e5,9e5] and GROUP BY key mod m, m log-uniform in [16,1e5]) over
    row windows of log-uniform size [64K, 64M] of one resident columnar table
    (key int4 Zipf(1.1) over 10,000, v int8, w float8, x int4, a int4); one
    persistent graph-captured launch against one launch per fragment."""
    from pgstrom import datagen, kds, workloads as W, xpuhip
    r = np.random.default_rng(20261202 + 3)
    log("generating %d-row base table for %d fragments" % (base_rows, nfrag))
    cols = datagen.preagg_table(base_rows, seed_workload=3)
    cols = [kds.Column("a", "int4", r.integers(0, 1_000_000, size=base_rows, dtype=np.int32))] + cols
    base = kds.build_arrow_kds(cols)
    sizes = np.exp(r.uniform(np.log(64 << 10), np.log(64 << 20), size=nfrag)).astype(np.int64)
    sizes = np.minimum(sizes, base_rows)
    windows = [(int(r.integers(0, (base_rows - n) // 64 + 1)) * 64, int(n)) for n in sizes]
    stride = 1024
    hdrs = kds.window_headers(base, windows, nfrag * stride, stride)
    dev = xpuhip.DeviceBuffer(len(hdrs) + len(base))
    xpuhip.lib().xpuhip_memcpy_htod(dev.ptr, hdrs.ctypes.data, len(hdrs))
    xpuhip.lib().xpuhip_memcpy_htod(dev.ptr + len(hdrs), base.ctypes.data, len(base))
    del base
    sessions = []
    for i in range(nfrag):
        k = int(r.integers(100_000, 900_001))
        m = int(round(np.exp(r.uniform(np.log(16), np.log(100_000)))))
        s = xpuhip.Session(W.fragment_query(cols, k=k, m=m))
        s.set_tuning(jit=xpuhip.JIT_ON)
        sessions.append(s)
    total = int(sizes.sum())
    tasks = [(s, dev.ptr + i * stride, n) for i, (s, (_, n)) in enumerate(zip(sessions, windows))]

    def plain():
        for s, p, n in tasks:
            rc, _, _ = s.exec_device(_DevPtr(p), n)
            assert rc == 0

    def batched():
        xpuhip.batch_exec(tasks, graph=True)

    log("autotuning the batch")
    tuned = xpuhip.batch_autotune(tasks)
    out = {"tuning": tuned}
    for name, fn in (("plain", plain), ("batched", batched)):
        for s in sessions:
            s.reset()
        for _ in range(warmup):
            fn()
        xpuhip.lib().xpuhip_device_synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        xpuhip.lib().xpuhip_device_synchronize()
        wall = (time.perf_counter() - t0) / steps
        out[name] = dict(rows_per_s=total / wall, ms_per_step=wall * 1e3)
    for s in sessions:
        s.close()
    dev.free()
    out["fragments"] = nfrag
    out["rows_per_step"] = total
    out["speedup"] = out["batched"]["rows_per_s"] / out["plain"]["rows_per_s"]
    return out


def bench_served(nreq, nback, min_rows, max_rows, table_rows, slots=8, max_batch=32):
    """Served form against the one-query-at-a-time form on the same
    open-loop Poisson arrivals of nreq queries (60% scans, 25% pre-aggs, 15%
    star joins over row windows of log-uniform size of tables resident in
    HBM).  Served: nback backends with sessions (and dimension hash tables)
    kept across their queries, each running its queries one at a time,
    through the request loop with `slots` slots and coalesced group-by
    fragments.  One query at a time: each query opens its own session,
    attaches its dimension tables, runs, reads its result and closes.
    Loads: 0.8x and 3x the one-query-at-a-time capacity."""
    from pgstrom import serving, xpuhip
    log("served: generating resident tables (%d rows each)" % table_rows)
    res = serving.Resident(scan_rows=table_rows // 2, preagg_rows=table_rows, fact_rows=table_rows,
                           dim_sizes=(2_000_000, 200_000, 2_556), nwin=nreq + 8)
    reqs = serving.make_requests(res, nreq, min_rows, max_rows)
    rows = sum(q["n"] for q in reqs)
    dsts = [xpuhip.DeviceBuffer(serving.dst_bytes(q["n"])) if q["kind"] == "scan" else None
            for q in reqs]
    backends = serving.Backends(res, nback)

    def summary(st, span, stats):
        assert all(x.rc == 0 for x in st), [x.err for x in st if x.rc][:3]
        lat = np.array([x.total_ms for x in st])
        per_kind = {}
        for kind in serving.KINDS:
            ex = [x.exec_ms for x, q in zip(st, reqs) if q["kind"] == kind]
            rk = sum(q["n"] for q in reqs if q["kind"] == kind)
            per_kind[kind] = dict(requests=len(ex), exec_ms_mean=float(np.mean(ex)) if ex else 0.0,
                                  rows=rk)
        out = dict(rows_per_s=rows / span, req_per_s=len(reqs) / span, makespan_s=span,
                   p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                   per_kind=per_kind)
        if stats:
            out.update(launches=stats["launches"], coalesced=stats["coalesced"])
        return out

    def served(arrivals):
        srv = xpuhip.Server(hbm_budget=200 << 30, slots=slots, max_batch=max_batch)
        st, span = serving.run(srv, backends, reqs, arrivals, dsts)
        stats = srv.stats()
        srv.close()
        return summary(st, span, stats)

    def plain(arrivals):
        st, span = serving.run_plain(res, reqs, arrivals, dsts)
        return summary(st, span, None)

    # warm: specialised kernels of every kind compile once
    served([0.0] * len(reqs))
    # capacities: every query arrives at once
    cap = plain([0.0] * len(reqs))
    cap_served = served([0.0] * len(reqs))
    mu = cap["req_per_s"]
    rng = np.random.default_rng(20261202 + 5)
    out = {"requests": nreq, "backends": nback, "rows": rows, "slots": slots, "max_batch": max_batch,
           "plain_capacity_req_per_s": mu, "plain_capacity_rows_per_s": cap["rows_per_s"],
           "served_capacity_req_per_s": cap_served["req_per_s"],
           "served_capacity_rows_per_s": cap_served["rows_per_s"],
           "served_vs_plain_capacity": cap_served["rows_per_s"] / cap["rows_per_s"]}
    for load in (0.8, 3.0):
        gaps = rng.exponential(1.0 / (load * mu), size=len(reqs))
        arr = list(np.cumsum(gaps) - gaps[0])
        out["load_%.1fx" % load] = {"plain": plain(arr), "served": served(arr)}
    for d in dsts:
        if d is not None:
            d.free()
    backends.close()
    res.free()
    hi = out["load_3.0x"]
    out["served_vs_plain_rows_per_s_at_3x"] = hi["served"]["rows_per_s"] / hi["plain"]["rows_per_s"]
    out["served_p99_at_3x_vs_plain_p99_at_0.8x"] = (hi["served"]["p99_ms"] /
                                                    out["load_0.8x"]["plain"]["p99_ms"])
    return out


def traffic_of(name):
    """HBM bytes per launch of the dominant kernel, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/ (null if absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    return t.get("traffic_bytes"), t.get("source")


def copy_probe(nbytes):
    """A plain device-to-device copy of known size, for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters (run under --pmc)."""
    from pgstrom import xpuhip
    xpuhip.init(0)
    a = xpuhip.DeviceBuffer(nbytes)
    b = xpuhip.DeviceBuffer(nbytes)
    for _ in range(3):
        xpuhip.lib().xpuhip_memcpy_dtod(b.ptr, a.ptr, nbytes)
    xpuhip.lib().xpuhip_device_synchronize()
    a.free()
    b.free()
    print(json.dumps({"copy_probe_bytes": nbytes}))


def host_cpu():
    """nproc, the CPUs this process may run on, the CPU model, and whether a
    PostgreSQL server binary is installed."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    import glob
    import shutil
    pg = shutil.which("postgres") or shutil.which("pg_ctl") or \
        (sorted(glob.glob("/usr/lib/postgresql/*/bin/postgres")) or [None])[-1]
    return dict(nproc=os.cpu_count() or 1, usable=usable, model=model, postgres=pg)


def cpu_baseline(cols_fn, sample_rows, min_seconds=10.0):
    """The oracle (own CPU restatement of the reference's group-by, OpenMP)
    on a bounded sample of the same table and query: repeated passes over one
    sample chunk with every CPU this process may use, for at least
    min_seconds, and one pass on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from pgstrom import datagen, kds, workloads as W
    cols = datagen.preagg_table(sample_rows, fragment=1)
    src = kds.build_arrow_kds(cols)
    blob = W.preagg_query(cols)
    host = host_cpu()
    ncores = host["usable"]
    o = oracle.OracleSession(blob)
    passes = 0
    t0 = time.perf_counter()
    while True:
        rc, _, _ = o.exec(src, nthreads=ncores)
        assert rc == 0
        passes += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    o.close()
    o = oracle.OracleSession(blob)
    t1 = time.perf_counter()
    rc, _, _ = o.exec(src, nthreads=1)
    dt1 = time.perf_counter() - t1
    o.close()
    assert rc == 0
    return dict(value=passes * sample_rows / dt, unit="rows/s", cores=ncores, kind="port",
                sample="%d passes over %d rows of the same pre-aggregation table and query, "
                       "%.1f s" % (passes, sample_rows, dt),
                single_core_rows_per_s=sample_rows / dt1, nproc=host["nproc"],
                cpu_model=host["model"], postgres_installed=host["postgres"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--only", default="all",
                    choices=["all", "preagg", "batch", "served", "join", "scan", "heap", "pinned",
                             "sort"])
    ap.add_argument("--served-requests", type=int, default=300)
    ap.add_argument("--served-table-rows", type=int, default=50_000_000)
    ap.add_argument("--served-slots", type=int, default=8)
    ap.add_argument("--served-max-batch", type=int, default=32)
    ap.add_argument("--batch-frags", type=int, default=256)
    ap.add_argument("--batch-base-rows", type=int, default=1 << 26)
    ap.add_argument("--cpu-sample", type=int, default=20_000_000,
                    help="rows of the CPU baseline sample (0: skip)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device copy of this many bytes (counter scaling)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is described in DESIGN.md, not run)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from pgstrom import xpuhip
    xpuhip.init(0)

    if args.only == "batch":
        print(json.dumps(bench_batch(args.batch_frags, args.batch_base_rows, 1, 3)))
        return
    if args.only == "join":
        st = bench_star(200_000_000, args.warmup, args.steps)
        print(json.dumps(join_line(st)))
        return
    if args.only == "scan":
        print(json.dumps(scan_line(bench_scan(1_000_000, args.warmup, args.steps))))
        return
    if args.only == "pinned":
        print(json.dumps(bench_pinned(args.warmup, args.steps)))
        return
    if args.only == "sort":
        print(json.dumps(bench_sort(3_000_000, args.warmup, args.steps)))
        return
    if args.only == "heap":
        print(json.dumps(heap_line(bench_heap(1_000_000, args.warmup, args.steps))))
        return
    if args.only == "served":
        print(json.dumps(bench_served(args.served_requests, 32, 1_000_000, 50_000_000,
                                      args.served_table_rows, args.served_slots,
                                      args.served_max_batch)))
        return
    pa = bench_preagg(args.rows, args.warmup, args.steps)
    kms = pa["kernel_ms"]
    # dominant kernel: the scan + group-by pass, which reads every column byte
    achieved = pa["algo_bytes"] / (pa["pass1_ms"] / 1e3)
    traffic, traffic_src = traffic_of("preagg_c1_traffic.json")
    others = {}
    if args.only == "all":
        others["scan_1M"] = scan_line(bench_scan(1_000_000, 3, 20))
        others["scan_1M_heap_pages"] = heap_line(bench_heap(1_000_000, 3, 20))
        others["batched_256"] = bench_batch(args.batch_frags, args.batch_base_rows, 1, 3)
        others["served"] = bench_served(args.served_requests, 32, 1_000_000, 50_000_000,
                                        args.served_table_rows)
        others["star_join_200M"] = join_line(bench_star(200_000_000, 1, 5))
        others["pinned_inner_build"] = bench_pinned(2, 10)
        others["gpu_sort_1M_groups"] = bench_sort(3_000_000, 2, 10)
    cpu = None
    if args.cpu_sample > 0:
        log("cpu baseline")
        cpu = cpu_baseline(None, args.cpu_sample)

    line = {
        "metric": "rows/s & %HBM-peak per scan/join/preagg; p99 latency, batched vs plain",
        "value": pa["rows"] / (pa["wall_ms"] / 1e3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pa["wall_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+f64",
        "data": "synthetic (seeded numpy PCG64), resident in HBM",
        "config": {"workload": "GpuPreAgg 100M rows: key int4 Zipf(1.1) over 10,000, v int8, "
                               "w float8, x int4; GROUP BY key count(*), sum(v), avg(w), min(x), max(x)",
                   "rows": pa["rows"], "groups": pa["groups"], "tuning": pa["tuning"]},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "xpu_jit_kernel" if pa["jit"] else "xpu_scan_agg_kernel",
                     "kernel_ms": pa["pass1_ms"], "spill_pass_ms": pa["spill_ms"],
                     "all_passes_ms": kms, "algo_bytes_per_launch": pa["algo_bytes"],
                     "algo_bytes_per_row": 24, "spilled_row_frac": pa["spill_frac"],
                     "lds_table_entries": pa["lds_nents"]},
        "cpu_baseline": cpu,
        "workloads": others,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
