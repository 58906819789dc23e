"""The served form: request streams of many backends through the GPU
service's request loop (xpuhip_server), against the one-query-at-a-time
service on the same requests.

A request is one query fragment of a backend: a scan (WHERE a < 500000 AND
b > 0.0 projecting (c, d)), a pre-aggregation (GROUP BY key: count, sum,
avg, min, max) or a star join + pre-aggregation, over a row window of a
table resident in HBM (window_headers), with the backend's session (dimension
hash tables attached once) kept open across its requests."""
import threading
import time

import numpy as np

from . import datagen, kds, workloads as W, xpuhip

KINDS = ("scan", "preagg", "join")


class Resident:
    """Base tables in device memory and KDS headers of row windows over them."""

    STRIDE = 2048

    def __init__(self, scan_rows, preagg_rows, fact_rows, dim_sizes, nwin, seed=0):
        self.seed = seed
        self.tables = {}
        self.scan_cols = datagen.scan_table(scan_rows, fragment=seed)
        self.preagg_cols = datagen.preagg_table(preagg_rows, fragment=seed)
        self.fact, self.dims = datagen.star_tables(fact_rows, dim_sizes=dim_sizes)
        self.kmrels = W.star_inner(self.dims)
        self.nwin = nwin
        for kind, cols in (("scan", self.scan_cols), ("preagg", self.preagg_cols),
                           ("join", self.fact)):
            base = kds.build_arrow_kds(cols)
            hdr_bytes = nwin * self.STRIDE
            dev = xpuhip.DeviceBuffer(hdr_bytes + len(base))
            xpuhip.lib().xpuhip_memcpy_htod(dev.ptr + hdr_bytes, base.ctypes.data, len(base))
            self.tables[kind] = dict(dev=dev, base=base, rows=len(cols[0]), used=0,
                                     hdr_bytes=hdr_bytes)

    def window(self, kind, r0, n):
        """Device address of a KDS over rows [r0, r0 + n) of a table."""
        t = self.tables[kind]
        i = t["used"]
        assert i < self.nwin
        t["used"] += 1
        h = kds.window_headers(t["base"], [(r0, n)], t["hdr_bytes"] - i * self.STRIDE,
                               self.STRIDE)
        xpuhip.lib().xpuhip_memcpy_htod(t["dev"].ptr + i * self.STRIDE, h.ctypes.data, len(h))
        return t["dev"].ptr + i * self.STRIDE

    def blob(self, kind):
        if kind == "scan":
            return W.scan_query(self.scan_cols)
        if kind == "preagg":
            return W.preagg_query(self.preagg_cols)
        return W.star_query(self.fact)

    def free(self):
        for t in self.tables.values():
            t["dev"].free()


def make_requests(res, nreq, min_rows, max_rows, mix=(0.60, 0.25, 0.15), seed=1):
    """nreq requests: kind by the mix, log-uniform window sizes (64-row
    aligned starts) within each table."""
    r = np.random.default_rng(20261202 + 4 + seed)
    out = []
    for _ in range(nreq):
        kind = KINDS[int(r.choice(3, p=mix))]
        rows = res.tables[kind]["rows"]
        hi = min(max_rows, rows)
        n = int(np.exp(r.uniform(np.log(min_rows), np.log(hi))))
        r0 = int(r.integers(0, (rows - n) // 64 + 1)) * 64
        out.append(dict(kind=kind, r0=r0, n=n, src=res.window(kind, r0, n)))
    return out


class Backends:
    """Sessions of nback backends (one per kind each, kept open) and one
    kds_dst per scan request."""

    def __init__(self, res, nback, tuning=None):
        self.sess = []
        for _ in range(nback):
            d = {}
            for kind in KINDS:
                s = xpuhip.Session(res.blob(kind))
                if kind == "join":
                    s.attach_inner(res.kmrels)
                if tuning:
                    s.set_tuning(**tuning)
                d[kind] = s
            self.sess.append(d)

    def close(self):
        for d in self.sess:
            for s in d.values():
                s.close()


def dst_bytes(n):
    """kds_dst large enough for every row of an n-row scan window."""
    return 4096 + n * 96


class Status:
    """Outcome of one query: return code, times (ms) and, for a group-by,
    its kds_final as read back by the backend."""

    def __init__(self, rc, queue_ms, exec_ms, total_ms, batched=0, result=None, err=""):
        self.rc = rc
        self.queue_ms = queue_ms
        self.exec_ms = exec_ms
        self.total_ms = total_ms
        self.batched = batched
        self.result = result
        self.err = err


def _wait_until(t):
    delay = t - time.perf_counter()
    if delay > 0:
        time.sleep(delay)


def run(server, backends, reqs, arrivals, dsts, keep_results=False):
    """The served form, open loop: request i arrives at arrivals[i] seconds
    at backend i % nback.  A backend runs its queries one at a time, as a
    PostgreSQL backend does: it submits a query once it has arrived and the
    backend's previous query is done, waits for it, and reads a group-by's
    result back (and clears its session for the next query).  Returns the
    per-request statuses in request order and the makespan."""
    nback = len(backends.sess)
    done = [None] * len(reqs)
    shape_of = {k: i + 1 for i, k in enumerate(KINDS)}
    t0 = time.perf_counter()

    def backend(b):
        for i in range(b, len(reqs), nback):
            q = reqs[i]
            _wait_until(t0 + arrivals[i])
            s = backends.sess[b][q["kind"]]
            if q["kind"] == "scan":
                d = dsts[i]
                tk = server.submit(s, q["src"], q["n"], d.ptr, d.nbytes,
                                   hbm_bytes=q["n"] * 40 + d.nbytes, shape=shape_of["scan"])
            else:
                tk = server.submit(s, q["src"], q["n"], hbm_bytes=q["n"] * 28,
                                   shape=shape_of[q["kind"]], batchable=(q["kind"] == "preagg"))
            st = server.wait(tk)
            result = None
            if st.rc == 0 and q["kind"] != "scan":
                result = s.final()
                s.reset()
            total = (time.perf_counter() - (t0 + arrivals[i])) * 1e3
            done[i] = Status(st.rc, st.queue_ms, st.exec_ms, total, st.batched,
                             result if keep_results else None,
                             st.errbuf.message.decode(errors="replace") if st.rc else "")

    threads = [threading.Thread(target=backend, args=(b,)) for b in range(nback)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    return done, time.perf_counter() - t0


def run_plain(res, reqs, arrivals, dsts, keep_results=False, tuning=None):
    """The one-query-at-a-time form, open loop: queries run one after
    another in arrival order, each as its own session (opened, its dimension
    tables attached, run, its group-by result read back, closed), with no
    state kept from one query to the next.  Returns the statuses in request
    order and the makespan."""
    done = [None] * len(reqs)
    blobs = {k: res.blob(k) for k in KINDS}
    t0 = time.perf_counter()
    for i, q in enumerate(reqs):
        _wait_until(t0 + arrivals[i])
        start = time.perf_counter()
        s = xpuhip.Session(blobs[q["kind"]])
        if tuning:
            s.set_tuning(**tuning)
        if q["kind"] == "join":
            s.attach_inner(res.kmrels)
        t1 = time.perf_counter()
        if q["kind"] == "scan":
            d = dsts[i]
            rc, _, _ = s.exec_device(_Addr(q["src"]), q["n"], d, d.nbytes)
        else:
            rc, _, _ = s.exec_device(_Addr(q["src"]), q["n"])
        result = s.final() if rc == 0 and q["kind"] != "scan" else None
        s.close()
        end = time.perf_counter()
        done[i] = Status(rc, (start - (t0 + arrivals[i])) * 1e3, (end - t1) * 1e3,
                         (end - (t0 + arrivals[i])) * 1e3, 0, result if keep_results else None)
    return done, time.perf_counter() - t0


class _Addr:
    """A device address where the session API takes a device buffer."""

    def __init__(self, ptr):
        self.ptr = ptr
