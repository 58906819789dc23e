"""The served form: requests of many backends admitted by the GPU service's
request loop (continuous batching, coalesced group-by fragments, work of
admitted requests overlapping on the device) give, request by request, the
result of running each alone."""
import numpy as np
import pytest

from pgstrom import abi, kds, serving, xpuhip
from helpers import assert_groups_equal

pytestmark = pytest.mark.gpu

C1_ACTIONS = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__PAVG_FP,
              abi.KAGG_ACTION__PMIN_INT32, abi.KAGG_ACTION__PMAX_INT32, abi.KAGG_ACTION__VREF]
C2_ACTIONS = [abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF, abi.KAGG_ACTION__VREF]


@pytest.fixture(scope="module")
def resident(gpu):
    res = serving.Resident(scan_rows=400_000, preagg_rows=1_500_000, fact_rows=1_500_000,
                           dim_sizes=(20_000, 2_000, 2_556), nwin=128)
    yield res
    res.free()


def results_of(reqs, backends, dsts):
    out = []
    for i, q in enumerate(reqs):
        s = backends.sess[i][q["kind"]]
        if q["kind"] == "scan":
            out.append([r for r in kds.parse_row_kds(dsts[i].to_host())])
        else:
            out.append(s.final())
    return out


def run_form(resident, reqs, slots, max_batch, burst):
    backends = serving.Backends(resident, len(reqs))
    dsts = [xpuhip.DeviceBuffer(serving.dst_bytes(q["n"])) if q["kind"] == "scan" else None
            for q in reqs]
    srv = xpuhip.Server(hbm_budget=200 << 30, slots=slots, max_batch=max_batch)
    if burst:
        arrivals = [0.0] * len(reqs)
        st, _ = serving.run(srv, backends, reqs, arrivals, dsts)
    else:
        st = []
        for i, q in enumerate(reqs):
            one, _ = serving.run(srv, _Single(backends, i), [q], [0.0], [dsts[i]])
            st += one
    stats = srv.stats()
    srv.close()
    assert all(s.rc == 0 for s in st), [s.errbuf.message for s in st if s.rc]
    got = results_of(reqs, backends, dsts)
    backends.close()
    for d in dsts:
        if d is not None:
            d.free()
    return got, st, stats


class _Single:
    """The sessions of request i alone (its own backend)."""

    def __init__(self, backends, i):
        self.sess = [backends.sess[i]]


def test_served_equals_alone(resident):
    reqs = serving.make_requests(resident, 40, 2_000, 600_000, seed=7)
    alone, _, _ = run_form(resident, reqs, slots=1, max_batch=1, burst=False)
    mix, st, stats = run_form(resident, reqs, slots=8, max_batch=16, burst=True)
    assert stats["coalesced"] > 0, stats     # group-by fragments shared launches
    for q, a, b in zip(reqs, alone, mix):
        if q["kind"] == "scan":
            assert a == b
        else:
            assert_groups_equal(b, a, C1_ACTIONS if q["kind"] == "preagg" else C2_ACTIONS)
