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


def run_served(resident, reqs, slots, max_batch):
    backends = serving.Backends(resident, len(reqs))
    dsts = [xpuhip.DeviceBuffer(serving.dst_bytes(q["n"])) if q["kind"] == "scan" else None
            for q in reqs]
    srv = xpuhip.Server(hbm_budget=200 << 30, slots=slots, max_batch=max_batch)
    st, _ = serving.run(srv, backends, reqs, [0.0] * len(reqs), dsts, keep_results=True)
    stats = srv.stats()
    srv.close()
    backends.close()
    return st, dsts, stats


def run_alone(resident, reqs):
    dsts = [xpuhip.DeviceBuffer(serving.dst_bytes(q["n"])) if q["kind"] == "scan" else None
            for q in reqs]
    st, _ = serving.run_plain(resident, reqs, [0.0] * len(reqs), dsts, keep_results=True)
    return st, dsts


def test_served_equals_alone(resident):
    """Queries of many backends arriving at once (so group-by fragments of
    different backends share launches) give each query the result of
    running it alone, in its own session."""
    reqs = serving.make_requests(resident, 40, 2_000, 600_000, seed=7)
    alone, adst = run_alone(resident, reqs)
    mix, mdst, stats = run_served(resident, reqs, slots=8, max_batch=16)
    assert all(s.rc == 0 for s in alone), [s.err for s in alone if s.rc]
    assert all(s.rc == 0 for s in mix), [s.err for s in mix if s.rc]
    assert stats["coalesced"] > 0, stats     # group-by fragments shared launches
    for q, a, b, da, db in zip(reqs, alone, mix, adst, mdst):
        if q["kind"] == "scan":
            assert kds.parse_row_kds(da.to_host()) == kds.parse_row_kds(db.to_host())
        else:
            assert_groups_equal(b.result, a.result, C1_ACTIONS if q["kind"] == "preagg" else C2_ACTIONS)
    for d in adst + mdst:
        if d is not None:
            d.free()
