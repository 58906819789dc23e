"""xpuhip_task_exec_device_noalloc, the entry point that allocates nothing,
against xpuhip_task_exec_device on the same device-resident chunks: kds_dst
bytes equal for a scan + projection (C0's query), group states equal for a
scan + group-by (C1's query), on two calls in a row (the second with a
smaller chunk, so the shared scratch holds the first call's contents), and a
chunk beyond the scratch's limit is refused before any launch."""
import numpy as np
import pytest

from pgstrom import abi, datagen, kds, workloads as W, xpuhip
from helpers import assert_groups_equal, run_oracle_groupby

pytestmark = pytest.mark.gpu

C1_ACTIONS = [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__PAVG_FP,
              abi.KAGG_ACTION__PMIN_INT32, abi.KAGG_ACTION__PMAX_INT32, abi.KAGG_ACTION__VREF]


def _scan_once(blob, src, n, noalloc, dst_length=16 << 20):
    s = xpuhip.Session(blob)
    d_src = xpuhip.DeviceBuffer(len(src), src)
    d_dst = xpuhip.DeviceBuffer(dst_length)
    run = s.exec_device_noalloc if noalloc else s.exec_device
    rc, res, pos = run(d_src, n, d_dst, dst_length)
    assert rc == 0 and pos == 0
    out = d_dst.to_host()
    d_src.free()
    d_dst.free()
    s.close()
    return out, res.nitems_out


@pytest.mark.parametrize("nrows", [[300_000, 20_000], [70_000, 1]])
def test_noalloc_scan_projection(gpu, nrows):
    for i, n in enumerate(nrows):
        print("noalloc scan: call %d, %d rows" % (i, n), flush=True)
        cols = datagen.scan_table(n, fragment=500 + i)
        src = kds.build_arrow_kds(cols)
        blob = W.scan_query(cols)
        exp, nexp = _scan_once(blob, src, n, noalloc=False)
        got, ngot = _scan_once(blob, src, n, noalloc=True)
        assert ngot == nexp
        assert np.array_equal(got, exp)


def test_noalloc_scan_resume(gpu):
    """kds_dst too small for the chunk: 'd' and a resume position, as the
    allocating entry point returns them."""
    cols = datagen.scan_table(60_000, fragment=510)
    src = kds.build_arrow_kds(cols)
    blob = W.scan_query(cols)
    outs = {}
    for noalloc in (False, True):
        s = xpuhip.Session(blob)
        d_src = xpuhip.DeviceBuffer(len(src), src)
        d_dst = xpuhip.DeviceBuffer(256 << 10)
        run = s.exec_device_noalloc if noalloc else s.exec_device
        pos, bufs = 0, []
        for _ in range(100):
            rc, _, pos = run(d_src, 60_000, d_dst, 256 << 10, pos)
            bufs.append(d_dst.to_host())
            if rc == 0:
                break
            assert rc == abi.ERRCODE_SUSPEND_NO_SPACE and pos > 0
        outs[noalloc] = bufs
        d_src.free()
        d_dst.free()
        s.close()
    assert len(outs[True]) == len(outs[False]) > 1
    for a, b in zip(outs[True], outs[False]):
        # header, row index and the tuples it addresses; bytes past them may
        # hold what tiles beyond the resume point wrote before the stop
        ha, _ = kds.kds_colmeta(memoryview(bytes(a)))
        hb, _ = kds.kds_colmeta(memoryview(bytes(b)))
        assert (ha.nitems, ha.usage, ha.length) == (hb.nitems, hb.usage, hb.length)
        assert kds.parse_row_kds(a) == kds.parse_row_kds(b)


def _groupby_once(blob, src, n, noalloc):
    s = xpuhip.Session(blob)
    d_src = xpuhip.DeviceBuffer(len(src), src)
    run = s.exec_device_noalloc if noalloc else s.exec_device
    rc, res, _ = run(d_src, n)
    assert rc == 0
    fin = s.final()
    d_src.free()
    s.close()
    return fin, res


@pytest.mark.parametrize("nrows", [[2_000_000, 50_000]])
def test_noalloc_groupby(gpu, nrows):
    for i, n in enumerate(nrows):
        print("noalloc group-by: call %d, %d rows" % (i, n), flush=True)
        cols = datagen.preagg_table(n, fragment=40 + i)
        src = kds.build_arrow_kds(cols)
        blob = W.preagg_query(cols)
        exp, rexp = _groupby_once(blob, src, n, noalloc=False)
        got, rgot = _groupby_once(blob, src, n, noalloc=True)
        assert rgot.nitems_in == rexp.nitems_in == n
        assert_groups_equal(got, exp, C1_ACTIONS)
        if n <= 50_000:
            assert_groups_equal(got, run_oracle_groupby(blob, [src]), C1_ACTIONS)


def test_noalloc_refuses_beyond_scratch(gpu):
    """A projection chunk whose tile totals exceed the scratch array is
    refused on the host, before any launch."""
    cols = datagen.scan_table(1000, fragment=520)
    src = kds.build_arrow_kds(cols)
    s = xpuhip.Session(W.scan_query(cols))
    d_src = xpuhip.DeviceBuffer(len(src), src)
    d_dst = xpuhip.DeviceBuffer(1 << 20)
    with pytest.raises(xpuhip.XpuError) as e:
        s.exec_device_noalloc(d_src, 4_000_000_000, d_dst, 1 << 20)
    assert "exceeds the no-alloc scratch" in str(e.value)
    d_src.free()
    d_dst.free()
    s.close()


def test_noalloc_refuses_joins(gpu):
    """A session with joins is refused (the entry point runs scans only)."""
    fact, dims = datagen.star_tables(1000, dim_sizes=(100, 50, 20))
    s = xpuhip.Session(W.star_query(fact))
    with pytest.raises(xpuhip.XpuError) as e:
        s.exec_device_noalloc(None, 0)
    assert "scan + projection and scan + group-by sessions only" in str(e.value)
    s.close()
