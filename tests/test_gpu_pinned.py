"""Pinned inner buffers on the GPU: the inner side of a join built by GPU
scans whose projections stay on the device as KDS_FORMAT_HASH buffers
(DEVTASK__PINNED_HASH_RESULTS), named by the join's kern_multirels chunks
(pinned_buffer, buffer_id).  The buffers hold exactly the oracle's
projection rows with their join-key hashes, and the join over them gives the
results of the join over the host-built inner tables and of the oracle."""
import numpy as np
import pytest

from pgstrom import abi, datagen, hashing, kds, workloads as W, xpuhip
from helpers import skip_interpreter_join_groupby, assert_groups_equal, run_gpu_groupby, run_oracle_groupby, run_oracle_projection

pytestmark = pytest.mark.gpu

C2_ACTIONS = [abi.KAGG_ACTION__PSUM_INT64, abi.KAGG_ACTION__VREF, abi.KAGG_ACTION__VREF]
DIM_SIZES = (200_000, 20_000, 2_556)


def _split(cols, parts):
    n = len(cols[0])
    edges = np.linspace(0, n, parts + 1).astype(int)
    return [datagen.take_rows(cols, (np.arange(n) >= a) & (np.arange(n) < b))
            for a, b in zip(edges[:-1], edges[1:])]


def _pinned_dims(dims, jit, parts=2):
    """One session per dimension (two for the first: parallel workers of one
    plan share its buffer), each chunk of the dimension through them."""
    sessions = []
    for i, dim in enumerate(dims):
        blob = W.dim_scan_pinned(dim, i)
        workers = [xpuhip.Session(blob) for _ in range(2 if i == 0 else 1)]
        for w in workers:
            w.set_tuning(jit=jit)
        for j, chunk in enumerate(_split(dim, parts)):
            rc, _, res, _ = workers[j % len(workers)].exec(kds.build_arrow_kds(chunk))
            assert rc == 0
        sessions.append((blob, workers))
    return sessions


def test_pinned_buffer_contents(gpu, jit):
    """Each pinned buffer holds the rows the oracle projects from the same
    scan, each item carrying HashValue(pk) in kern_tupitem.hash."""
    _, dims = datagen.star_tables(10, dim_sizes=DIM_SIZES)
    sessions = _pinned_dims(dims, jit)
    try:
        for i, (blob, workers) in enumerate(sessions):
            got = kds.parse_row_kds(workers[0].final())
            exp_rows = []
            for chunk in _split(dims[i], 2):
                exp, _ = run_oracle_projection(blob, kds.build_arrow_kds(chunk), 32 << 20)
                exp_rows += [r[0] for r in kds.parse_row_kds(exp)]
            assert sorted(r[0] for r in got) == sorted(exp_rows)
            assert sorted(r[1] for r in got) == list(range(len(got)))
            # the hashes: parse the buffer's items directly
            raw = memoryview(bytes(workers[0].final()))
            hd, cms = kds.kds_colmeta(raw)
            ridx = np.frombuffer(raw, dtype=np.uint64, count=hd.nitems,
                                 offset=abi.kds_head_length(hd.nr_colmeta))
            pks = np.array([kds.decode_tupitem(cms, raw, hd.length - int(o))[0][0] for o in ridx],
                           dtype=np.int32)
            hv = [kds.decode_tupitem(cms, raw, hd.length - int(o))[2] for o in ridx]
            assert hv == [int(h) for h in hashing.hash_keys_np([(pks, 4)])]
    finally:
        for _, workers in sessions:
            for w in workers:
                w.close()


@pytest.mark.parametrize("left_outer", [False, True])
def test_star_join_over_pinned_inners(gpu, jit, left_outer):
    """C2's star join with its three dimensions built on the GPU: the same
    groups as with the host-built inner tables, and as the oracle's."""
    skip_interpreter_join_groupby(jit)
    fact, dims = datagen.star_tables(300_000, dim_sizes=DIM_SIZES)
    src = kds.build_arrow_kds(fact)
    blob = W.star_query(fact, left_outer=left_outer)
    sessions = _pinned_dims(dims, jit)
    try:
        got = run_gpu_groupby(blob, [src], kmrels=W.star_inner_pinned(left_outer=left_outer),
                              tuning=dict(jit=jit))
    finally:
        for _, workers in sessions:
            for w in workers:
                w.close()
    km = W.star_inner(dims)
    if left_outer:
        km = kds.build_multirels([(k, True) for k, _ in _inner_kds(dims)])
    host = run_gpu_groupby(blob, [src], kmrels=km, tuning=dict(jit=jit))
    exp = run_oracle_groupby(blob, [src], kmrels=km)
    assert_groups_equal(got, exp, C2_ACTIONS)
    assert_groups_equal(host, exp, C2_ACTIONS)


def _inner_kds(dims):
    """(KDS_FORMAT_HASH bytes, left_outer) of each dimension, as star_inner
    builds them."""
    coldefs = [("pk", "int4"), ("category", "int2"), ("txt", "text")]
    out = []
    for i, (pk, cat, txt) in enumerate(dims):
        mod, rem = W.dim_filter_mod(i)
        sel = np.nonzero(pk.values % mod == rem)[0]
        hashes = hashing.hash_keys_np([(pk.values[sel], 4)])
        rows = [(int(pk.values[j]), int(cat.values[j]), bytes(txt.values[j])) for j in sel]
        out.append((kds.build_hash_kds(coldefs, rows, [int(h) for h in hashes]), False))
    return out


def test_pinned_buffer_full_is_an_error(gpu):
    """A pinned buffer too small for the projected rows fails the task."""
    _, dims = datagen.star_tables(10, dim_sizes=(200_000, 10, 10))
    s = xpuhip.Session(W.dim_scan_pinned(dims[0], 0, pinned_length=1 << 20))
    with pytest.raises(xpuhip.XpuError) as ei:
        s.exec(kds.build_arrow_kds(dims[0]))
    s.close()
    assert "pinned inner buffer is full" in str(ei.value)


def test_missing_pinned_buffer_is_an_error(gpu):
    """A kern_multirels chunk naming a buffer no open session holds is refused."""
    fact, _ = datagen.star_tables(1000, dim_sizes=(100, 100, 100))
    s = xpuhip.Session(W.star_query(fact))
    with pytest.raises(xpuhip.XpuError) as ei:
        s.attach_inner(W.star_inner_pinned())
    s.close()
    assert "was not found" in str(ei.value)
