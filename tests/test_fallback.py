"""CPU fallback on the oracle (fallback.c / HandleErrorIfCpuFallback,
xpu_common.cu:1898-1955) against closed forms: which rows fall back, the
bytes of their kern_fallbackitems, the rows that stay on the GPU path, and
the error without a fallback buffer.  Compressed and external texts are
carried as they are (xpu_text_datum_heap_read with length -1)."""
import pytest

from pgstrom import abi, kds, workloads as W
from helpers import rows_of
import fallbackcases as F
import oracle


def run(blob, src, nthreads=3):
    o = oracle.OracleSession(blob)
    rc, dst, res = o.exec(src, dst_length=4 << 20, nthreads=nthreads)
    fb = o.fallback()
    err = o.err.describe() if rc else ""
    o.close()
    return rc, dst, fb, err


def fb_items(fb):
    return [] if fb is None else [it[:4] for it in kds.parse_fallback_kds(fb)]


@pytest.mark.parametrize("nthreads", [1, 4])
def test_textlen_falls_back(nthreads):
    rs = F.rows()
    cols, src = F.chunk(rs)
    rc, dst, fb, err = run(F.q_textlen(), src, nthreads)
    assert rc == 0, err
    exp_rows = [(a, c, d) for a, g, c, d in rs
                if g is not None and not F.is_raw(g) and len(g) > 5]
    assert [tuple(r) for r in rows_of(dst)] == exp_rows
    items = kds.parse_fallback_kds(fb)
    assert [it[:4] for it in items] == F.expected_fallback(rs)
    assert [it[4] for it in items] == list(range(len(items)))     # row ids: their index


def test_textlen_single_byte_encoding_reads_the_header():
    """Under a single-byte encoding textlen of a compressed or external datum
    is its header's size: nothing falls back."""
    rs = F.rows()
    cols, src = F.chunk(rs)
    rc, dst, fb, err = run(F.q_textlen(encoding="SQL_ASCII"), src)
    assert rc == 0, err
    exp = [(a, c, d) for a, g, c, d in rs
           if g is not None and (F.raw_textlen(g) if F.is_raw(g) else len(g)) > 5]
    assert [tuple(r) for r in rows_of(dst)] == exp
    assert fb is None


def test_compare_falls_back_projection_copies():
    """g <> 'zz' needs the bytes (fallback); a projected g of a row that
    passes is copied as it is."""
    rs = F.rows()
    cols, src = F.chunk(rs)
    rc, dst, fb, err = run(F.q_texteq(), src)
    assert rc == 0, err
    assert fb_items(fb) == F.expected_fallback(rs)
    exp = [(a, g, d) for a, g, c, d in rs if g is not None and not F.is_raw(g) and g != b"zz"]
    assert [tuple(r) for r in rows_of(dst)] == exp


def test_groupby_text_key_falls_back():
    rs = F.rows()
    cols, src = F.chunk(rs)
    o = oracle.OracleSession(F.q_groupby_g())
    rc, _, _ = o.exec(src, nthreads=2)
    assert rc == 0, o.err.describe()
    fb = o.fallback()
    groups = W.decode_groups(o.final(), [abi.KAGG_ACTION__NROWS_ANY, abi.KAGG_ACTION__PSUM_INT64,
                                         abi.KAGG_ACTION__VREF])
    o.close()
    assert fb_items(fb) == F.expected_fallback(rs)
    exp = {}
    for a, g, c, d in rs:
        if F.is_raw(g):
            continue
        n, s = exp.get((g,), (0, 0))
        exp[(g,)] = (n + 1, s + c)
    assert {k: (v[0], v[1][1]) for k, v in groups.items()} == exp


def test_without_fallback_buffer_is_an_error():
    rs = F.rows()
    cols, src = F.chunk(rs)
    rc, dst, fb, err = run(F.q_textlen(fallback=False), src)
    assert rc == abi.ERRCODE_SUSPEND_FALLBACK
    assert "text datum is compressed or external" in err


def test_fallback_store_empties_when_taken():
    rs = F.rows(130)
    cols, src = F.chunk(rs)
    o = oracle.OracleSession(F.q_textlen())
    o.exec(src, dst_length=1 << 20)
    o.exec(src, dst_length=1 << 20)
    items = kds.parse_fallback_kds(o.fallback())
    assert len(items) == 2 * len(F.expected_fallback(rs))
    assert o.fallback() is None
    o.close()
