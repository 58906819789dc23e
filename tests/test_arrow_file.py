"""The Arrow IPC chunk producer (xpuhip_arrow_open / xpuhip_arrow_load_batch:
arrowFdwLoadRecordBatch, arrow_fdw.c:4018-4100) on files pyarrow writes: each
record batch's chunk gives the oracle the rows the same data gives through
the test-side Arrow builder, for Int, FloatingPoint, Utf8, LargeUtf8, Bool,
Decimal128, FixedSizeBinary, NULLs and an empty batch; unreferenced fields
are not copied; type mismatches and compressed files are refused.  Host
code only: no GPU."""
import numpy as np
import pytest

from pgstrom import abi, codegen as cg, kds, xpuhip
from helpers import run_oracle_projection
import arrowfiles as AF


def _query(quals=True):
    plan = cg.Plan(AF.COLDEFS)
    v = {n: plan.scan_var(n) for n, _ in AF.COLDEFS}
    if quals:
        plan.scan_quals = cg.Or(cg.Func("int4gt", v["a"], cg.Const("int4", 0)),
                                cg.NullTest(v["a"], True))
    plan.set_projection([v[n] for n, _ in AF.COLDEFS], AF.COLDEFS)
    return plan.build()


@pytest.fixture(scope="module")
def arrow_file(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("arrow") / "t.arrow")
    batches = AF.write_file(path)
    return path, batches


def test_batches_give_the_builders_rows(arrow_file):
    path, batches = arrow_file
    f = xpuhip.ArrowFile(open(path, "rb").read())
    assert f.nbatches == len(batches)
    assert [x[0] for x in f.fields] == ["a", "b", "c", "t", "lt", "f", "d", "fb", "unused", "s"]
    blob = _query()
    for i, rows in enumerate(batches):
        chunk = f.batch(i, AF.COLDEFS)
        hd = abi.kern_data_store_head.from_buffer(chunk)
        assert hd.nitems == len(rows["a"]) and hd.format == abi.KDS_FORMAT_ARROW
        got, _ = run_oracle_projection(blob, chunk, 4 << 20)
        exp, _ = run_oracle_projection(blob, kds.build_arrow_kds(AF.columns(rows)), 4 << 20)
        assert [r[0] for r in kds.parse_row_kds(got)] == [r[0] for r in kds.parse_row_kds(exp)]
    f.close()


def test_unreferenced_fields_are_not_copied(arrow_file):
    path, batches = arrow_file
    f = xpuhip.ArrowFile(open(path, "rb").read())
    full = f.batch(2, AF.COLDEFS)
    names = [n for n, _ in AF.COLDEFS]
    part = f.batch(2, AF.COLDEFS, [n if n in ("a", "s") else None for n in names])
    assert len(part) < len(full) // 4
    hd, cms = kds.kds_colmeta(memoryview(bytes(part)))
    for cm, n in zip(cms, names):
        assert (cm.values_length > 0) == (n in ("a", "s"))
    f.close()


def test_type_mismatch_and_compression_are_refused(tmp_path):
    path = str(tmp_path / "c.arrow")
    AF.write_file(path, sizes=(100,))
    f = xpuhip.ArrowFile(open(path, "rb").read())
    bad = [("a", "float8")] + AF.COLDEFS[1:]
    with pytest.raises(xpuhip.XpuError, match="cannot be read from Arrow field"):
        f.batch(0, bad)
    f.close()
    path2 = str(tmp_path / "z.arrow")
    AF.write_file(path2, sizes=(100,), compression="lz4")
    f = xpuhip.ArrowFile(open(path2, "rb").read())
    with pytest.raises(xpuhip.XpuError, match="compressed"):
        f.batch(0, AF.COLDEFS)
    f.close()
    with pytest.raises(xpuhip.XpuError, match="not an Arrow IPC file"):
        xpuhip.ArrowFile(b"PAR1" + b"\0" * 64)
