"""The drop-in boundary: struct layouts, opcode numbers and exported symbols.

tests/golden/abi_layout.json is dumped from the reference headers
(oracle/ref_driver.cc compiled against src/xpu_common.h by
tests/golden/gen_abi_layout.py); every fact this library declares must match."""
import ctypes
import json
import os
import re

from pgstrom import abi, xpuhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "abi_layout.json")))


def test_library_layout_matches_reference():
    mine = json.loads(xpuhip.abi_layout_json())
    assert len(mine) > 500
    bad = {k: (v, GOLDEN.get(k)) for k, v in mine.items() if GOLDEN.get(k) != v}
    assert not bad, bad


def test_python_mirror_layout():
    g = GOLDEN
    assert ctypes.sizeof(abi.kern_colmeta) == g["sizeof:kern_colmeta"]
    assert ctypes.sizeof(abi.ArrowTypeOptions) == g["sizeof:ArrowTypeOptions"]
    assert abi.KDS_COLMETA_OFFSET == g["offsetof:kern_data_store.colmeta"]
    assert abi.SESSION_POFFSET_OFFSET == g["offsetof:kern_session_info.poffset"]
    assert ctypes.sizeof(abi.kern_session_info_head) == g["offsetof:kern_session_info.nparams"] + 4
    assert abi.KEXP_U_OFFSET == g["offsetof:kern_expression.u.data"]
    assert abi.KMRELS_CHUNKS_OFFSET == g["offsetof:kern_multirels.chunks"]
    assert ctypes.sizeof(abi.kern_multirels_chunk) == g["sizeof:kern_multirels.chunks[0]"]
    assert ctypes.sizeof(abi.kern_varslot_desc) == g["sizeof:kern_varslot_desc"]
    assert ctypes.sizeof(abi.kern_aggregate_desc) == g["sizeof:kern_aggregate_desc"]
    assert abi.MINIMAL_TUPLE_OFFSET == g["const:MINIMAL_TUPLE_OFFSET"]
    assert abi.KERN_EXPRESSION_MAGIC == g["const:KERN_EXPRESSION_MAGIC"]
    for f, _ in abi.kern_exec_results._fields_:
        key = "offsetof:kern_exec_results." + f
        if key in g:
            assert getattr(abi.kern_exec_results, f).offset == g[key], f
    for f, _ in abi.kern_session_info_head._fields_:
        key = "offsetof:kern_session_info." + f
        if key in g:
            assert getattr(abi.kern_session_info_head, f).offset == g[key], f


def test_python_opcode_table():
    from pgstrom import opcodes
    for name, (num, *_rest) in opcodes.FUNCS.items():
        assert GOLDEN["func:" + name] == num, name
    for name, num in opcodes.SPECIAL.items():
        assert GOLDEN["func:" + name] == num, name


def test_exports_match_header():
    """Every entry point include/xpuhip.h declares is exported by the .so."""
    hdr = open(os.path.join(ROOT, "include", "xpuhip.h")).read()
    declared = set(re.findall(r"\b(xpuhip_[a-z_]+)\s*\(", hdr))
    lib = xpuhip.lib()
    missing = [d for d in declared if not hasattr(lib, d)]
    assert not missing, missing
    assert declared == set(xpuhip.EXPORTS)
