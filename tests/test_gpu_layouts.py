"""Projection tuples written on the GPU against the reference's own
kern_form_minimal_tuple bytes (tests/golden/layouts.json), with the
interpreter and the specialised kernel."""
import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from pgstrom import kds  # noqa: E402
from helpers import project_all_query, run_gpu_projection  # noqa: E402
from test_layouts import GOLDEN, dst_tuples, external_tuple_chunk, tuple_chunk  # noqa: E402

pytestmark = pytest.mark.gpu


def test_gpu_projection_tuples_match_reference(gpu, jit):
    cols = tuple_chunk()
    blob = project_all_query(cols)
    outs = run_gpu_projection(blob, kds.build_arrow_kds(cols), 1 << 20, tuning=dict(jit=jit))
    assert len(outs) == 1
    assert [t.hex() for t in dst_tuples(outs[0])] == GOLDEN["tuples"]


def test_gpu_projection_external_tuples_match_reference(gpu, jit):
    """TOAST pointers and compressed datums on heap pages are projected as
    they are (the reference's bytes); nothing falls back."""
    cols, src = external_tuple_chunk()
    blob = project_all_query(cols)
    outs = run_gpu_projection(blob, src, 1 << 20, tuning=dict(jit=jit))
    assert len(outs) == 1
    assert [t.hex() for t in dst_tuples(outs[0])] == GOLDEN["tuples_external"]
