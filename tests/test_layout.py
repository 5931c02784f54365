"""BVA layout (layout.rs:34-155) against the reference's llama8b fixture
(qafold/f65521/tests/fixtures/llama8b-layout.json, copied to tests/golden),
for both the oracle and the product library's host layout."""
import ctypes
import json
import os

from conftest import GOLDEN

KEYS = ["tensor_index", "row_offset", "col_offset", "row_len", "col_len", "physical_row_len",
        "physical_col_len", "dense_offset"]


def fixture():
    with open(os.path.join(GOLDEN, "llama8b-layout.json")) as f:
        return json.load(f)


def test_oracle_layout_matches_fixture(oracle):
    fx = fixture()
    shapes = fx["shapes"]
    flat = (ctypes.c_uint64 * (2 * len(shapes)))(*[v for s in shapes for v in s])
    cap = 1024
    b = (ctypes.c_uint64 * (8 * cap))()
    nb, act, lg = ctypes.c_size_t(), ctypes.c_uint64(), ctypes.c_uint64()
    assert oracle.orc_layout(flat, ctypes.c_size_t(len(shapes)), b, ctypes.c_size_t(cap), ctypes.byref(nb),
                             ctypes.byref(act), ctypes.byref(lg)) == 0
    assert (len(shapes), nb.value, act.value, lg.value) == (291, 489, 8053329920, 8589934592)
    for i, ref in enumerate(fx["buckets"]):
        assert [b[8 * i + j] for j in range(8)] == [ref[k] for k in KEYS]


def test_product_layout_matches_fixture():
    import parcomm_amd as pa
    fx = fixture()
    p = pa.Parameters([pa.Shape(r, c) for r, c in fx["shapes"]], pa.Config(chunk_log=26))
    assert (p.active, p.logical, p.n_buckets) == (8053329920, 8589934592, 489)
    for got, ref in zip(p.buckets(), fx["buckets"]):
        assert list(got) == [ref[k] for k in KEYS]


def test_intervals_cap_and_rounding(oracle):
    # layout.rs:34-54: at most four dyadic intervals, the last one rounded up.
    flat = (ctypes.c_uint64 * 2)(128256, 7)
    b = (ctypes.c_uint64 * (8 * 64))()
    nb, act, lg = ctypes.c_size_t(), ctypes.c_uint64(), ctypes.c_uint64()
    assert oracle.orc_layout(flat, ctypes.c_size_t(1), b, ctypes.c_size_t(64), ctypes.byref(nb), ctypes.byref(act),
                             ctypes.byref(lg)) == 0
    rows = sorted({(b[8 * i + 1], b[8 * i + 3], b[8 * i + 5]) for i in range(nb.value)})
    assert rows == [(0, 65536, 65536), (65536, 32768, 32768), (98304, 16384, 16384), (114688, 16384, 13568)]
