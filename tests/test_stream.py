"""Model input (parcomm_amd/stream.py): the converter's stream description
is checked against the layout as the reference CLI checks it
(main.rs:116-128, 146-153), and safetensors headers are read back."""
import json

import numpy as np
import pytest


def test_description_checks_follow_the_cli(tmp_path):
    import parcomm_amd as pa
    from parcomm_amd.stream import check_description, read_description
    d = {"format": "parcomm-dense-bva-stream-description", "storage_dtype": "i8", "shapes": [[32, 32]],
         "stream_bytes": 1024, "layout": {"logical_entries": 1024}}  # the reference's cli_stream.py case
    path = tmp_path / "input.json"
    path.write_text(json.dumps(d))
    shapes, dtype, desc = read_description(path)
    assert shapes == [pa.Shape(32, 32)] and dtype == "i8"
    p = pa.Parameters(shapes, pa.Config(chunk_log=10))
    check_description(p, desc)
    for bad in ({"stream_bytes": 1023}, {"layout": {"logical_entries": 2048}}):
        with pytest.raises(ValueError, match="layout mismatch"):
            check_description(p, {**d, **bad})
    path.write_text(json.dumps({**d, "format": "other"}))
    with pytest.raises(ValueError):
        read_description(path)


def test_safetensors_header_round_trip(tmp_path):
    from parcomm_amd.stream import safetensors_index, write_safetensors
    a = np.arange(-60, 60, dtype=np.int8).reshape(8, 15)
    b = np.arange(7, dtype=np.int16)
    path = tmp_path / "m.safetensors"
    write_safetensors(path, {"w": a, "bias": b})
    idx = safetensors_index(path)
    assert idx["w"][:2] == (np.int8, (8, 15)) and idx["bias"][:2] == (np.int16, (7,))
    raw = path.read_bytes()
    off, n = idx["w"][2], idx["w"][3]
    assert np.array_equal(np.frombuffer(raw[off:off + n], dtype=np.int8).reshape(8, 15), a)
