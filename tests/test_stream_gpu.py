"""Model input into HBM on the GPU: the dense BVA stream (qf_stream_bva, the
reference CLI's input, main.rs:145-173) and safetensors checkpoints scattered
by the device Layout::encode give the same BVA as the host scatter, and
commits of it equal the oracle's roots; short and long streams are rejected."""
import json

import numpy as np
import pytest

from oracle_ref import layout_encode, oracle_commit, synthetic_int8, synthetic_int16

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("field", ["f65521", "koalabear"])
def test_bva_stream_and_safetensors_input(tmp_path, oracle, oracle_kb, field):
    import parcomm_amd as pa
    from parcomm_amd.stream import load_bva_stream, load_safetensors, write_safetensors
    f = pa.KOALABEAR if field == "koalabear" else pa.F65521
    shapes = [(300, 200), (64, 33), (1, 7)]
    p = pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=14), field=f)
    gen = synthetic_int16 if field == "koalabear" else synthetic_int8
    mats = [gen(r * c, 5 + i) for i, (r, c) in enumerate(shapes)]
    want = layout_encode(p, mats)
    item = want.dtype.itemsize
    desc = {"format": "parcomm-dense-bva-stream-description", "storage_dtype": "i16" if item == 2 else "i8",
            "shapes": [list(s) for s in shapes], "stream_bytes": want.nbytes,
            "layout": {"logical_entries": p.logical}}
    stream = tmp_path / "bva.bin"
    stream.write_bytes(want.tobytes())
    print("dense BVA stream", field, want.nbytes, "bytes", flush=True)
    got = load_bva_stream(p, desc, str(stream))
    assert np.array_equal(got.to_host(p.entry_dtype), want)
    roots, _ = oracle_commit(oracle_kb if item == 2 else oracle, 14, want, p.blocks)
    assert p.commit(got).statement.roots == roots
    stream.write_bytes(want.tobytes()[:-1])
    with pytest.raises(pa.QfError, match="ends early"):
        load_bva_stream(p, desc, str(stream))
    stream.write_bytes(want.tobytes() + b"\0")
    with pytest.raises(pa.QfError, match="excess BVA input"):
        load_bva_stream(p, desc, str(stream))
    # two safetensors shards, the matrices in their own row-major shapes
    write_safetensors(tmp_path / "a.safetensors", {"layer.q": mats[0].reshape(300, 200)})
    write_safetensors(tmp_path / "b.safetensors", {"layer.k": mats[1].reshape(64, 33), "norm": mats[2].reshape(1, 7)})
    got2 = load_safetensors(p, [tmp_path / "a.safetensors", tmp_path / "b.safetensors"], ["layer.q", "layer.k", "norm"])
    assert np.array_equal(got2.to_host(p.entry_dtype), want)
