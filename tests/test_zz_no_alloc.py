"""qf_commit_roots_noalloc (no device allocation, fixed per-device workspace)
against qf_commit_device, through the C ABI: two calls in a row, the second
with a smaller BVA, bit-exact roots in both fields; a block too large for the
workspace is rejected with the documented error."""
import ctypes

import numpy as np
import pytest

from oracle_ref import layout_encode, synthetic_int8, synthetic_int16

pytestmark = pytest.mark.gpu


def roots_noalloc(pa, p, buf):
    L = pa.lib(p.field)
    L.qf_commit_roots_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p,
                                          ctypes.c_size_t]
    out = ctypes.create_string_buffer(32 * p.blocks)
    rc = L.qf_commit_roots_noalloc(p._h, buf.ptr, p.active, out, p.blocks)
    if rc:
        raise pa.QfError(rc, L.qf_last_error().decode())
    return [out.raw[32 * i:32 * i + 32] for i in range(p.blocks)]


@pytest.mark.parametrize("field", ["f65521", "koalabear"])
def test_noalloc_roots_match_commit_device(field):
    import parcomm_amd as pa
    f = pa.KOALABEAR if field == "koalabear" else pa.F65521
    gen = synthetic_int16 if field == "koalabear" else synthetic_int8
    # a larger then a smaller input: 4 blocks at chunk_log 22 (the largest block
    # the workspace takes in both fields), then 3 blocks at chunk_log 16
    for shapes, chunk_log in [([(3000, 4000), (5, 7)], 22), ([(300, 400), (9, 1)], 16)]:
        p = pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log), field=f)
        values = layout_encode(p, [gen(r * c, 7 + r) for r, c in shapes])
        buf = pa.DeviceBuffer.from_host(values, f)
        print("noalloc", field, shapes, "chunk_log", chunk_log, "blocks", p.blocks, flush=True)
        got = roots_noalloc(pa, p, buf)
        want = p.commit_recomputed(buf)
        assert got == want.statement.roots
        want.close()
        assert roots_noalloc(pa, p, buf) == got  # again, on the workspace the call left
        buf.close()


def test_noalloc_rejects_blocks_beyond_the_workspace():
    import parcomm_amd as pa
    p = pa.Parameters([pa.Shape(4096, 4096)], pa.Config(chunk_log=24))
    buf = pa.DeviceBuffer(p.active)
    with pytest.raises(pa.QfError) as e:
        roots_noalloc(pa, p, buf)
    assert e.value.status == 5 and "fixed workspace" in str(e.value)
    buf.close()
