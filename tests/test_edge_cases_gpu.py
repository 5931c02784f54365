"""Edge cases of the reference's own tests, in both fields, against the CPU
oracle (whole-proof byte equality) and the verifier:

- packing widths 1..6 with every entry at the type's minimum (-128, -32768)
  and query coordinates 0, 1 and -1 (end_to_end.rs:170-192);
- a 65523 x 1 matrix, more rows than the F65521 characteristic
  (end_to_end.rs:299-310);
- full-range entries with both extremes planted, in shapes whose terms run on
  the MFMA lane projections (packed_log 3 and 4), with the vector-ALU form
  (QF_PROJECT_MFMA=0) and the per-matrix form giving the same bytes;
- the device Layout::encode (layout.rs:115-135) against the host scatter."""
import os

import numpy as np
import pytest

from oracle_ref import (layout_encode, oracle_commit, oracle_prove, oracle_prove_points, shapes_flat,
                        synthetic_int16)

pytestmark = pytest.mark.gpu

FIELDS = ["f65521", "koalabear"]


def setup(field, shapes, chunk_log, packed_log=4):
    import parcomm_amd as pa
    f = pa.KOALABEAR if field == "koalabear" else pa.F65521
    p = pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log, packed_log=packed_log),
                      field=f)
    return pa, p


def scalar(field, v):
    """The 16-byte canonical form of the base-field constant v (mod p)."""
    import parcomm_amd as pa
    info = pa.field_info(pa.KOALABEAR if field == "koalabear" else pa.F65521)
    c = np.zeros(info["lanes"], dtype=info["coef"])
    c[0] = v % info["p"]
    return c.tobytes()


def oracle_for(field, oracle, oracle_kb):
    return oracle_kb if field == "koalabear" else oracle


def full_range(field, n, seed):
    if field == "koalabear":
        return synthetic_int16(n, seed)
    rng = np.random.default_rng(seed)
    v = rng.integers(-128, 128, size=n).astype(np.int8)
    v[:4] = [-128, 127, -1, 0]
    v[-2:] = [127, -128]
    return v


def prove_and_compare(pa, p, oracle, shapes, chunk_log, packed_log, values, points=None):
    committed = p.commit(values)
    if points is None:
        claims = p.benchmark_claims(committed)
        want, _, _ = oracle_prove(oracle, shapes_flat(shapes), len(shapes), chunk_log, packed_log, values)
    else:
        claims = p.evaluate_claims(committed, points)
        want, _, _ = oracle_prove_points(oracle, shapes_flat(shapes), len(shapes), chunk_log, packed_log, values,
                                         points)
    proof, _ = p.prove(committed, claims)
    assert len(proof) == len(want) and proof == want
    assert p.verify(committed.statement, claims, proof)
    return committed, claims, proof


@pytest.mark.parametrize("field", FIELDS)
@pytest.mark.parametrize("packed_log", [1, 2, 3, 4, 5, 6])
def test_minimum_entries_and_zero_one_coordinates(oracle, oracle_kb, field, packed_log):
    pa, p = setup(field, [(5, 7)], 30, packed_log)
    low = -32768 if field == "koalabear" else -128
    values = layout_encode(p, [np.full(35, low, dtype=p.entry_dtype)])
    zero, one, minus = scalar(field, 0), scalar(field, 1), scalar(field, -1)
    points = [([zero, one, minus], [one, zero, minus])]
    print("minimum entries", field, "packed_log", packed_log, flush=True)
    prove_and_compare(pa, p, oracle_for(field, oracle, oracle_kb), [(5, 7)], 30, packed_log, values, points)


@pytest.mark.parametrize("field", FIELDS)
def test_rows_above_the_characteristic(oracle, oracle_kb, field):
    shapes = [(65523, 1)]
    pa, p = setup(field, shapes, 10)
    assert p.blocks == 64
    e = np.arange(65523, dtype=np.int64)
    if field == "koalabear":
        mat = ((e * 73) % 65536).astype(np.uint16).view(np.int16)
    else:
        mat = ((e * 73) % 256).astype(np.uint8).view(np.int8)
    values = layout_encode(p, [mat])
    print("65523 x 1", field, flush=True)
    prove_and_compare(pa, p, oracle_for(field, oracle, oracle_kb), shapes, 10, 4, values)


@pytest.mark.parametrize("field", FIELDS)
@pytest.mark.parametrize("shapes,chunk_log,packed_log", [
    ([(1024, 4096), (3, 5)], 22, 4),   # 16-lane terms of 2^10..2^18 rows: MFMA EO projection, ragged end
    ([(2048, 1024), (7, 3)], 20, 3),   # 8-lane terms of 2^13 rows and more: the packed_log 3 MFMA form
    ([(1024, 4096), (5, 9)], 22, 5),   # EO projection at stride 64: F65521's four-record MFMA rows
])
def test_full_range_entries_on_the_mfma_projections(oracle, oracle_kb, field, shapes, chunk_log, packed_log):
    pa, p = setup(field, shapes, chunk_log, packed_log)
    mats = [full_range(field, r * c, 90 + i) for i, (r, c) in enumerate(shapes)]
    values = layout_encode(p, mats)
    orc = oracle_for(field, oracle, oracle_kb)
    print("full range", field, shapes, chunk_log, packed_log, "blocks", p.blocks, flush=True)
    committed, claims, proof = prove_and_compare(pa, p, orc, shapes, chunk_log, packed_log, values)
    os.environ["QF_PROJECT_MFMA"] = "0"
    try:
        valu, _ = p.prove(committed, claims)
    finally:
        del os.environ["QF_PROJECT_MFMA"]
    assert valu == proof
    pa.set_launch_mode(pa.LAUNCH_PER_MATRIX, p.field)
    try:
        per_matrix, _ = p.prove(committed, claims)
    finally:
        pa.set_launch_mode(pa.LAUNCH_BATCHED, p.field)
    assert per_matrix == proof


@pytest.mark.parametrize("field", FIELDS)
def test_device_encode_matches_host_scatter(oracle, oracle_kb, field):
    shapes = [(300, 200), (7, 1), (1, 7), (64, 33), (513, 129), (9, 1)]
    pa, p = setup(field, shapes, 14)
    mats = [full_range(field, r * c, 40 + i) for i, (r, c) in enumerate(shapes)]
    want = layout_encode(p, mats)
    got = p.encode(mats)  # host matrices, staged one at a time
    assert np.array_equal(got.to_host(p.entry_dtype), want)
    dev = [pa.DeviceBuffer.from_host(m, p.field) for m in mats]
    got_dev = p.encode(dev)  # device matrices, one launch
    assert np.array_equal(got_dev.to_host(p.entry_dtype), want)
    stream = pa.DeviceBuffer(want.nbytes, p.field)  # zeroed; one matrix at a time, in reverse order
    for i in reversed(range(len(mats))):
        p.encode_matrix(i, dev[i] if i % 2 else mats[i], stream)
    assert np.array_equal(stream.to_host(p.entry_dtype), want)
    committed = p.commit(got_dev)
    roots, _ = oracle_commit(oracle_for(field, oracle, oracle_kb), 14, want, p.blocks)
    assert committed.statement.roots == roots
    with pytest.raises(pa.QfError) as e:
        p.encode(mats[:-1])
    assert e.value.status == 5 and "matrix count mismatch" in str(e.value)
    with pytest.raises(pa.QfError) as e:
        p.encode(mats[:-1] + [mats[-1][:-1]])
    assert e.value.status == 5 and "matrix length mismatch" in str(e.value)
