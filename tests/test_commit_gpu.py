"""Commitment parity: the HIP commit path against the CPU oracle, bit-exact on
roots and codewords (protocol.rs:131-185, packed.rs:220-258,
qafold_short_pc.rs:461-549)."""
import numpy as np
import pytest

from oracle_ref import DEMO_SHAPES, demo_bva, layout_encode, oracle_commit, synthetic_int8

pytestmark = pytest.mark.gpu


def params(shapes, chunk_log):
    import parcomm_amd as pa
    return pa.Parameters([pa.Shape(r, c) for r, c in shapes], pa.Config(chunk_log=chunk_log))


def test_demo_roots_and_codewords_match_oracle(oracle):
    p = params(DEMO_SHAPES, 10)
    values = demo_bva(lambda mats: layout_encode(p, mats))
    print("demo chunk_log 10 blocks", p.blocks, flush=True)
    committed = p.commit(values)
    roots, cws = oracle_commit(oracle, 10, values, p.blocks, want_codeword=True)
    assert committed.statement.roots == roots
    for b in range(p.blocks):
        assert np.array_equal(committed.codeword(b), cws[b]), b
    # all-zero trailing block shares the first zero root (protocol.rs:169-175)
    assert roots[3] == committed.statement.roots[3]


@pytest.mark.parametrize("chunk_log,rows,cols,seed", [(12, 64, 64, 1), (14, 100, 163, 2), (16, 256, 256, 3),
                                                      (18, 1000, 300, 4)])
def test_random_roots_and_codewords_match_oracle(oracle, chunk_log, rows, cols, seed):
    p = params([(rows, cols)], chunk_log)
    mat = synthetic_int8(rows * cols, seed)
    values = layout_encode(p, [mat])
    print("random", chunk_log, rows, cols, "blocks", p.blocks, flush=True)
    committed = p.commit(values)
    roots, cws = oracle_commit(oracle, p.block_len.bit_length() - 1, values, p.blocks, want_codeword=True)
    assert committed.statement.roots == roots
    for b in range(p.blocks):
        assert np.array_equal(committed.codeword(b), cws[b]), b


def test_cfg1_2048_square_chunk22_root_matches_oracle(oracle):
    p = params([(2048, 2048)], 22)
    values = layout_encode(p, [synthetic_int8(2048 * 2048, 2048)])
    print("cfg1 2048x2048 chunk 22", flush=True)
    committed = p.commit(values)
    roots, _ = oracle_commit(oracle, 22, values, p.blocks)
    assert committed.statement.roots == roots


def test_recompute_mode_gives_identical_roots():
    p = params(DEMO_SHAPES, 10)
    values = demo_bva(lambda mats: layout_encode(p, mats))
    assert p.commit(values).statement.roots == p.commit_recomputed(values).statement.roots


def test_nonzero_padding_is_rejected():
    import parcomm_amd as pa
    p = params([(31, 31)], 10)
    values = layout_encode(p, [np.ones(31 * 31, dtype=np.int8)])
    b = next(b for b in p.buckets() if b[3] > b[5] or b[4] > b[6])
    (m, ro, co, rows, cols, pr, pc, off) = b
    bad = values.copy()
    bad[off + pc if cols > pc else off + pr * cols] = 1
    with pytest.raises(pa.QfError) as e:
        p.commit(bad)
    assert e.value.status == 5
    with pytest.raises(pa.QfError):
        p.commit(values[:-1])


def test_device_resident_input_matches_host_input():
    import parcomm_amd as pa
    p = params([(300, 500)], 14)
    values = layout_encode(p, [synthetic_int8(300 * 500, 9)])
    d = pa.DeviceBuffer.from_host(values)
    assert p.commit(d).statement.roots == p.commit(values).statement.roots


def test_synthetic_bva_respects_padding_and_distribution(oracle):
    p = params([(1000, 300), (77, 5)], 18)
    buf = p.synthetic_bva(seed=11)
    values = buf.to_host()
    for (m, ro, co, rows, cols, pr, pc, off) in p.buckets():
        view = values[off:off + rows * cols].reshape(rows, cols)
        assert not view[pr:, :].any() and not view[:, pc:].any()
    phys = np.concatenate([values[off:off + rows * cols].reshape(rows, cols)[:pr, :pc].ravel()
                           for (m, ro, co, rows, cols, pr, pc, off) in p.buckets()])
    assert abs(phys.astype(np.float64).std() - 20.0) < 0.5 and abs(phys.mean()) < 0.5
    committed = p.commit(buf)
    roots, _ = oracle_commit(oracle, 18, values, p.blocks)
    assert committed.statement.roots == roots
