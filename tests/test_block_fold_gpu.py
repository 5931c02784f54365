"""The F65521 block fold on the matrix cores (packed_batch.rs's fold of the
PCS blocks into one by the powers of the batch challenge): whole proofs against
the oracle's bytes, and against the vector-ALU form (QF_BLOCK_FOLD_MFMA=0), for
1 to 128 blocks (1 to 32 contraction steps of four blocks, the last one
partial), ragged ends and entries all at -128 (the largest digit sums). The same
switch covers the native opening's column folds (every case) and its top-level
projection through 32-row groups (message halves of at least 2^10 rows: the
chunk_log-20 case)."""
import numpy as np
import pytest

from oracle_ref import oracle_prove, shapes_flat
from test_edge_cases_gpu import full_range, layout_encode, setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shapes,chunk_log,packed_log,fill", [
    ([(500, 1000)], 12, 2, "full"),             # 128 blocks: 32 steps
    ([(301, 701), (9, 3)], 14, 4, "full"),      # 14 blocks, the last step partial, ragged end
    ([(40, 50)], 12, 1, "full"),                # one block
    ([(512, 1024)], 12, 3, "min"),              # 128 blocks of -128
    ([(1000, 999)], 13, 3, "full"),             # 128 blocks at packed_log 3
    ([(1500, 1500)], 20, 4, "full"),            # 3 blocks; top-level projection in 32-row groups (2^12 rows)
])
def test_block_fold_matches_the_oracle(oracle, shapes, chunk_log, packed_log, fill, monkeypatch):
    pa, p = setup("f65521", shapes, chunk_log, packed_log)
    if fill == "min":
        mats = [np.full(r * c, -128, dtype=np.int8) for r, c in shapes]
    else:
        mats = [full_range("f65521", r * c, 11 + i) for i, (r, c) in enumerate(shapes)]
    values = layout_encode(p, mats)
    print("block fold", shapes, chunk_log, packed_log, fill, flush=True)
    committed = p.commit(values)
    claims = p.benchmark_claims(committed)
    proof, _ = p.prove(committed, claims)
    monkeypatch.setenv("QF_BLOCK_FOLD_MFMA", "0")
    valu, _ = p.prove(committed, claims)
    want, _, _ = oracle_prove(oracle, shapes_flat(shapes), len(shapes), chunk_log, packed_log, values)
    assert proof == want
    assert valu == want
