"""EO rounds that project their outputs on the query's terms instead of reading
a materialized query (eo.rs:97-181 with tbv.rs:393-428's query fold). The
threshold QF_EO_OTF_MIN=2048 keeps these rounds on down to 4096-element levels,
so small shapes run every kind of record over many levels: bucket terms of
whole chunks, pieces (bucket terms shorter than a chunk, column-padding terms
whose zero factors sit at or below the pair bit), and items (terms of at most
one element). Proofs must equal the oracle's bytes, with the threshold forced
and with the default one (where only the first round reads no query)."""
import numpy as np
import pytest

from oracle_ref import oracle_prove, shapes_flat
from test_edge_cases_gpu import full_range, layout_encode, oracle_for, prove_and_compare, scalar, setup

pytestmark = pytest.mark.gpu

FIELDS = ["f65521", "koalabear"]


@pytest.mark.parametrize("field", FIELDS)
@pytest.mark.parametrize("shapes,chunk_log,packed_log", [
    ([(500, 1000), (3, 5), (7, 1)], 20, 2),                    # column padding pairs data with padding
    ([(2048, 1024), (64, 64), (4096, 1), (33, 17)], 22, 4),     # chunked buckets, short buckets, items
    ([(700, 1500), (1, 9)], 21, 1),                            # stride 2
    ([(1024, 2048), (100, 300)], 22, 6),                       # stride 64
])
def test_on_the_fly_rounds_match_the_oracle(oracle, oracle_kb, field, shapes, chunk_log, packed_log, monkeypatch):
    pa, p = setup(field, shapes, chunk_log, packed_log)
    mats = [full_range(field, r * c, 7 + i) for i, (r, c) in enumerate(shapes)]
    values = layout_encode(p, mats)
    print("on-the-fly EO", field, shapes, chunk_log, packed_log, flush=True)
    committed = p.commit(values)
    claims = p.benchmark_claims(committed)
    default, _ = p.prove(committed, claims)
    monkeypatch.setenv("QF_EO_OTF_MIN", "2048")
    forced, _ = p.prove(committed, claims)
    want, _, _ = oracle_prove(oracle_for(field, oracle, oracle_kb), shapes_flat(shapes), len(shapes), chunk_log,
                              packed_log, values)
    assert forced == want
    assert default == want


@pytest.mark.parametrize("field", FIELDS)
def test_on_the_fly_rounds_with_zero_and_one_coordinates(oracle, oracle_kb, field, monkeypatch):
    """Claim coordinates 0, 1 and -1 put zero entries into bucket factors (bits
    whose pair lies on one side only) at every level of the rounds."""
    shapes = [(1024, 512), (5, 7)]
    pa, p = setup(field, shapes, 20, 3)
    mats = [full_range(field, r * c, 30 + i) for i, (r, c) in enumerate(shapes)]
    values = layout_encode(p, mats)
    rng = np.random.default_rng(5)
    special = [scalar(field, 0), scalar(field, 1), scalar(field, -1)]
    points = []
    for r, c in shapes:
        nr, nc = max(1, (r - 1).bit_length()), max(1, (c - 1).bit_length())
        points.append(([special[int(k)] for k in rng.integers(0, 3, nr)],
                       [special[int(k)] for k in rng.integers(0, 3, nc)]))
    monkeypatch.setenv("QF_EO_OTF_MIN", "2048")
    print("on-the-fly EO with 0/1/-1 coordinates", field, flush=True)
    prove_and_compare(pa, p, oracle_for(field, oracle, oracle_kb), shapes, 20, 3, values, points)
