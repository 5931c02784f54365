"""Synthetic workload shapes of the benchmarks: Llama-3.1-shaped weight
matrices (rows, cols), in the order the model stores them: embedding and
lm_head first, then per layer q, k, v, o, gate, up, down."""


def llama8b_shapes():
    """Llama-3.1-8B: 32 layers, hidden 4096, MLP 14336, 8 KV heads; 226 matrices."""
    layer = [(4096, 4096), (1024, 4096), (1024, 4096), (4096, 4096), (14336, 4096), (14336, 4096), (4096, 14336)]
    return [(128256, 4096), (128256, 4096)] + layer * 32


def llama70b_shapes():
    """Llama-3.1-70B: 80 layers, hidden 8192, MLP 28672, 8 KV heads; 562 matrices."""
    layer = [(8192, 8192), (1024, 8192), (1024, 8192), (8192, 8192), (28672, 8192), (28672, 8192), (8192, 28672)]
    return [(128256, 8192), (128256, 8192)] + layer * 80


# Proof sizes the reference published for these configurations
# (results/reference/qafold-measurements.json:17, 85, 160): 8B with i8 entries
# (F65521), 8B with i16 entries (KoalaBear), both chunk_log 30 / packed_log 4,
# and 70B with i8 entries, chunk_log 30 / packed_log 6.
PUBLISHED_PROOF_BYTES = {"f65521-8b": 9287797, "koalabear-8b": 10404021, "f65521-70b": 81645677}
