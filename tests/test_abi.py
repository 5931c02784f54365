"""The product C ABI: every declared symbol is exported, parameters match the
oracle, and invalid configurations are rejected (no GPU needed)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "parcomm_amd.h")).read()
    return sorted(set(re.findall(r"\b(qf_[a-z_]+)\s*\(", text)))


@pytest.mark.parametrize("field", [0, 1], ids=["f65521", "koalabear"])
def test_every_declared_symbol_is_exported(field):
    import parcomm_amd as pa
    product = pa.lib(field)
    names = declared_symbols()
    assert len(names) >= 10
    for name in names:
        assert hasattr(product, name), name
    assert product.qf_field() == field


@pytest.mark.parametrize("field", [0, 1], ids=["f65521", "koalabear"])
@pytest.mark.parametrize("chunk_log", [10, 12, 14, 17])
def test_parameter_digest_matches_oracle(oracle, oracle_kb, chunk_log, field):
    import parcomm_amd as pa
    orc = oracle_kb if field else oracle
    h = orc.orc_pcs_new(chunk_log, 100, bytes([0x51] * 32))
    d = (ctypes.c_uint8 * 32)()
    orc.orc_pcs_parameter_digest(ctypes.c_void_p(h), d)
    orc.orc_pcs_free(h)
    p = pa.Parameters([pa.Shape(1, 1 << chunk_log)], pa.Config(chunk_log=chunk_log), field=field)
    assert p.parameter_digest() == bytes(d)


def test_invalid_configs_are_rejected():
    import parcomm_amd as pa
    for cfg in [pa.Config(chunk_log=9), pa.Config(chunk_log=31), pa.Config(packed_log=0), pa.Config(packed_log=7),
                pa.Config(query_rf_bits=31), pa.Config(query_rf_bits=121)]:
        with pytest.raises(pa.QfError) as e:
            pa.Parameters([pa.Shape(4, 4)], cfg)
        assert e.value.status == 5
    with pytest.raises(pa.QfError):
        pa.Parameters([pa.Shape(0, 4)])
    with pytest.raises(pa.QfError):
        pa.Parameters([pa.Shape(1 << 31, 1)])


def test_chunk_log_is_capped_by_logical_length():
    import parcomm_amd as pa
    p = pa.Parameters([pa.Shape(31, 37), pa.Shape(65, 17), pa.Shape(1, 7), pa.Shape(9, 1)], pa.Config(chunk_log=30))
    assert (p.active, p.logical, p.block_len, p.blocks) == (2305, 4096, 4096, 1)
    p = pa.Parameters([pa.Shape(31, 37), pa.Shape(65, 17), pa.Shape(1, 7), pa.Shape(9, 1)], pa.Config(chunk_log=10))
    assert (p.block_len, p.blocks) == (1024, 4)
