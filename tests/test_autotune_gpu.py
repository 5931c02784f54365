"""The tile-pass autotuner: its report after a pass has run, and the shapes
recorded in QF_TUNE_FILE by one process being the ones a later process
launches without timing them again. The picks never change results: every
other GPU test runs with the autotuner on."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# One 2^20-element WHT pass (mode 2) over 4 vectors, then the report as JSON.
CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import parcomm_amd as pa
L = pa.lib(pa.F65521)
x = pa.DeviceBuffer.from_host(np.zeros((4 << 20, 8), dtype=np.uint16))
o = pa.DeviceBuffer(4 << 24)
rc = L.qf_test_tile_pass(2, x.ptr, o.ptr, None, None, 4, 1 << 20, 1 << 20, 20, 0, 12, 0)
assert rc == 0, L.qf_last_error()
print(json.dumps(pa.autotune_report()))
"""


def run_child(tune_file):
    env = dict(os.environ, QF_TUNE_FILE=str(tune_file))
    env.pop("QF_AUTOTUNE", None)
    out = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_report_lists_timed_candidates():
    import parcomm_amd as pa
    x = pa.DeviceBuffer.from_host(np.zeros((1 << 16, 8), dtype=np.uint16))
    o = pa.DeviceBuffer(1 << 20)
    L = pa.lib(pa.F65521)
    assert L.qf_test_tile_pass(2, x.ptr, o.ptr, None, None, 1, 1 << 16, 1 << 16, 16, 0, 12, 0) == 0, L.qf_last_error()
    rep = pa.autotune_report()
    if os.environ.get("QF_AUTOTUNE") == "0":
        pytest.skip("autotuner switched off for this run")
    mine = [e for e in rep if e["mode"] == 2 and e["L"] == 16 and e["s"] == 0 and e["t"] == 12 and e["vectors"] == 1]
    assert mine, rep
    e = mine[0]
    assert len(e["tried"]) >= 4 and all(t["ms"] > 0 for t in e["tried"])
    assert e["best"] in [t["shape"] for t in e["tried"]]


def test_tune_file_is_recorded_and_reused(tmp_path):
    f = tmp_path / "tune.txt"
    first = run_child(f)
    assert f.exists() and f.read_text().strip(), "no picks recorded"
    e1 = [e for e in first if e["L"] == 20 and e["vectors"] == 4]
    assert e1 and e1[0]["tried"], first
    second = run_child(f)
    e2 = [e for e in second if e["L"] == 20 and e["vectors"] == 4]
    assert e2 and e2[0]["tried"] == [], "the second process timed the pass again instead of reading the file"
    assert e2[0]["best"] == e1[0]["best"]
