"""Builds the product libraries with hipcc for gfx950: one per field backend,
from the same sources (the reference ships one crate per field):
parcomm_amd/libparcomm_amd.so (F65521 / Ext8, i8 entries) and
parcomm_amd/libparcomm_amd_kb.so (KoalaBear / Ext4, i16 entries, -DQF_KB).

Run as `python -m parcomm_amd.build` or through __graft_entry__.build().
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libparcomm_amd.so")
OUT_KB = os.path.join(HERE, "libparcomm_amd_kb.so")
TARGETS = ((OUT, []), (OUT_KB, ["-DQF_KB"]))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def sources():
    return sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".cpp", ".hip"))
    )


def build(verbose=False):
    """Compiles both libraries concurrently; returns their paths."""
    srcs = sources()
    procs = []
    for out, extra in TARGETS:
        cmd = [HIPCC] + FLAGS + extra + ["-o", out + ".tmp"] + srcs
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((out, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for out, proc in procs:
        log, _ = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(log)
            failed.append(os.path.basename(out))
    if failed:
        raise RuntimeError("hipcc failed building " + ", ".join(failed))
    for out, _ in procs:
        os.replace(out + ".tmp", out)
    return [out for out, _ in TARGETS]


if __name__ == "__main__":
    print(build(verbose=True))
