"""Builds the product libraries with hipcc for gfx950: one per field backend,
from the same sources (the reference ships one crate per field):
parcomm_amd/libparcomm_amd.so (F65521 / Ext8, i8 entries) and
parcomm_amd/libparcomm_amd_kb.so (KoalaBear / Ext4, i16 entries, -DQF_KB).

Every source is compiled to an object in a temporary directory, all of them
in parallel, and each library is linked from its objects.

Run as `python -m parcomm_amd.build` or through __graft_entry__.build().
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libparcomm_amd.so")
OUT_KB = os.path.join(HERE, "libparcomm_amd_kb.so")
TARGETS = ((OUT, []), (OUT_KB, ["-DQF_KB"]))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def sources():
    return sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".cpp", ".hip"))
    )


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return proc.returncode, proc.stdout


def build(verbose=False, jobs=None):
    """Compiles both libraries (objects in parallel); returns their paths."""
    srcs = sources()
    jobs = jobs or min(16, os.cpu_count() or 4)
    tmp = tempfile.mkdtemp(prefix="parcomm_amd_build_")
    try:
        compiles = []  # (target index, object, command)
        for ti, (out, extra) in enumerate(TARGETS):
            for src in srcs:
                obj = os.path.join(tmp, f"{ti}_{os.path.basename(src)}.o")
                compiles.append((ti, obj, [HIPCC] + FLAGS + extra + ["-c", src, "-o", obj]))
        # the largest sources first, so the pool finishes together
        compiles.sort(key=lambda c: -os.path.getsize(c[2][-3]))
        with ThreadPoolExecutor(jobs) as pool:
            results = list(pool.map(lambda c: _run(c[2], verbose), compiles))
        failed = [c[2][-3] for c, (rc, _) in zip(compiles, results) if rc != 0]
        for (rc, log) in results:
            if rc != 0:
                sys.stderr.write(log)
        if failed:
            raise RuntimeError("hipcc failed on " + ", ".join(sorted(set(os.path.basename(f) for f in failed))))
        for ti, (out, extra) in enumerate(TARGETS):
            objs = [c[1] for c in compiles if c[0] == ti]
            rc, log = _run([HIPCC] + FLAGS + extra + ["-shared", "-o", out + ".tmp"] + objs, verbose)
            if rc != 0:
                sys.stderr.write(log)
                raise RuntimeError("hipcc failed linking " + os.path.basename(out))
        for out, _ in TARGETS:
            os.replace(out + ".tmp", out)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return [out for out, _ in TARGETS]


if __name__ == "__main__":
    print(build(verbose=True))
