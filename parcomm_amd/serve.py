"""Serving loop over a resident commitment: opening requests arrive as a
Poisson stream, are admitted while the device-memory budget allows, and run
concurrently, each on its worker thread's own HIP stream, so the kernels of
requests in flight overlap on the GPU. Latency is measured from arrival to the
finished proof.

A request picks a subset of the matrices and fresh public points for them; the
others keep the statement's benchmark points (main.rs:188-209). Every proof
still carries claims for all matrices, as Parameters::prove requires
(tbv.rs:123-141), so each response verifies against the unchanged public
commitment."""
import ctypes
import threading
import time
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import Claim, _check, field_info, lib, QF_FIELD_F65521


def device_memory(field=QF_FIELD_F65521):
    """(free, total) bytes of the current device."""
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib(field).qf_device_memory(ctypes.byref(free), ctypes.byref(total)), field)
    return free.value, total.value


def request_bytes(params):
    """Device memory one opening of these parameters holds at its peak: the
    packed even/odd witness and query (3 x logical/s elements of 16 bytes),
    the batched folded codeword and tree, and the sumcheck vectors."""
    s = 1 << params.config.packed_log
    eo = 3 * (params.logical // s) * 16
    n_values = params.block_len // field_info(params.field)["lanes"]
    folded = 2 * n_values * 16 + 2 * (n_values // 16) * 32
    return eo + folded + 2 * n_values * 16


def _element(rng, field):
    """A uniform extension element in its 16-byte canonical form."""
    info = field_info(field)
    c = rng.integers(0, info["p"], size=info["lanes"], dtype=np.uint64)
    return c.astype(info["coef"]).tobytes()


@dataclass
class Request:
    rid: int
    arrival: float
    subset: List[int]
    claims: List[Claim] = None
    proof: bytes = b""
    done: float = 0.0


def request_claims(params, committed, base_claims, subset, rng):
    """Claims of one request: fresh points on `subset`, benchmark points elsewhere."""
    points = []
    chosen = set(subset)
    for m, c in enumerate(base_claims):
        if m in chosen:
            points.append(([_element(rng, params.field) for _ in c.row], [_element(rng, params.field) for _ in c.col]))
        else:
            points.append((list(c.row), list(c.col)))
    return params.evaluate_claims(committed, points)


def serve(params, committed, base_claims, n_requests=32, rate=50.0, seed=1, workers=None, verify=False):
    """Runs a Poisson stream of opening requests; returns (stats, requests)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_requests)
    sizes = rng.integers(1, len(base_claims) + 1, size=n_requests)
    reqs = [Request(i, 0.0, sorted(rng.choice(len(base_claims), size=int(k), replace=False).tolist()))
            for i, k in enumerate(sizes)]
    per = request_bytes(params)
    free, _ = device_memory(params.field)
    budget_slots = max(1, int((free * 0.9) // per))
    slots = max(1, min(workers or 4, budget_slots))
    admit = threading.Semaphore(slots)
    ready = []
    cv = threading.Condition()
    t0 = time.perf_counter()

    def producer():
        t = t0
        for r, g in zip(reqs, gaps):
            t += g
            delay = t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            r.arrival = time.perf_counter()
            with cv:
                ready.append(r)
                cv.notify()

    def worker():
        while True:
            with cv:
                while not ready and not stop[0]:
                    cv.wait()
                if not ready:
                    return
                r = ready.pop(0)
            with admit:
                rrng = np.random.default_rng(seed * 1_000_003 + r.rid)
                r.claims = request_claims(params, committed, base_claims, r.subset, rrng)
                r.proof, _ = params.prove(committed, r.claims)
                r.done = time.perf_counter()
            if verify:
                params.verify(committed.statement, r.claims, r.proof)

    stop = [False]
    threads = [threading.Thread(target=worker) for _ in range(slots)]
    for th in threads:
        th.start()
    prod = threading.Thread(target=producer)
    prod.start()
    prod.join()
    with cv:
        stop[0] = True
        cv.notify_all()
    for th in threads:
        th.join()
    lat = np.array([r.done - r.arrival for r in reqs]) * 1e3
    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    stats = {
        "requests": n_requests,
        "arrival_rate_per_s": rate,
        "concurrency": slots,
        "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "mean_ms": round(float(lat.mean()), 2),
        "openings_per_s": round(n_requests / span, 3),
        "request_device_bytes": per,
        "seed": seed,
    }
    return stats, reqs
