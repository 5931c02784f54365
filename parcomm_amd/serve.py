"""Serving loop over a resident commitment: opening requests arrive as a
Poisson stream and wait in a queue; a request is admitted as soon as a slot is
free and the device-memory budget has room for it, and runs on its worker
thread's own HIP stream, so the kernels of the requests in flight overlap on
the GPU. Latency is measured from arrival to the finished proof.

A request picks a subset of the matrices and fresh public points for them; the
others keep the statement's benchmark points (main.rs:188-209). Every proof
still carries claims for all matrices, as Parameters::prove requires
(tbv.rs:123-141), so each response verifies against the unchanged public
commitment."""
import collections
import ctypes
import threading
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import Claim, _check, field_info, lib, QF_FIELD_F65521


def device_memory(field=QF_FIELD_F65521):
    """(free, total) bytes of the current device; free counts the blocks the
    library's caching allocator holds idle."""
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib(field).qf_device_memory(ctypes.byref(free), ctypes.byref(total)), field)
    return free.value, total.value


def request_bytes(params):
    """Device memory one opening of these parameters holds at its peak: the
    even/odd witness of the first two binary levels (rows/2 and rows/4
    elements of 16 bytes, rows = logical/s) and the query materialized once
    the rounds are short (n + n/2 elements, n at most rows/2 and 2^22), the
    batched folded codeword and tree, and the sumcheck vectors."""
    s = 1 << params.config.packed_log
    rows = params.logical // s
    eo = 12 * rows + 24 * min(rows // 2, 1 << 22)
    n_values = params.block_len // field_info(params.field)["lanes"]
    folded = 2 * n_values * 16 + 2 * (n_values // 16) * 32
    return eo + folded + 2 * n_values * 16


@dataclass
class Request:
    rid: int
    arrival: float
    subset: List[int]
    claims: List[Claim] = None
    proof: bytes = b""
    admitted: float = 0.0
    done: float = 0.0
    error: Optional[str] = None


def request_claims(params, committed, base_claims, subset, rng):
    """Claims of one request: fresh points on `subset`, benchmark points
    elsewhere. The fresh coordinates come from one draw of uniform
    coefficients (16-byte canonical elements)."""
    info = field_info(params.field)
    chosen = set(subset)
    total = sum(len(c.row) + len(c.col) for m, c in enumerate(base_claims) if m in chosen)
    raw = rng.integers(0, info["p"], size=(total, info["lanes"]), dtype=np.uint64).astype(info["coef"]).tobytes()
    elems = [raw[16 * i:16 * i + 16] for i in range(total)]
    points, k = [], 0
    for m, c in enumerate(base_claims):
        if m in chosen:
            nr, nc = len(c.row), len(c.col)
            points.append((elems[k:k + nr], elems[k + nr:k + nr + nc]))
            k += nr + nc
        else:
            points.append((list(c.row), list(c.col)))
    return params.evaluate_claims(committed, points)


class _Admission:
    """Slots and a device-memory budget. The budget is the free device memory
    measured whenever no request is in flight (the commitment stays resident);
    a request is admitted while the bytes of the admitted requests plus its
    own fit in `fraction` of it, and at least one request always runs."""

    def __init__(self, field, slots, per_request, fraction):
        self.field, self.slots, self.per, self.fraction = field, slots, per_request, fraction
        self.cv = threading.Condition()
        self.in_flight = 0
        self.max_in_flight = 0
        self.budget = device_memory(field)[0]

    def fits(self):
        if self.in_flight >= self.slots:
            return False
        if self.in_flight == 0:
            self.budget = device_memory(self.field)[0]
            return True
        return (self.in_flight + 1) * self.per <= self.fraction * self.budget

    def acquire(self):
        with self.cv:
            while not self.fits():
                self.cv.wait()
            self.in_flight += 1
            self.max_in_flight = max(self.max_in_flight, self.in_flight)

    def release(self):
        with self.cv:
            self.in_flight -= 1
            self.cv.notify_all()


def serve(params, committed, base_claims, n_requests=32, rate=50.0, seed=1, workers=None, verify=False,
          memory_fraction=0.9):
    """Runs a Poisson stream of `n_requests` opening requests at `rate` per
    second with at most `workers` (default 6) in flight; returns (stats,
    requests). A request that fails keeps its error message and is counted
    in stats["failed"]; latencies cover the completed requests only."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_requests)
    sizes = rng.integers(1, len(base_claims) + 1, size=n_requests)
    reqs = [Request(i, 0.0, sorted(rng.choice(len(base_claims), size=int(k), replace=False).tolist()))
            for i, k in enumerate(sizes)]
    per = request_bytes(params)
    slots = max(1, workers or 6)
    adm = _Admission(params.field, slots, per, memory_fraction)
    queue = collections.deque()
    qcv = threading.Condition()
    stop = [False]
    t0 = time.perf_counter()

    def producer():
        t = t0
        for r, g in zip(reqs, gaps):
            t += g
            delay = t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            r.arrival = time.perf_counter()
            with qcv:
                queue.append(r)
                qcv.notify()

    def worker():
        while True:
            with qcv:
                while not queue and not stop[0]:
                    qcv.wait()
                if not queue:
                    return
                r = queue.popleft()
            adm.acquire()
            r.admitted = time.perf_counter()
            try:
                rrng = np.random.default_rng(seed * 1_000_003 + r.rid)
                r.claims = request_claims(params, committed, base_claims, r.subset, rrng)
                r.proof, _ = params.prove(committed, r.claims)
                r.done = time.perf_counter()
            except Exception as e:  # recorded on the request; the loop keeps serving
                r.error = f"{type(e).__name__}: {e}"
            finally:
                adm.release()
            if verify and r.error is None:
                try:
                    params.verify(committed.statement, r.claims, r.proof)
                except Exception as e:
                    r.error = f"verification failed: {e}"

    threads = [threading.Thread(target=worker) for _ in range(slots)]
    for th in threads:
        th.start()
    prod = threading.Thread(target=producer)
    prod.start()
    prod.join()
    with qcv:
        stop[0] = True
        qcv.notify_all()
    for th in threads:
        th.join()
    done = [r for r in reqs if r.error is None]
    failed = [r for r in reqs if r.error is not None]
    stats = {
        "requests": n_requests,
        "completed": len(done),
        "failed": len(failed),
        "arrival_rate_per_s": rate,
        "concurrency": slots,
        "max_in_flight": adm.max_in_flight,
        "request_device_bytes": per,
        "seed": seed,
    }
    if failed:
        stats["first_error"] = failed[0].error
    if done:
        lat = np.array([r.done - r.arrival for r in done]) * 1e3
        wait = np.array([r.admitted - r.arrival for r in done]) * 1e3
        service = np.array([r.done - r.admitted for r in done]) * 1e3
        span = max(r.done for r in done) - min(r.arrival for r in reqs)
        stats.update({
            "p50_ms": round(float(np.percentile(lat, 50)), 2),
            "p99_ms": round(float(np.percentile(lat, 99)), 2),
            "mean_ms": round(float(lat.mean()), 2),
            "mean_queue_ms": round(float(wait.mean()), 2),
            "mean_service_ms": round(float(service.mean()), 2),
            "openings_per_s": round(len(done) / span, 3),
        })
    return stats, reqs
