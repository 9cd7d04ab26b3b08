"""Continuous-batching serving loop over libphpfdtd_hip.

Simulation requests (spec, medium, step count) arrive over time; every
request's fields and lattice state stay on the device from admission to
completion.  One engine thread owns a ``phpfdtd_batch`` (include/phpfdtd.h):
each kernel of a time step is one launch over the tiles of every member, so
many small domains fill the GPU the way one large domain does.

Engine cycle:
  1. arrivals join a FIFO queue;
  2. admission: the head of the queue is admitted when its device footprint
     (phpfdtd_footprint_bytes) fits the HBM budget beside the live requests
     and a member slot is free; its context is built on a worker thread (host
     precompute and uploads overlap the GPU work);
  3. built contexts join the batch at the next chunk boundary;
  4. the batch runs a chunk of steps that ends when the next member finishes
     (at most ``max_chunk``); one chunk stays queued ahead of the host, so the
     GPU does not wait for bookkeeping;
  5. finished members leave the batch at that boundary; their results are
     read and their contexts destroyed on a worker thread, which returns
     their footprint to the budget.

A member computes exactly what it computes alone: the batched kernels run the
same per-tile code on the member's own arguments.
"""
import collections
import concurrent.futures as cf
import os
import sys
import threading
import time

import numpy as np

from . import _abi
from .sim import Batch, Simulation, footprint_bytes

HBM_BYTES = 288e9


class Request:
    """One simulation request: spec + medium (php_fdtd_amd.sim form), run for
    spec["grid"]["tsteps"] steps, monitor trace returned."""

    def __init__(self, rid, spec, medium, arrival=0.0):
        self.id = rid
        self.spec = spec
        self.medium = medium
        self.steps = int(spec["grid"]["tsteps"])
        self.cells = int(medium.size)
        self.arrival = float(arrival)  # seconds after the stream starts
        self.footprint = 0
        self.t_arrival = self.t_admit = self.t_start = self.t_done = None
        self.rows = None
        self.error = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


def default_workers():
    """Build threads: the host cores this process may use, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 4
    return max(2, min(16, n))


def _build(req, precision):
    sim = Simulation(req.spec, req.medium, precision)
    m = req.spec.get("monitor")
    if m:
        sim.set_monitor(m["x"], m["y"], m["z"], m.get("atom", 0), m.get("dir", 0), ref_compat=True)
    return sim


class Server:
    def __init__(self, precision=_abi.FP32, hbm_budget=0.85 * HBM_BYTES, max_members=64, max_chunk=256,
                 workers=None, progress=None):
        self.precision = precision
        self.progress = progress  # seconds between status lines on stderr (None: quiet)
        self.budget = float(hbm_budget)
        self.max_members = int(max_members)
        self.max_chunk = int(max_chunk)
        # context builds (host precompute, uploads) run in parallel: with short
        # requests they, not the GPU, bound the batch size
        self.workers = int(workers) if workers else default_workers()
        self.lib = _abi.load_library()
        self._lock = threading.Lock()
        self._in_use = 0.0
        self.trace = []  # (time, members, chunk) per batch run

    def _finish(self, req, sim):
        try:
            req.rows = sim.monitor(req.steps)
            d = sim.sync()
            req.diverged = bool(d.diverged)
        except Exception as e:  # reported per request
            req.error = str(e)
        finally:
            sim.close()
            req.t_done = time.perf_counter()
            with self._lock:
                self._in_use -= req.footprint

    def serve(self, requests):
        """Serve ``requests`` (Request objects, arrival offsets in seconds);
        blocks until every request is done; returns the requests."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        batch = Batch(self.precision, self.lib)
        pool = cf.ThreadPoolExecutor(self.workers)
        pending = collections.deque()
        building = []   # (req, future)
        active = []     # [req, sim, steps left]
        finishing = []
        nxt = 0
        t0 = time.perf_counter()
        last = t0
        ndone = 0
        try:
            while nxt < len(reqs) or pending or building or active:
                now = time.perf_counter()
                while nxt < len(reqs) and t0 + reqs[nxt].arrival <= now:
                    r = reqs[nxt]
                    r.t_arrival = t0 + r.arrival
                    pending.append(r)
                    nxt += 1
                # admission: FIFO, by device footprint and member slots
                while pending and len(active) + len(building) < self.max_members:
                    r = pending[0]
                    if not r.footprint:
                        r.footprint = footprint_bytes(r.spec, r.medium, self.precision, self.lib)
                    with self._lock:
                        if r.footprint > self.budget:
                            r.error = "footprint %.3g B exceeds the HBM budget" % r.footprint
                            r.t_done = time.perf_counter()
                            pending.popleft()
                            continue
                        if self._in_use + r.footprint > self.budget:
                            break
                        self._in_use += r.footprint
                    pending.popleft()
                    r.t_admit = time.perf_counter()
                    building.append((r, pool.submit(_build, r, self.precision)))
                # built contexts join at this chunk boundary
                still = []
                for r, fut in building:
                    if fut.done():
                        try:
                            sim = fut.result()
                        except Exception as e:
                            r.error = str(e)
                            r.t_done = time.perf_counter()
                            with self._lock:
                                self._in_use -= r.footprint
                            continue
                        batch.add(sim)
                        r.t_start = time.perf_counter()
                        active.append([r, sim, r.steps])
                    else:
                        still.append((r, fut))
                building = still
                if not active:
                    # idle: wait for a build or the next arrival
                    wait = [f for _, f in building]
                    until = (t0 + reqs[nxt].arrival - time.perf_counter()) if nxt < len(reqs) else None
                    if wait:
                        cf.wait(wait, timeout=until if until is None or until > 0 else 0,
                                return_when=cf.FIRST_COMPLETED)
                    elif until is not None and until > 0:
                        time.sleep(until)
                    continue
                chunk = min(self.max_chunk, min(a[2] for a in active))
                batch.run(chunk)
                self.trace.append((time.perf_counter() - t0, len(active), chunk))
                for a in active:
                    a[2] -= chunk
                done = [a for a in active if a[2] == 0]
                ndone += len(done)
                if self.progress and time.perf_counter() - last >= self.progress:
                    last = time.perf_counter()
                    print("serve: %.0f s, %d done, %d running, %d building, %d queued" % (
                        last - t0, ndone, len(active) - len(done), len(building), len(pending) + len(reqs) - nxt),
                          file=sys.stderr, flush=True)
                if done:
                    active = [a for a in active if a[2] > 0]
                    for r, sim, _ in done:
                        batch.remove(sim)
                        finishing.append(pool.submit(self._finish, r, sim))
                batch.wait(1)
            batch.sync()
            for f in finishing:
                f.result()
        finally:
            pool.shutdown(wait=True)
            # after an error: release the contexts still held (a normal end holds none)
            for r, fut in building:
                if not fut.exception():
                    fut.result().close()
            for r, sim, _ in active:
                sim.close()
            batch.close()
        return reqs


def run_alone(req, precision):
    """One request on its own context (graph replays), end to end: build, run,
    read the monitor, destroy.  Returns (rows, seconds)."""
    t = time.perf_counter()
    sim = _build(req, precision)
    sim.run(req.steps)
    rows = sim.monitor(req.steps)
    sim.close()
    return rows, time.perf_counter() - t


def summarize(reqs):
    """Throughput and time-to-result of a served stream."""
    ok = [r for r in reqs if r.error is None and r.t_done is not None]
    lat = np.array([r.latency for r in ok])
    first = min(r.t_arrival for r in reqs)
    last = max(r.t_done for r in ok)
    span = last - first
    cell_updates = float(sum(r.cells * r.steps for r in ok))
    return dict(requests=len(reqs), completed=len(ok), makespan_s=round(span, 4),
                solves_per_s=round(len(ok) / span, 4), gcell_updates_per_s=round(cell_updates / span / 1e9, 3),
                p50_s=round(float(np.percentile(lat, 50)), 4), p99_s=round(float(np.percentile(lat, 99)), 4),
                mean_s=round(float(lat.mean()), 4))
