"""Request loop: continuous batching of t-SNE jobs on one device.

Requests arrive over time (a Poisson stream in the benchmark).  Preparer threads (four by
default) build arriving jobs' affinities (kNN, bandwidth search, symmetrisation) and optimiser
state, each job on its own stream, so that the builds of several requests overlap on the device;
a job is admitted only while the context's HBM budget has room for its footprint;
the stepping thread joins prepared jobs to the running batch at the next step boundary, advances
the batch a few epochs at a time (one graph replay per epoch for all members) and hands finished
jobs back.  Affinity builds of arriving jobs therefore overlap the epochs of running ones on the
device, and a job's result is the same as its fit alone (the batch computes each member exactly
as a single job)."""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import Batch, Context, TSNE, job_footprint
from . import _lib


@dataclass
class Request:
    rid: int
    x: np.ndarray
    d_out: int = 2
    method: str = "fitsne"
    theta: float = 0.5
    epochs: int = 1000
    perplexity: float = 30.0
    arrival: float = 0.0          # seconds after the start of the stream
    y0: Optional[np.ndarray] = None
    max_boxes: int = 1000         # FIt-SNE grid capacity per axis (sizes the job's grids)


@dataclass
class Result:
    rid: int
    n: int
    d_out: int
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    prepare_s: float = 0.0        # affinities + optimiser state on the preparing thread
    embedding: Optional[np.ndarray] = None
    kl: Optional[float] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def _make_job(req: Request, ctx: Context) -> TSNE:
    t = TSNE(req.x, d_out=req.d_out, ctx=ctx).perplexity(req.perplexity).epochs(req.epochs)
    t._set(max_boxes=req.max_boxes)
    if req.y0 is not None:
        t.initial_embedding(req.y0)
    return t


def _collect(job: TSNE, res: Result, want_kl: bool, t0: float):
    res.embedding = job.embedding()
    if want_kl:
        res.kl = job.kl_divergence()
    res.finished = time.perf_counter() - t0


def serve(requests: List[Request], ctx: Optional[Context] = None, step_epochs: int = 25,
          hbm_budget: int = 0, want_kl: bool = True, preparers: int = 4,
          graph: bool = True) -> List[Result]:
    """Serves the requests with continuous batching; returns one Result per request (times in
    seconds from the start of the stream).  `preparers` threads build arriving jobs'
    affinities concurrently (each job on its own stream), taking the requests in arrival
    order.  `graph=False` launches the batch's epochs plainly instead of replaying a captured
    graph (same kernels and results; for tracing tools that cannot follow graphs)."""
    ctx = ctx or Context(0, hbm_budget)
    reqs = sorted(requests, key=lambda r: r.arrival)
    results = {r.rid: Result(r.rid, r.x.shape[0], r.d_out, r.arrival) for r in reqs}
    ready: "queue.Queue" = queue.Queue()
    budget = hbm_budget
    reserved = {"bytes": 0}
    lock = threading.Condition()
    errors = []
    stats = {"step_s": 0.0, "steps": 0}
    t0 = time.perf_counter()

    pending = iter(reqs)
    take = threading.Lock()

    def preparer():
        try:
            while True:
                with take:
                    r = next(pending, None)
                if r is None:
                    break
                delay = r.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                need = job_footprint(r.x.shape[0], r.x.shape[1], r.d_out, r.method,
                                     perplexity=r.perplexity, max_boxes=r.max_boxes)
                with lock:  # HBM admission: wait until running jobs leave room
                    while budget and reserved["bytes"] + need > budget and reserved["bytes"] > 0:
                        lock.wait(0.05)
                    reserved["bytes"] += need
                p0 = time.perf_counter()
                job = _make_job(r, ctx)
                job.prepare(r.method, r.theta)
                results[r.rid].prepare_s = time.perf_counter() - p0
                ready.put((r, job, need))
        except Exception as e:  # surfaced by the stepping loop
            errors.append(e)
        finally:
            ready.put(None)

    nprep = max(1, int(preparers))
    threads = [threading.Thread(target=preparer, daemon=True) for _ in range(nprep)]
    for th in threads:
        th.start()
    batch = Batch(ctx, graph=graph)
    live = {}
    feeding = True
    running_preparers = nprep
    while feeding or len(batch):
        # join whatever has been prepared; block only when nothing is running
        while True:
            try:
                item = ready.get(block=not len(batch) and feeding, timeout=None)
            except queue.Empty:
                break
            if item is None:
                running_preparers -= 1
                feeding = running_preparers > 0
                if not feeding:
                    break
                continue
            r, job, need = item
            batch.add(job, r.method, r.theta)
            results[r.rid].admitted = time.perf_counter() - t0
            live[job._h.value] = (r, job, need)
            if ready.empty():
                break
        if errors:
            raise errors[0]
        if len(batch):
            s0 = time.perf_counter()
            batch.step(step_epochs)
            stats["step_s"] += time.perf_counter() - s0
            stats["steps"] += 1
            for job in batch.take_finished():
                r, _, need = live.pop(job._h.value)
                _collect(job, results[r.rid], want_kl, t0)
                with lock:
                    reserved["bytes"] -= need
                    lock.notify_all()
    for th in threads:
        th.join()
    if errors:
        raise errors[0]
    batch.close()
    serve.last_stats = dict(stats, span_s=time.perf_counter() - t0,
                            prepare_s=sum(r.prepare_s for r in results.values()))
    return [results[r.rid] for r in requests]


def serve_plain(requests: List[Request], ctx: Optional[Context] = None,
                want_kl: bool = True) -> List[Result]:
    """The same stream served one job at a time (first come, first served), each job a plain
    fit: the baseline of the batched loop."""
    ctx = ctx or Context(0)
    out = []
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda q: q.arrival):
        delay = r.arrival - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        res = Result(r.rid, r.x.shape[0], r.d_out, r.arrival)
        res.admitted = time.perf_counter() - t0
        job = _make_job(r, ctx)
        if r.method == "fitsne":
            job.fit_sne()
        else:
            job.barnes_hut(r.theta)
        _collect(job, res, want_kl, t0)
        out.append(res)
        del job
    by = {q.rid: q for q in out}
    return [by[r.rid] for r in requests]


def latency_summary(results: List[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {"jobs": len(results), "jobs_per_s": len(results) / span if span > 0 else None,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "span_s": span}
