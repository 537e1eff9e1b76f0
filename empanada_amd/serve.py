"""Request loop over the batched engine (include/empanada_batch.h): jobs arrive over time, are
admitted by the engine's HBM budget as earlier jobs finish and free memory, and share every step
with the jobs already resident (continuous batching: a step runs one round of every resident
job, whatever round each is at). Reports latency (arrival to last record) and throughput.

The engine owns admission, device memory and the kernels; this loop only submits jobs when their
arrival time has passed and steps the engine."""
import time

import numpy as np

from . import native


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """n arrival times (seconds) of a Poisson process of `rate` jobs per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


class Server:
    """One engine serving a stream of jobs. collect=True keeps every job's records (output.bin
    bytes via output(job))."""

    def __init__(self, hbm_budget=0, max_active=0, device=0, collect=False, tuning=None):
        self.engine = native.Batch(hbm_budget=hbm_budget, max_active=max_active, device=device, collect=collect,
                                   tuning=tuning)

    def close(self):
        self.engine.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, jobs, arrivals):
        """jobs: list of (reads, min_ovlp); arrivals: non-decreasing seconds from the start. Returns
        a dict with per-job latency (seconds), the job ids, and the wall time of the whole run."""
        b = self.engine
        order = np.argsort(arrivals, kind="stable")
        t0 = b.clock_ms()
        ids = [None] * len(jobs)
        k, active, queued = 0, 0, 0
        while k < len(jobs) or active or queued:
            now = (b.clock_ms() - t0) * 1e-3
            while k < len(jobs) and arrivals[order[k]] <= now:
                j = order[k]
                ids[j] = b.submit(*jobs[j])
                k += 1
                queued += 1
            if active or queued:
                active, queued = b.step()
            elif k < len(jobs):
                time.sleep(max(0.0, arrivals[order[k]] - now))
        wall = (b.clock_ms() - t0) * 1e-3
        lat = np.array([(b.info(ids[j])["done_ms"] - t0) * 1e-3 - arrivals[j] for j in range(len(jobs))])
        status = [b.status.get(ids[j]) for j in range(len(jobs))]
        return dict(latency=lat, ids=ids, wall=wall, status=status)

    def output(self, job: int) -> bytes:
        return self.engine.output(job)


def latency_stats(lat: np.ndarray):
    return dict(p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3), max_ms=float(lat.max() * 1e3))
