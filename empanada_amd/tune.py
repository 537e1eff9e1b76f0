"""Autotuner of the batched engine (emp_batch_set_tuning), with a JSON cache.

What it tunes, by measurement on the GPU:

- the join/dedup kernel variant, which fixes the fine-bucket capacity: 512 entries per side with
  256 threads, or 1,024 with 512 threads (one choice per engine: every job of a step shares the
  launch);
- per class of jobs (read length x read count), the mean fine-bucket load (how many buckets the
  second partition pass fans a job out to, hence the LDS work per join work-group) and the mean
  coarse-bucket load (the split of the fan-out between the two partition passes: with a job of x
  entries, pass 1 writes x / coarse_load runs and pass 2 fans each out to coarse_load / fine_load).

Method: a calibration sample of the batched workload's distribution (by default sets disjoint from
the ones a benchmark measures) is split into the classes; for each kernel variant, coordinate
descent per class over the fine load, then the coarse load, each candidate timed as the wall time
of the class's jobs through one engine that has run the whole sample once (best of `repeats`); the
variant with the lower total wins. Outputs do not depend on the tuning, so every candidate is
valid.

The cache (JSON) holds the chosen tuning with the times that chose it; load() returns it when its
library version and class grid match, so a later run reuses it without measuring."""
import json
import os
import time

from . import gen, native

# Classes: read length at most (the workload's lengths 100, 150, 250) x reads at most.
READ_LEN_MAX = (100, 150, 255)
N_MAX = (32768, 262144, 1 << 23)
CAPS = (512, 1024)
FINE = {512: (128, 192, 256, 320, 384), 1024: (256, 384, 512, 640, 768)}
COARSE = (8192, 16384, 32768, 65536)
DEFAULT = dict(bucket_cap=512, classes=[])  # the engine's built-in loads (256 / 16384) for every job
FORMAT = 2  # of the cache file


def class_of(L, n):
    for li, lm in enumerate(READ_LEN_MAX):
        if L <= lm:
            for ni, nm in enumerate(N_MAX):
                if n <= nm:
                    return li, ni
    return None


def _grid():
    return [dict(read_len_max=lm, n_max=nm) for lm in READ_LEN_MAX for nm in N_MAX]


def _time_jobs(b, jobs, tuning, repeats):
    """Best wall time (s) of the jobs (device-resident: (address, n, L, min_ovlp)) through engine b
    under `tuning`."""
    b.set_tuning(tuning)
    best = None
    for _ in range(repeats):
        b.reset()
        t = time.perf_counter()
        for addr, n, L, o in jobs:
            b.submit_device(addr, n, L, o)
        b.run()
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    return best


def tune(sample, budget=160 << 30, repeats=2, log=None):
    """Tuning for jobs like `sample` (a list of (reads, min_ovlp)), timed with the reads resident in
    HBM (the engine's own work, not the host copies). Returns the cache dict."""
    from . import hipmem
    dev = hipmem.DeviceJobs(sample)
    try:
        return _tune(dev.jobs(), len(sample), budget, repeats, log)
    finally:
        dev.free()


def _tune(sample, n_sample, budget, repeats, log):
    by_class = {}
    for job in sample:
        k = class_of(job[2], job[1])
        if k is not None:
            by_class.setdefault(k, []).append(job)
    grid = _grid()
    t_start = time.perf_counter()
    results = {}
    with native.Batch(collect=False, hbm_budget=budget) as b:
        for cap in CAPS:  # a warm-up pass of the whole sample per variant: allocations, code objects
            _time_jobs(b, sample, dict(bucket_cap=cap, classes=[]) if cap == 512 else
                       dict(bucket_cap=cap, classes=[dict(read_len_max=255, n_max=1 << 23, fine_load=512,
                                                          coarse_load=16384)]), 1)
        for cap in CAPS:
            classes, total = [], 0.0
            for gi, g in enumerate(grid):
                k = (gi // len(N_MAX), gi % len(N_MAX))
                jobs = by_class.get(k)
                fine, coarse = FINE[cap][len(FINE[cap]) // 2], 16384
                if not jobs:  # no sample: the variant's middle loads
                    classes.append(dict(g, fine_load=fine, coarse_load=coarse, jobs=0, s=None))
                    continue

                def run(f, c):
                    t = dict(bucket_cap=cap, classes=[dict(g, fine_load=f, coarse_load=max(c, f))])
                    return _time_jobs(b, jobs, t, repeats)
                best = run(fine, coarse)
                trials = {"%d/%d" % (fine, coarse): round(best, 4)}
                for f in FINE[cap]:
                    if f != fine:
                        dt = run(f, coarse)
                        trials["%d/%d" % (f, coarse)] = round(dt, 4)
                        if dt < best:
                            best, fine = dt, f
                for c in COARSE:
                    if c != coarse:
                        dt = run(fine, c)
                        trials["%d/%d" % (fine, c)] = round(dt, 4)
                        if dt < best:
                            best, coarse = dt, c
                if log:
                    log("cap %d class L<=%d n<=%d: %d jobs, best fine %d coarse %d, %.4f s" %
                        (cap, g["read_len_max"], g["n_max"], len(jobs), fine, coarse, best))
                classes.append(dict(g, fine_load=fine, coarse_load=max(coarse, fine), jobs=len(jobs),
                                    s=round(best, 4), trials=trials))
                total += best
            results[cap] = (total, classes)
        b.set_tuning(None)
    cap = min(results, key=lambda c: results[c][0])
    return {
        "format": FORMAT,
        "library": native.version(),
        "arch": "gfx950",
        "grid": {"read_len_max": list(READ_LEN_MAX), "n_max": list(N_MAX)},
        "bucket_cap": cap,
        "classes": [{k: v for k, v in c.items() if k in ("read_len_max", "n_max", "fine_load", "coarse_load")}
                    for c in results[cap][1]],
        "measured": {str(c): {"total_s": round(results[c][0], 4), "classes": results[c][1]} for c in results},
        "sample_jobs": n_sample,
        "timed": "reads resident in HBM",
        "tuning_s": round(time.perf_counter() - t_start, 1),
    }


def load(path):
    """The cached tuning at `path` if it fits this library and class grid, else None."""
    try:
        with open(path) as f:
            c = json.load(f)
    except (OSError, ValueError):
        return None
    if c.get("format") != FORMAT or c.get("library") != native.version() or c.get("grid") != {"read_len_max": list(READ_LEN_MAX),
                                                                  "n_max": list(N_MAX)}:
        return None
    return c


def load_or_tune(path, sample_fn, budget=160 << 30, log=None):
    """The cached tuning, or a fresh one (sample_fn() gives the calibration jobs) written to path."""
    c = load(path)
    if c is not None:
        return c, True
    c = tune(sample_fn(), budget=budget, log=log)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(c, f, indent=1)
    os.replace(tmp, path)
    return c, False


def calibration_sample(count=256, first=4096):
    """Sets of the batched workload's distribution that a 4,096-set run does not measure."""
    return [gen.batched_set(i) for i in range(first, first + count)]


def main():
    import argparse
    ap = argparse.ArgumentParser(description="tune the batched engine and write the JSON cache")
    ap.add_argument("--cache", required=True)
    ap.add_argument("--sets", type=int, default=256, help="calibration sets (i = 4096 ...)")
    ap.add_argument("--budget", type=int, default=160, help="HBM budget, GiB")
    ap.add_argument("--force", action="store_true", help="measure even when the cache fits")
    a = ap.parse_args()
    if a.force and os.path.exists(a.cache):
        os.unlink(a.cache)
    c, reused = load_or_tune(a.cache, lambda: calibration_sample(a.sets), budget=a.budget << 30,
                             log=lambda m: print(m, flush=True))
    print(json.dumps({"reused": reused, "bucket_cap": c["bucket_cap"], "tuning_s": c.get("tuning_s"),
                      "totals": {k: v["total_s"] for k, v in c["measured"].items()}}), flush=True)


if __name__ == "__main__":
    main()
