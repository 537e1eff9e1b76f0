"""Benchmark of the overlap path on one MI355X.

Workload (default): 50 Mbp i.i.d. genome, 15,000,000 reads of 100 bp, 50 % reverse-complemented,
0.5 % substitutions, minimum overlap 63 (seeds 1100 / 1101). One step = the whole path over the
reads, which are resident in HBM when the timed region starts: fingerprints + 2-bit columns,
partition, dedup, every overlap round, edge compaction and delivery of every output.bin record to
host memory. Metric: Gbases/s = n * L / step time.

Prints one JSON line with the roofline of the dominant kernel (HIP events recorded by the library
on its stream around every launch, algorithmic bytes per launch from the live per-round counts)
and the oracle timed on the host cores over a bounded sample of the same workload.

Two more measurements ride on the same line (skip them with --batched-sets 0 / --served-jobs 0):

- ``batched``: read sets of the batched workload (genome log-uniform over 5 kbp-5 Mbp, L from
  {100, 150, 250}, coverage 10-50x, -o 31) run through one batched engine, against the plain
  single-job path run once per set on the same GPU; jobs/s of both, reads in host memory.
- ``served``: the same distribution arriving as a Poisson stream at stepped offered loads, served
  by the engine with continuous batching; achieved jobs/s and p50/p99 latency per load.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def cpu_baseline(workload, scale, min_ovlp, check_gpu):
    """The oracle over a bounded sample (same genome/read model and coverage, `scale` of the size)."""
    import oracle_lib
    from empanada_amd import gen, native

    reads, _ = gen.workload_reads(workload, scale)
    n, L = reads.shape
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    rc, want, summ, _ = oracle_lib.run_reads(reads.tobytes(), n, L, min_ovlp=min_ovlp)
    dt = time.perf_counter() - t0
    assert rc == 0, rc
    same = None
    if check_gpu:
        with native.Context(min_ovlp=min_ovlp) as c:
            c.add_reads(reads)
            same = c.run() == want
    return {
        "value": n * L / dt / 1e9,
        "unit": "Gbases/s",
        "cores": threads,
        "kind": "port",
        "sample": "%.0f kbp genome, %d reads x %d bp (%g%% of the workload, same coverage), %.2f s" % (
            gen.WORKLOADS[workload]["G"] * scale / 1e3, n, L, 100 * scale, dt),
    }, same


def traffic_of(kernel, workload):
    """HBM bytes per launch of `kernel` measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a
    1 GiB copy's counters) and recorded in profiles/roofline_traffic.json; None when not measured."""
    try:
        with open(os.path.join(ROOT, "profiles", "roofline_traffic.json")) as f:
            t = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return t["traffic_bytes_per_launch"] if t and t.get("workload") == workload else None


def batched_bench(n_sets):
    """The first n_sets sets of the batched workload: one batched engine vs the plain path per set."""
    from empanada_amd import gen, native
    t = time.perf_counter()
    sets = [gen.batched_set(i) for i in range(n_sets)]
    t_gen = time.perf_counter() - t
    reads = sum(r.shape[0] for r, _ in sets)
    bases = sum(r.size for r, _ in sets)
    with native.Batch(collect=False) as b:
        for r, o in sets[:8]:  # warm-up: allocations, code objects
            b.submit(r, o)
        b.run()
        b.reset()
        t = time.perf_counter()
        for r, o in sets:
            b.submit(r, o)
        b.run()
        t_b = time.perf_counter() - t
        edges_b = sum(b.info(j)["n_edges"] for j in range(n_sets))
    with native.Context(min_ovlp=sets[0][1]) as c:
        c.add_reads(sets[0][0])
        c.run(collect=False)
        edges_p = 0
        t = time.perf_counter()
        for r, o in sets:
            c.reset()
            c.add_reads(r)
            c.run(collect=False)
            edges_p += c.summary()["n_edges"]
        t_p = time.perf_counter() - t
    return {
        "workload": "%d sets: genome log-uniform 5 kbp-5 Mbp, L in {100,150,250}, coverage 10-50x, 50%% RC, "
                    "0.5%% substitutions, -o 31 (seeds 1300+i, 1300+i+4096)" % n_sets,
        "reads": reads, "bases": bases,
        "batched_s": round(t_b, 4), "plain_s": round(t_p, 4),
        "batched_jobs_per_s": round(n_sets / t_b, 2), "plain_jobs_per_s": round(n_sets / t_p, 2),
        "speedup": round(t_p / t_b, 2),
        "batched_gbases_per_s": round(bases / t_b / 1e9, 3),
        "edges_equal": edges_b == edges_p,
        "reads_in": "host memory (both paths copy them to HBM inside the timed region)",
        "gen_s": round(t_gen, 2),
    }


def served_bench(n_jobs, loads):
    """Poisson arrivals of batched-workload sets at each offered load (jobs/s)."""
    from empanada_amd import gen, serve
    jobs = [gen.batched_set(i) for i in range(n_jobs)]
    out = []
    with serve.Server() as srv:
        srv.run(jobs[:4], [0.0] * 4)  # warm-up
        for k, rate in enumerate(loads):
            srv.engine.reset()
            arr = serve.poisson_arrivals(n_jobs, rate, 1400 + k)
            res = srv.run(jobs, arr)
            st = serve.latency_stats(res["latency"])
            out.append(dict(offered_jobs_per_s=round(rate, 2), achieved_jobs_per_s=round(n_jobs / res["wall"], 2),
                            p50_ms=round(st["p50_ms"], 2), p99_ms=round(st["p99_ms"], 2),
                            all_ok=all(s == 0 for s in res["status"])))
    return {"jobs_per_load": n_jobs, "arrivals": "Poisson, seed 1400 + load index", "loads": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("incore", "ooc"), default="incore",
                    help="incore: 50 Mbp, 15 M x 100 bp, -o 63; ooc: 3.1 Gbp with repeats, 413 M x 150 bp, -o 75")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the workload (genome and reads)")
    ap.add_argument("--cpu-scale", type=float, default=None,
                    help="fraction of the workload the oracle runs on (default 1.0 in-core, 0.01 ooc)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true",
                    help="also compare the full-size output with the oracle (in-core), or the pipelined rounds' "
                         "output with the synchronous rounds' (3.1 Gbp)")
    ap.add_argument("--batched-sets", type=int, default=256, help="sets of the batched workload (0: skip)")
    ap.add_argument("--served-jobs", type=int, default=96, help="jobs per offered load when serving (0: skip)")
    ap.add_argument("--served-loads", default="0.5,1,2,4,8",
                    help="offered loads as multiples of the plain path's jobs/s")
    ap.add_argument("--audit", action="store_true",
                    help="after the timed steps, one more step with the collision audit on (reported apart)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES",
                    help="only run three device-to-device copies of BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibrate:
        from empanada_amd import hipmem
        a, b = hipmem.DeviceBuffer(args.copy_calibrate), hipmem.DeviceBuffer(args.copy_calibrate)
        for _ in range(3):
            hipmem.copy_d2d(b, a, args.copy_calibrate)
        hipmem.synchronize()
        print(json.dumps({"copy_calibrate_bytes": args.copy_calibrate, "copies": 3}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this path is measured on a single MI355X (--gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    from empanada_amd import gen, hipmem, native

    ooc = args.workload == "ooc"
    cpu_scale = args.cpu_scale if args.cpu_scale is not None else (0.01 if ooc else 1.0)
    wl = gen.WORKLOADS[args.workload]
    t_gen = time.perf_counter()
    reads, min_ovlp = gen.workload_reads(args.workload, args.scale)
    n, L = reads.shape
    t_gen = time.perf_counter() - t_gen
    # 2n entries must stay below M = reduce_mem / 16 (in-core order): -r 96g for the 3.1 Gbp case
    ctx = native.Context(min_ovlp=min_ovlp, profile=False, reduce_mem=(96 << 30) if ooc else (1 << 30))
    stream = ctx.stream
    dev = {"reads": None}

    def upload():
        if dev["reads"] is None:
            dev["reads"] = hipmem.DeviceBuffer.from_array(reads)
            hipmem.synchronize()

    def step():
        ctx.reset()
        ctx.add_reads_device(dev["reads"].address, n, L)
        if ooc:  # the 3.1 Gbp reads leave HBM once mapped (62 GB)
            dev["reads"].free()
            dev["reads"] = None
        ctx.run(collect=False)

    for _ in range(args.warmup):
        upload()
        step()
    hipmem.synchronize()

    def timed(profile):
        """Per step: reads uploaded (outside the timed region), then the step bracketed by device
        synchronisation and HIP events on the library stream."""
        ctx.set_profile(profile)
        ctx.kernel_clear()
        ev0, ev1 = hipmem.Event(), hipmem.Event()
        wall = gpu = 0.0
        for _ in range(args.steps):
            upload()
            hipmem.synchronize()
            t0 = time.perf_counter()
            ev0.record(stream)
            step()
            ev1.record(stream)
            ev1.synchronize()
            hipmem.synchronize()
            wall += time.perf_counter() - t0
            gpu += ev0.elapsed_ms(ev1)
        return wall * 1e3 / args.steps, gpu / args.steps

    # the metric: a timed region without per-launch events
    ms_per_step, gpu_ms_region = timed(False)
    # the roofline: the same steps again with HIP events around every launch on the library stream
    ms_per_step_prof, _ = timed(True)
    kern = ctx.kernels()
    summ = ctx.summary()
    rounds = ctx.rounds()

    dom = max(kern, key=lambda k: kern[k]["ms"])
    kd = kern[dom]
    avg_ms = kd["ms"] / max(kd["launches"], 1)
    bytes_per_launch = kd["bytes"] / max(kd["launches"], 1)
    achieved = kd["bytes"] / (kd["ms"] * 1e-3) / 1e9 if kd["ms"] > 0 else 0.0
    value = n * L / (ms_per_step * 1e-3) / 1e9

    audit = None
    if args.audit:  # not part of the metric: the audit adds a sort and string checks per list
        upload()
        ctx.set_profile(False)
        ctx.set_audit(True)
        hipmem.synchronize()
        t0 = time.perf_counter()
        step()
        hipmem.synchronize()
        t_audit = time.perf_counter() - t0
        ctx.set_audit(False)
        au = ctx.audits()
        coll = [a["strings"] - a["keys"] for a in au]
        audit = {
            "seconds": round(t_audit, 3),
            "lists": len(au),
            "collisions": sum(coll),
            "max_per_list": max(coll) if coll else 0,
            "by_list": {name: {"lists": sum(1 for a in au if a["list"] == k),
                               "entries": sum(a["entries"] for a in au if a["list"] == k),
                               "collisions": sum(a["strings"] - a["keys"] for a in au if a["list"] == k)}
                        for k, name in ((0, "reads"), (1, "suffixes"), (2, "prefixes"))},
            "expected_per_list_round_0": round(au[1]["entries"] ** 2 / 2 / 2.0**58, 4) if len(au) > 1 else None,
        }

    same_full = same_sync = None
    if args.check and not ooc:
        import oracle_lib
        rc, want, _, _ = oracle_lib.run_reads(reads.tobytes(), n, L, min_ovlp=min_ovlp)
        ctx.reset()
        ctx.add_reads(reads)
        same_full = rc == 0 and ctx.run() == want

    cpu, same_sample = (None, None)
    if not args.no_cpu:
        cpu, same_sample = cpu_baseline(args.workload, cpu_scale, min_ovlp, check_gpu=True)
    ctx.close()
    if args.check and ooc:
        # full size, where the oracle would take hours: the back-to-back rounds against the rounds
        # with the host between them (counters read after every join), byte for byte; one
        # context at a time (each holds about 110 GB)
        upload()
        outs = []
        for sync in (False, True):
            if sync:
                os.environ["EMPANADA_SYNC_ROUNDS"] = "1"
            with native.Context(min_ovlp=min_ovlp, reduce_mem=96 << 30) as cc:
                cc.add_reads_device(dev["reads"].address, n, L)
                outs.append(cc.run())
            os.environ.pop("EMPANADA_SYNC_ROUNDS", None)
        same_sync = outs[0] == outs[1]
        del outs
    if dev["reads"] is not None:
        dev["reads"].free()
    del reads

    batched = batched_bench(args.batched_sets) if args.batched_sets > 0 and not ooc else None
    served = None
    if args.served_jobs > 0 and batched:
        loads = [float(x) * batched["plain_jobs_per_s"] for x in args.served_loads.split(",")]
        served = served_bench(args.served_jobs, loads)

    line = {
        "metric": "Gbases/s end-to-end overlap detection on 1 GPU",
        "value": round(value, 4),
        "unit": "Gbases/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": ("%s: %d bp genome%s, %d x %d bp reads, 50%% RC, 0.5%% substitutions, -o %d%s" % (
                args.workload, int(wl["G"] * args.scale),
                " (10% repeats: 300 bp and 6 kbp families, 16 copies, 1% divergence)" if ooc else "",
                n, L, min_ovlp, ", -r 96g" if ooc else "")),
            "reads": n, "read_len": L, "min_ovlp": min_ovlp,
            "dedup": summ["n_dedup"], "rounds": summ["n_rounds"], "edges": summ["n_edges"],
            "records": summ["n_records"],
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_of(dom, args.workload),
            "alg_bytes_per_launch": round(bytes_per_launch),
            "avg_launch_ms": round(avg_ms, 4),
        },
        "kernels": {k: {"launches": v["launches"] // max(args.steps, 1), "ms_per_step": round(v["ms"] / args.steps, 3),
                        "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
                    for k, v in kern.items()},
        "gpu_event_ms_per_step": round(gpu_ms_region, 3),
        "ms_per_step_profiled": round(ms_per_step_prof, 3),
        "cpu_baseline": cpu,
        "parity_sample_vs_oracle": same_sample,
        "parity_full_vs_oracle": same_full,
        "parity_full_pipelined_vs_synchronous": same_sync,
        "gen_s": round(t_gen, 2),
        "batched": batched,
        "served": served,
    }
    if audit is not None:
        line["audit"] = audit
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
