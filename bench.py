"""Benchmark of the overlap path on one MI355X.

Workload (default): 50 Mbp i.i.d. genome, 15,000,000 reads of 100 bp, 50 % reverse-complemented,
0.5 % substitutions, minimum overlap 63 (seeds 1100 / 1101), as FASTQ (3.2 GB). One step is the
whole job from FASTQ bytes already in HBM to output.bin written and closed: the FASTQ parsed on the
device in blocks of map_mem/16 bytes (read_block's rules), fingerprints + 2-bit columns, partition,
dedup, every overlap round, edge compaction, the records to host memory and into the file. Each step
writes a new output.bin (the previous step's file is removed before the step, untimed).
Metric: Gbases/s = n * L / step time.

Beside it, on the same line:

- ``e2e_from_page_cache``: the reference's t_e2e, FASTQ file open (page-cached) to output.bin
  closed, through emp_add_fastq_file (the file streamed over PCIe while blocks are parsed) and
  emp_run_to_file: the PCIe-inclusive rate;
- ``overwrite_existing``: the headline step onto the previous step's output.bin (written over in
  place and cut to the new size: the file fopen "wb" would leave);
- ``reads_in_hbm``: the overlap path alone over kept reads already in HBM, records to host memory;
- ``roofline`` of the dominant kernel (HIP events recorded by the library on its stream around
  every launch, algorithmic bytes per launch from the live per-round counts);
- ``cpu_baseline``: the oracle on the host cores over the same FASTQ file, file to file, with the
  host described from lscpu; its output.bin is compared with the GPU's (``parity_full_vs_oracle``).

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

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
T0 = time.perf_counter()


def progress(what):
    """One line on stderr per finished phase (the JSON line alone goes to stdout)."""
    print("[bench %7.1f s] %s" % (time.perf_counter() - T0, what), file=sys.stderr, flush=True)


def host_cpu():
    """The host's CPU from lscpu: model, sockets, cores per socket, threads per core."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus"}
    for ln in out.splitlines():
        k, _, v = ln.partition(":")
        if k.strip() in keys:
            v = v.strip()
            info[keys[k.strip()]] = int(v) if v.isdigit() else v
    return info


def cpu_baseline(fq_path, n, L, min_ovlp, reduce_mem, gpu_out):
    """The oracle on the host cores, FASTQ file to output.bin (the t_e2e boundaries), over the whole
    workload; returns the baseline and whether its output.bin equals the GPU's."""
    import oracle_lib
    host = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or host.get("cpus") or os.cpu_count() or 1
    out_path = fq_path + ".oracle.bin"
    t0 = time.perf_counter()
    rc, summ = oracle_lib.run_file(fq_path, out_path, min_ovlp=min_ovlp, reduce_mem=reduce_mem)
    dt = time.perf_counter() - t0
    assert rc == 0, rc
    with open(out_path, "rb") as f:
        same = f.read() == gpu_out
    os.unlink(out_path)
    return {
        "value": n * L / dt / 1e9,
        "unit": "Gbases/s",
        "cores": threads,
        "kind": "port",
        "host": host,
        "sample": "the whole workload: FASTQ file (page-cached) to output.bin, %d reads x %d bp, %.2f s "
                  "(FASTQ parse single-threaded, overlap path on %d OpenMP threads)" % (n, L, dt, threads),
    }, same


def cpu_baseline_sample(workload, scale, min_ovlp, reduce_mem):
    """The oracle over a bounded sample of a workload too large for it (same genome/read model and
    coverage, `scale` of the size), reads in memory; its output is compared with the GPU's."""
    import oracle_lib
    from empanada_amd import gen, native
    reads, _ = gen.workload_reads(workload, scale)
    n, L = reads.shape
    host = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or host.get("cpus") or os.cpu_count() or 1
    t0 = time.perf_counter()
    rc, want, _, _ = oracle_lib.run_reads(reads.tobytes(), n, L, min_ovlp=min_ovlp, reduce_mem=reduce_mem)
    dt = time.perf_counter() - t0
    assert rc == 0, rc
    with native.Context(min_ovlp=min_ovlp, reduce_mem=reduce_mem) as c:
        c.add_reads(reads)
        same = c.run() == want
    return {
        "value": n * L / dt / 1e9, "unit": "Gbases/s", "cores": threads, "kind": "port", "host": host,
        "sample": "%.0f kbp genome, %d reads x %d bp (%g%% of the workload, same coverage), kept reads in memory "
                  "to records, %.2f s" % (gen.WORKLOADS[workload]["G"] * scale / 1e3, n, L, 100 * scale, dt),
    }, same


def ooc_bench(steps, cpu_scale, check_sync, no_cpu):
    """The 3.1 Gbp workload (413 M x 150 bp reads, 10 % repeats, -o 75 -r 96g): the reads in
    pageable host memory are streamed into HBM inside the timed region (packed to 2-bit codes by
    host threads into pinned stages, DMA of a quarter of the bytes, expanded and mapped on the
    device while the next stage is packed), then the overlap path with every record delivered to
    host memory."""
    from empanada_amd import gen, hipmem, native
    n, L = gen.workload_size("ooc")
    t = time.perf_counter()
    reads, o = gen.workload_reads("ooc", 1.0)
    t_gen = time.perf_counter() - t
    res = {"workload": "3.1 Gbp genome with 10%% repeats (300 bp and 6 kbp families, 16 copies, 1%% divergence), "
                       "%d x %d bp reads, 50%% RC, 0.5%% substitutions, -o %d -r 96g; reads in pageable host memory "
                       "streamed into HBM inside the timed region (2-bit codes over PCIe through pinned stages), "
                       "records to host memory" % (n, L, o),
           "reads": n, "read_len": L, "gen_s": round(t_gen, 2)}
    with native.Context(min_ovlp=o, reduce_mem=96 << 30) as c:
        def step():
            c.reset()
            c.add_reads(reads)
            c.run(collect=False)
        step()  # warm-up: allocations (about 110 GB of HBM)
        hipmem.synchronize()
        wall = 0.0
        for _ in range(steps):
            t0 = time.perf_counter()
            step()
            hipmem.synchronize()
            wall += time.perf_counter() - t0
        ms = wall * 1e3 / steps
        summ = c.summary()
        c.set_profile(True)
        c.kernel_clear()
        step()
        kern = c.kernels()
        c.set_profile(False)
    kj = kern.get("k_join", {"ms": 0, "bytes": 0, "launches": 1})
    res.update({"steps": steps, "ms_per_step": round(ms, 1), "gbases_per_s": round(n * L / (ms * 1e-3) / 1e9, 3),
                "h2d_bytes_per_step": n * L // 4,  # 2-bit codes (a batch with other bytes would go as ASCII)
                "dedup": summ["n_dedup"], "rounds": summ["n_rounds"], "edges": summ["n_edges"],
                "k_join_roofline_frac": round(kj["bytes"] / (kj["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if kj["ms"] else None,
                "kernels_ms": {k: round(v["ms"], 2) for k, v in kern.items()}})
    if check_sync:
        outs = []
        for sync in (False, True):
            if sync:
                os.environ["EMPANADA_SYNC_ROUNDS"] = "1"
            with native.Context(min_ovlp=o, reduce_mem=96 << 30) as cc:
                cc.add_reads(reads)
                outs.append(cc.run())
            os.environ.pop("EMPANADA_SYNC_ROUNDS", None)
        res["parity_full_pipelined_vs_synchronous"] = outs[0] == outs[1]
        del outs
    del reads
    if not no_cpu:
        cpu, same = cpu_baseline_sample("ooc", cpu_scale, o, 96 << 30)
        res["cpu_baseline"] = cpu
        res["parity_sample_vs_oracle"] = same
        res["speedup_vs_cpu"] = round(res["gbases_per_s"] / cpu["value"], 1)
    return res


def traffic_of(kernel, workload):
    """HBM bytes per launch of `kernel` measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a
    1 GiB copy's counters) and recorded in profiles/roofline_traffic.json; None when not measured."""
    try:
        with open(os.path.join(ROOT, "profiles", "roofline_traffic.json")) as f:
            t = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return t["traffic_bytes_per_launch"] if t and t.get("workload") == workload else None


def batched_bench(sets, plain_stride, hbm_budget, tuning_cache):
    """The sets of the batched workload through one batched engine (admission by HBM budget),
    against the plain single-job path run once per set on every plain_stride-th set, on the same
    GPU, with the reads (a) resident in HBM and (b) in host memory (copied to HBM inside the timed
    region). With resident reads the engine also runs untuned (its built-in bucket loads) and with
    every kernel launched one by one instead of replaying each round step's captured HIP graph; the
    tuning is the autotuner's choice, from its JSON cache or measured now on calibration sets of the
    same distribution that the run does not measure."""
    from empanada_amd import hipmem, native, tune
    n_sets = len(sets)
    reads = sum(r.shape[0] for r, _ in sets)
    bases = sum(r.size for r, _ in sets)
    t = time.perf_counter()
    tuning, reused = tune.load_or_tune(tuning_cache, tune.calibration_sample, budget=hbm_budget)
    t_tune = time.perf_counter() - t
    t = time.perf_counter()
    dev = hipmem.DeviceJobs(sets)
    t_up = time.perf_counter() - t
    djobs = dev.jobs()

    def run_batch(tuning, graph, device):
        with native.Batch(collect=False, hbm_budget=hbm_budget, tuning=tuning, graph=graph) as b:
            def submit_all(k):
                for i in range(k):
                    if device:
                        b.submit_device(*djobs[i])
                    else:
                        b.submit(*sets[i])
            submit_all(n_sets)  # warm-up pass: buffers sized for the whole workload, code objects
            b.run()
            b.reset()
            t = time.perf_counter()
            submit_all(n_sets)
            b.run()
            dt = time.perf_counter() - t
            return dt, [b.info(j)["n_edges"] for j in range(n_sets)], b.graph_stats()

    sub = list(range(0, n_sets, plain_stride))

    big = max(sub, key=lambda i: sets[i][0].size)

    def run_plain(device):
        with native.Context(min_ovlp=sets[0][1]) as c:
            c.add_reads(sets[big][0])  # warm-up: buffers sized for the largest set, code objects
            c.run(collect=False)
            edges = []
            t = time.perf_counter()
            for i in sub:
                c.reset()
                if device:
                    addr, n, L, _ = djobs[i]
                    c.add_reads_device(addr, n, L)
                else:
                    c.add_reads(sets[i][0])
                c.run(collect=False)
                edges.append(c.summary()["n_edges"])
            return time.perf_counter() - t, edges

    t_b, edges_b, (captures, replays) = run_batch(tuning, True, True)
    progress("batched, tuned: %.3f s" % t_b)
    t_u, edges_u, _ = run_batch(None, True, True)
    progress("batched, untuned: %.3f s" % t_u)
    t_n, edges_n, _ = run_batch(None, False, True)
    progress("batched, launches without graphs: %.3f s" % t_n)
    t_p, edges_p = run_plain(True)
    progress("plain path on %d sets: %.3f s" % (len(sub), t_p))
    dev.free()
    t_bh, edges_bh, _ = run_batch(tuning, True, False)
    progress("batched, reads from host memory: %.3f s" % t_bh)
    t_ph, edges_ph = run_plain(False)
    progress("plain path, reads from host memory: %.3f s" % t_ph)
    jb, jp = n_sets / t_b, len(sub) / t_p
    return {
        "workload": "%d sets: genome log-uniform 5 kbp-5 Mbp, L in {100,150,250}, coverage 10-50x, 50%% RC, "
                    "0.5%% substitutions, -o 31 (seeds 1300+i, 1300+i+4096)" % n_sets,
        "reads": reads, "bases": bases, "hbm_budget": hbm_budget,
        "reads_in": "HBM (uploaded before the timed region, %.1f s); host_reads: host memory, copied to HBM "
                    "inside the timed region by both paths" % t_up,
        "batched_s": round(t_b, 3), "batched_jobs_per_s": round(jb, 2),
        "batched_gbases_per_s": round(bases / t_b / 1e9, 3),
        "plain_sets": len(sub), "plain_stride": plain_stride, "plain_s": round(t_p, 3),
        "plain_jobs_per_s": round(jp, 2),
        "speedup": round(jb / jp, 2),
        "untuned_jobs_per_s": round(n_sets / t_u, 2),
        "tuned_vs_untuned": round(t_u / t_b, 3),
        "untuned_no_graph_jobs_per_s": round(n_sets / t_n, 2),
        "graph_vs_no_graph": round(t_n / t_u, 3),
        "graph_captures": captures, "graph_replays": replays,
        "host_reads": {"batched_jobs_per_s": round(n_sets / t_bh, 2), "plain_jobs_per_s": round(len(sub) / t_ph, 2),
                       "speedup": round((n_sets / t_bh) / (len(sub) / t_ph), 2),
                       "batched_gbases_per_s": round(bases / t_bh / 1e9, 3)},
        "tuning": {"cache": os.path.relpath(tuning_cache, ROOT), "reused_from_cache": reused,
                   "seconds": round(t_tune, 1), "bucket_cap": tuning["bucket_cap"],
                   "classes": tuning["classes"],
                   "calibration": "%d sets of the same distribution not measured here (i = 4096..), reads in HBM" %
                                  tuning.get("sample_jobs", 0)},
        "edges_equal_tuned_untuned_no_graph_host": edges_b == edges_u == edges_n == edges_bh,
        "edges_equal_on_plain_sets": [edges_b[i] for i in sub] == edges_p == edges_ph,
        "note": "plain path on every %d-th set (same distribution), batched on all of them" % plain_stride,
    }


def served_bench(jobs, loads, bound_ms, tuning):
    """Poisson arrivals of batched-workload sets at each offered load (jobs/s)."""
    from empanada_amd import serve
    n_jobs = len(jobs)
    out = []
    with serve.Server(tuning=tuning) as srv:
        srv.run(jobs[:4], [0.0] * 4)  # warm-up
        for k, rate in enumerate(loads):
            srv.engine.reset()
            arr = serve.poisson_arrivals(n_jobs, rate, 1400 + k)
            res = srv.run(jobs, arr)
            st = serve.latency_stats(res["latency"])
            out.append(dict(offered_jobs_per_s=round(rate, 2), achieved_jobs_per_s=round(n_jobs / res["wall"], 2),
                            p50_ms=round(st["p50_ms"], 2), p99_ms=round(st["p99_ms"], 2),
                            all_ok=all(s == 0 for s in res["status"])))
            progress("served load %.1f jobs/s: p99 %.1f ms" % (rate, st["p99_ms"]))
    within = [x["offered_jobs_per_s"] for x in out if x["p99_ms"] <= bound_ms]
    return {"jobs_per_load": n_jobs, "arrivals": "Poisson, seed 1400 + load index", "loads": out,
            "p99_bound_ms": bound_ms,
            "max_offered_load_within_bound": max(within) if within else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("incore", "ooc"), default="incore",
                    help="incore: 50 Mbp, 15 M x 100 bp, -o 63; ooc: 3.1 Gbp with repeats, 413 M x 150 bp, -o 75")
    ap.add_argument("--scale", type=float, default=1.0, help="fraction of the workload (genome and reads)")
    ap.add_argument("--cpu-scale", type=float, default=None,
                    help="fraction of the 3.1 Gbp workload the oracle runs on (default 0.01); the in-core "
                         "baseline is the whole workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true",
                    help="3.1 Gbp: also compare the pipelined rounds' output with the synchronous rounds' at full "
                         "size (in-core: the full-size output is always compared with the oracle's)")
    ap.add_argument("--ooc-steps", type=int, default=1,
                    help="timed steps of the 3.1 Gbp workload, its reads streamed from host memory (0: skip)")
    ap.add_argument("--batched-sets", type=int, default=4096, help="sets of the batched workload (0: skip)")
    ap.add_argument("--plain-stride", type=int, default=8,
                    help="the plain path's comparison runs every n-th set of the batched workload")
    ap.add_argument("--batched-budget", type=int, default=160, help="the batched engine's HBM budget, GiB")
    ap.add_argument("--tuning-cache", default=os.path.join(ROOT, "empanada_amd", "tuning_gfx950.json"),
                    help="the batched engine's autotuner cache (JSON; tuned on the GPU when absent or stale)")
    ap.add_argument("--served-jobs", type=int, default=1000, help="jobs per offered load when serving (0: skip)")
    ap.add_argument("--served-loads", default="1,2,4,8",
                    help="offered loads as multiples of the plain path's jobs/s")
    ap.add_argument("--p99-bound-ms", type=float, default=250.0, help="the declared p99 latency bound")
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
    wl = gen.WORKLOADS[args.workload]
    t_gen = time.perf_counter()
    reads, min_ovlp = gen.workload_reads(args.workload, args.scale)
    n, L = reads.shape
    t_gen = time.perf_counter() - t_gen
    progress("workload generated: %d x %d bp" % (n, L))
    # 2n entries must stay below M = reduce_mem / 16 (in-core order): -r 96g for the 3.1 Gbp case
    reduce_mem = (96 << 30) if ooc else (1 << 30)
    map_mem = 1 << 30
    ctx = native.Context(min_ovlp=min_ovlp, profile=False, reduce_mem=reduce_mem, map_mem=map_mem)
    stream = ctx.stream
    tmp = os.environ.get("TMPDIR", "/tmp")
    fq_path = os.path.join(tmp, "emp_bench_%s.fastq" % args.workload)
    out_path = os.path.join(tmp, "emp_bench_%s.output.bin" % args.workload)
    dev = {"reads": None, "fq": None}
    fq_size = 0
    if not ooc:
        t = time.perf_counter()
        gen.write_fastq(fq_path, reads)
        fq_size = os.path.getsize(fq_path)
        with open(fq_path, "rb") as f:
            dev["fq"] = hipmem.DeviceBuffer.from_array(np.frombuffer(f.read(), np.uint8))
        t_gen += time.perf_counter() - t

    def upload():
        if dev["reads"] is None:
            dev["reads"] = hipmem.DeviceBuffer.from_array(reads)
            hipmem.synchronize()

    def step_fastq():
        """FASTQ in HBM -> output.bin closed (the metric)."""
        ctx.reset()
        ctx.add_fastq_device(dev["fq"].address, fq_size)
        ctx.run_to_file(out_path)

    def step_file():
        """FASTQ file (page cache) -> output.bin closed: the PCIe-inclusive t_e2e."""
        ctx.reset()
        ctx.add_fastq_file(fq_path)
        ctx.run_to_file(out_path)

    def step_reads():
        """Kept reads in HBM -> records in host memory: the overlap path alone."""
        upload()
        ctx.reset()
        ctx.add_reads_device(dev["reads"].address, n, L)
        if ooc:  # the 3.1 Gbp reads leave HBM once mapped (62 GB)
            dev["reads"].free()
            dev["reads"] = None
        ctx.run(collect=False)

    def timed(step, steps, profile, before=None):
        """Mean wall and HIP-event milliseconds of `steps` runs of step, each bracketed by device
        synchronisation and events on the library stream."""
        ctx.set_profile(profile)
        ctx.kernel_clear()
        ev0, ev1 = hipmem.Event(), hipmem.Event()
        wall = gpu = 0.0
        for _ in range(steps):
            if before:
                before()
            hipmem.synchronize()
            t0 = time.perf_counter()
            ev0.record(stream)
            step()
            ev1.record(stream)
            ev1.synchronize()
            hipmem.synchronize()
            wall += time.perf_counter() - t0
            gpu += ev0.elapsed_ms(ev1)
        return wall * 1e3 / steps, gpu / steps

    def fresh_output():
        """Removes the previous step's output.bin (outside the timed region): every step writes a
        new file, as a first run of the tool does. Writing onto an existing file is measured apart
        (overwrite_existing)."""
        if os.path.exists(out_path):
            os.unlink(out_path)

    main_step = step_reads if ooc else step_fastq
    before = upload if ooc else fresh_output
    for _ in range(args.warmup):
        main_step()
    hipmem.synchronize()
    # the metric: a timed region without per-launch events
    ms_per_step, gpu_ms_region = timed(main_step, args.steps, False, before)
    summ = ctx.summary()
    # the roofline: the same steps again with HIP events around every launch on the library stream
    progress("%d timed steps: %.3f ms per step" % (args.steps, ms_per_step))
    ms_per_step_prof, _ = timed(main_step, args.steps, True, before)
    progress("%d profiled steps: %.3f ms per step" % (args.steps, ms_per_step_prof))
    kern = ctx.kernels()
    rounds = ctx.rounds()
    value = n * L / (ms_per_step * 1e-3) / 1e9

    e2e = reads_hbm = overwrite = None
    gpu_out = None
    if not ooc:
        with open(out_path, "rb") as f:
            gpu_out = f.read()
        file_steps = max(1, min(args.steps, 5))
        # the headline step onto an output.bin left by the step before
        ms_over, _ = timed(step_fastq, file_steps, False)
        overwrite = {"ms_per_step": round(ms_over, 3),
                     "gbases_per_s": round(n * L / (ms_over * 1e-3) / 1e9, 4), "steps": file_steps,
                     "path": "as the metric, with output.bin already there (written over in place and "
                             "cut to the new size: the file fopen \"wb\" leaves)"}
        step_file()  # warm-up of the streaming buffers
        ms_file, _ = timed(step_file, file_steps, False, fresh_output)
        with open(out_path, "rb") as f:
            same_file = f.read() == gpu_out
        e2e = {"ms_per_step": round(ms_file, 3), "gbases_per_s": round(n * L / (ms_file * 1e-3) / 1e9, 4),
               "steps": file_steps, "fastq_bytes": fq_size,
               "pcie_gbytes_per_s": round(fq_size / (ms_file * 1e-3) / 1e9, 2),
               "output_equal_to_hbm_step": same_file,
               "path": "emp_add_fastq_file + emp_run_to_file: FASTQ file open (page-cached) to output.bin closed"}
        step_reads()
        ms_reads, _ = timed(step_reads, args.steps, False)
        reads_hbm = {"ms_per_step": round(ms_reads, 3), "gbases_per_s": round(n * L / (ms_reads * 1e-3) / 1e9, 4),
                     "path": "emp_add_reads_device + emp_run: kept reads in HBM to records in host memory"}

    dom = max((k for k in kern if k.startswith("k_")), key=lambda k: kern[k]["ms"])
    kd = kern[dom]
    avg_ms = kd["ms"] / max(kd["launches"], 1)
    bytes_per_launch = kd["bytes"] / max(kd["launches"], 1)
    achieved = kd["bytes"] / (kd["ms"] * 1e-3) / 1e9 if kd["ms"] > 0 else 0.0

    audit = None
    if args.audit:  # not part of the metric: the audit adds a sort and string checks per list
        ctx.set_profile(False)
        ctx.set_audit(True)
        hipmem.synchronize()
        t0 = time.perf_counter()
        step_reads()
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

    ctx.close()
    for k in ("reads", "fq"):
        if dev[k] is not None:
            dev[k].free()
            dev[k] = None
    cpu, same_full = (None, None)
    if not args.no_cpu and not ooc:
        cpu, same_full = cpu_baseline(fq_path, n, L, min_ovlp, reduce_mem, gpu_out)
        progress("CPU baseline: %.4f Gbases/s, output equal: %s" % (cpu["value"], same_full))
    same_sample = same_sync = None
    if ooc and not args.no_cpu:
        cpu, same_sample = cpu_baseline_sample(args.workload, args.cpu_scale or 0.01, min_ovlp, reduce_mem)
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
    for pth in (fq_path, out_path):
        if os.path.exists(pth):
            os.unlink(pth)
    del reads, gpu_out

    ooc_line = None
    if args.ooc_steps > 0 and not ooc:
        ooc_line = ooc_bench(args.ooc_steps, args.cpu_scale or 0.01, args.check, args.no_cpu)
        progress("3.1 Gbp workload: %.1f ms per step" % ooc_line["ms_per_step"])

    batched = served = None
    if (args.batched_sets > 0 or args.served_jobs > 0) and not ooc:
        t = time.perf_counter()
        sets = [gen.batched_set(i) for i in range(max(args.batched_sets, args.served_jobs))]
        t_sets = time.perf_counter() - t
        plain_rate = None
        tuning = None
        if args.batched_sets > 0:
            progress("batched workload generated: %d sets" % len(sets))
            batched = batched_bench(sets[:args.batched_sets], args.plain_stride, args.batched_budget << 30,
                                    args.tuning_cache)
            batched["gen_s"] = round(t_sets, 2)
            plain_rate = batched["plain_jobs_per_s"]
            tuning = batched["tuning"]
        if args.served_jobs > 0:
            if plain_rate is None:
                plain_rate = batched_bench(sets[:256], 4, args.batched_budget << 30,
                                           args.tuning_cache)["plain_jobs_per_s"]
            loads = [float(x) * plain_rate for x in args.served_loads.split(",")]
            served = served_bench(sets[:args.served_jobs], loads, args.p99_bound_ms, tuning)
            progress("served: %d loads" % len(loads))
        del sets

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
            "workload": ("%s: %d bp genome%s, %d x %d bp reads, 50%% RC, 0.5%% substitutions, -o %d%s; %s" % (
                args.workload, int(wl["G"] * args.scale),
                " (10% repeats: 300 bp and 6 kbp families, 16 copies, 1% divergence)" if ooc else "",
                n, L, min_ovlp, ", -r 96g" if ooc else "",
                "kept reads in HBM to records in host memory" if ooc else
                "%d bytes of FASTQ in HBM to output.bin closed" % fq_size)),
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
        "e2e_from_page_cache": e2e,
        "overwrite_existing": overwrite,
        "reads_in_hbm": reads_hbm,
        "cpu_baseline": cpu,
        "parity_sample_vs_oracle": same_sample,
        "parity_full_vs_oracle": same_full,
        "parity_full_pipelined_vs_synchronous": same_sync,
        "gen_s": round(t_gen, 2),
        "ooc": ooc_line,
        "batched": batched,
        "served": served,
    }
    if audit is not None:
        line["audit"] = audit
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
