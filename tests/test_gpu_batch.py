"""GPU parity of the batched engine (include/empanada_batch.h) through the C ABI.

Every job's records must be byte for byte the output.bin the oracle writes for that job alone
(and so what a single-job context writes), whatever else shares the engine: jobs of other sizes,
read lengths and minimum overlaps, jobs admitted in later steps, an HBM budget that admits them in
waves, jobs that fail."""
import random

import numpy as np
import pytest

import oracle_lib as ol
from empanada_amd import gen, native

pytestmark = pytest.mark.gpu


def reads_of(strings):
    return np.frombuffer("".join(strings).encode(), dtype=np.uint8).reshape(len(strings), -1)


def oracle_out(r, o):
    n, L = r.shape if r.size else (0, 1)
    rc, want, summ, _ = ol.run_reads(r.tobytes(), n, L, min_ovlp=o)
    assert rc == 0
    return want, summ


# (genome length, reads, read length, min overlap, P(reverse complement), P(substitution), seeds)
JOBS = [
    (20_000, 3000, 100, 31, 0.5, 0.005, (101, 102)),
    (5_000, 1500, 150, 31, 0.5, 0.0, (103, 104)),
    (8_000, 1200, 250, 31, 0.5, 0.005, (105, 106)),
    (200_000, 40000, 100, 63, 0.5, 0.005, (107, 108)),
    (1_000, 600, 100, 31, 0.5, 0.0, (109, 110)),       # 60x: large equal-key groups
    (50_000, 7000, 150, 75, 0.0, 0.0, (111, 112)),
    (3_000, 16, 100, 31, 0.5, 0.0, (113, 114)),
    (100_000, 9000, 250, 100, 0.5, 0.01, (115, 116)),
]


def make_jobs(spec):
    out = []
    for G, n, L, o, p_rc, p_sub, seeds in spec:
        g = gen.genome(G, seeds[0])
        out.append((gen.reads(g, n, L, p_rc, p_sub, seeds[1]), o))
    return out


def test_batch_jobs_match_the_oracle_and_the_single_path():
    jobs = make_jobs(JOBS)
    with native.Batch() as b:
        ids = [b.submit(r, o) for r, o in jobs]
        b.run()
        for (r, o), j in zip(jobs, ids):
            want, summ = oracle_out(r, o)
            assert b.status[j] == 0
            got = b.output(j)
            assert got == want, "job %d differs from the oracle" % j
            info = b.info(j)
            assert info["n_dedup"] == summ["n_dedup"] and info["n_edges"] == summ["n_edges"]
    r, o = jobs[3]
    with native.Context(min_ovlp=o) as c:
        c.add_reads(r)
        assert c.run() == oracle_out(r, o)[0]


def test_jobs_admitted_in_later_steps_and_budget_waves():
    jobs = make_jobs(JOBS[:6])
    wants = [oracle_out(r, o)[0] for r, o in jobs]
    # continuous: a new job every other step while earlier ones are mid-run
    with native.Batch() as b:
        ids, k = [], 0
        while k < len(jobs) or len(b.finished) < len(jobs):
            if k < len(jobs):
                ids.append(b.submit(*jobs[k]))
                k += 1
            b.step()
            b.step()
        for j, want in zip(ids, wants):
            assert b.output(j) == want
        starts = [b.info(j)["admit_ms"] for j in ids]
        assert starts == sorted(starts)
    # a budget that holds the largest job alone: admitted in waves, same outputs
    need = [native.job_bytes(r.shape[0], r.shape[1]) for r, _ in jobs]
    budget = max(need)
    assert sum(need) > budget
    with native.Batch(hbm_budget=budget) as b:
        ids = [b.submit(r, o) for r, o in jobs]
        b.run()
        for j, want in zip(ids, wants):
            assert b.output(j) == want
        spans = [(b.info(j)["admit_ms"], b.info(j)["done_ms"]) for j in ids]
        for t, _ in spans:  # the resident jobs never need more than the budget
            assert sum(c for c, (a, d) in zip(need, spans) if a <= t < d) <= budget
        assert max(a for a, _ in spans) > min(d for _, d in spans)  # a later wave
    # at most 3 resident jobs
    with native.Batch(max_active=3) as b:
        ids = [b.submit(r, o) for r, o in jobs]
        b.run()
        assert all(b.output(j) == w for j, w in zip(ids, wants))


def test_edge_jobs():
    rng = random.Random(9)
    s = "".join(rng.choice("ACGT") for _ in range(60))
    common = "".join(rng.choice("ACGT") for _ in range(30))
    big_group = ["".join(rng.choice("ACGT") for _ in range(10)) + common for _ in range(700)] + \
                [common + "".join(rng.choice("ACGT") for _ in range(10)) for _ in range(700)]
    cases = [
        (np.zeros((0, 100), dtype=np.uint8), 31),                       # empty
        (reads_of(["ACGTTGCAACGTTGCA"]), 4),                            # one read
        (reads_of(["ACGTACGT", "ACGTACGT", "TTTTTTTT", "AAAAAAAA"]), 2),  # palindrome, duplicate, RC pair
        (reads_of(["acgtacgtgg", "CGTACGTGGA", "gtacgtggat"]), 3),     # lowercase
        (reads_of([s] * 3000 + ["".join(rng.choice("ACGT") for _ in range(60)) for _ in range(200)]), 20),
        (reads_of(big_group), 29),                                      # 700 x 700 key group
        (reads_of(["ACGTACGTAC"] * 3), 10),                             # min overlap = read length
    ]
    with native.Batch() as b:
        ids = [b.submit(r, o) for r, o in cases]
        b.run()
        for (r, o), j in zip(cases, ids):
            assert b.status[j] == 0
            assert b.output(j) == oracle_out(r, o)[0], "case %d" % j


def test_a_failing_job_does_not_disturb_the_others():
    jobs = make_jobs(JOBS[:2])
    bad = reads_of(["ACGTACGTAC", "ACGRACGTAC"])
    with native.Batch() as b:
        j0 = b.submit(*jobs[0])
        jb = b.submit(bad, 3)
        j1 = b.submit(*jobs[1])
        b.run()
        assert b.status[jb] == -5
        assert b.output(j0) == oracle_out(*jobs[0])[0]
        assert b.output(j1) == oracle_out(*jobs[1])[0]


def test_host_reads_cross_as_codes():
    """Jobs from host memory cross PCIe as 2-bit codes, expanded on the device: a job of 700,001 x
    101 bp reads (two 64 MB stages, the second ending inside a 4-byte word and inside a packed
    byte) with every 5th read in lower case gives the oracle's records for the upper-case reads; the
    same job with one stray byte in its second stage crosses that stage as it is and fails with
    EMP_ERR_BASE; a small clean job beside them is undisturbed."""
    g = gen.genome(2_000_000, 121)
    r = gen.reads(g, 700_001, 101, 0.5, 0.005, 122)
    want, _ = oracle_out(r, 51)
    low = r.copy()
    low[::5] |= 0x20
    bad = r.copy()
    bad[690_000, 50] = ord("N")
    small = make_jobs(JOBS[:1])[0]
    with native.Batch() as b:
        j0 = b.submit(low, 51)
        j1 = b.submit(bad, 51)
        j2 = b.submit(*small)
        b.run()
        assert b.status[j0] == 0 and b.output(j0) == want
        assert b.status[j1] == -5
        assert b.status[j2] == 0 and b.output(j2) == oracle_out(*small)[0]


def test_fastq_submission_and_reuse():
    g = gen.genome(30_000, 201)
    r = gen.reads(g, 4000, 100, 0.5, 0.005, 202)
    fq = gen.fastq_bytes(r)
    want = oracle_out(r, 31)[0]
    with native.Batch() as b:
        j = b.submit_fastq(fq, 31)
        b.run()
        assert b.output(j) == want
        b.reset()
        j = b.submit(r, 31)
        b.run()
        assert b.output(j) == want


def test_served_jobs_are_the_same_alone_and_in_the_mix():
    """Jobs from the batched-set distribution (small ones, so the oracle stays quick) arrive as a
    Poisson stream faster than they finish; each job's output equals the oracle's and the output of
    the job run alone."""
    from empanada_amd import serve
    picks = [i for i in range(400) if gen.batched_set_params(i)["n"] <= 12000][:16]
    jobs = [gen.batched_set(i) for i in picks]
    arrivals = serve.poisson_arrivals(len(jobs), rate=2000.0, seed=1400)
    with serve.Server(collect=True) as srv:
        res = srv.run(jobs, arrivals)
        assert all(s == 0 for s in res["status"])
        outs = [srv.output(j) for j in res["ids"]]
    overlapped = sum(1 for a in arrivals if a < res["wall"] / 2)
    assert overlapped > 1
    for (r, o), got in zip(jobs, outs):
        assert got == oracle_out(r, o)[0]
    for (r, o), got in list(zip(jobs, outs))[:3]:
        with native.Batch() as b:
            j = b.submit(r, o)
            b.run()
            assert b.output(j) == got


def test_graph_replays_equal_launches_and_the_oracle():
    """Round steps replayed from captured graphs (grids above the step's needs, tables updated
    between replays, admission steps launched eagerly in between) against the same jobs with every
    kernel launched one by one, and against the oracle."""
    jobs = make_jobs(JOBS)
    wants = [oracle_out(r, o)[0] for r, o in jobs]
    outs = {}
    for graph in (True, False):
        with native.Batch(graph=graph) as b:
            ids, k = [], 0
            while k < len(jobs) or len(b.finished) < len(jobs):
                if k < len(jobs):  # one job admitted per step: tables change under the graph
                    ids.append(b.submit(*jobs[k]))
                    k += 1
                b.step()
            outs[graph] = [b.output(j) for j in ids]
            captures, replays = b.graph_stats()
        if graph:
            assert captures >= 1 and replays > 2 * captures
        else:
            assert (captures, replays) == (0, 0)
    assert outs[True] == outs[False] == wants


@pytest.mark.parametrize("tuning", [
    dict(bucket_cap=1024, classes=[dict(read_len_max=255, n_max=1 << 23, fine_load=768, coarse_load=65536)]),
    dict(bucket_cap=1024, classes=[dict(read_len_max=100, n_max=5000, fine_load=256, coarse_load=8192),
                                   dict(read_len_max=255, n_max=1 << 23, fine_load=512, coarse_load=16384)]),
    dict(bucket_cap=512, classes=[dict(read_len_max=150, n_max=1 << 23, fine_load=128, coarse_load=65536),
                                  dict(read_len_max=255, n_max=1 << 23, fine_load=384, coarse_load=8192)]),
])
def test_tunings_give_the_oracles_records(tuning):
    """Every kernel variant and bucket load the autotuner may pick leaves the records unchanged."""
    jobs = make_jobs(JOBS)
    with native.Batch(tuning=tuning) as b:
        ids = [b.submit(r, o) for r, o in jobs]
        b.run()
        for (r, o), j in zip(jobs, ids):
            assert b.status[j] == 0
            assert b.output(j) == oracle_out(r, o)[0], "job %d" % j


def test_tuning_is_checked():
    with native.Batch() as b:
        for bad in (dict(bucket_cap=768), dict(bucket_cap=512, classes=[dict(read_len_max=100, n_max=10,
                                                                             fine_load=16, coarse_load=64)]),
                    dict(bucket_cap=512, classes=[dict(read_len_max=100, n_max=10, fine_load=600,
                                                       coarse_load=1000)])):
            with pytest.raises(native.EmpError) as e:
                b.set_tuning(bad)
            assert e.value.code == -1
        r, o = make_jobs(JOBS[:1])[0]
        b.submit(r, o)
        with pytest.raises(native.EmpError):  # not while a job is queued
            b.set_tuning(None)
        b.run()
        b.set_tuning(None)
