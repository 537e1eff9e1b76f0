"""GPU parity when the lists exceed M = reduce_mem / 16 (2n > M): cq_sort_by_key's chunked sort and
FIFO merges (cqutils.cuh:179-247) decide the order of equal keys, hence which id of a key the
dedup keeps (cq_unique_by_key, cqutils.cuh:271-306) and the order of every record's edges. The
library's out-of-core path (tie keys from the host planner, dedup and joins in (key, tie) order on
the device) against the oracle, byte for byte, with per-round counts and the record framing."""
import random

import numpy as np
import pytest

import bruteforce as bf
import oracle_lib as ol
from empanada_amd import gen, native

pytestmark = pytest.mark.gpu

EO_PARTITION, EMP_PARTITION = -6, -9  # the oracle's and the library's partition error


def _reads_of(strings):
    return np.frombuffer("".join(strings).encode(), dtype=np.uint8).reshape(len(strings), -1)


def _dense_reads(seed, n, G, L):
    """n reads of a G-base random genome, G close to n, on random strands: duplicate reads and
    equal keys in every round, spread over many sort chunks."""
    rng = random.Random(seed)
    g = "".join(rng.choice("ACGT") for _ in range(G))
    out = []
    for _ in range(n):
        p = rng.randrange(0, G - L + 1)
        out.append(g[p:p + L] if rng.random() < 0.5 else bf.revcomp(g[p:p + L]))
    return out


def _min_id_dedup(strings):
    keys = {}
    for k, s in enumerate(strings):
        for o, key in ((0, ol.key_of(s)), (1, ol.key_of_rc(s))):
            keys.setdefault(key, []).append(2 * k + o)
    return sorted(min(v) for v in keys.values())


def _compare(r, o, M):
    """Runs the oracle and the library on reads r (n x L) at buffer M; returns the oracle's rc."""
    n, L = r.shape
    rc, want, osum, orounds = ol.run_reads(r.tobytes(), n, L, min_ovlp=o, reduce_mem=16 * M)
    print("n=%d L=%d o=%d M=%d oracle rc=%d" % (n, L, o, M, rc), flush=True)
    with native.Context(min_ovlp=o, reduce_mem=16 * M) as c:
        c.add_reads(r)
        if rc == EO_PARTITION:
            with pytest.raises(native.EmpError) as e:
                c.run()
            assert e.value.code == EMP_PARTITION
            return rc
        assert rc == 0
        got = c.run()
        summ, rounds = c.summary(), c.rounds()
    assert summ["n_dedup"] == osum["n_dedup"]
    assert [(x["s"], x["p"], x["edges"], x["s_after"], x["p_after"], x["records"]) for x in rounds] == \
        [(x["s"], x["p"], x["edges"], x["s_after"], x["p_after"], x["records"]) for x in orounds]
    assert got == want
    return rc


@pytest.mark.parametrize("seed,n,G,M", [(0, 200, 120, 40), (3, 200, 120, 40), (1, 300, 200, 64), (2, 300, 200, 64),
                                        (0, 600, 400, 96), (3, 600, 400, 96)])
def test_dense_reads_dedup_keeps_the_merge_orders_first(seed, n, G, M):
    """Dense reads at M well below 2n (5 to 12 sort chunks): the dedup keeps ids other than the
    smallest of each key (checked on the oracle's dedup first), and the rounds sort lists of
    several chunks with many equal keys."""
    reads = _dense_reads(seed, n, G, 24)
    rc, kept = ol.dedup(reads, M)
    assert rc == 0 and sorted(kept) != _min_id_dedup(reads)
    assert _compare(_reads_of(reads), 8, M) == 0


@pytest.mark.parametrize("G,M,seed", [(5000, 6, 0), (20000, 8, 0), (2000, 12, 0), (2000, 4, 1)])
def test_tiny_buffers(G, M, seed):
    """M of 4 to 12 over 300 reads: 50 to 150 sort chunks, windows of 2 to 6 entries, thousands of
    records; the last case stops with a partition error in both."""
    reads = _dense_reads(seed, 300, G, 24)
    assert _compare(_reads_of(reads), 12, M) == (EO_PARTITION if M == 4 else 0)


@pytest.mark.parametrize("M", [16384, 29999, 45000, 59999])
def test_generated_reads_vs_oracle(M):
    """30,000 x 100 bp reads (both strands, 0.5 % substitutions) at M from 2n / 4 to 2n - 1: two to
    four sort chunks in the dedup and in the early rounds, the later rounds in core."""
    g = gen.genome(2 * 10**5, 7)
    r = gen.reads(g, 30000, 100, 0.5, 0.005, 8)
    assert _compare(r, 31, M) == 0


def test_repeat_workload_sample_out_of_core():
    """A sample of the 3.1 Gbp workload (300 bp and 6 kbp repeat families, 1 % divergence, 150 bp
    reads, -o 75) at M = 2n / 3: repeat copies give equal keys across chunks."""
    r, o = gen.workload_reads("ooc", 0.0005)
    assert _compare(r, o, (2 * r.shape[0]) // 3) == 0


def test_search_partition_error_out_of_core():
    """A key group of more than M/2 suffixes and prefixes in a list over M: both stop."""
    rng = random.Random(21)
    b = "".join(rng.choice("ACGT") for _ in range(45))
    reads = [bf.revcomp(b) + "".join(rng.choice("ACGT") for _ in range(10)) + b for _ in range(200)]
    assert _compare(_reads_of(reads), 31, 300) == EO_PARTITION


def test_more_than_65535_chunks_is_an_error():
    """M = 2 with 70,000 reads: 70,000 chunks, more than the 16-bit chunk ids hold."""
    g = gen.genome(10**5, 3)
    r = gen.reads(g, 70000, 50, 0.5, 0.0, 4)
    with native.Context(min_ovlp=30, reduce_mem=32) as c:
        c.add_reads(r)
        with pytest.raises(native.EmpError) as e:
            c.run()
    assert e.value.code == -6
