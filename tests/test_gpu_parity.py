"""GPU parity: the HIP path through the C ABI against the oracle, bit for bit.

Stage level: fingerprints (fwd + reverse complement), 2-bit columns, NextSfx / PrevPfx including
the u64 wrap. Pipeline level: output.bin of whole runs (FASTQ and kept-read entry points) on seeded
inputs, edge cases the reference's code paths define, and error paths."""
import random

import numpy as np
import pytest

import bruteforce as bf
import oracle_lib as ol
from empanada_amd import gen, native

pytestmark = pytest.mark.gpu

FQ = lambda seqs: b"".join(b"@r%d\n%s\n+\n%s\n" % (i, s, b"I" * len(s)) for i, s in enumerate(seqs))


def reads_of(strings):
    return np.frombuffer("".join(strings).encode(), dtype=np.uint8).reshape(len(strings), -1)


def run_gpu_fastq(data, **kw):
    with native.Context(**kw) as c:
        c.add_fastq(data)
        return c.run(), c.summary(), c.rounds()


# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("L", [1, 2, 15, 16, 31, 100, 150, 250, 255])
def test_stage_fingerprints(L):
    rng = random.Random(L)
    n = 777
    strs = ["".join(rng.choice("ACGTacgt") for _ in range(L)) for _ in range(n)]
    keys = native.stage_fingerprints(reads_of(strs))
    for k in range(n):
        assert int(keys[2 * k]) == ol.key_of(strs[k])
        assert int(keys[2 * k + 1]) == ol.key_of_rc(strs[k])


def test_stage_fingerprints_rejects_non_acgt():
    with pytest.raises(native.EmpError):
        native.stage_fingerprints(reads_of(["ACGN"]))


@pytest.mark.parametrize("n,L", [(1, 7), (16, 100), (1000, 33), (70001, 20), (1048577, 9)])
def test_stage_columns(n, L):
    """The map kernel's column store, the one the rounds gather from: base j of read k at bits
    2(8(k%2) + j%8) of word (k / 2^20, j / 8, (k % 2^20) / 2); columns past L zero."""
    rng = np.random.default_rng(n)
    r = rng.choice(np.frombuffer(b"ACGTacgt", dtype=np.uint8), size=(n, L))
    store = native.stage_columns(r)
    codes = ((r >> 1) & 3).astype(np.uint32)  # A0 C1 T2 G3 (blocks.cuh:33-40), lowercase alike
    G8 = (L + 7) // 8
    pad = np.zeros((store.shape[0] << 20, G8 * 8), dtype=np.uint32)
    pad[:n, :L] = codes
    want = np.zeros(store.shape, dtype=np.uint32)
    for half in (0, 1):
        rows = pad[half::2].reshape(store.shape[0], 1 << 19, G8, 8)  # (chunk, word, group, column)
        for c in range(8):
            want |= rows[:, :, :, c].transpose(0, 2, 1) << np.uint32(2 * (8 * half + c))
    valid = np.zeros((store.shape[0] << 20,), dtype=bool)
    valid[:n] = True
    vmask = np.where(valid[0::2], 0xFFFF, 0).astype(np.uint32) | np.where(valid[1::2], 0xFFFF0000, 0).astype(np.uint32)
    vmask = vmask.reshape(store.shape[0], 1, 1 << 19)
    assert np.array_equal(store & vmask, want & vmask)


def test_stage_roll_matches_oracle_and_wrap():
    rng = random.Random(3)
    keys, bases = [], []
    for _ in range(3000):
        keys.append((rng.randrange(bf.P1) << 32) | rng.randrange(bf.P2))
        bases.append(rng.choice("ACGTacgt"))
    # the u64-wrap branch of PrevPfx: a half below ENCODING(base)
    for h in range(5):
        for b in "ACGT":
            keys.append((h << 32) | h)
            bases.append(b)
            keys.append((h << 32) | (bf.P2 - 1))
            bases.append(b)
    k = np.array(keys, dtype=np.uint64)
    bb = np.frombuffer("".join(bases).encode(), dtype=np.uint8)
    for length in (2, 64, 100, 255):
        out = native.stage_roll(k, bb, 0, length)
        assert all(int(out[i]) == ol.next_sfx(keys[i], bases[i], length) for i in range(len(keys)))
    out = native.stage_roll(k, bb, 1, 0)
    assert all(int(out[i]) == ol.prev_pfx(keys[i], bases[i]) for i in range(len(keys)))


# ---------------------------------------------------------------------------------------------
def _random_reads(rng, n, L, glen, p_rc, p_sub):
    g = "".join(rng.choice("ACGT") for _ in range(glen))
    out = []
    for _ in range(n):
        st = rng.randrange(glen - L + 1)
        s = g[st:st + L]
        if rng.random() < p_rc:
            s = bf.revcomp(s)
        s = "".join((rng.choice([b for b in "ACGT" if b != c]) if rng.random() < p_sub else c) for c in s)
        out.append(s)
    return out


@pytest.mark.parametrize("seed,n,L,o,glen,p_rc,p_sub", [
    (1, 40, 20, 8, 300, 0.0, 0.0),
    (2, 80, 24, 10, 400, 0.5, 0.0),
    (3, 120, 30, 12, 500, 0.5, 0.01),
    (4, 150, 32, 15, 250, 0.5, 0.005),
    (5, 60, 16, 16, 200, 0.5, 0.0),
    (6, 200, 28, 6, 150, 0.5, 0.02),
    (7, 500, 40, 10, 120, 0.5, 0.0),     # extreme coverage: large equal-key groups
    (8, 300, 7, 3, 60, 0.5, 0.0),        # one 8-column group per read: first and last the same
    (9, 300, 9, 4, 80, 0.5, 0.01),       # a full group and a one-column group
])
def test_pipeline_small_vs_oracle_and_bruteforce(seed, n, L, o, glen, p_rc, p_sub):
    rng = random.Random(seed)
    reads = _random_reads(rng, n, L, glen, p_rc, p_sub)
    data = FQ([r.encode() for r in reads])
    rc, want, _, orounds = ol.run_fastq(data, min_ovlp=o)
    assert rc == 0
    got, summ, rounds = run_gpu_fastq(data, min_ovlp=o)
    assert got == want
    if bf.wrap_free(reads, o):
        assert got == bf.encode(bf.overlap(reads, o)[0])
    for a, b in zip(rounds, orounds):
        assert (a["s"], a["p"], a["edges"], a["s_after"], a["p_after"]) == \
               (b["s"], b["p"], b["edges"], b["s_after"], b["p_after"])


@pytest.mark.parametrize("G,n,L,o,p_rc,p_sub,seeds", [
    (10**6, 20000, 100, 31, 0.0, 0.0, (1000, 1001)),      # the small workload
    (2 * 10**5, 30000, 100, 31, 0.5, 0.005, (7, 8)),     # 15x, both strands, errors
    (3 * 10**5, 20000, 150, 75, 0.5, 0.005, (9, 10)),
    (5 * 10**4, 8000, 250, 31, 0.5, 0.0, (11, 12)),      # 40x, long reads
    (10**5, 50000, 64, 20, 0.5, 0.01, (13, 14)),
])
def test_pipeline_generated_vs_oracle(G, n, L, o, p_rc, p_sub, seeds):
    g = gen.genome(G, seeds[0])
    r = gen.reads(g, n, L, p_rc, p_sub, seeds[1])
    rc, want, osum, _ = ol.run_reads(r.tobytes(), n, L, min_ovlp=o)
    assert rc == 0
    with native.Context(min_ovlp=o) as c:
        c.add_reads(r)
        got = c.run()
        s = c.summary()
    assert got == want
    assert s["n_dedup"] == osum["n_dedup"] and s["n_edges"] == osum["n_edges"]


def test_prefix_wrap_takes_the_fallback_slot():
    """Reads whose prefix fingerprints reach PrevPfx's u64 wrap at rounds of every phase of the
    column-group cycle: the wrapped key is outside the prefix's computed target buckets, so the
    entry takes the fallback slot. Output and per-round list sizes equal the oracle's."""
    rng = random.Random(77)
    g = gen.genome(2 * 10**5, 31)
    base = gen.reads(g, 20000, 100, 0.5, 0.005, 32)
    crafted = [bf.wrap_read(100, ln, rng) for ln in (70, 75, 80, 81, 82, 83, 99)]
    assert not bf.wrap_free(crafted, 63)
    r = np.concatenate([base, np.frombuffer("".join(crafted).encode(), np.uint8).reshape(-1, 100)])
    rc, want, osum, orounds = ol.run_reads(r.tobytes(), r.shape[0], 100, min_ovlp=63)
    assert rc == 0
    with native.Context(min_ovlp=63) as c:
        c.add_reads(r)
        got, rounds = c.run(), c.rounds()
    assert got == want
    assert [(x["s"], x["p"], x["edges"], x["s_after"], x["p_after"]) for x in rounds] == \
           [(x["s"], x["p"], x["edges"], x["s_after"], x["p_after"]) for x in orounds]


def test_match_overflow_bucket_with_a_wrapped_prefix_survivor():
    """A bucket with more matched records than the LDS join holds (400 reads ending in C and 400
    starting with C, key(C) in bucket 0) that also holds surviving prefixes whose next step takes
    PrevPfx's u64 wrap (the fallback slot of the next round's buckets). The whole bucket goes to
    the large-bucket path, which writes its survivors once: per-round list sizes and output equal
    the oracle's."""
    rng = random.Random(91)
    L, ln = 40, 30
    C = bf.wrap_read(ln, ln, rng)  # key(C) has H1 < 4: bucket 0 at any bucket count
    reads = ["".join(rng.choice("ACGT") for _ in range(L - ln)) + C for _ in range(400)]
    reads += [C + "".join(rng.choice("ACGT") for _ in range(L - ln)) for _ in range(400)]
    reads += [bf.wrap_read(L, ln, rng) for _ in range(5)]  # length-30 prefixes with H1 < 4, unmatched
    g = gen.genome(2 * 10**5, 93)
    r = np.concatenate([reads_of(reads), gen.reads(g, 5000, L, 0.5, 0.0, 94)])
    rc, want, osum, orounds = ol.run_reads(r.tobytes(), r.shape[0], L, min_ovlp=25)
    assert rc == 0
    with native.Context(min_ovlp=25) as c:
        c.add_reads(r)
        got, rounds = c.run(), c.rounds()
    assert rounds[L - 1 - ln]["big_buckets"] > 0
    assert [(x["s"], x["p"], x["edges"], x["s_after"], x["p_after"]) for x in rounds] == \
           [(x["s"], x["p"], x["edges"], x["s_after"], x["p_after"]) for x in orounds]
    assert got == want


def test_pool_overflow_in_pipelined_rounds(monkeypatch, tmp_path):
    """Edges of one round beyond the whole first edge pool (sub-ranges and overflow region): 6,000
    key groups of 30 x 30 reads sharing a 30-mer (10.8 M edges with their reverse complements at
    overlap 30) among 2 M reads, on the back-to-back rounds. Buckets whose edges find no room
    count none, nothing is copied for them, and the run restarts on the synchronous rounds with a
    larger pool: the output equals the oracle's."""
    monkeypatch.delenv("EMPANADA_SYNC_ROUNDS", raising=False)
    rng = np.random.default_rng(17)
    L, ln, k, groups = 40, 30, 30, 6000
    A = np.frombuffer(b"ACGT", np.uint8)
    common = np.repeat(rng.choice(A, size=(groups, 1, ln)), k, axis=1)
    sfx = np.concatenate([rng.choice(A, size=(groups, k, L - ln)), common], axis=2).reshape(-1, L)
    pfx = np.concatenate([common, rng.choice(A, size=(groups, k, L - ln))], axis=2).reshape(-1, L)
    g = gen.genome(10**8, 61)
    r = np.ascontiguousarray(np.concatenate([gen.reads(g, 2_000_000, L, 0.5, 0.0, 62), sfx, pfx]))
    n = r.shape[0]
    rc, want, osum, orounds = ol.run_reads(r.tobytes(), n, L, min_ovlp=29)
    assert rc == 0
    # the first pool: 256 sub-ranges of max(4096, 2n/256) edges and an overflow region of
    # max(2^20, 2n/8)
    first_pool = 256 * max(4096, 2 * n // 256) + max(1 << 20, 2 * n // 8)
    assert orounds[L - 1 - ln]["edges"] > first_pool
    with native.Context(min_ovlp=29) as c:
        c.add_reads(r)
        got = c.run()
        assert c.summary()["n_edges"] == osum["n_edges"]
    assert got == want
    _check_file_output(r, 29, want, tmp_path)


def _check_file_output(r, o, want, tmp_path):
    """emp_run_to_file writes the records of the back-to-back rounds while later rounds run; when
    the run falls back to the synchronous rounds those records are not written twice."""
    out = tmp_path / "output.bin"
    with native.Context(min_ovlp=o) as c:
        c.add_reads(r)
        c.run_to_file(str(out))
    assert out.read_bytes() == want


@pytest.mark.parametrize("case", ["genome", "groups600", "groups3000"])
def test_file_output_with_and_without_fallback(case, tmp_path):
    if case == "genome":  # no fallback: every record written during the rounds
        g = gen.genome(5 * 10**5, 71)
        r, o = gen.reads(g, 100_000, 100, 0.5, 0.005, 72), 31
    else:  # a large key group in round 10: the first rounds are written, then the run falls back
        r, o = reads_of(_shared_overlap_reads(int(case[6:]), 5)), 29
    rc, want, _, _ = ol.run_reads(r.tobytes(), r.shape[0], r.shape[1], min_ovlp=o)
    assert rc == 0
    _check_file_output(r, o, want, tmp_path)


def test_file_output_over_a_longer_file_spanning_reservations(tmp_path):
    """An output.bin of more than 64 MB (the file's blocks are reserved 64 MB at a time on ext4 and
    xfs, the last reservation cut back to the bytes written), written where a longer file was: the
    bytes of emp_run's records and nothing after them."""
    g = gen.genome(27 * 10**6, 91)
    r = gen.reads(g, 8_000_000, 100, 0.5, 0.005, 92)
    out = tmp_path / "output.bin"
    with native.Context(min_ovlp=63) as c:
        c.add_reads(r)
        want = c.run()
        assert len(want) > (64 << 20)
        out.write_bytes(b"\xff" * (len(want) + 12345))
        c.reset()
        c.add_reads(r)
        c.run_to_file(str(out))
    assert out.stat().st_size == len(want)
    assert out.read_bytes() == want


def test_pipelined_and_synchronous_rounds_agree(monkeypatch):
    """2,000,000 x 100 bp reads (4 M entries): the back-to-back rounds (each join clears the counts
    it read; bucket sets in multiples of 256 buckets) and the rounds with the host between them
    (powers of two) give the oracle's bytes and the same per-round counts."""
    g = gen.genome(5 * 10**6, 21)
    r = gen.reads(g, 2_000_000, 100, 0.5, 0.005, 22)
    rc, want, osum, orounds = ol.run_reads(r.tobytes(), r.shape[0], 100, min_ovlp=40)
    assert rc == 0
    got = {}
    for mode in ("pipelined", "sync"):
        if mode == "sync":
            monkeypatch.setenv("EMPANADA_SYNC_ROUNDS", "1")
        else:
            monkeypatch.delenv("EMPANADA_SYNC_ROUNDS", raising=False)
        with native.Context(min_ovlp=40) as c:
            c.add_reads(r)
            got[mode] = (c.run(), c.rounds())
    for mode, (out, rounds) in got.items():
        assert out == want, mode
        assert [(x["s"], x["p"], x["edges"], x["s_after"], x["p_after"]) for x in rounds] == \
               [(x["s"], x["p"], x["edges"], x["s_after"], x["p_after"]) for x in orounds], mode
    # bucket counts: powers of two on the synchronous rounds; on the pipelined ones multiples of 256
    # holding at most 1,792 entries per bucket on average, never growing from one round to the next,
    # and round 0's sized from the 2n entries before the dedup (4,000,000 / 1,792 -> 2,304 buckets)
    sync_nb = [x["buckets"] for x in got["sync"][1]]
    pipe_nb = [x["buckets"] for x in got["pipelined"][1]]
    assert all(nb & (nb - 1) == 0 for nb in sync_nb)
    assert pipe_nb[0] == 2304 and any(nb & (nb - 1) for nb in pipe_nb)
    assert all(nb % 256 == 0 for nb in pipe_nb)
    assert all(b <= a for a, b in zip(pipe_nb, pipe_nb[1:]))
    assert all(max(x["s"], x["p"]) <= 1792 * x["buckets"] for x in got["pipelined"][1])


def test_fastq_and_read_entry_points_agree_and_block_size_is_irrelevant():
    g = gen.genome(3 * 10**5, 21)
    r = gen.reads(g, 30000, 100, 0.5, 0.005, 22)
    fq = gen.fastq_bytes(r)
    outs = set()
    for block in (1 << 26, 1 << 16, 12345):
        with native.Context(min_ovlp=40) as c:
            c.add_fastq(fq, block=block)
            outs.add(c.run())
    with native.Context(min_ovlp=40) as c:
        c.add_reads(r)
        outs.add(c.run())
    assert len(outs) == 1
    rc, want, _, _ = ol.run_fastq(fq, min_ovlp=40)
    assert outs == {want}


def test_many_batches_and_device_entry_point():
    """More reads than one staging batch (2^20): batches continue the read ids and columns."""
    g = gen.genome(4 * 10**6, 31)
    r = gen.reads(g, (1 << 20) + 4321, 40, 0.5, 0.002, 32)
    rc, want, _, _ = ol.run_reads(r.tobytes(), r.shape[0], 40, min_ovlp=25)
    assert rc == 0
    with native.Context(min_ovlp=25) as c:
        c.add_reads(r)
        assert c.run() == want
    from empanada_amd import hipmem
    d = hipmem.DeviceBuffer.from_array(r)
    with native.Context(min_ovlp=25) as c:
        c.add_reads_device(d.address, r.shape[0], 40)
        assert c.run() == want
    d.free()
    # pinned host memory: copied straight into HBM batch by batch, in two calls whose first one
    # ends off a multiple of 16 (its last reads are staged), then pageable memory again
    h = hipmem.HostBuffer(r.nbytes)
    hr = h.array(np.uint8, r.shape)
    hr[:] = r
    cut = 700_003
    with native.Context(min_ovlp=25) as c:
        c.add_reads(hr[:cut])
        c.add_reads(hr[cut:cut + 200_000])
        c.add_reads(r[cut + 200_000:])
        assert c.run() == want
    h.free()


def test_reset_and_rerun_is_deterministic():
    g = gen.genome(2 * 10**5, 41)
    r = gen.reads(g, 20000, 100, 0.5, 0.005, 42)
    with native.Context(min_ovlp=31) as c:
        c.add_reads(r)
        a = c.run()
        c.reset()
        c.add_reads(r)
        b = c.run()
        c.reset()
        r2 = gen.reads(g, 5000, 60, 0.5, 0.0, 43)
        c.add_reads(r2)
        d = c.run()
    assert a == b
    assert d == ol.run_reads(r2.tobytes(), 5000, 60, min_ovlp=31)[1]


def test_edge_cases():
    assert run_gpu_fastq(b"")[0] == b""
    assert run_gpu_fastq(FQ([b"NNNN", b"ACGN"]))[0] == b""
    single = FQ([b"ACGTTGCAACGTTGCA"])
    assert run_gpu_fastq(single, min_ovlp=4)[0] == ol.run_fastq(single, min_ovlp=4)[1]
    pal = FQ([b"ACGTACGT", b"ACGTACGT", b"TTTTTTTT", b"AAAAAAAA"])  # palindrome, duplicate, RC pair
    assert run_gpu_fastq(pal, min_ovlp=2)[0] == ol.run_fastq(pal, min_ovlp=2)[1]
    lower = FQ([b"acgtacgtgg", b"CGTACGTGGA", b"gtacgtggat"])
    assert run_gpu_fastq(lower, min_ovlp=3)[0] == ol.run_fastq(lower, min_ovlp=3)[1]


def test_error_paths():
    with pytest.raises(native.EmpError) as e:
        run_gpu_fastq(FQ([b"ACGT", b"ACG"]))
    assert e.value.code == -4
    with pytest.raises(native.EmpError) as e:
        run_gpu_fastq(FQ([b"ACGT", b"ACRT"]))
    assert e.value.code == -5
    with pytest.raises(native.EmpError) as e:
        run_gpu_fastq(b"xx\nACGT\n+\nIIII\n")
    assert e.value.code == -3
    with pytest.raises(native.EmpError) as e:  # reduce_mem / 16 below 2
        run_gpu_fastq(FQ([b"ACGT"]), reduce_mem=16)
    assert e.value.code == -1
    # 2n = 300 > M = 200: the out-of-core merge order, as the oracle restates it
    rng = random.Random(5)
    reads = _random_reads(rng, 150, 30, 600, 0.5, 0.0)
    data = FQ([x.encode() for x in reads])
    rc, want, _, _ = ol.run_fastq(data, min_ovlp=12, reduce_mem=16 * 200)
    assert rc == 0
    assert run_gpu_fastq(data, min_ovlp=12, reduce_mem=16 * 200)[0] == want


@pytest.mark.parametrize("m", [300, 320, 450, 540])
def test_search_partitions_records_vs_oracle(m):
    """Lists reaching M/2 = m/2 split each round into key-aligned records (cqutils.cuh:107-137)."""
    rng = random.Random(5)
    reads = _random_reads(rng, 150, 30, 600, 0.5, 0.0)
    data = FQ([r.encode() for r in reads])
    rc, want, osum, orounds = ol.run_fastq(data, min_ovlp=12, reduce_mem=16 * m)
    assert rc == 0 and osum["n_records"] > len(orounds)
    got, summ, rounds = run_gpu_fastq(data, min_ovlp=12, reduce_mem=16 * m)
    assert got == want
    assert summ["n_records"] == osum["n_records"]
    assert [r["records"] for r in rounds] == [r["records"] for r in orounds]


@pytest.mark.parametrize("m", [60000, 61000, 90000, 119999])
def test_search_partitions_generated_vs_oracle(m):
    """30,000 x 100 bp reads (both strands, 0.5 % substitutions) at M from 2n to 4n: two or three
    records per round while a list holds M/2 entries, cuts inside buckets."""
    g = gen.genome(2 * 10**5, 7)
    r = gen.reads(g, 30000, 100, 0.5, 0.005, 8)
    rc, want, osum, orounds = ol.run_reads(r.tobytes(), 30000, 100, min_ovlp=31, reduce_mem=16 * m)
    assert rc == 0
    with native.Context(min_ovlp=31, reduce_mem=16 * m) as c:
        c.add_reads(r)
        got = c.run()
        summ, rounds = c.summary(), c.rounds()
    assert got == want
    assert summ["n_records"] == osum["n_records"]
    assert [x["records"] for x in rounds] == [x["records"] for x in orounds]


def test_search_partitions_with_large_buckets():
    """A 3000 x 3000 key group (a bucket over capacity, spilled) in rounds framed at M = 2n."""
    reads = _shared_overlap_reads(3000, 3000)
    r = reads_of(reads)
    m = 2 * r.shape[0]
    rc, want, osum, _ = ol.run_reads(r.tobytes(), r.shape[0], 40, min_ovlp=29, reduce_mem=16 * m)
    assert rc == 0 and osum["n_records"] > 11
    with native.Context(min_ovlp=29, reduce_mem=16 * m) as c:
        c.add_reads(r)
        got = c.run()
        rounds = c.rounds()
    assert any(x["big_buckets"] > 0 for x in rounds)
    assert got == want


def test_search_partition_that_cannot_form_is_an_error():
    """Reads rc(B) + 10 random bases + B: at overlap 45 every suffix is B and every prefix rc(B), so
    a window of M/2 equal keys leaves no cut (the reference loops forever); the oracle and the
    library both stop with a partition error."""
    rng = random.Random(21)
    b = "".join(rng.choice("ACGT") for _ in range(45))
    reads = [bf.revcomp(b) + "".join(rng.choice("ACGT") for _ in range(10)) + b for _ in range(200)]
    data = FQ([x.encode() for x in reads])
    assert ol.run_fastq(data, min_ovlp=31, reduce_mem=16 * 400)[0] == -6
    with pytest.raises(native.EmpError) as e:
        run_gpu_fastq(data, min_ovlp=31, reduce_mem=16 * 400)
    assert e.value.code == -9
    assert run_gpu_fastq(data, min_ovlp=31)[0] == ol.run_fastq(data, min_ovlp=31)[1]


def _shared_overlap_reads(k, seed):
    """k reads ending in one common 30-mer and k reads starting with it: a single key group of k
    suffixes and k prefixes at overlap 30 (k*k edges)."""
    rng = random.Random(seed)
    common = "".join(rng.choice("ACGT") for _ in range(30))
    a = ["".join(rng.choice("ACGT") for _ in range(10)) + common for _ in range(k)]
    b = [common + "".join(rng.choice("ACGT") for _ in range(10)) for _ in range(k)]
    return a + b


@pytest.mark.parametrize("k", [600, 3000])
def test_large_key_groups_take_the_global_path(k):
    """600 x 600 overflows the LDS match records; 3000 x 3000 overflows a bucket (spill)."""
    reads = _shared_overlap_reads(k, k)
    r = reads_of(reads)
    rc, want, osum, _ = ol.run_reads(r.tobytes(), r.shape[0], 40, min_ovlp=29)
    assert rc == 0 and osum["n_edges"] >= k * k
    with native.Context(min_ovlp=29) as c:
        c.add_reads(r)
        got = c.run()
        rounds = c.rounds()
    assert any(x["big_buckets"] > 0 for x in rounds)
    assert got == want


def test_identical_reads_dedup_big_bucket():
    rng = random.Random(77)
    s = "".join(rng.choice("ACGT") for _ in range(60))
    reads = [s] * 5000 + ["".join(rng.choice("ACGT") for _ in range(60)) for _ in range(300)]
    r = reads_of(reads)
    rc, want, osum, _ = ol.run_reads(r.tobytes(), r.shape[0], 60, min_ovlp=20)
    assert rc == 0
    with native.Context(min_ovlp=20) as c:
        c.add_reads(r)
        got = c.run()
        assert c.summary()["n_dedup"] == osum["n_dedup"]
    assert got == want


def test_edge_pool_growth_keeps_earlier_rounds():
    """Reads of a genome give edges in every round; a 1500 x 1500 group sharing a 25-base overlap
    adds 2.25M edges in a later round, more than the pool's first overflow region holds, so the
    pool grows after earlier rounds have written their edges."""
    rng = random.Random(5)
    genome = "".join(rng.choice("ACGT") for _ in range(4000))
    reads = [genome[p:p + 40] for p in (rng.randrange(0, 4000 - 40) for _ in range(3000))]
    common = "".join(rng.choice("ACGT") for _ in range(25))
    reads += ["".join(rng.choice("ACGT") for _ in range(15)) + common for _ in range(1500)]
    reads += [common + "".join(rng.choice("ACGT") for _ in range(15)) for _ in range(1500)]
    r = reads_of(reads)
    rc, want, osum, _ = ol.run_reads(r.tobytes(), r.shape[0], 40, min_ovlp=20)
    assert rc == 0 and osum["n_edges"] > (1 << 21)
    with native.Context(min_ovlp=20) as c:
        c.add_reads(r)
        got = c.run()
        rounds = c.rounds()
    big = max(range(len(rounds)), key=lambda i: rounds[i]["edges"])
    assert rounds[big]["edges"] >= 1500 * 1500 and sum(x["edges"] for x in rounds[:big]) > 0
    assert got == want


def test_repeat_genome_sample_vs_oracle():
    """A 0.05 % sample of the 3.1 Gbp workload: 1.55 Mbp genome with 300 bp and 6 kbp repeat
    families (16 copies, 1 % divergence), 150 bp reads at 20x, -o 75: repeat copies give equal keys
    across distant reads."""
    r, o = gen.workload_reads("ooc", 0.0005)
    rc, want, osum, _ = ol.run_reads(r.tobytes(), r.shape[0], r.shape[1], min_ovlp=o)
    assert rc == 0
    with native.Context(min_ovlp=o) as c:
        c.add_reads(r)
        got = c.run()
        assert c.summary()["n_edges"] == osum["n_edges"]
    assert got == want


@pytest.mark.parametrize("L", [2, 37, 100, 150, 255])
@pytest.mark.parametrize("first", ["clean", "dropped", "longer"])
def test_fastq_clean_records_copied_by_the_classify_pass(L, first, tmp_path):
    """FASTQ whose records are all kept reads of one length: the classify pass writes the rows
    itself (rows of L bytes, unaligned to 4 bytes for most L), over one range or three (the later
    ones with the context's length set). With record 0 dropped ('N') or longer than the rest, the
    length taken from record 0 is wrong or record 0 is no read: the rank pass and the copy run
    instead. Against the oracle, through the device, host-block and file entry points."""
    rng = random.Random(1000 + L)
    G = 3000 + L
    g = "".join(rng.choice("ACGT") for _ in range(G))
    seqs = []
    for _ in range(2500):
        p = rng.randrange(0, G - L + 1)
        s = g[p:p + L]
        seqs.append(s if rng.random() < 0.5 else bf.revcomp(s))
    if first == "dropped":
        seqs[0] = "N" + seqs[0][1:]
    elif first == "longer":
        seqs.insert(0, seqs[0] + "A")
    data = FQ([s.encode() for s in seqs])
    o = max(1, (2 * L) // 3)
    rc, want, osum, _ = ol.run_fastq(data, min_ovlp=o)
    if first == "longer":
        assert rc == -3  # READLEN: the other lengths come after record 0's
    else:
        assert rc == 0
    from empanada_amd import hipmem
    dbuf = hipmem.DeviceBuffer.from_array(np.frombuffer(data, np.uint8))
    path = tmp_path / "in.fastq"
    path.write_bytes(data)
    try:
        for how in ("device", "device3", "host", "file"):
            try:
                with native.Context(min_ovlp=o, map_mem=16 * 8192) as c:
                    if how == "host":
                        c.add_fastq(data, block=8192)
                    elif how == "file":
                        c.add_fastq_file(str(path))
                    else:
                        c.add_fastq_device(dbuf.address, len(data), pieces=3 if how == "device3" else 1)
                    got, code = c.run(), 0
                    if code == 0:
                        assert c.summary()["n_reads"] == osum["n_reads"], how
            except native.EmpError as e:
                got, code = None, e.code
            assert code == EO2EMP[rc], (how, code)
            if code == 0:
                assert got == want, how
    finally:
        dbuf.free()


EO2EMP = {0: 0, -2: -3, -3: -4, -4: -5}  # oracle status -> C ABI status


@pytest.mark.parametrize("seed", range(8))
def test_fastq_parser_rules_and_error_order(seed, tmp_path):
    """The device FASTQ parser against the oracle's reader: dropped reads (a byte in [64,127]
    with bit 3 set), empty sequence lines, lowercase, blocks cut mid-record, and which error a
    block reports when it holds a bad base and a length mismatch (the first in file order)."""
    rng = random.Random(seed)
    L = rng.choice([20, 33, 64])
    seqs = []
    for _ in range(rng.randrange(60, 300)):
        u = rng.random()
        if u < 0.08:
            seqs.append("")
        elif u < 0.25:
            s = [rng.choice("ACGT") for _ in range(L)]
            s[rng.randrange(L)] = rng.choice("NnKkMmHX")
            seqs.append("".join(s))
        else:
            seqs.append("".join(rng.choice("ACGTacgt") for _ in range(L)))
    if seed % 4 == 1:
        seqs.insert(rng.randrange(len(seqs)), "ACGR" + "A" * (L - 4))   # bad base
    if seed % 4 == 2:
        seqs.insert(rng.randrange(len(seqs)), "A" * (L + 1))            # other length
    if seed % 4 == 3:
        k = rng.randrange(len(seqs) // 2)
        seqs.insert(k, "ACGR" + "A" * (L - 4))
        seqs.insert(k + rng.randrange(1, 5), "A" * (L - 1))
    data = FQ([s.encode() for s in seqs])
    from empanada_amd import hipmem
    dbuf = hipmem.DeviceBuffer.from_array(np.frombuffer(data, np.uint8))
    path = tmp_path / "in.fastq"
    path.write_bytes(data)
    for block in (1 << 20, 997, 4103):
        rc_o, want, _, _ = ol.run_fastq(data, min_ovlp=L // 2, map_mem=16 * block)
        # host blocks one by one (emp_add_fastq), the whole buffer in HBM in one call or three
        # (emp_add_fastq_device: the block chain on the device), the file (emp_add_fastq_file)
        for how in ("host", "device", "device3", "file"):
            try:
                with native.Context(min_ovlp=L // 2, map_mem=16 * block) as c:
                    if how == "host":
                        c.add_fastq(data, block=block)
                    elif how == "file":
                        c.add_fastq_file(str(path))
                    else:
                        c.add_fastq_device(dbuf.address, len(data), pieces=3 if how == "device3" else 1)
                    got, rc_g = c.run(), 0
            except native.EmpError as e:
                got, rc_g = None, e.code
            assert rc_g == EO2EMP[rc_o], (block, how, rc_o, rc_g)
            if rc_g == 0:
                assert got == want, (block, how)
    dbuf.free()


def test_context_reused_across_read_lengths():
    """One context, reset between runs whose read lengths go 100, 250, 100, 150 bp (pageable arrays
    through the pinned stages, which grow for 250 bp and are kept for the shorter lengths after
    it): every run gives the oracle's output for its own reads."""
    with native.Context(min_ovlp=31) as c:
        for k, L in enumerate((100, 250, 100, 150)):
            g = gen.genome(50_000, 600 + k)
            r = gen.reads(g, 8000 + 16 * k + 3, L, 0.5, 0.005, 700 + k)
            rc, want, summ, _ = ol.run_reads(r.tobytes(), r.shape[0], L, min_ovlp=31)
            assert rc == 0 and summ["n_edges"] > 0
            c.reset()
            c.add_reads(r)
            assert c.run() == want, L
