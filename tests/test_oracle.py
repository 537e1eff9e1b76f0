"""Pins the oracle: known answers, rolling-update invariants, the FASTQ reader, and agreement with
a string-level brute force on small read sets. CPU only."""
import random
import struct

import pytest

import bruteforce as bf
import oracle_lib as ol

P1, P2 = bf.P1, bf.P2


def is_prime(p):
    if p < 2:
        return False
    i = 2
    while i * i <= p:
        if p % i == 0:
            return False
        i += 1
    return True


def test_primes_and_inverses():
    assert is_prime(P1) and is_prime(P2)
    assert P1 == 2**29 - 3 and P2 == 2**29 - 33
    # Fermat inverses (overlap.cu:121-122) and the values they must take.
    assert ol.lib().eo_modulo_exp(5, P1 - 2, P1) == 107374182 == pow(5, -1, P1)
    assert ol.lib().eo_modulo_exp(7, P2 - 2, P2) == 76695840 == pow(7, -1, P2)
    assert (5 * 107374182) % P1 == 1 and (7 * 76695840) % P2 == 1


def test_modulo_tables():
    m1, m2 = ol.modulos()
    assert m1 == [pow(5, i, P1) for i in range(256)]
    assert m2 == [pow(7, i, P2) for i in range(256)]
    assert (m1[99], m1[149], m1[255]) == (418017119, 104577844, 121899516)
    assert (m2[99], m2[149], m2[255]) == (510103779, 377358276, 526359340)


def test_known_keys():
    assert ol.key_of("A") == 0x0000000100000001
    assert ol.key_of("ACGT") == 0x000000C6000001D8
    assert ol.key_of("GATTACA") == 0x0001094700079053
    # homopolymer: geometric series, e(T) = 3
    h1 = 3 * (5**12 - 1) // 4 % P1
    h2 = 3 * (7**12 - 1) // 6 % P2
    assert ol.key_of("T" * 12) == (h1 << 32) | h2 == 0x0AE9F7BC1C80A59C
    # lower case encodes like upper case
    assert ol.key_of("gattaca") == ol.key_of("GATTACA")


def test_key_matches_polynomial_definition():
    rng = random.Random(7)
    for _ in range(300):
        n = rng.randint(1, 255)
        s = "".join(rng.choice("ACGT") for _ in range(n))
        assert ol.key_of(s) == bf.key(s)
        assert ol.key_of_rc(s) == bf.key(bf.revcomp(s))


def test_next_sfx_invariant():
    rng = random.Random(11)
    for _ in range(500):
        n = rng.randint(2, 255)
        s = "".join(rng.choice("ACGT") for _ in range(n))
        assert ol.next_sfx(bf.key(s), s[0], n) == bf.key(s[1:])


def test_prev_pfx_invariant_and_wrap():
    rng = random.Random(13)
    checked = 0
    for _ in range(500):
        n = rng.randint(2, 255)
        s = "".join(rng.choice("ACGT") for _ in range(n))
        k = bf.key(s)
        if (k >> 32) >= 4 and (k & 0xFFFFFFFF) >= 4:
            assert ol.prev_pfx(k, s[-1]) == bf.key(s[:-1])
            checked += 1
    assert checked > 450
    inv1, inv2 = pow(5, -1, P1), pow(7, -1, P2)
    # the u64 wrap (overlap.cu:87-88): h < e gives ((INV * (h - e mod 2^64)) mod 2^64) mod P
    for h, base in [(1, "G"), (0, "A"), (2, "T"), (1, "C"), (0, "G"), (3, "G")]:
        assert h < bf.enc(base)
        e = bf.enc(base)
        want1 = ((inv1 * ((h - e) % 2**64)) % 2**64) % P1
        want2 = ((inv2 * ((h - e) % 2**64)) % 2**64) % P2
        got = ol.prev_pfx((h << 32) | h, base)
        assert got == (want1 << 32) | want2
        true_step = ((h - e) * inv1) % P1
        assert want1 != true_step  # the wrap really differs from the modular inverse step
    assert ol.prev_pfx((1 << 32) | 100, "G") >> 32 == 214748939


def test_wrap_read_steps_through_the_wrap():
    """bf.wrap_read's prefix of length ln has H1 below its last base's encoding, so the oracle's
    PrevPfx step from it takes the wrap and leaves the true key of the shorter prefix."""
    rng = random.Random(5)
    for ln in (40, 63, 80, 99):
        r = bf.wrap_read(100, ln, rng)
        k = bf.key(r[:ln])
        assert (k >> 32) < bf.enc(r[ln - 1])
        assert ol.prev_pfx(k, r[ln - 1]) >> 32 != bf.key(r[:ln - 1]) >> 32


FQ = lambda seqs: b"".join(b"@r%d\n%s\n+\n%s\n" % (i, s, b"I" * len(s)) for i, s in enumerate(seqs))


@pytest.mark.parametrize("block", [10**6, 50, 64, 101, 250])
def test_reader_rules(block):
    rng = random.Random(block)
    seqs = []
    for i in range(60):
        s = bytearray(rng.choice(b"ACGT") for _ in range(20))
        if i % 7 == 3:
            s[rng.randrange(20)] = ord("N")
        if i % 11 == 5:
            s[rng.randrange(20)] = ord("n")
        if i % 5 == 2:
            s = s.lower()
        seqs.append(bytes(s))
    data = FQ(seqs)
    rc, reads = ol.parse(data, block)
    assert rc == 0
    assert reads == bf.kept_reads(data)
    assert len(reads) == len([s for s in seqs if b"N" not in s and b"n" not in s])


def test_reader_edge_cases():
    # empty sequence line vanishes; last record without newline is kept
    data = b"@a\n\n+\n\n@b\nACGT\n+\nIIII\n@c\nTTGA\n+\nIIII"
    rc, reads = ol.parse(data, 1 << 20)
    assert rc == 0 and reads == ["ACGT", "TTGA"] == bf.kept_reads(data)
    # IUPAC bytes with bit 3 set drop the read (K, M, Y, H); others are rejected
    rc, reads = ol.parse(FQ([b"ACGK", b"ACGT", b"AYGT", b"HHHH"]), 1 << 20)
    assert rc == 0 and reads == ["ACGT"]
    assert ol.parse(FQ([b"ACGT", b"ACRT"]), 1 << 20)[0] == -4  # R survives the N filter: unsupported
    assert ol.parse(FQ([b"ACGT", b"ACG"]), 1 << 20)[0] == -3   # unequal lengths
    assert ol.parse(b"xx\nACGT\n+\nIIII\n", 1 << 20)[0] == -2  # first byte not '@'
    assert ol.parse(FQ([b"A" * 40]), 16)[0] == -2              # record longer than a block
    assert ol.parse(FQ([b"A" * 256]), 1 << 20)[0] == -3        # longer than 255
    assert ol.parse(b"@a\r\nACGT\r\n+\r\nIIII\r\n", 1 << 20)[0] == -4  # CRLF rejected


def _random_reads(rng, n, L, genome_len, p_rc, p_sub):
    g = "".join(rng.choice("ACGT") for _ in range(genome_len))
    out = []
    for _ in range(n):
        st = rng.randrange(genome_len - L + 1)
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
    (4, 150, 32, 15, 250, 0.5, 0.005),   # high coverage: duplicates and large key groups
    (5, 60, 16, 16, 200, 0.5, 0.0),      # o == L: no rounds, empty output
    (6, 200, 28, 6, 150, 0.5, 0.02),     # short overlaps, many matches per key
])
def test_oracle_matches_bruteforce(seed, n, L, o, glen, p_rc, p_sub):
    rng = random.Random(seed)
    reads = _random_reads(rng, n, L, glen, p_rc, p_sub)
    assert bf.wrap_free(reads, o)
    data = FQ([r.encode() for r in reads])
    rc, out, summ, rounds = ol.run_fastq(data, min_ovlp=o)
    assert rc == 0
    recs, stats = bf.overlap(reads, o)
    assert out == bf.encode(recs)
    assert len(rounds) == len(stats)
    for a, b in zip(rounds, stats):
        for f in ("s", "p", "s_after", "p_after", "edges", "ovlp"):
            assert a[f] == b[f], (f, a, b)


def test_block_size_does_not_change_output():
    rng = random.Random(99)
    reads = _random_reads(rng, 100, 25, 400, 0.5, 0.01)
    data = FQ([r.encode() for r in reads])
    outs = {ol.run_fastq(data, min_ovlp=10, map_mem=m * 16)[1] for m in (10**6, 97, 160, 333)}
    assert len(outs) == 1


def test_search_partitions_split_records():
    """Lists reaching M/2 split a round into key-aligned records (cqutils.cuh:107-137); the
    concatenated edges equal the single-record run."""
    rng = random.Random(5)
    reads = _random_reads(rng, 150, 30, 600, 0.5, 0.0)
    data = FQ([r.encode() for r in reads])
    rc, big, _, _ = ol.run_fastq(data, min_ovlp=12)
    assert rc == 0
    rc, small, _, rounds = ol.run_fastq(data, min_ovlp=12, reduce_mem=16 * 320)
    assert rc == 0
    a, b = bf.decode(big), bf.decode(small)
    assert len(b) > len(a)
    merged = {}
    for ell, e in b:
        merged.setdefault(ell, []).extend(e)
    assert [(ell, merged[ell]) for ell, _ in a] == a


# Search partitions worked by hand from cqutils.cuh:113-135 (window w = min(rest, M/2) per side;
# unless both windows are below M/2, cut at m = min(last of A's window, last of B's window) with
# lower_bound on each side). Each case notes the steps.
SEARCH_CASES = [
    # M = 6, H = 3. (0,0): windows 10 20 30 | 15 25 35, m = 30 -> lb 2 | 2.  (2,2): 30 40 50 |
    # 35 45 55, m = 50 -> 2 | 2.  (4,4): 50 60 70 | 55 65 75, m = 70 -> 2 | 2.  (6,6): 70 80 |
    # 75 85, both below 3: the rest.  upper_bound would cut (3, 2) first; windows of M = 6 would
    # give (5, 5), (3, 3).
    ([10, 20, 30, 40, 50, 60, 70, 80], [15, 25, 35, 45, 55, 65, 75, 85], 6, 0,
     [(2, 2), (2, 2), (2, 2), (2, 2)]),
    # M = 6. (0,0): 1 2 3 | 7 8 9, m = 3 -> 2 | 0 (a partition without prefixes: no record).
    # (2,0): 3 4 5 | 7 8 9, m = 5 -> 2 | 0.  (4,0): 5 6 7 | 7 8 9, m = 7 -> 2 | 0.  (6,0): 7 8 |
    # 7 8 9 (B's window is full), m = min(8, 9) = 8 -> 1 | 1.  (7,1): 8 | 8 9, both below 3: 1 | 2.
    # upper_bound would cut (3, 0) first.
    ([1, 2, 3, 4, 5, 6, 7, 8], [7, 8, 9], 6, 0, [(2, 0), (2, 0), (2, 0), (1, 1), (1, 2)]),
    # M = 7, H = 3 (integer half). (0,0): 1 1 2 | 1 3 3, m = min(2, 3) = 2 -> 2 | 1.  (2,1): 2 4 4 |
    # 3 3 5, m = min(4, 5) = 4 -> 1 | 2.  (3,3): 4 4 6 | 5 6, m = 6 -> 2 | 1 (B's window 5 6 is
    # below 3 but A's is not).  (5,4): 6 | 6, both below 3: 1 | 1.
    ([1, 1, 2, 4, 4, 6], [1, 3, 3, 5, 6], 7, 0, [(2, 1), (1, 2), (2, 1), (1, 1)]),
    # M = 4, H = 2. (0,0): 5 6 | 1 2, m = 2 -> 0 | 1.  (0,1): 5 6 | 2, m = min(6, 2) = 2 -> 0 | 0:
    # no cut, the reference pushes empty partitions forever.
    ([5, 6, 7, 8, 9, 10, 11], [1, 2], 4, -6, None),
    # one side empty from the start: one partition of the other side alone
    ([3, 4, 5], [], 4, 0, [(3, 0)]),
    ([], [3, 4, 5], 4, 0, [(0, 3)]),
    # both windows below H from the start: one partition
    ([1, 2, 3], [2, 3, 4], 8, 0, [(3, 3)]),
]


@pytest.mark.parametrize("a,b,M,rc,parts", SEARCH_CASES)
def test_search_partitions_hand_worked(a, b, M, rc, parts):
    got_rc, got = ol.search_partitions(a, b, M)
    assert got_rc == rc
    if rc == 0:
        assert got == parts


def test_search_partitions_of_a_run_match_the_records():
    """The records of a framed run are the partitions with both sides non-empty: per round, the
    oracle's record count equals the partitions computed from its own sorted lists' keys, which are
    rebuilt here from the reads (the suffix and prefix strings of the live reads)."""
    rng = random.Random(5)
    reads = _random_reads(rng, 150, 30, 600, 0.5, 0.0)
    data = FQ([r.encode() for r in reads])
    M = 320
    rc, out, _, rounds = ol.run_fastq(data, min_ovlp=12, reduce_mem=16 * M)
    assert rc == 0
    # round 0: every kept read (deduplicated by key, both orientations) is live on both sides
    oriented = {}
    for k, r in enumerate(reads):
        for rid, s in ((2 * k, r), (2 * k + 1, bf.revcomp(r))):
            key = ol.key_of(s)
            if key not in oriented or rid < oriented[key][0]:
                oriented[key] = (rid, s)
    live = list(oriented.values())
    L = len(reads[0])
    sfx = sorted(ol.key_of(s[1:]) for _, s in live)
    pfx = sorted(ol.key_of(s[:L - 1]) for _, s in live)
    prc, parts = ol.search_partitions(sfx, pfx, M)
    assert prc == 0
    assert rounds[0]["records"] == sum(1 for na, nb in parts if na and nb) > 1


# Merge partitions worked by hand from cqutils.cuh:43-82 (windows of M/2; case 1: A's window last
# <= B's first -> A's window alone; case 2: B's window last <= A's first -> B's window alone; case
# 3: the side with the larger window last is cut by upper_bound at the other's; the rest once
# both windows are short; a side left over alone at the end).
MERGE_CASES = [
    # (0,0) 1 2 | 3 4: case 1 (2,0). (2,0) 5 6 | 3 4: case 2 (0,2). (2,2) 5 6 | 7 8: case 1 (2,0).
    # A done: (0,2).
    ([1, 2, 5, 6], [3, 4, 7, 8], 4, [(2, 0), (0, 2), (2, 0), (0, 2)]),
    # (0,0) 1 3 | 2 3: case 3 with equal window lasts, no cut (2,2). (2,2) 5 7 | 8 9: case 1 (2,0).
    # A done: (0,2).
    ([1, 3, 5, 7], [2, 3, 8, 9], 4, [(2, 2), (2, 0), (0, 2)]),
    # (0,0) 1 4 | 2 3: case 3, 3 < 4: A cut at upper_bound(3) = 1 -> (1,2). (1,2) 4 6 | 5 8: case 3,
    # 6 < 8: B cut at upper_bound(6) = 1 -> (2,1). (3,3) 9 | 8: both short -> (1,1).
    ([1, 4, 6, 9], [2, 3, 5, 8], 4, [(1, 2), (2, 1), (1, 1)]),
    # (0,0) 5 5 | 3 5: case 2 (B's last 5 <= A's first 5) -> (0,2): B's 5 goes before A's 5s.
    # (0,2) 5 5 | 9 9: case 1 (2,0). (2,2) 6 7 | 9 9: case 1 (2,0). A done: (0,2).
    ([5, 5, 6, 7], [3, 5, 9, 9], 4, [(0, 2), (2, 0), (2, 0), (0, 2)]),
]


@pytest.mark.parametrize("a,b,M,parts", MERGE_CASES)
def test_merge_partitions_hand_worked(a, b, M, parts):
    assert ol.merge_partitions(a, b, M) == (0, parts)


def test_cq_sort_two_chunks_hand_worked():
    """Keys 5 3 5 7 | 5 9 5 6 with values 0..7, M = 4. Chunk 0 sorts to 3(1) 5(0) 5(2) 7(3), chunk 1
    to 5(4) 5(6) 6(7) 9(5). Merge with windows of 2: (0,0) 3 5 | 5 5 case 1 -> 3(1) 5(0); (2,0)
    5 7 | 5 5 case 2 -> 5(4) 5(6); (2,2) 5 7 | 6 9 case 3 cuts B at upper_bound(7) -> [5(2) 7(3)
    6(7)] sorted stably -> 5(2) 6(7) 7(3); then 9(5). Key 5 ends 0, 4, 6, 2: chunk 1's pair ahead of
    chunk 0's second, where one stable sort would give 0, 2, 4, 6."""
    rc, k, v = ol.cq_sort([5, 3, 5, 7, 5, 9, 5, 6], list(range(8)), 4)
    assert rc == 0
    assert k == [3, 5, 5, 5, 5, 6, 7, 9]
    assert v == [1, 0, 4, 6, 2, 7, 3, 5]
    assert ol.cq_sort([5, 3, 5, 7, 5, 9, 5, 6], list(range(8)), 8)[2] == [1, 0, 2, 4, 6, 7, 3, 5]


def test_cq_sort_three_chunks_fifo_order_hand_worked():
    """Keys 4 4 | 2 4 | 1 4 with values 0..5, M = 2: chunks 4(0) 4(1), 2(2) 4(3), 1(4) 4(5). The FIFO
    merges chunks 0 and 1 first (windows of 1: case 2 -> 2(2); case 1 -> 4(0); case 1 -> 4(1); the
    rest 4(3)), then chunk 2 with that run, chunk 2 as A: case 1 -> 1(4); case 2 -> 2(2); case 1 ->
    4(5); the rest 4(0) 4(1) 4(3). Chunk 2's 4 comes first."""
    rc, k, v = ol.cq_sort([4, 4, 2, 4, 1, 4], list(range(6)), 2)
    assert rc == 0
    assert k == [1, 2, 4, 4, 4, 4]
    assert v == [4, 2, 5, 0, 1, 3]


def test_cq_sort_distinct_keys_is_the_sorted_order():
    rng = random.Random(4)
    keys = rng.sample(range(1 << 40), 3000)
    for M in (2, 7, 64, 999, 3000):
        rc, k, v = ol.cq_sort(keys, list(range(3000)), M)
        assert rc == 0 and k == sorted(keys) and [keys[i] for i in v] == k


def _edge_sets(out):
    """Per overlap length: the sorted edge list of all records of that length."""
    d = {}
    for ell, e in bf.decode(out):
        d.setdefault(ell, []).extend(e)
    return {ell: sorted(e) for ell, e in d.items()}


def test_out_of_core_runs_keep_the_edge_sets():
    """2n > M: the chunked sort and merges reorder equal keys, the chunked regeneration reorders the
    lists, the records split by M/2; the edges of every overlap length are the in-core run's (the
    same pairs), and with no equal keys in a list the concatenated records are the in-core ones."""
    rng = random.Random(12)
    # no two oriented reads equal (the dedup's choice between equal reads is what the order decides)
    seen, reads = set(), []
    for r in _random_reads(rng, 500, 40, 1500, 0.5, 0.01):
        if r not in seen and bf.revcomp(r) not in seen and r != bf.revcomp(r):
            seen.update((r, bf.revcomp(r)))
            reads.append(r)
    assert len(reads) > 300
    data = FQ([r.encode() for r in reads])
    rc, ref, rsum, _ = ol.run_fastq(data, min_ovlp=20)
    assert rc == 0
    for m in (700, 301, 128):
        rc, out, summ, rounds = ol.run_fastq(data, min_ovlp=20, reduce_mem=16 * m)
        assert rc == 0, m
        assert summ["n_dedup"] == rsum["n_dedup"] and summ["n_edges"] == rsum["n_edges"]
        assert _edge_sets(out) == _edge_sets(ref), m
    # error-free reads of one strand, distinct start positions: no list holds a key twice
    g = "".join(rng.choice("ACGT") for _ in range(20000))
    starts = rng.sample(range(20000 - 40), 600)
    reads = [g[p:p + 40] for p in starts]
    data = FQ([r.encode() for r in reads])
    rc, ref, _, _ = ol.run_fastq(data, min_ovlp=25)
    rc2, out, _, _ = ol.run_fastq(data, min_ovlp=25, reduce_mem=16 * 500)
    assert rc == rc2 == 0
    merged = {}
    for ell, e in bf.decode(out):
        merged.setdefault(ell, []).extend(e)
    assert [(ell, merged[ell]) for ell, _ in bf.decode(ref)] == bf.decode(ref)


def test_out_of_core_dedup_keeps_the_first_of_the_merged_order():
    """Read 0 and read 3 equal (X), M = 4 entries: chunk 0 holds ids 0-3 (X and read 1), chunk 1 ids
    4-7 (read 2 and X again). Read 1's keys both exceed key(X) < key(rc X), and read 2 has one key
    below key(X) and one above, so chunk 1's two smallest keys are key(read 2 or its rc) and key(X)
    (id 6), and chunk 0's smallest is key(X) (id 0): the first merge window is case 2 (B's window
    last, key(X), is not above A's first, key(X)), which puts id 6 ahead of id 0. The dedup keeps
    id 6 where the in-core order (M = 8) keeps id 0."""
    rng = random.Random(3)
    KEY = ol.key_of
    while True:  # X with key(X) < key(rc X), read 1 above key(X) on both strands, read 2 across it
        x = "".join(rng.choice("ACGT") for _ in range(20))
        if KEY(x) < KEY(bf.revcomp(x)):
            break
    while True:
        r1 = "".join(rng.choice("ACGT") for _ in range(20))
        if min(KEY(r1), KEY(bf.revcomp(r1))) > KEY(x):
            break
    while True:
        r2 = "".join(rng.choice("ACGT") for _ in range(20))
        k2 = sorted((KEY(r2), KEY(bf.revcomp(r2))))
        if k2[0] < KEY(x) < k2[1] and k2[1] != KEY(bf.revcomp(x)):
            break
    reads = [x, r1, r2, x]
    rc, incore = ol.dedup(reads, 8)
    rc2, ooc = ol.dedup(reads, 4)
    assert rc == rc2 == 0 and len(incore) == len(ooc) == 6
    assert 0 in incore and 6 not in incore
    assert 6 in ooc and 0 not in ooc
    # key order: read 2's smaller key (id 4 or 5), then key(X): id 6 out of core, id 0 in core
    assert ooc[0] in (4, 5) and ooc[1] == 6
    assert incore[0] in (4, 5) and incore[1] == 0


def test_search_partition_without_a_cut_is_an_error():
    """rc(B) + 10 random bases + B: at overlap 45 all 400 suffix keys are key(B) and all prefix keys
    key(rc B); with M = 400 the first window is one key on both sides, lower_bound cuts nothing and
    the reference's loop (cqutils.cuh:113-131) never ends. The oracle reports it; with M large the
    same reads run (one record per round) and end with no edge at 45 (B is not its own rc)."""
    rng = random.Random(21)
    b = "".join(rng.choice("ACGT") for _ in range(45))
    assert bf.revcomp(b) != b
    reads = [bf.revcomp(b) + "".join(rng.choice("ACGT") for _ in range(10)) + b for _ in range(200)]
    data = FQ([x.encode() for x in reads])
    assert ol.run_fastq(data, min_ovlp=31, reduce_mem=16 * 400)[0] == -6
    rc, out, summ, rounds = ol.run_fastq(data, min_ovlp=31)
    assert rc == 0 and summ["n_records"] == 69
    r45 = [r for r in rounds if r["ovlp"] == 45][0]
    assert r45["s"] == 400 and r45["edges"] == 0


def test_empty_and_trivial():
    rc, out, summ, _ = ol.run_fastq(b"")
    assert rc == 0 and out == b""
    rc, out, _, _ = ol.run_fastq(FQ([b"NNNN"]))
    assert rc == 0 and out == b""
    # a single read: its own suffix/prefix pairs are self edges, so every record is empty
    rc, out, _, rounds = ol.run_fastq(FQ([b"ACGTTGCAACGTTGCA"]), min_ovlp=4)
    assert rc == 0
    recs = bf.decode(out)
    assert all(len(e) == 0 for _, e in recs) and len(recs) == 12
    assert [r[0] for r in recs] == list(range(15, 3, -1))


def test_record_layout():
    rng = random.Random(3)
    reads = _random_reads(rng, 50, 20, 200, 0.5, 0.0)
    rc, out, summ, rounds = ol.run_fastq(FQ([r.encode() for r in reads]), min_ovlp=8)
    assert rc == 0
    off = 0
    for r in rounds:
        n, ell = struct.unpack_from("<Qi", out, off)
        assert n == r["edges"] and ell == r["ovlp"]
        off += 12 + 8 * n
    assert off == len(out)
