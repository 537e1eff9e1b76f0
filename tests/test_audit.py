"""The collision audit (FPCHECK, overlap.cu:215-219,274-284,295-359): distinct fingerprints and
distinct strings per list. CPU tests pin the oracle against the string-level brute force and a
real dual-hash collision (tests/golden/collision_pair.json, from make_collision.py); GPU tests
compare the library's audit with the oracle's."""
import json
import os
import random

import numpy as np
import pytest

import bruteforce as bf
import oracle_lib as ol
from empanada_amd import gen

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "collision_pair.json")) as f:
    PAIR = json.load(f)
CA, CB = PAIR["a"], PAIR["b"]


def rnd(rng, n):
    return "".join(rng.choice("ACGT") for _ in range(n))


def reads_array(strs):
    return np.frombuffer("".join(strs).encode(), dtype=np.uint8).reshape(len(strs), len(strs[0]))


def oracle_audit(strs, o, **kw):
    r = reads_array(strs)
    rc, out, summ, au = ol.run_reads_audit(r.tobytes(), r.shape[0], r.shape[1], min_ovlp=o, **kw)
    assert rc == 0
    return out, summ, au


def collision_sets():
    """Read sets around the planted pair: (name, reads, min_ovlp)."""
    rng = random.Random(31)
    x, y = rnd(rng, 7), rnd(rng, 7)
    # flanks that differ next to the pair, so that no longer string collides as well
    dedup = [CA, CB, CA] + [rnd(rng, 24) for _ in range(20)]                        # full reads collide
    sfx = [x + "A" + CA, y + "C" + CB] + [rnd(rng, 32) for _ in range(30)]            # suffixes at 24
    pfx = [CA + "G" + x, CB + "T" + y] + [rnd(rng, 32) for _ in range(30)]            # prefixes at 24
    return [("dedup", dedup, 10), ("suffix", sfx, 20), ("prefix", pfx, 20)]


def test_collision_fixture_is_a_real_collision():
    assert CA != CB and len(CA) == len(CB) == 24
    assert bf.key(CA) == bf.key(CB) == PAIR["key"]
    assert ol.key_of(CA) == ol.key_of(CB) == PAIR["key"]
    d = [bf.enc(x) - bf.enc(y) for x, y in zip(CA, CB)]
    assert d == PAIR["difference"]


@pytest.mark.parametrize("name", ["dedup", "suffix", "prefix"])
def test_oracle_audit_counts_the_planted_collision(name):
    reads, o = dict((n, (r, o)) for n, r, o in collision_sets())[name]
    _, _, au = oracle_audit(reads, o)
    assert au == bf.audit(reads, o)
    hit = [a for a in au if a["strings"] != a["keys"]]
    assert len(hit) == 1 and hit[0]["strings"] == hit[0]["keys"] + 1
    want = {"dedup": (0, len(reads[0])), "suffix": (1, 24), "prefix": (2, 24)}[name]
    assert (hit[0]["list"], hit[0]["ovlp"]) == want


def test_false_overlap_from_the_collision_is_in_the_output():
    """A suffix CA and a prefix CB share a key: the hash join reports the overlap (as the
    reference does), and the audit of both lists shows no collision inside either list."""
    rng = random.Random(32)
    reads = [rnd(rng, 8) + CA, CB + rnd(rng, 8)] + [rnd(rng, 32) for _ in range(10)]
    out, _, au = oracle_audit(reads, 20)
    recs = bf.decode(out)
    assert (24, [(0, 2)]) in recs
    assert all(a["strings"] == a["keys"] for a in au)


@pytest.mark.parametrize("seed,n,L,o,glen,p_rc,p_sub", [
    (1, 40, 20, 8, 300, 0.0, 0.0),
    (3, 120, 30, 12, 500, 0.5, 0.01),
    (7, 500, 40, 10, 120, 0.5, 0.0),
    (8, 300, 24, 6, 60, 0.5, 0.02),
])
def test_oracle_audit_matches_bruteforce(seed, n, L, o, glen, p_rc, p_sub):
    rng = random.Random(seed)
    g = rnd(rng, glen)
    reads = []
    for _ in range(n):
        st = rng.randrange(glen - L + 1)
        s = g[st:st + L]
        if rng.random() < p_rc:
            s = bf.revcomp(s)
        reads.append("".join((rng.choice([b for b in "ACGT" if b != c]) if rng.random() < p_sub else c) for c in s))
    if not bf.wrap_free(reads, o):
        pytest.skip("the PrevPfx wrap is outside the brute force")
    _, _, au = oracle_audit(reads, o)
    assert au == bf.audit(reads, o)
    assert all(a["strings"] == a["keys"] for a in au)  # no collision in these sets


# ---------------------------------------------------------------------------------------------
# GPU: the library's audit against the oracle's, through the C ABI.
def gpu_audit(strs, o, audit=True, **kw):
    from empanada_amd import native
    r = reads_array(strs) if isinstance(strs, list) else strs
    with native.Context(min_ovlp=o, **kw) as c:
        c.set_audit(audit)
        c.add_reads(r)
        out = c.run()
        return out, c.audits()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["dedup", "suffix", "prefix"])
def test_gpu_audit_finds_the_planted_collision(name):
    reads, o = dict((n, (r, o)) for n, r, o in collision_sets())[name]
    want_out, _, want = oracle_audit(reads, o)
    out, got = gpu_audit(reads, o)
    assert out == want_out
    assert got == want
    assert sum(a["strings"] - a["keys"] for a in got) == 1


@pytest.mark.gpu
@pytest.mark.parametrize("G,n,L,o,p_rc,p_sub,seeds", [
    (2 * 10**5, 30000, 100, 31, 0.5, 0.005, (7, 8)),     # 15x, both strands, errors
    (3 * 10**4, 20000, 150, 75, 0.5, 0.0, (9, 10)),      # 100x: many equal suffixes per key
    (5 * 10**4, 8000, 250, 31, 0.5, 0.0, (11, 12)),
])
def test_gpu_audit_generated_vs_oracle(G, n, L, o, p_rc, p_sub, seeds):
    g = gen.genome(G, seeds[0])
    r = gen.reads(g, n, L, p_rc, p_sub, seeds[1])
    rc, want_out, _, want = ol.run_reads_audit(r.tobytes(), n, L, min_ovlp=o)
    assert rc == 0
    out, got = gpu_audit(r, o)
    assert out == want_out
    assert got == want
    assert gpu_audit(r, o, audit=False) == (want_out, [])


@pytest.mark.gpu
def test_gpu_audit_large_buckets_vs_oracle():
    """5,000 identical reads (one dedup key group over a bucket's capacity) and a 3000 x 3000
    suffix/prefix key group (spilled buckets in the rounds)."""
    rng = random.Random(77)
    s = rnd(rng, 40)
    reads = [s] * 5000 + [rnd(rng, 40) for _ in range(300)]
    common = rnd(rng, 30)
    reads += [rnd(rng, 10) + common for _ in range(3000)] + [common + rnd(rng, 10) for _ in range(3000)]
    r = reads_array(reads)
    rc, want_out, _, want = ol.run_reads_audit(r.tobytes(), r.shape[0], 40, min_ovlp=29)
    assert rc == 0
    out, got = gpu_audit(r, 29)
    assert out == want_out
    assert got == want
