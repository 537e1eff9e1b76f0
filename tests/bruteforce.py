"""String-level restatement of the overlap rules, for small read sets only.

Independent of the oracle's arithmetic: kept reads come from a whole-file line parser, keys from
the polynomial definition H(s) = sum e(s_i) R^(|s|-1-i) mod P (Python integers), suffix/prefix
matches from string equality, and the round structure from the rules of the reference:

- kept reads: FASTQ line i with i % 4 == 1; dropped if a byte in [64,127] has bit 3 set
  (preprocess.cu:119-121,150-155); empty lines vanish.
- oriented read 2k = read_k, 2k+1 = reverse complement (fingerprint.cu:112-113).
- dedup keeps the smallest id per distinct key (overlap.cu:212-213,222).
- round t (t < L - o while suffixes remain): suffixes and prefixes of length L-1-t; an edge
  (s, p) for every suffix/prefix pair with equal strings and s//2 != p//2, ordered by
  (key, s&1, s, p&1, p) (overlap.cu:413-415,557-601); a record (n, L-1-t, from[], to[]) when
  both lists are non-empty; every s that is a `from` and every p that is a `to` leaves its list
  (overlap.cu:424-429,694-710).

Hash collisions and the u64 wrap of PrevPfx (overlap.cu:87-88) are outside this model; callers
keep inputs small enough that neither occurs (checked by `wrap_free`).
"""
import struct

P1, P2, R1, R2 = 536870909, 536870879, 5, 7
REVC = {ord("A"): "T", ord("C"): "G", ord("G"): "C", ord("T"): "A",
        ord("a"): "T", ord("c"): "G", ord("g"): "C", ord("t"): "A"}


def enc(c):
    return ((ord(c) >> 1) & 3) + 1


def key(s):
    n = len(s)
    h1 = sum(enc(c) * pow(R1, n - 1 - i, P1) for i, c in enumerate(s)) % P1
    h2 = sum(enc(c) * pow(R2, n - 1 - i, P2) for i, c in enumerate(s)) % P2
    return (h1 << 32) | h2


def revcomp(s):
    return "".join(REVC[ord(c)] for c in reversed(s))


def kept_reads(fastq: bytes):
    lines = fastq.split(b"\n")
    out = []
    for i, ln in enumerate(lines):
        if i % 4 != 1 or len(ln) == 0:
            continue
        if any(64 <= c < 128 and (c & 8) for c in ln):
            continue
        out.append(ln.decode("ascii"))
    return out


def rot(r):
    return ((r & 1) << 31) | (r >> 1)


def overlap(reads, min_ovlp):
    """Returns (records, stats); records = list of (ovlp, [(from, to), ...])."""
    if not reads:
        return [], []
    L = len(reads[0])
    assert all(len(r) == L for r in reads)
    orient = []
    for r in reads:
        orient.append(r)
        orient.append(revcomp(r))
    first = {}
    for rid, s in enumerate(orient):
        k = key(s)
        if k not in first:
            first[k] = rid
    S = set(first.values())
    P = set(S)
    records, stats = [], []
    t = 0
    while t < L - min_ovlp and S:
        ell = L - 1 - t
        sg, pg = {}, {}
        for r in S:
            sg.setdefault(orient[r][t + 1:], []).append(r)
        for r in P:
            pg.setdefault(orient[r][:ell], []).append(r)
        edges = []
        for w in sorted(set(sg) & set(pg), key=key):
            for s in sorted(sg[w], key=rot):
                for p in sorted(pg[w], key=rot):
                    if s >> 1 != p >> 1:
                        edges.append((s, p))
        st = dict(s=len(S), p=len(P), edges=len(edges), ovlp=ell)
        if S and P:
            records.append((ell, edges))
        S -= {e[0] for e in edges}
        P -= {e[1] for e in edges}
        st.update(s_after=len(S), p_after=len(P))
        stats.append(st)
        t += 1
    return records, stats


def encode(records):
    out = bytearray()
    for ell, edges in records:
        out += struct.pack("<Qi", len(edges), ell)
        out += struct.pack("<%dI" % len(edges), *[e[0] for e in edges])
        out += struct.pack("<%dI" % len(edges), *[e[1] for e in edges])
    return bytes(out)


def decode(data: bytes):
    recs, off = [], 0
    while off < len(data):
        n, ell = struct.unpack_from("<Qi", data, off)
        off += 12
        fr = struct.unpack_from("<%dI" % n, data, off)
        off += 4 * n
        to = struct.unpack_from("<%dI" % n, data, off)
        off += 4 * n
        recs.append((ell, list(zip(fr, to))))
    assert off == len(data)
    return recs


def wrap_free(reads, min_ovlp):
    """True when no prefix fingerprint that the rounds step through has a half below 5, so the
    rolling prefix update never takes the u64-wrap branch and equals the true prefix key."""
    if not reads:
        return True
    L = len(reads[0])
    for r in reads:
        for s in (r, revcomp(r)):
            for ln in range(min_ovlp + 1, L + 1):
                k = key(s[:ln])
                if (k >> 32) < 5 or (k & 0xFFFFFFFF) < 5:
                    return False
    return True


def wrap_read(L, ln, rng):
    """A read of length L whose prefix of length ln has H1 < e(its last base): stepping that
    prefix back takes PrevPfx's u64-wrap branch (overlap.cu:87-88). The first ln - 13 bases are
    random; the next 12 are solved for by base-5 digits (H is linear in them), the last one is G
    (e = 4): H1(prefix[:ln-1]) * 5 + 4 = v (mod P1) for a random v < 4."""
    dec = {1: "A", 2: "C", 3: "T", 4: "G"}
    k, inv5, p5k = 12, pow(5, -1, P1), pow(5, 12, P1)
    while True:
        a = "".join(rng.choice("ACGT") for _ in range(ln - 1 - k))
        tgt = (rng.randrange(4) - 4) * inv5 % P1
        x = (tgt - (key(a) >> 32) * p5k) % P1
        digits = []
        for _ in range(k):
            digits.append(x % 5)
            x //= 5
        if x == 0 and all(1 <= d <= 4 for d in digits):
            pre = a + "".join(dec[d] for d in reversed(digits)) + "G"
            assert (key(pre) >> 32) < 4
            return pre + "".join(rng.choice("ACGT") for _ in range(L - ln))


def audit(reads, min_ovlp):
    """The collision audit from strings: for the 2n oriented reads, then per round for the live
    suffix and prefix lists, (list, ovlp, entries, distinct keys, distinct strings). The live lists
    follow string-equality matching, so callers use inputs where no key collision joins a suffix
    to a prefix (a collision inside one list is what the audit is for)."""
    if not reads:
        return []
    L = len(reads[0])
    orient = []
    for r in reads:
        orient.append(r)
        orient.append(revcomp(r))

    def rec(lst, ell, strs):
        return dict(list=lst, ovlp=ell, entries=len(strs), keys=len({key(w) for w in strs}), strings=len(set(strs)))

    out = [rec(0, L, orient)]
    first = {}
    for rid, s in enumerate(orient):
        first.setdefault(key(s), rid)
    S = set(first.values())
    P = set(S)
    t = 0
    while t < L - min_ovlp and S:
        ell = L - 1 - t
        out.append(rec(1, ell, [orient[r][t + 1:] for r in S]))
        out.append(rec(2, ell, [orient[r][:ell] for r in P]))
        sg, pg = {}, {}
        for r in S:
            sg.setdefault(orient[r][t + 1:], []).append(r)
        for r in P:
            pg.setdefault(orient[r][:ell], []).append(r)
        used_s, used_p = set(), set()
        for w in set(sg) & set(pg):
            for a in sg[w]:
                for b in pg[w]:
                    if a >> 1 != b >> 1:
                        used_s.add(a)
                        used_p.add(b)
        S -= used_s
        P -= used_p
        t += 1
    return out
