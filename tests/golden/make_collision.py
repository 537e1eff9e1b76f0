"""Writes collision_pair.json: two different 24-base strings with the same dual Rabin-Karp key.

The key of a string s of length n is (H1(s), H2(s)) with H_j(s) = sum e(s_i) R_j^(n-1-i) mod P_j,
e(A,C,T,G) = (1,2,3,4), (R1,P1) = (5, 2^29-3), (R2,P2) = (7, 2^29-33). Two strings collide when
their difference d_i = e(s_i) - e(s'_i), each in [-3, 3], satisfies
sum d_i R_j^(n-1-i) = 0 mod P_j for both j. Such a d is a short vector of the lattice spanned by
the rows (unit vector i | W R1^(n-1-i) | W R2^(n-1-i)) and (0 | W P1 | 0), (0 | 0 | W P2): the
weight W forces the last two coordinates of any short vector to zero, and LLL reduction
(Lenstra, Lenstra, Lovasz 1982) finds one. The result is checked exactly before it is written.
The pair differs in its first and last base, so no proper prefix or suffix of it collides.
Run from the repository root: python3 tests/golden/make_collision.py
"""
import json
import os
import random

import numpy as np

P1, P2, R1, R2 = 536870909, 536870879, 5, 7
ENC = {"A": 1, "C": 2, "T": 3, "G": 4}
BASE = {v: k for k, v in ENC.items()}


def lll(rows, delta=0.99):
    """Textbook LLL on integer rows; Gram-Schmidt in float64 (the result is verified exactly)."""
    B = [list(r) for r in rows]
    n = len(B)

    def gs():
        Bf = np.array(B, dtype=np.float64)
        Q = np.zeros_like(Bf)
        mu = np.zeros((n, n))
        nrm = np.zeros(n)
        for i in range(n):
            v = Bf[i].copy()
            for j in range(i):
                mu[i, j] = Bf[i] @ Q[j] / nrm[j]
                v -= mu[i, j] * Q[j]
            Q[i] = v
            nrm[i] = v @ v
        return mu, nrm

    mu, nrm = gs()
    k = 1
    while k < n:
        for j in range(k - 1, -1, -1):
            q = int(round(mu[k, j]))
            if q:
                B[k] = [a - q * b for a, b in zip(B[k], B[j])]
                mu, nrm = gs()
        if nrm[k] >= (delta - mu[k, k - 1] ** 2) * nrm[k - 1]:
            k += 1
        else:
            B[k], B[k - 1] = B[k - 1], B[k]
            mu, nrm = gs()
            k = max(k - 1, 1)
    return B


def key(s):
    n = len(s)
    h1 = sum(ENC[c] * pow(R1, n - 1 - i, P1) for i, c in enumerate(s)) % P1
    h2 = sum(ENC[c] * pow(R2, n - 1 - i, P2) for i, c in enumerate(s)) % P2
    return (h1 << 32) | h2


def main(n=24, W=1 << 16, seed=2024):
    rows = []
    for i in range(n):
        r = [0] * (n + 2)
        r[i] = 1
        r[n] = W * pow(R1, n - 1 - i, P1)
        r[n + 1] = W * pow(R2, n - 1 - i, P2)
        rows.append(r)
    rows.append([0] * n + [W * P1, 0])
    rows.append([0] * n + [0, W * P2])
    # d[0] and d[n-1] non-zero: the pair differs in its first and last base, so no shorter
    # prefix or suffix of the pair collides as well
    cands = [r[:n] for r in lll(rows) if r[n] == 0 and r[n + 1] == 0 and r[0] != 0 and r[n - 1] != 0
             and max(abs(x) for x in r[:n]) <= 3]
    d = min(cands, key=lambda v: sum(x * x for x in v))
    rng = random.Random(seed)
    a, b = [], []
    for x in d:  # e(a_i) - e(b_i) = x with both codes in 1..4
        lo = rng.randint(1, 4 - abs(x))
        hi = lo + abs(x)
        a.append(BASE[hi if x >= 0 else lo])
        b.append(BASE[lo if x >= 0 else hi])
    a, b = "".join(a), "".join(b)
    assert a != b and key(a) == key(b)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "collision_pair.json")
    with open(out, "w") as f:
        json.dump({"a": a, "b": b, "key": key(a), "difference": d}, f)
        f.write("\n")
    print(a, b, hex(key(a)))


if __name__ == "__main__":
    main()
