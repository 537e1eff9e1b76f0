"""The out-of-core order's host planner (emp_stage_ooc_ties, CPU only) against the oracle's
cq_sort_by_key (cqutils.cuh:179-247): the entries sorted by (key, tie) must come out in the
oracle's order, id for id, on lists with many equal keys spread over many chunks, where the
merge partitions (cqutils.cuh:36-84) reorder equal keys. The planner is the product's own code
(the overlap phase runs it on lists read back from the device); only the device kernels are
left to the GPU tests."""
import numpy as np
import pytest

from empanada_amd import native
from tests import oracle_lib as ol


def _keys(rng, n, distinct, bucket_bits):
    """n 58-bit keys drawn from `distinct` values; their high halves share few buckets."""
    hi = rng.integers(0, 1 << 29, size=distinct, dtype=np.uint64)
    if bucket_bits > 1:  # crowd the values into two buckets so that groups hold many keys
        top = rng.integers(0, 2, size=distinct, dtype=np.uint64) << np.uint64(29 - bucket_bits)
        hi = (hi & np.uint64((1 << (29 - bucket_bits)) - 1)) | top
    lo = rng.integers(0, 1 << 29, size=distinct, dtype=np.uint64)
    vals = (hi << np.uint64(32)) | lo
    return vals[rng.integers(0, distinct, size=n)]


def _order(keys, rids, ties):
    o = np.lexsort((ties, keys))
    return [int(x) for x in np.asarray(rids)[o]]


def _oracle_order(keys_in, rids_in, M):
    rc, _, vals = ol.cq_sort([int(k) for k in keys_in], [int(r) for r in rids_in], M)
    assert rc == 0
    return vals


CASES = [  # (n, M, distinct keys, bucket bits)
    (40, 4, 3, 1), (41, 5, 4, 1), (64, 8, 2, 3), (100, 7, 5, 8), (257, 16, 9, 8), (500, 30, 6, 4),
    (999, 64, 40, 8), (1200, 100, 3, 12), (2000, 2, 50, 8), (2000, 3, 7, 2), (3000, 256, 20, 16),
    (4000, 1000, 5, 8), (4097, 4096, 2, 8),
]


@pytest.mark.parametrize("n,M,distinct,bb", CASES)
@pytest.mark.parametrize("seed", [1, 2])
def test_dedup_order_matches_cq_sort(n, M, distinct, bb, seed):
    """The dedup's list: ids 0..n-1 in input order (overlap.cu:212-213)."""
    rng = np.random.default_rng(seed * 1000 + n)
    keys = _keys(rng, n, distinct, bb)
    rids = np.arange(n, dtype=np.uint32)
    ties, _ = native.stage_ooc_ties(keys, rids, M, False, bb)
    assert _order(keys, rids, ties) == _oracle_order(keys, rids, M)


@pytest.mark.parametrize("n,M,distinct,bb", CASES)
@pytest.mark.parametrize("seed", [3, 4])
def test_round_order_matches_cq_sort(n, M, distinct, bb, seed):
    """A round's list: distinct ids, input order = id order cut into chunks of M, each chunk's
    even ids then its odd ids (regenerate_fingerprints, overlap.cu:613-659)."""
    rng = np.random.default_rng(seed * 1000 + n)
    rids = np.sort(rng.choice(3 * n, size=n, replace=False)).astype(np.uint32)
    keys = _keys(rng, n, distinct, bb)
    k_in, r_in = [], []
    for lo in range(0, n, M):
        for parity in (0, 1):
            for i in range(lo, min(n, lo + M)):
                if rids[i] & 1 == parity:
                    k_in.append(keys[i])
                    r_in.append(rids[i])
    perm = rng.permutation(n)  # the planner takes the entries in any order
    ties, _ = native.stage_ooc_ties(keys[perm], rids[perm], M, True, bb)
    assert _order(keys[perm], rids[perm], ties) == _oracle_order(k_in, r_in, M)


def test_case2_puts_b_first():
    """Hand-worked: M = 4 (windows of 2), chunks [k2 k1 k1 k1] [k1 k0 k1 k1]. Sorted runs
    A = [k1(1) k1(2) k1(3) k2(0)], B = [k0(5) k1(4) k1(6) k1(7)] (ids in brackets). Partition 1:
    a's window ends k1, b's first is k0: not case 1; b's window ends k1 <= a's first k1: case 2,
    B's [k0 k1(4)] alone. Partition 2: a's window [k1(1) k1(2)] ends k1 <= b's next k1: case 1.
    Then both windows are full again: a = [k1(3) k2(0)], b = [k1(6) k1(7)]: la = k2 > fb = k1,
    lb = k1 <= fa = k1: case 2, B's window alone. The rest: A's [k1(3) k2(0)].
    Order: 5 4 1 2 6 7 3 0, where the default tie order would put 1 2 3 before 4 6 7."""
    k0, k1, k2 = (np.uint64(v << 32 | 9) for v in (1, 2, 3))
    keys = np.array([k2, k1, k1, k1, k1, k0, k1, k1], dtype=np.uint64)
    rids = np.arange(8, dtype=np.uint32)
    ties, _ = native.stage_ooc_ties(keys, rids, 4, False, 8)
    assert _order(keys, rids, ties) == [5, 4, 1, 2, 6, 7, 3, 0]
    assert _oracle_order(keys, rids, 4) == [5, 4, 1, 2, 6, 7, 3, 0]


def test_in_core_ties_are_the_input_order():
    rng = np.random.default_rng(5)
    keys = _keys(rng, 64, 4, 8)
    rids = rng.choice(200, size=64, replace=False).astype(np.uint32)
    ties, _ = native.stage_ooc_ties(keys, rids, 64, True, 8)
    assert [int(t) for t in ties] == [int(((r & 1) << 31) | (r >> 1)) for r in rids]
    ties, _ = native.stage_ooc_ties(keys, np.arange(64, dtype=np.uint32), 1 << 20, False, 8)
    assert [int(t) for t in ties] == list(range(64))


def test_unique_windows():
    """cq_get_partitions (cqutils.cuh:86-105) over the merged dedup list: a window of M entries
    holding one key that is not the last window never ends."""
    ka, kb = np.uint64(5 << 32), np.uint64(7 << 32)
    # merged: [a a a a b b]; M = 4: the first window is all a and 6 - 0 > 4: stuck
    keys = np.array([ka, ka, kb, ka, kb, ka], dtype=np.uint64)
    _, rc = native.stage_ooc_ties(keys, np.arange(6, dtype=np.uint32), 4, False, 8)
    assert rc == native.ERRORS_BY_NAME["EMP_ERR_PARTITION"]
    # merged: [a a a b b b b]; M = 4: window [a a a b] cuts at 3, then [b b b b] is the last
    keys = np.array([kb, ka, kb, ka, kb, ka, kb], dtype=np.uint64)
    _, rc = native.stage_ooc_ties(keys, np.arange(7, dtype=np.uint32), 4, False, 8)
    assert rc == 0
    # merged: [a a a b b b b b]; M = 4: [a a a b] -> 3, then [b b b b] with 5 left: stuck
    keys = np.array([kb, ka, kb, ka, kb, ka, kb, kb], dtype=np.uint64)
    _, rc = native.stage_ooc_ties(keys, np.arange(8, dtype=np.uint32), 4, False, 8)
    assert rc == native.ERRORS_BY_NAME["EMP_ERR_PARTITION"]


def test_too_many_chunks():
    keys = np.zeros(65536 * 2 + 1, dtype=np.uint64)
    rids = np.arange(keys.size, dtype=np.uint32)
    with pytest.raises(native.EmpError) as e:
        native.stage_ooc_ties(keys, rids, 2, False, 8)
    assert e.value.code == native.ERRORS_BY_NAME["EMP_ERR_OUT_OF_CORE"]
