"""Synthetic genomes, reads and FASTQ for the overlap path (bindings to libempgen.so).

The workloads, with their seeds (genome, reads):

- ``small``     1 Mbp i.i.d. genome; 20,000 x 100 bp forward, error-free reads; -o 31. (1000, 1001)
- ``incore``    50 Mbp; 15,000,000 x 100 bp, 50 % reverse-complemented, 0.5 % substitutions; -o 63.
                (1100, 1101)
- ``ooc``       3.1 Gbp with 10 % repeats (300 bp and 6 kbp families, 16 copies, 1 % divergence);
                150 bp reads at 20x, both strands, 0.5 % substitutions; -o 75. (1200, 1201)
- ``batched``   4,096 read sets; set i: genome log-uniform over [5 kbp, 5 Mbp], L from {100,150,250},
                coverage uniform over [10, 50], 50 % RC, 0.5 % substitutions; -o 31.
                (1300 + i, 1300 + i + 4096)
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libempgen.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("libempgen.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(_LIB_PATH)
        L.empgen_genome.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]
        L.empgen_reads.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.empgen_add_repeats.restype = ctypes.c_uint64
        L.empgen_add_repeats.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double,
                                         ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, ctypes.c_void_p,
                                         ctypes.c_uint64, ctypes.c_uint64]
        L.empgen_fastq_size.restype = ctypes.c_uint64
        L.empgen_fastq_size.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.empgen_fastq.restype = ctypes.c_uint64
        L.empgen_fastq.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p]
        L.empgen_write_fastq.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int]
        _lib = L
    return _lib


def genome(G: int, seed: int) -> np.ndarray:
    g = np.empty(G, dtype=np.uint8)
    _load().empgen_genome(seed, G, g.ctypes.data)
    return g


def add_repeats(g: np.ndarray, seed: int, frac: float, len_f: int, copies: int, div: float,
                used: np.ndarray, unit: int) -> int:
    """Repeat copies of length len_f on free slots of `used` (one byte per `unit` bases, shared by
    every family size: copies never overlap)."""
    return _load().empgen_add_repeats(seed, g.ctypes.data, g.size, frac, len_f, copies, div, used.ctypes.data,
                                      used.size, unit)


def reads(g: np.ndarray, n: int, L: int, p_rc: float, p_sub: float, seed: int, out: np.ndarray = None) -> np.ndarray:
    """Row-major (n, L) uint8 array of ASCII reads (into `out`, a C-contiguous (n, L) uint8 array, when given)."""
    r = np.empty((n, L), dtype=np.uint8) if out is None else out
    assert r.shape == (n, L) and r.dtype == np.uint8 and r.flags["C_CONTIGUOUS"]
    _load().empgen_reads(seed, g.ctypes.data, g.size, n, L, p_rc, p_sub, r.ctypes.data)
    return r


def fastq_bytes(r: np.ndarray) -> bytes:
    n, L = r.shape
    size = _load().empgen_fastq_size(n, L)
    buf = np.empty(size, dtype=np.uint8)
    w = _load().empgen_fastq(np.ascontiguousarray(r).ctypes.data, n, L, buf.ctypes.data)
    assert w == size
    return buf.tobytes()


def write_fastq(path: str, r: np.ndarray) -> None:
    n, L = r.shape
    if _load().empgen_write_fastq(path.encode(), np.ascontiguousarray(r).ctypes.data, n, L) != 0:
        raise OSError("could not write " + path)


WORKLOADS = {
    "small": dict(G=10**6, n=20000, L=100, p_rc=0.0, p_sub=0.0, o=31, seeds=(1000, 1001)),
    "incore": dict(G=5 * 10**7, n=15_000_000, L=100, p_rc=0.5, p_sub=0.005, o=63, seeds=(1100, 1101)),
    # 10 % repeats: 5 % in 300 bp families and 5 % in 6 kbp families, 16 copies each, 1 % per-copy
    # divergence (repeat seeds 1207 and 1208); 20x coverage
    "ooc": dict(G=3_100_000_000, n=413_333_333, L=150, p_rc=0.5, p_sub=0.005, o=75, seeds=(1200, 1201),
                repeats=((0.05, 300, 16, 0.01, 1207), (0.05, 6000, 16, 0.01, 1208))),
}


_M64 = (1 << 64) - 1


def _splitmix64(state: int):
    """One SplitMix64 step: (new state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & _M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return state, z ^ (z >> 31)


def batched_set_params(i: int, base: int = 1300, n_sets: int = 4096):
    """Parameters of read set i of the batched workload: genome length log-uniform over
    [5 kbp, 5 Mbp], read length from {100, 150, 250}, coverage uniform over [10, 50], all drawn by
    SplitMix64 from seed base + i. Returns dict(G, L, n, cov, o, seeds)."""
    st = base + i
    st, a = _splitmix64(st)
    st, b = _splitmix64(st)
    st, c = _splitmix64(st)
    u = (a >> 11) * 2.0 ** -53
    G = int(round(5e3 * 1000.0 ** u))
    L = (100, 150, 250)[b % 3]
    cov = 10.0 + 40.0 * ((c >> 11) * 2.0 ** -53)
    n = max(1, int(round(G * cov / L)))
    return dict(G=G, L=L, n=n, cov=cov, o=31, seeds=(base + i, base + i + n_sets))


def batched_set(i: int, base: int = 1300, n_sets: int = 4096):
    """Reads of set i of the batched workload (50 % reverse-complemented, 0.5 % substitutions).
    Returns (reads, min_ovlp)."""
    p = batched_set_params(i, base, n_sets)
    g = genome(p["G"], p["seeds"][0])
    return reads(g, p["n"], p["L"], 0.5, 0.005, p["seeds"][1]), p["o"]


def workload_size(name: str, scale: float = 1.0):
    """(reads, read length) of a named workload at `scale`."""
    w = WORKLOADS[name]
    return max(int(w["n"] * scale), 1), w["L"]


def workload_reads(name: str, scale: float = 1.0, out: np.ndarray = None):
    """Reads of a named workload; scale < 1 shrinks genome and read count together (same
    coverage, same per-base work), for bounded samples. Returns (reads, min_ovlp); the reads go
    into `out` when given (shape workload_size(name, scale))."""
    w = WORKLOADS[name]
    G = max(int(w["G"] * scale), 10 * w["L"])
    n = max(int(w["n"] * scale), 1)
    g = genome(G, w["seeds"][0])
    fams = w.get("repeats", ())
    if fams:
        # one occupancy map for every family size (cells of their common divisor), the longest
        # families placed first: copies of different sizes never overlap
        unit = math.gcd(*[f[1] for f in fams])
        used = np.zeros(G // unit, dtype=np.uint8)
        for frac, len_f, copies, div, seed in sorted(fams, key=lambda f: -f[1]):
            add_repeats(g, seed, frac, len_f, copies, div, used, unit)
    return reads(g, n, w["L"], w["p_rc"], w["p_sub"], w["seeds"][1], out=out), w["o"]
