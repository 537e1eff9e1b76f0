"""ctypes bindings to the oracle (oracle/build/libemp_oracle.so). Test infrastructure only."""
import ctypes
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "build", "libemp_oracle.so")
BIN = os.path.join(ORACLE_DIR, "build", "emp_oracle")


class Opts(ctypes.Structure):
    _fields_ = [("min_ovlp", ctypes.c_int), ("map_mem", ctypes.c_uint64), ("reduce_mem", ctypes.c_uint64)]


class Summary(ctypes.Structure):
    _fields_ = [("n_reads", ctypes.c_uint64), ("read_len", ctypes.c_int), ("n_dedup", ctypes.c_uint64),
                ("n_rounds", ctypes.c_int), ("n_records", ctypes.c_uint64), ("n_edges", ctypes.c_uint64)]


class Round(ctypes.Structure):
    _fields_ = [("s", ctypes.c_uint64), ("p", ctypes.c_uint64), ("s_after", ctypes.c_uint64),
                ("p_after", ctypes.c_uint64), ("edges", ctypes.c_uint64), ("ovlp", ctypes.c_int),
                ("records", ctypes.c_int)]


class Audit(ctypes.Structure):
    _fields_ = [("list", ctypes.c_int32), ("ovlp", ctypes.c_int32), ("entries", ctypes.c_uint64),
                ("keys", ctypes.c_uint64), ("strings", ctypes.c_uint64)]


class Buf(ctypes.Structure):
    _fields_ = [("data", ctypes.POINTER(ctypes.c_ubyte)), ("size", ctypes.c_size_t), ("cap", ctypes.c_size_t)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or not os.path.exists(BIN):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        L = ctypes.CDLL(LIB)
        L.eo_key_of.restype = ctypes.c_uint64
        L.eo_key_of.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.eo_key_of_rc.restype = ctypes.c_uint64
        L.eo_key_of_rc.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.eo_next_sfx.restype = ctypes.c_uint64
        L.eo_next_sfx.argtypes = [ctypes.c_uint64, ctypes.c_char, ctypes.c_int]
        L.eo_prev_pfx.restype = ctypes.c_uint64
        L.eo_prev_pfx.argtypes = [ctypes.c_uint64, ctypes.c_char]
        L.eo_modulo_exp.restype = ctypes.c_uint32
        L.eo_modulo_exp.argtypes = [ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint]
        L.eo_modulos.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
        L.eo_run_mem.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(Opts), ctypes.POINTER(Buf),
                                 ctypes.POINTER(Summary), ctypes.POINTER(Round)]
        L.eo_overlap_reads.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_uint64, ctypes.POINTER(Buf), ctypes.POINTER(Summary),
                                       ctypes.POINTER(Round)]
        L.eo_overlap_reads_audit.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_uint64, ctypes.POINTER(Buf), ctypes.POINTER(Summary),
                                             ctypes.POINTER(Round), ctypes.POINTER(Audit), ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int)]
        L.eo_parse_fastq_mem.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_int)]
        L.eo_search_partitions.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_size_t)]
        L.eo_run_file.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Opts), ctypes.POINTER(Summary),
                                  ctypes.POINTER(Round)]
        L.eo_merge_partitions.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t, ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t,
                                          ctypes.POINTER(ctypes.c_size_t)]
        L.eo_cq_sort_by_key.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32),
                                        ctypes.c_size_t, ctypes.c_size_t]
        L.eo_dedup.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t,
                               ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)]
        L.eo_buf_free.argtypes = [ctypes.POINTER(Buf)]
        L.eo_strerror.restype = ctypes.c_char_p
        L.eo_strerror.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def key_of(s: str) -> int:
    return lib().eo_key_of(s.encode(), len(s))


def key_of_rc(s: str) -> int:
    return lib().eo_key_of_rc(s.encode(), len(s))


def next_sfx(key: int, base: str, length: int) -> int:
    return lib().eo_next_sfx(key, base.encode(), length)


def prev_pfx(key: int, base: str) -> int:
    return lib().eo_prev_pfx(key, base.encode())


def modulos():
    m1 = (ctypes.c_uint32 * 256)()
    m2 = (ctypes.c_uint32 * 256)()
    lib().eo_modulos(m1, m2)
    return list(m1), list(m2)


def search_partitions(a, b, M):
    """cq_get_search_partitions over key-sorted lists a, b: (rc, [(na, nb), ...])."""
    A = (ctypes.c_uint64 * max(len(a), 1))(*a)
    B = (ctypes.c_uint64 * max(len(b), 1))(*b)
    cap = 4 + 2 * (len(a) + len(b))
    parts = (ctypes.c_size_t * (2 * cap))()
    n = ctypes.c_size_t()
    rc = lib().eo_search_partitions(A, len(a), B, len(b), M, parts, cap, ctypes.byref(n))
    return rc, [(parts[2 * i], parts[2 * i + 1]) for i in range(n.value)]


def merge_partitions(a, b, M):
    """cq_get_merge_partitions over key-sorted runs a, b: (rc, [(na, nb), ...])."""
    A = (ctypes.c_uint64 * max(len(a), 1))(*a)
    B = (ctypes.c_uint64 * max(len(b), 1))(*b)
    cap = 4 + 4 * (len(a) + len(b))
    parts = (ctypes.c_size_t * (2 * cap))()
    n = ctypes.c_size_t()
    rc = lib().eo_merge_partitions(A, len(a), B, len(b), M, parts, cap, ctypes.byref(n))
    return rc, [(parts[2 * i], parts[2 * i + 1]) for i in range(n.value)]


def cq_sort(keys, vals, M):
    """cq_sort_by_key of (key, value) pairs with device buffer M: (rc, keys, values)."""
    K = (ctypes.c_uint64 * max(len(keys), 1))(*keys)
    V = (ctypes.c_uint32 * max(len(vals), 1))(*vals)
    rc = lib().eo_cq_sort_by_key(K, V, len(keys), M)
    return rc, list(K)[:len(keys)], list(V)[:len(vals)]


def dedup(reads, M):
    """The kept ids of the dedup (key order) over equal-length read strings with buffer M."""
    n, L = len(reads), len(reads[0]) if reads else 0
    ids = (ctypes.c_uint32 * max(2 * n, 1))()
    nd = ctypes.c_uint64()
    rc = lib().eo_dedup("".join(reads).encode(), n, L, M, ids, ctypes.byref(nd))
    return rc, list(ids)[:nd.value]


def parse(fastq: bytes, block_bytes: int):
    """Returns (rc, reads list)."""
    ptr = ctypes.c_void_p()
    n = ctypes.c_uint64()
    L = ctypes.c_int()
    rc = lib().eo_parse_fastq_mem(fastq, len(fastq), block_bytes, ctypes.byref(ptr), ctypes.byref(n),
                                  ctypes.byref(L))
    if rc:
        return rc, None
    data = ctypes.string_at(ptr, n.value * L.value) if n.value else b""
    ctypes.CDLL(None).free(ptr)
    return 0, [data[i * L.value:(i + 1) * L.value].decode() for i in range(n.value)]


def run_fastq(fastq: bytes, min_ovlp=31, map_mem=1 << 30, reduce_mem=1 << 30):
    """Returns (rc, output bytes, summary dict, list of round dicts)."""
    o = Opts(min_ovlp, map_mem, reduce_mem)
    b = Buf()
    s = Summary()
    rr = (Round * 256)()
    rc = lib().eo_run_mem(fastq, len(fastq), ctypes.byref(o), ctypes.byref(b), ctypes.byref(s), rr)
    out = ctypes.string_at(b.data, b.size) if b.size else b""
    lib().eo_buf_free(ctypes.byref(b))
    summ = {f: getattr(s, f) for f, _ in Summary._fields_}
    rounds = [{f: getattr(rr[i], f) for f, _ in Round._fields_} for i in range(min(s.n_rounds, 256))]
    return rc, out, summ, rounds


def run_file(in_path: str, out_path: str, min_ovlp=31, map_mem=1 << 30, reduce_mem=1 << 30):
    """FASTQ file -> output.bin file (the reference's CLI path). Returns (rc, summary dict)."""
    o = Opts(min_ovlp, map_mem, reduce_mem)
    s = Summary()
    rc = lib().eo_run_file(in_path.encode(), out_path.encode(), ctypes.byref(o), ctypes.byref(s), None)
    return rc, {f: getattr(s, f) for f, _ in Summary._fields_}


def run_reads(reads: bytes, n: int, L: int, min_ovlp=31, reduce_mem=1 << 30):
    b = Buf()
    s = Summary()
    rr = (Round * 256)()
    rc = lib().eo_overlap_reads(reads, n, L, min_ovlp, reduce_mem, ctypes.byref(b), ctypes.byref(s), rr)
    out = ctypes.string_at(b.data, b.size) if b.size else b""
    lib().eo_buf_free(ctypes.byref(b))
    summ = {f: getattr(s, f) for f, _ in Summary._fields_}
    rounds = [{f: getattr(rr[i], f) for f, _ in Round._fields_} for i in range(min(s.n_rounds, 256))]
    return rc, out, summ, rounds


def run_reads_audit(reads: bytes, n: int, L: int, min_ovlp=31, reduce_mem=1 << 30):
    """run_reads plus the collision audit: (rc, output bytes, summary, list of audit dicts)."""
    b = Buf()
    s = Summary()
    cap = 2 * 256 + 1
    au = (Audit * cap)()
    na = ctypes.c_int(0)
    rc = lib().eo_overlap_reads_audit(reads, n, L, min_ovlp, reduce_mem, ctypes.byref(b), ctypes.byref(s), None,
                                      au, cap, ctypes.byref(na))
    out = ctypes.string_at(b.data, b.size) if b.size else b""
    lib().eo_buf_free(ctypes.byref(b))
    summ = {f: getattr(s, f) for f, _ in Summary._fields_}
    return rc, out, summ, [{f: getattr(au[i], f) for f, _ in Audit._fields_} for i in range(na.value)]
