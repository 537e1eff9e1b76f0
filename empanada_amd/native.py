"""ctypes bindings to libempanada_hip.so (the C ABI of include/empanada.h and
include/empanada_stages.h). Loading fails loudly when the library is not built: there is no
fallback path."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# EMPANADA_LIB selects another build of the library (a tuning variant); the default is the in-tree one.
LIB_PATH = os.environ.get("EMPANADA_LIB") or os.path.join(_HERE, "libempanada_hip.so")
CLI_PATH = os.path.join(_HERE, "bin", "empanada")

EMP_OK = 0
ERRORS = {
    -1: "EMP_ERR_ARG", -2: "EMP_ERR_HIP", -3: "EMP_ERR_FORMAT", -4: "EMP_ERR_READLEN", -5: "EMP_ERR_BASE",
    -6: "EMP_ERR_OUT_OF_CORE", -7: "EMP_ERR_NOMEM", -8: "EMP_ERR_TOO_MANY", -9: "EMP_ERR_PARTITION",
    -10: "EMP_ERR_IO", -11: "EMP_ERR_CAPACITY",
}
ERRORS_BY_NAME = {v: k for k, v in ERRORS.items()}

# Every symbol the headers in include/ declare.
EXPORTS = [
    "emp_opts_default", "emp_ctx_create", "emp_ctx_destroy", "emp_last_error", "emp_add_fastq",
    "emp_add_fastq_device", "emp_add_fastq_file", "emp_add_reads",
    "emp_add_reads_device", "emp_run", "emp_run_to_file", "emp_reset", "emp_get_summary", "emp_get_round",
    "emp_kernel_count", "emp_kernel_info", "emp_kernel_clear", "emp_set_profile", "emp_stream", "emp_version",
    "emp_set_audit", "emp_audit_count", "emp_get_audit",
    "emp_stage_fingerprints", "emp_stage_columns", "emp_stage_roll", "emp_stage_ooc_ties",
    "emp_stage_fingerprints_noalloc",
    "emp_batch_opts_default", "emp_batch_create", "emp_batch_destroy", "emp_batch_last_error", "emp_batch_submit",
    "emp_batch_submit_fastq", "emp_batch_job_bytes", "emp_batch_step", "emp_batch_run", "emp_batch_job_info",
    "emp_batch_reset", "emp_batch_clock_ms", "emp_batch_kernel_count", "emp_batch_kernel_info",
    "emp_batch_kernel_clear", "emp_batch_set_profile", "emp_batch_stream", "emp_batch_set_tuning",
    "emp_batch_graph_stats",
]


class Opts(ctypes.Structure):
    _fields_ = [("min_ovlp", ctypes.c_int32), ("map_mem", ctypes.c_uint64), ("reduce_mem", ctypes.c_uint64),
                ("base_blocksz", ctypes.c_uint64), ("kv_chunksz", ctypes.c_uint64), ("hbm_budget", ctypes.c_uint64),
                ("device", ctypes.c_int32), ("profile", ctypes.c_int32), ("stream", ctypes.c_void_p)]


class RoundStats(ctypes.Structure):
    _fields_ = [("s", ctypes.c_uint64), ("p", ctypes.c_uint64), ("s_after", ctypes.c_uint64),
                ("p_after", ctypes.c_uint64), ("edges", ctypes.c_uint64), ("ovlp", ctypes.c_int32),
                ("records", ctypes.c_int32), ("bucket_bits", ctypes.c_int32), ("big_buckets", ctypes.c_int32),
                ("buckets", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class Audit(ctypes.Structure):
    _fields_ = [("list", ctypes.c_int32), ("ovlp", ctypes.c_int32), ("entries", ctypes.c_uint64),
                ("keys", ctypes.c_uint64), ("strings", ctypes.c_uint64)]


class Summary(ctypes.Structure):
    _fields_ = [("n_reads", ctypes.c_uint64), ("read_len", ctypes.c_int32), ("n_rounds", ctypes.c_int32),
                ("n_dedup", ctypes.c_uint64), ("n_records", ctypes.c_uint64), ("n_edges", ctypes.c_uint64),
                ("map_ms", ctypes.c_double), ("overlap_ms", ctypes.c_double)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("device", ctypes.c_int32), ("profile", ctypes.c_int32),
                ("stream", ctypes.c_void_p), ("max_active", ctypes.c_uint32), ("no_graph", ctypes.c_int32)]


class BatchClass(ctypes.Structure):
    _fields_ = [("read_len_max", ctypes.c_int32), ("n_max", ctypes.c_uint32), ("fine_load", ctypes.c_uint32),
                ("coarse_load", ctypes.c_uint32)]


MAX_CLASSES = 16


class BatchTuning(ctypes.Structure):
    _fields_ = [("bucket_cap", ctypes.c_uint32), ("n_classes", ctypes.c_uint32),
                ("classes", BatchClass * MAX_CLASSES)]


class JobInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("read_len", ctypes.c_int32),
                ("min_ovlp", ctypes.c_int32), ("rounds", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("n_reads", ctypes.c_uint64), ("n_dedup", ctypes.c_uint64), ("n_edges", ctypes.c_uint64),
                ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double)]


JOB_SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32,
                            ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32))
JOB_DONE = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32)

SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32,
                        ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32))

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libempanada_hip.so is not built (run __graft_entry__.build()); "
                               "the HIP path has no fallback")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i32, cp = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_char_p
        L.emp_opts_default.argtypes = [ctypes.POINTER(Opts)]
        L.emp_ctx_create.argtypes = [ctypes.POINTER(Opts), ctypes.POINTER(vp)]
        L.emp_ctx_destroy.argtypes = [vp]
        L.emp_last_error.restype = cp
        L.emp_last_error.argtypes = [vp]
        L.emp_add_fastq.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
        L.emp_add_fastq_device.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
        L.emp_add_fastq_file.argtypes = [vp, cp]
        L.emp_add_reads.argtypes = [vp, vp, u64, i32]
        L.emp_add_reads_device.argtypes = [vp, vp, u64, i32]
        L.emp_run.argtypes = [vp, SINK, vp]
        L.emp_run_to_file.argtypes = [vp, cp]
        L.emp_reset.argtypes = [vp]
        L.emp_get_summary.argtypes = [vp, ctypes.POINTER(Summary)]
        L.emp_get_round.argtypes = [vp, i32, ctypes.POINTER(RoundStats)]
        L.emp_kernel_count.restype = i32
        L.emp_kernel_count.argtypes = [vp]
        L.emp_kernel_info.argtypes = [vp, i32, ctypes.POINTER(cp), ctypes.POINTER(u64),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.emp_kernel_clear.argtypes = [vp]
        L.emp_set_profile.argtypes = [vp, i32]
        L.emp_set_audit.argtypes = [vp, i32]
        L.emp_audit_count.restype = i32
        L.emp_audit_count.argtypes = [vp]
        L.emp_get_audit.argtypes = [vp, i32, ctypes.POINTER(Audit)]
        L.emp_stream.restype = vp
        L.emp_stream.argtypes = [vp]
        L.emp_version.restype = cp
        L.emp_stage_fingerprints.argtypes = [i32, vp, u64, i32, vp]
        L.emp_stage_columns.argtypes = [i32, vp, u64, i32, vp]
        L.emp_stage_roll.argtypes = [i32, vp, vp, u64, i32, i32, vp]
        L.emp_stage_ooc_ties.argtypes = [vp, vp, u64, u64, i32, i32, vp]
        L.emp_stage_fingerprints_noalloc.argtypes = [vp, u64, i32, vp, vp, vp]
        L.emp_batch_opts_default.argtypes = [ctypes.POINTER(BatchOpts)]
        L.emp_batch_create.argtypes = [ctypes.POINTER(BatchOpts), ctypes.POINTER(vp)]
        L.emp_batch_destroy.argtypes = [vp]
        L.emp_batch_last_error.restype = cp
        L.emp_batch_last_error.argtypes = [vp]
        L.emp_batch_submit.argtypes = [vp, vp, u64, i32, i32, i32, ctypes.POINTER(i32)]
        L.emp_batch_submit_fastq.argtypes = [vp, vp, ctypes.c_size_t, i32, ctypes.POINTER(i32)]
        L.emp_batch_job_bytes.restype = u64
        L.emp_batch_job_bytes.argtypes = [u64, i32]
        L.emp_batch_step.argtypes = [vp, JOB_SINK, JOB_DONE, vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.emp_batch_run.argtypes = [vp, JOB_SINK, JOB_DONE, vp]
        L.emp_batch_job_info.argtypes = [vp, i32, ctypes.POINTER(JobInfo)]
        L.emp_batch_reset.argtypes = [vp]
        L.emp_batch_clock_ms.restype = ctypes.c_double
        L.emp_batch_clock_ms.argtypes = [vp]
        L.emp_batch_kernel_count.restype = i32
        L.emp_batch_kernel_count.argtypes = [vp]
        L.emp_batch_kernel_info.argtypes = [vp, i32, ctypes.POINTER(cp), ctypes.POINTER(u64),
                                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.emp_batch_set_tuning.argtypes = [vp, ctypes.POINTER(BatchTuning)]
        L.emp_batch_graph_stats.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.emp_batch_kernel_clear.argtypes = [vp]
        L.emp_batch_set_profile.argtypes = [vp, i32]
        L.emp_batch_stream.restype = vp
        L.emp_batch_stream.argtypes = [vp]
        _lib = L
    return _lib


class EmpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Context:
    """One overlap job on one GPU (wraps emp_ctx)."""

    def __init__(self, min_ovlp=31, map_mem=1 << 30, reduce_mem=1 << 30, hbm_budget=0, device=0, profile=False,
                 stream=None):
        o = Opts()
        lib().emp_opts_default(ctypes.byref(o))
        o.min_ovlp, o.map_mem, o.reduce_mem = min_ovlp, map_mem, reduce_mem
        o.hbm_budget, o.device, o.profile = hbm_budget, device, 1 if profile else 0
        o.stream = stream
        self._h = ctypes.c_void_p()
        rc = lib().emp_ctx_create(ctypes.byref(o), ctypes.byref(self._h))
        if rc:
            raise EmpError(rc, "emp_ctx_create failed")

    def _check(self, rc):
        if rc:
            raise EmpError(rc, lib().emp_last_error(self._h).decode(errors="replace"))

    def close(self):
        if self._h:
            lib().emp_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        return lib().emp_stream(self._h)

    def add_fastq(self, data: bytes, block: int = None):
        """Feeds FASTQ bytes through the block reader (block = map_mem / 16 by default)."""
        if block is None:
            block = 64 << 20
        off = 0
        buf = (ctypes.c_char * len(data)).from_buffer_copy(data) if data else None
        base = ctypes.addressof(buf) if buf is not None else 0
        while off < len(data):
            got = min(block, len(data) - off)
            final = 1 if got < block or off + got == len(data) else 0
            used = ctypes.c_size_t()
            self._check(lib().emp_add_fastq(self._h, base + off, got, final, ctypes.byref(used)))
            if final:
                break
            off += used.value

    def add_fastq_device(self, ptr: int, size: int, pieces: int = 1):
        """FASTQ bytes already in device memory, parsed in blocks of map_mem / 16 bytes. With pieces > 1
        the bytes are handed over in that many prefixes (each call resumes where the last stopped)."""
        off = 0
        for k in range(pieces):
            end = size if k == pieces - 1 else min(size, max(off, size * (k + 1) // pieces))
            used = ctypes.c_size_t()
            last = 1 if end == size else 0
            self._check(lib().emp_add_fastq_device(self._h, ptr + off, end - off, last, ctypes.byref(used)))
            off += used.value
            if last:
                break

    def add_fastq_file(self, path: str):
        """A whole FASTQ file, streamed by the library (blocks of map_mem / 16 bytes)."""
        self._check(lib().emp_add_fastq_file(self._h, path.encode()))

    def add_reads(self, reads: np.ndarray):
        r = np.ascontiguousarray(reads, dtype=np.uint8)
        n, L = r.shape
        self._check(lib().emp_add_reads(self._h, r.ctypes.data, n, L))

    def add_reads_device(self, ptr: int, n: int, L: int):
        self._check(lib().emp_add_reads_device(self._h, ptr, n, L))

    def run(self, collect=True):
        """Runs the overlap phase. Returns the output.bin bytes (collect=True) or None."""
        chunks = []

        @SINK
        def sink(user, n, ovlp, fr, to):
            if collect:
                chunks.append(np.array([n], dtype=np.uint64).tobytes())
                chunks.append(np.array([ovlp], dtype=np.int32).tobytes())
                if n:
                    chunks.append(np.ctypeslib.as_array(fr, shape=(n,)).tobytes())
                    chunks.append(np.ctypeslib.as_array(to, shape=(n,)).tobytes())
            return 0

        # nothing collected: no sink (the library still copies the records to host memory)
        self._check(lib().emp_run(self._h, sink if collect else SINK(), None))
        return b"".join(chunks) if collect else None

    def run_to_file(self, path: str):
        self._check(lib().emp_run_to_file(self._h, path.encode()))

    def reset(self):
        self._check(lib().emp_reset(self._h))

    def summary(self):
        s = Summary()
        self._check(lib().emp_get_summary(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in Summary._fields_}

    def rounds(self):
        out = []
        i = 0
        while True:
            r = RoundStats()
            if lib().emp_get_round(self._h, i, ctypes.byref(r)):
                break
            out.append({f: getattr(r, f) for f, _ in RoundStats._fields_})
            i += 1
        return out

    def kernels(self):
        out = {}
        for i in range(lib().emp_kernel_count(self._h)):
            name, launches = ctypes.c_char_p(), ctypes.c_uint64()
            ms, by = ctypes.c_double(), ctypes.c_double()
            lib().emp_kernel_info(self._h, i, ctypes.byref(name), ctypes.byref(launches), ctypes.byref(ms),
                                  ctypes.byref(by))
            out[name.value.decode()] = dict(launches=launches.value, ms=ms.value, bytes=by.value)
        return out

    def kernel_clear(self):
        lib().emp_kernel_clear(self._h)

    def set_profile(self, on: bool):
        lib().emp_set_profile(self._h, 1 if on else 0)

    def set_audit(self, on: bool):
        """Collision audit (FPCHECK) for later runs: see audits()."""
        lib().emp_set_audit(self._h, 1 if on else 0)

    def audits(self):
        """Audit records of the last run: list (0 before dedup, 1 suffixes, 2 prefixes), ovlp,
        entries, distinct keys, distinct strings."""
        out = []
        for i in range(lib().emp_audit_count(self._h)):
            a = Audit()
            self._check(lib().emp_get_audit(self._h, i, ctypes.byref(a)))
            out.append({f: getattr(a, f) for f, _ in Audit._fields_})
        return out


def records_bytes(records):
    """output.bin bytes of a list of (n, ovlp, from, to) records (overlap.cu:148-161)."""
    out = []
    for n, ovlp, fr, to in records:
        out.append(np.array([n], dtype=np.uint64).tobytes())
        out.append(np.array([ovlp], dtype=np.int32).tobytes())
        if n:
            out.append(fr.tobytes())
            out.append(to.tobytes())
    return b"".join(out)


class Batch:
    """Many overlap jobs sharing one engine (wraps emp_batch). Each job's output is collected as
    its output.bin bytes."""

    def __init__(self, hbm_budget=0, max_active=0, device=0, profile=False, stream=None, collect=True, graph=True,
                 tuning=None):
        o = BatchOpts()
        lib().emp_batch_opts_default(ctypes.byref(o))
        o.hbm_budget, o.max_active, o.device, o.profile, o.stream = hbm_budget, max_active, device, \
            1 if profile else 0, stream
        o.no_graph = 0 if graph else 1
        self._h = ctypes.c_void_p()
        rc = lib().emp_batch_create(ctypes.byref(o), ctypes.byref(self._h))
        if rc:
            raise EmpError(rc, "emp_batch_create failed")
        self.collect = collect
        self.records = {}      # job -> list of (n, ovlp, from, to)
        self.status = {}       # job -> status of finished jobs
        self.finished = []     # job ids in completion order
        self._keep = {}        # job -> host reads kept alive until admission

        @JOB_SINK
        def sink(user, job, n, ovlp, fr, to):
            if self.collect:
                a = np.ctypeslib.as_array(fr, shape=(n,)).copy() if n else None
                b = np.ctypeslib.as_array(to, shape=(n,)).copy() if n else None
                self.records.setdefault(job, []).append((n, ovlp, a, b))
            return 0

        @JOB_DONE
        def done(user, job, status):
            self.status[job] = status
            self.finished.append(job)
            self._keep.pop(job, None)

        self._sink, self._done = sink, done
        if tuning is not None:
            self.set_tuning(tuning)

    def _check(self, rc):
        if rc:
            raise EmpError(rc, lib().emp_batch_last_error(self._h).decode(errors="replace"))

    def close(self):
        if self._h:
            lib().emp_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        return lib().emp_batch_stream(self._h)

    def submit(self, reads: np.ndarray, min_ovlp: int) -> int:
        r = np.ascontiguousarray(reads, dtype=np.uint8)
        n, L = r.shape if r.ndim == 2 else (0, 0)
        job = ctypes.c_int32()
        self._check(lib().emp_batch_submit(self._h, r.ctypes.data if n else None, n, L, min_ovlp, 0,
                                           ctypes.byref(job)))
        self._keep[job.value] = r
        return job.value

    def submit_device(self, ptr: int, n: int, L: int, min_ovlp: int) -> int:
        job = ctypes.c_int32()
        self._check(lib().emp_batch_submit(self._h, ptr, n, L, min_ovlp, 1, ctypes.byref(job)))
        return job.value

    def submit_fastq(self, data: bytes, min_ovlp: int) -> int:
        job = ctypes.c_int32()
        self._check(lib().emp_batch_submit_fastq(self._h, data, len(data), min_ovlp, ctypes.byref(job)))
        return job.value

    def _sink_arg(self):
        # nothing collected: no sink (the engine still copies each step's records to host memory)
        return self._sink if self.collect else JOB_SINK()

    def step(self):
        """One step; returns (active, queued)."""
        a, q = ctypes.c_int32(), ctypes.c_int32()
        self._check(lib().emp_batch_step(self._h, self._sink_arg(), self._done, None, ctypes.byref(a),
                                         ctypes.byref(q)))
        return a.value, q.value

    def run(self):
        self._check(lib().emp_batch_run(self._h, self._sink_arg(), self._done, None))

    def output(self, job: int) -> bytes:
        return records_bytes(self.records.get(job, []))

    def info(self, job: int):
        i = JobInfo()
        self._check(lib().emp_batch_job_info(self._h, job, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in JobInfo._fields_ if f != "pad"}

    def clock_ms(self) -> float:
        return lib().emp_batch_clock_ms(self._h)

    def reset(self):
        self._check(lib().emp_batch_reset(self._h))
        self.records.clear()
        self.status.clear()
        self.finished.clear()

    def kernels(self):
        out = {}
        for i in range(lib().emp_batch_kernel_count(self._h)):
            name, launches = ctypes.c_char_p(), ctypes.c_uint64()
            ms, by = ctypes.c_double(), ctypes.c_double()
            lib().emp_batch_kernel_info(self._h, i, ctypes.byref(name), ctypes.byref(launches), ctypes.byref(ms),
                                        ctypes.byref(by))
            out[name.value.decode()] = dict(launches=launches.value, ms=ms.value, bytes=by.value)
        return out

    def kernel_clear(self):
        lib().emp_batch_kernel_clear(self._h)

    def set_tuning(self, tuning):
        """tuning: None (defaults) or dict(bucket_cap=512|1024, classes=[dict(read_len_max, n_max,
        fine_load, coarse_load), ...]), as empanada_amd/tune.py writes it."""
        if tuning is None:
            self._check(lib().emp_batch_set_tuning(self._h, None))
            return
        t = BatchTuning()
        t.bucket_cap = int(tuning.get("bucket_cap", 512))
        cls = tuning.get("classes", [])
        if len(cls) > MAX_CLASSES:
            raise ValueError("at most %d tuning classes" % MAX_CLASSES)
        t.n_classes = len(cls)
        for i, c in enumerate(cls):
            t.classes[i] = BatchClass(int(c["read_len_max"]), int(c["n_max"]), int(c["fine_load"]),
                                      int(c["coarse_load"]))
        self._check(lib().emp_batch_set_tuning(self._h, ctypes.byref(t)))

    def graph_stats(self):
        """(graphs captured, round steps replayed from a graph)."""
        c, r = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(lib().emp_batch_graph_stats(self._h, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def set_profile(self, on: bool):
        lib().emp_batch_set_profile(self._h, 1 if on else 0)


def job_bytes(n: int, L: int) -> int:
    return lib().emp_batch_job_bytes(n, L)


def version() -> str:
    return lib().emp_version().decode()


def stage_fingerprints(reads: np.ndarray, device=0) -> np.ndarray:
    r = np.ascontiguousarray(reads, dtype=np.uint8)
    n, L = r.shape
    out = np.empty(2 * n, dtype=np.uint64)
    rc = lib().emp_stage_fingerprints(device, r.ctypes.data, n, L, out.ctypes.data)
    if rc:
        raise EmpError(rc, "emp_stage_fingerprints")
    return out


def stage_columns(reads: np.ndarray, device=0) -> np.ndarray:
    """The column store (chunks of 2^20 reads, groups of 8 columns, 2 reads x 8 columns per word) as
    a (chunks, ceil(L/8), 2^19) uint32 array."""
    r = np.ascontiguousarray(reads, dtype=np.uint8)
    n, L = r.shape
    out = np.empty(((n + (1 << 20) - 1) >> 20, (L + 7) // 8, 1 << 19), dtype=np.uint32)
    rc = lib().emp_stage_columns(device, r.ctypes.data, n, L, out.ctypes.data)
    if rc:
        raise EmpError(rc, "emp_stage_columns")
    return out


def stage_ooc_ties(keys: np.ndarray, rids: np.ndarray, M: int, rounds: bool, bucket_bits: int):
    """Tie keys of the out-of-core order (host planner, no device work): (ties, status), status
    EMP_OK or EMP_ERR_PARTITION (the dedup's unique would not end; ties are still written)."""
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    r = np.ascontiguousarray(rids, dtype=np.uint32)
    if k.shape != r.shape:
        raise ValueError("keys and rids differ in shape")
    out = np.empty(k.size, dtype=np.uint64)
    rc = lib().emp_stage_ooc_ties(k.ctypes.data, r.ctypes.data, k.size, M, 1 if rounds else 0, bucket_bits,
                                  out.ctypes.data)
    if rc not in (EMP_OK, -9):
        raise EmpError(rc, "emp_stage_ooc_ties")
    return out, rc


def stage_roll(keys: np.ndarray, bases: np.ndarray, kind: int, length: int, device=0) -> np.ndarray:
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    b = np.ascontiguousarray(bases, dtype=np.uint8)
    out = np.empty_like(k)
    rc = lib().emp_stage_roll(device, k.ctypes.data, b.ctypes.data, k.size, kind, length, out.ctypes.data)
    if rc:
        raise EmpError(rc, "emp_stage_roll")
    return out
