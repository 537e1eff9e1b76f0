// Host side of libempanada_hip: contexts, staging, the round loop and the C ABI.
//
// Reference map: main() (main.cu:9-77) drives read_block / transpose / fingerprint / queue pushes
// per FASTQ block, then init_match + build_graph. Here a context replaces the globals and
// singletons (Config, Kmp/Vmp, M, d_buf: config.h:112-115, mempool.cuh:89-91, overlap.cu:41-43),
// the host CQueues become HBM-resident bucketed lists, and every Thrust stage is a kernel of
// emp_kernels.hip.
#include <errno.h>
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/uio.h>
#include <sys/vfs.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/empanada.h"
#include "../../include/empanada_stages.h"
#include "emp_fastq.h"
#include "emp_fastq_gpu.h"
#include "emp_hostcopy.h"
#include "emp_kernels.h"
#include "emp_ooc.h"
#include "emp_ooc_plan.h"
#include "emp_outfile.h"

using namespace emp;

namespace {

constexpr uint32_t STAGE_READS = 1u << 20;  // reads per host->device batch (multiple of 16)
constexpr uint32_t BIG_CAP = 1u << 16;      // bucket ids recorded for the large-bucket path

// Counters of one round's lists, filled by the launch that writes the round's bucket set (the
// dedup for round 0, round t-1's join for round t).
// One cache line per round: every work-group of a join adds to the next round's totals while it
// reads the joined round's (the bucket sizes), and a line shared by the two slows both down.
struct alignas(128) ListCtr {
    uint32_t total[2];  // entries per side
    uint32_t spill[2];  // records past a bucket's capacity, per side
    uint32_t n_big;     // buckets the writing launch handed to the large-bucket path
    uint32_t nx_ovf;    // a spill area overflowed
    uint32_t pad[26];
};
constexpr int MAX_ROUNDS = 256;  // L - o <= 255

struct DevCounters {
    unsigned long long pool[SUBS];  // edge-pool cursor per sub-range
    unsigned long long pool_ovf;    // cursor of the pool's overflow region
    unsigned long long pool_lost;   // edges of sub-range reservations abandoned (past the end)
    alignas(128) uint32_t abort;    // the dedup partition overflowed an intermediate bucket or its spill
    uint32_t spill;                 // records spilled by the dedup partition's final pass
    uint32_t in;                    // dedup partition input size, set by the host
    uint32_t too_many;              // a large bucket held 2^31 edges or more
    ListCtr r[MAX_ROUNDS + 1];      // r[t]: round t's lists
    // pipelined rounds stream their edges out: output position after the rounds copied so far,
    // and per round {edges, first position}
    alignas(128) unsigned long long obase;
    unsigned long long orec[MAX_ROUNDS][3];
};

double now_ms() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

uint32_t modpow(uint64_t a, uint64_t e, uint32_t p) {
    uint64_t r = 1;
    a %= p;
    while (e) {
        if (e & 1) r = r * a % p;
        a = a * a % p;
        e >>= 1;
    }
    return (uint32_t)r;
}

// Partition fan-out: smallest bits with ceil(X / 2^bits) <= BUCKET_TARGET.
int bucket_bits(uint64_t x) {
    int b = 0;
    while (b < 24 && ((x + (1ull << b) - 1) >> b) > BUCKET_TARGET) ++b;
    return b;
}

// Bucket bits of a round's bucket set: at least WIDE_MIN_BITS, so that h1's bits 21..28 of every
// entry are implied by its bucket and free to carry codes.
int round_bits(uint64_t x) { return std::max(WIDE_MIN_BITS, bucket_bits(x)); }

// Bucket count of a pipelined round's set: the smallest multiple of 2^WIDE_MIN_BITS with
// ceil(X / count) <= BUCKET_TARGET (key h in bucket floor(h count / 2^29); a multiple of 256
// keeps h1's bits 21..28 implied by the bucket). Powers of two alone leave a list just above one
// at half the target per bucket.
uint32_t round_buckets(uint64_t x) {
    const uint64_t unit = 1ull << WIDE_MIN_BITS;
    const uint64_t need = (x + BUCKET_TARGET - 1) / BUCKET_TARGET;
    return (uint32_t)std::max<uint64_t>(unit, (need + unit - 1) / unit * unit);
}

int ceil_log2(uint32_t x) {
    int b = 0;
    while ((1ull << b) < x) ++b;
    return b;
}

struct KStat {
    std::string name;
    uint64_t launches = 0;
    double ms = 0, bytes = 0;
};

struct Pending {
    int k;
    hipEvent_t a, b;
    double bytes;
};

}  // namespace

struct emp_ctx {
    emp_opts opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    uint64_t bytes = 0;
    std::vector<std::pair<void *, size_t>> allocs;

    // reads / columns / entries
    int L = 0;
    uint64_t n = 0;  // reads mapped (submitted to the device)
    uint32_t *d_cols = nullptr;  // column store (grouped layout of emp_device.h)
    size_t cols_chunks = 0;      // chunks allocated at the current read length
    size_t cols_bytes = 0;       // bytes allocated (kept when the read length changes)
    Entry *d_entries = nullptr;
    uint64_t entries_cap = 0;
    uint32_t *d_ends = nullptr;  // per read: first and last column groups (entries_cap / 2 of them)
    uint32_t *d_err = nullptr;

    // host staging (pinned, double-buffered)
    char *h_stage[2] = {nullptr, nullptr};
    uint8_t *d_batch[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_busy[2] = {false, false};
    uint32_t stage_n = 0;
    int stage_cur = 0;
    size_t stage_bytes = 0;

    // FASTQ blocks parsed on the device: the block, its newline positions and per-record
    // arrays, and the kept reads not yet mapped (fewer than 16 between calls)
    uint8_t *d_fq = nullptr;
    uint64_t fq_cap = 0;
    uint32_t *d_fq_cnt = nullptr;
    unsigned long long *d_fq_choff = nullptr;
    uint64_t fq_chunk_cap = 0, fq_choff_cap = 0;
    uint64_t *d_fq_nl = nullptr;
    uint64_t fq_nl_cap = 0;
    uint32_t *d_fq_kept = nullptr, *d_fq_len = nullptr;
    unsigned long long *d_fq_rank = nullptr;
    uint64_t fq_rec_cap = 0, fq_len_cap = 0, fq_rank_cap = 0;
    unsigned long long *d_fq_blk = nullptr;  // block totals of the parser's scans
    uint64_t fq_blk_cap = 0;
    FqResult *d_fq_res = nullptr, *h_fq_res = nullptr;
    uint8_t *d_acc = nullptr;
    uint64_t acc_bytes = 0, acc_n = 0;
    // kept reads in pinned host memory (emp_add_reads): two device batches they are copied into
    // on the copy stream, while the map kernel runs on the one before
    uint8_t *d_pin[2] = {nullptr, nullptr};
    uint64_t pin_cap[2] = {0, 0};
    // the same batches packed on the host (2-bit codes, a quarter of the bytes over PCIe), expanded
    // into d_pin on the device
    uint8_t *d_pack[2] = {nullptr, nullptr};
    uint64_t pack_cap[2] = {0, 0};
    hipEvent_t pin_ev[2] = {nullptr, nullptr}, pin_done[2] = {nullptr, nullptr};
    // large kept-read arrays in pageable memory: three pinned stages filled by parallel copies
    char *h_big[3] = {nullptr, nullptr, nullptr};
    hipEvent_t big_done[3] = {nullptr, nullptr, nullptr};
    // FASTQ files streamed into HBM (emp_add_fastq_file): the file's bytes, three pinned chunks
    // and the stream that copies them
    uint8_t *d_file = nullptr;
    uint64_t file_cap = 0;
    char *h_chunk[3] = {nullptr, nullptr, nullptr};
    hipStream_t cstream = nullptr;

    // overlap phase. Bucket set `set` holds round t's lists for t % 2 == set, per side (0 =
    // suffixes, 1 = prefixes): the dedup writes set 0, round t's join writes set (t + 1) % 2.
    Entry *d_bfine[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    uint32_t *d_bcnt[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    SpillRec *d_bspill[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    uint32_t bspill_cap[2][2] = {{0, 0}, {0, 0}};
    uint32_t nb_alloc = 0;
    // The dedup partition: intermediate buckets, and final buckets that live in bucket set 1's
    // suffix slots (free until round 0's join writes them), with their counts and spill area.
    Entry *d_inter[2] = {nullptr, nullptr};
    uint32_t *d_inter_cnt[2] = {nullptr, nullptr};
    uint64_t inter_cap = 0;  // entries per intermediate array
    Entry *d_dfine = nullptr;
    uint32_t *d_dcnt = nullptr;
    // the partition's payload (each entry's read's ends word) beside the intermediate and final slots
    uint32_t *d_inter_pay[2] = {nullptr, nullptr};
    uint32_t *d_dpay = nullptr;
    uint64_t dpay_cap = 0;
    SpillRec *d_dspill = nullptr;
    uint32_t dspill_cap = 0;
    double slack = 1.125;
    uint2 *d_pool = nullptr;  // edge pool: (from, to) pairs
    uint32_t *d_out_from = nullptr, *d_out_to = nullptr;
    uint32_t *h_from = nullptr, *h_to = nullptr;
    uint64_t pool_cap = 0;  // edges per sub-range of the pool
    uint64_t pool_ovf_cap = 0;  // edges of the pool's overflow region
    uint64_t out_cap = 0;   // edges the output arena holds
    uint64_t out_used = 0;  // edges in the arena not yet passed to the sink
    struct Record { uint64_t n, off; int32_t ovlp; };
    std::vector<Record> records;
    // per (round, bucket) slot: pool offset and count of the bucket's edges, output offset
    unsigned long long *d_edge_off = nullptr, *d_dst_off = nullptr, *d_eblk = nullptr;
    uint32_t *d_edge_cnt = nullptr;
    uint64_t eslots_cap = 0;
    std::vector<unsigned long long> pool_prev;  // edge-pool cursors after the last completed round
    // large buckets
    uint32_t *d_big = nullptr;
    unsigned long long *d_boff[2] = {nullptr, nullptr};
    uint64_t *d_bkey[2] = {nullptr, nullptr};
    uint32_t *d_brank[2] = {nullptr, nullptr}, *d_baux[2] = {nullptr, nullptr};
    uint64_t bscratch_cap = 0;
    DevCounters *d_ctr = nullptr, *h_ctr = nullptr;
    // the out-of-core order (2n > M): tie keys per id for each side, chunk per id, the id bitmap
    // of a round's list with its word counts and offsets, the planner's count table, the merge
    // tree's in-order indices, group reads, explicit ties, per-bucket scratch offsets and the
    // scratch ties of the global-memory sorts
    unsigned long long *d_tie[2] = {nullptr, nullptr};
    uint16_t *d_chunk = nullptr;
    uint64_t ooc_ids = 0;
    uint32_t *d_obm = nullptr, *d_opc = nullptr;
    unsigned long long *d_obase = nullptr, *d_oblk = nullptr;
    uint32_t *d_obcnt = nullptr, *d_oio = nullptr;
    uint32_t *d_ofids = nullptr;
    unsigned long long *d_ofoff = nullptr;
    uint64_t ofid_cap = 0, ofoff_cap = 0;
    OocRec *d_orec = nullptr;
    uint64_t orec_cap = 0;
    uint32_t *d_ovid = nullptr;
    unsigned long long *d_ovval = nullptr;
    uint64_t ovid_cap = 0, ovval_cap = 0;
    unsigned long long *d_ooff[2] = {nullptr, nullptr};
    uint32_t ooff_nb = 0;
    unsigned long long *d_btie[2] = {nullptr, nullptr};
    uint64_t btie_cap = 0;
    // pipelined rounds: each round's counter record read back (pinned) with an event per round
    bool pipeline = true;
    ListCtr *h_rec = nullptr, *d_rec_view = nullptr;  // pinned, and its device view
    std::vector<hipEvent_t> rec_ev;
    // their edges go straight to the pinned output arena (device view of h_from / h_to) on a
    // second stream, while later rounds run
    hipStream_t ostream = nullptr;
    hipStream_t dstream = nullptr;  // the arena's DMA to the pinned host arena (apart from ostream's scans)
    bool out_streamed = false;
    unsigned long long (*h_orec)[3] = nullptr;  // per round {edges, first position, overflow cursor}, pinned
    unsigned long long (*d_orec_view)[3] = nullptr;  // its device view
    std::vector<hipEvent_t> orec_ev;

    // emp_run_to_file: records written while the rounds run (a writer thread), and how many of
    // the run's records are already in the file
    OutFile *out_file = nullptr;
    uint64_t emitted = 0;
    std::unique_ptr<DirectIO> dio;   // its direct writes' AIO context and bounce buffers
    std::vector<hipEvent_t> dma_ev;  // per round: its edges are in the pinned host arena

    // statistics
    std::vector<emp_round_stats> rounds;
    emp_summary summary{};

    // kernel timing
    std::vector<KStat> kstats;
    // collision audit (emp_set_audit): one record per audited list, and its device counters
    bool audit = false;
    std::vector<emp_audit> audits;
    unsigned long long *d_audit = nullptr;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;
};

namespace {

int fail(emp_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                                   \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return fail((c), EMP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

int dalloc(emp_ctx *c, void **p, size_t sz) {
    if (sz == 0) sz = 256;
    if (c->opts.hbm_budget && c->bytes + sz > c->opts.hbm_budget)
        return fail(c, EMP_ERR_NOMEM, "HBM budget of %llu bytes exceeded (need %llu more)",
                    (unsigned long long)c->opts.hbm_budget, (unsigned long long)sz);
    hipError_t e = hipMalloc(p, sz);
    if (e != hipSuccess) return fail(c, EMP_ERR_NOMEM, "hipMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
    c->bytes += sz;
    c->allocs.push_back({*p, sz});
    return EMP_OK;
}

void dfree(emp_ctx *c, void *p) {
    if (!p) return;
    for (size_t i = 0; i < c->allocs.size(); ++i)
        if (c->allocs[i].first == p) {
            c->bytes -= c->allocs[i].second;
            c->allocs.erase(c->allocs.begin() + (long)i);
            break;
        }
    (void)hipFree(p);
}

int kstat_id(emp_ctx *c, const char *name) {
    for (size_t i = 0; i < c->kstats.size(); ++i)
        if (c->kstats[i].name == name) return (int)i;
    c->kstats.push_back(KStat{name});
    return (int)c->kstats.size() - 1;
}

hipEvent_t get_event(emp_ctx *c) {
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Brackets a launch with events when profiling. `bytes` is the launch's algorithmic traffic.
struct Timed {
    emp_ctx *c;
    int k = -1;
    hipEvent_t a = nullptr;
    double bytes;
    Timed(emp_ctx *c_, const char *name, double bytes_) : c(c_), bytes(bytes_) {
        if (c->opts.profile) {
            k = kstat_id(c, name);
            a = get_event(c);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~Timed() {
        if (k >= 0) {
            hipEvent_t b = get_event(c);
            (void)hipEventRecord(b, c->stream);
            c->pending.push_back(Pending{k, a, b, bytes});
        }
    }
};

// Call after a stream synchronisation: folds finished launch timings into the statistics.
void collect_timings(emp_ctx *c) {
    for (auto &p : c->pending) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        KStat &k = c->kstats[p.k];
        k.launches++;
        k.ms += ms;
        k.bytes += p.bytes;
        c->free_events.push_back(p.a);
        c->free_events.push_back(p.b);
    }
    c->pending.clear();
}

int sync(emp_ctx *c) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    collect_timings(c);
    return EMP_OK;
}

// Column store for n_total reads: one allocation of whole chunks, doubled (and copied) when it
// fills, so that a column gather is a single load.
int ensure_cols(emp_ctx *c, uint64_t n_total) {
    const size_t need = (size_t)((n_total + CHUNK_READS - 1) >> CHUNK_SHIFT);
    if (need <= c->cols_chunks) return EMP_OK;
    const size_t chunk_bytes = (size_t)col_groups8((uint32_t)c->L) * CHUNK_WWORDS * sizeof(uint32_t);
    if (need * chunk_bytes <= c->cols_bytes) {  // a run's first reads: the store kept from a run at another length fits
        c->cols_chunks = c->cols_bytes / chunk_bytes;
        return EMP_OK;
    }
    const size_t cap = std::max(need, c->cols_chunks * 2);
    void *p = nullptr;
    int rc = dalloc(c, &p, cap * chunk_bytes);
    if (rc) return rc;
    if (c->d_cols) {
        if (c->cols_chunks)
            HIPCHK(c, hipMemcpyAsync(p, c->d_cols, c->cols_chunks * chunk_bytes, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dfree(c, c->d_cols);
    }
    c->d_cols = (uint32_t *)p;
    c->cols_chunks = cap;
    c->cols_bytes = cap * chunk_bytes;
    return EMP_OK;
}

int ensure_entries(emp_ctx *c, uint64_t n_entries) {
    if (n_entries <= c->entries_cap) return EMP_OK;
    uint64_t cap = std::max<uint64_t>(c->entries_cap * 2, 1u << 16);
    while (cap < n_entries) cap *= 2;
    void *p = nullptr, *q = nullptr;
    int rc = dalloc(c, &p, cap * sizeof(Entry));
    if (rc) return rc;
    if ((rc = dalloc(c, &q, cap / 2 * sizeof(uint32_t)))) return rc;
    if (c->d_entries) {
        HIPCHK(c, hipMemcpyAsync(p, c->d_entries, 2 * c->n * sizeof(Entry), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(q, c->d_ends, c->n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dfree(c, c->d_entries);
        dfree(c, c->d_ends);
    }
    c->d_entries = (Entry *)p;
    c->d_ends = (uint32_t *)q;
    c->entries_cap = cap;
    return EMP_OK;
}

// A new read length: the column store's layout depends on it, so its chunks are counted again
// (ensure_cols) in the bytes kept; the host stages and their device batches (stage_bytes each) are
// kept too, and grown by stage_read when a longer length needs more.
void free_reads_storage(emp_ctx *c) {
    (void)hipStreamSynchronize(c->stream);
    c->cols_chunks = 0;
}

int set_len(emp_ctx *c, int L) {
    if (L <= 0 || L > MAX_RD_LEN) return fail(c, EMP_ERR_READLEN, "read length %d outside [1, 255]", L);
    if (c->L == L) return EMP_OK;
    if (c->n == 0 && c->stage_n == 0 && c->acc_n == 0) {
        // first read of a run: (re)size the per-length storage
        if (c->L != 0) free_reads_storage(c);
        c->L = L;
        return EMP_OK;
    }
    return fail(c, EMP_ERR_READLEN, "All reads must be of equal length (%d vs %d)", c->L, L);
}

// Runs the map kernel over n reads at device address d (4-byte aligned), appending at c->n.
int map_batch(emp_ctx *c, const uint8_t *d, uint64_t nreads) {
    if (nreads == 0) return EMP_OK;
    if (c->n % 16) return fail(c, EMP_ERR_ARG, "a batch follows one whose size is not a multiple of 16");
    if (2 * (c->n + nreads) >= (1ull << 32)) return fail(c, EMP_ERR_TOO_MANY, "more than 2^31 - 1 reads");
    int rc = ensure_cols(c, c->n + nreads);
    if (rc) return rc;
    rc = ensure_entries(c, 2 * (c->n + nreads));
    if (rc) return rc;
    int32_t Lpad;
    uint32_t rows;
    map_shape(c->L, &Lpad, &rows);
    uint64_t done = 0;
    while (done < nreads) {
        uint32_t m = (uint32_t)std::min<uint64_t>(nreads - done, 1u << 30);
        MapArgs a{};
        a.reads = d + done * (uint64_t)c->L;
        a.n = m;
        a.L = c->L;
        a.Lpad = Lpad;
        a.rows = rows;
        a.k0 = c->n + done;
        a.cols = c->d_cols;
        a.entries = c->d_entries;
        a.ends = c->d_ends;
        a.err = c->d_err;
        {
            Timed t(c, "k_map", (double)m * c->L * 1.25 + 24.0 * m);
            HIPCHK(c, launch_map(a, c->stream));
        }
        done += m;
    }
    c->n += nreads;
    return EMP_OK;
}

int stage_submit(emp_ctx *c) {
    if (c->stage_n == 0) return EMP_OK;
    int s = c->stage_cur;
    size_t bytes = (size_t)c->stage_n * c->L;
    if (!c->d_batch[s]) {
        void *p = nullptr;
        int rc = dalloc(c, &p, c->stage_bytes);
        if (rc) return rc;
        c->d_batch[s] = (uint8_t *)p;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_batch[s], c->h_stage[s], bytes, hipMemcpyHostToDevice, c->stream));
    int rc = map_batch(c, c->d_batch[s], c->stage_n);
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->stage_ev[s], c->stream));
    c->stage_busy[s] = true;
    c->stage_cur ^= 1;
    c->stage_n = 0;
    int t = c->stage_cur;
    if (c->stage_busy[t]) {
        HIPCHK(c, hipEventSynchronize(c->stage_ev[t]));
        c->stage_busy[t] = false;
    }
    return EMP_OK;
}

int acc_to_stage(emp_ctx *c);

int stage_read(emp_ctx *c, const char *seq, size_t len) {
    int rc = acc_to_stage(c);
    if (rc) return rc;
    rc = set_len(c, (int)std::min<size_t>(len, 1u << 20));
    if (rc) return rc;
    if (!c->h_stage[0] || c->stage_bytes < (size_t)STAGE_READS * c->L) {
        // first use, or reads longer than the stages were sized for (only at a run's first read:
        // the previous run's copies are waited for before its stages are freed)
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->stage_bytes = (size_t)STAGE_READS * c->L;
        for (int i = 0; i < 2; ++i) {
            if (c->h_stage[i]) (void)hipHostFree(c->h_stage[i]);
            c->h_stage[i] = nullptr;
            dfree(c, c->d_batch[i]);  // stage_submit allocates it again at stage_bytes
            c->d_batch[i] = nullptr;
            c->stage_busy[i] = false;
            HIPCHK(c, hipHostMalloc((void **)&c->h_stage[i], c->stage_bytes, hipHostMallocDefault));
            if (!c->stage_ev[i]) HIPCHK(c, hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming));
        }
    }
    memcpy(c->h_stage[c->stage_cur] + (size_t)c->stage_n * c->L, seq, len);
    if (++c->stage_n == STAGE_READS) return stage_submit(c);
    return EMP_OK;
}


// Grows a device buffer to at least `need` elements (contents dropped).
template <class T>
int ensure_dev(emp_ctx *c, T **p, uint64_t &cap, uint64_t need) {
    if (need <= cap) return EMP_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, *p);
    *p = nullptr;
    const uint64_t n = std::max<uint64_t>(need + need / 4, 1024);
    void *q = nullptr;
    int rc = dalloc(c, &q, n * sizeof(T));
    if (rc) return rc;
    *p = (T *)q;
    cap = n;
    return EMP_OK;
}

// Grows a device byte buffer to at least `need` bytes, keeping its contents.
int ensure_dev_keep(emp_ctx *c, uint8_t **p, uint64_t &cap, uint64_t need) {
    if (need <= cap) return EMP_OK;
    const uint64_t n = std::max<uint64_t>(need + need / 4, 1 << 20);
    void *q = nullptr;
    int rc = dalloc(c, &q, n);
    if (rc) return rc;
    if (*p && cap) HIPCHK(c, hipMemcpyAsync(q, *p, cap, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, *p);
    *p = (uint8_t *)q;
    cap = n;
    return EMP_OK;
}

// Parsed reads left in the device accumulation go to the host stage, ahead of reads added there
// (they are fewer than 16).
int acc_to_stage(emp_ctx *c) {
    if (c->acc_n == 0) return EMP_OK;
    const uint64_t m = c->acc_n;
    c->acc_n = 0;
    std::vector<char> tmp(m * (uint64_t)c->L);
    HIPCHK(c, hipMemcpyAsync(tmp.data(), c->d_acc, tmp.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint64_t i = 0; i < m; ++i) {
        int rc = stage_read(c, tmp.data() + i * (uint64_t)c->L, (size_t)c->L);
        if (rc) return rc;
    }
    return EMP_OK;
}

int read_result(emp_ctx *c) {
    HIPCHK(c, hipMemcpyAsync(c->h_fq_res, c->d_fq_res, sizeof(FqResult), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EMP_OK;
}

// read_block + the map phase on the device (preprocess.cu:195-237, fingerprint.cu:75-153) over a
// range of FASTQ bytes: the caller's host block (copied to HBM first) or bytes already in HBM. The
// range is parsed in blocks of B bytes, each resumed after the previous one's complete records, as
// the CLI's loop reads a file (main.cu:42-58); with `one_block` it is a single block of any size,
// final when `final` (emp_add_fastq's contract). Otherwise a block shorter than B is the last one
// only when `final` (the range ends the file); else parsing stops before it and *consumed says
// where. Kept reads join the device accumulation, which is mapped in multiples of 16 reads (the
// rest waits for the next range or the run). Three host synchronisations per range.
int add_fastq_range(emp_ctx *c, const char *data, size_t size, uint64_t B, bool one_block, bool final,
                    size_t *consumed, bool on_device) {
    *consumed = 0;
    if (size == 0) return EMP_OK;
    int rc;
    if (c->stage_n) {  // reads staged on the host come first
        const uint64_t b = (uint64_t)c->stage_n * c->L;
        if ((rc = ensure_dev(c, &c->d_acc, c->acc_bytes, c->acc_n * c->L + b))) return rc;
        HIPCHK(c, hipMemcpyAsync(c->d_acc + c->acc_n * c->L, c->h_stage[c->stage_cur], b, hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->acc_n += c->stage_n;
        c->stage_n = 0;
    }
    const uint64_t nch = fq_chunks(size);
    if (!on_device && (rc = ensure_dev(c, &c->d_fq, c->fq_cap, size))) return rc;
    const uint8_t *d = on_device ? reinterpret_cast<const uint8_t *>(data) : c->d_fq;
    if ((rc = ensure_dev(c, &c->d_fq_cnt, c->fq_chunk_cap, nch + 1))) return rc;
    if ((rc = ensure_dev(c, &c->d_fq_choff, c->fq_choff_cap, nch + 1))) return rc;
    if ((rc = ensure_dev(c, &c->d_fq_blk, c->fq_blk_cap, fq_scan_blocks(nch)))) return rc;
    if (!c->d_fq_res) {
        void *q = nullptr;
        if ((rc = dalloc(c, &q, sizeof(FqResult)))) return rc;
        c->d_fq_res = (FqResult *)q;
        HIPCHK(c, hipHostMalloc((void **)&c->h_fq_res, sizeof(FqResult), hipHostMallocDefault));
    }
    FqResult &R = *c->h_fq_res;
    R = FqResult{};
    R.first_bad = R.first_kept = R.first_len = R.first_odd = ~0ull;
    HIPCHK(c, hipMemcpyAsync(c->d_fq_res, c->h_fq_res, sizeof(FqResult), hipMemcpyHostToDevice, c->stream));
    if (!on_device) HIPCHK(c, hipMemcpyAsync(c->d_fq, data, size, hipMemcpyHostToDevice, c->stream));
    {
        Timed t(c, "k_fastq", 1.0 * size);
        HIPCHK(c, launch_fq_count(d, size, c->d_fq_cnt, c->d_fq_choff, c->d_fq_blk, c->stream));
    }
    unsigned long long nl = 0;
    HIPCHK(c, hipMemcpyAsync(&nl, c->d_fq_choff + nch, sizeof nl, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = ensure_dev(c, &c->d_fq_nl, c->fq_nl_cap, nl + 1))) return rc;
    {
        Timed t(c, "k_fastq", 1.0 * size + 8.0 * nl);  // the range again, a position per newline
        HIPCHK(c, launch_fq_frame(d, size, B, final ? 1 : 0, one_block ? 1 : 0, final ? 1 : 0, c->d_fq_choff,
                                  c->d_fq_nl, c->d_fq_res, c->stream));
    }
    if ((rc = read_result(c))) return rc;
    const uint64_t recs = R.records;
    const bool fresh = c->n == 0 && c->stage_n == 0 && c->acc_n == 0;  // the first kept read sets the length
    // Rows written by the classify pass itself: every record a kept read of the expected length
    // (the context's, or record 0's when the first kept read sets it), known once it has run.
    bool rows_done = false;
    if (recs) {
        if ((rc = ensure_dev(c, &c->d_fq_kept, c->fq_rec_cap, recs))) return rc;
        if ((rc = ensure_dev(c, &c->d_fq_len, c->fq_len_cap, recs))) return rc;
        if ((rc = ensure_dev(c, &c->d_fq_rank, c->fq_rank_cap, recs + 1))) return rc;
        if ((rc = ensure_dev(c, &c->d_fq_blk, c->fq_blk_cap, fq_scan_blocks(recs)))) return rc;
        const uint64_t Lt = fresh ? R.len0 : (uint64_t)c->L;
        if (Lt >= 1 && Lt <= (uint64_t)FQ_ROW_MAX) {
            if ((rc = ensure_dev_keep(c, &c->d_acc, c->acc_bytes, (c->acc_n + recs) * Lt))) return rc;
            {
                // two newline positions and the sequence line per record (about a quarter of a
                // record), kept flag and length written, the row written
                Timed t(c, "k_fastq", 0.25 * R.use + 24.0 * recs + (double)Lt * recs);
                HIPCHK(c, launch_fq_classify_copy(d, size, c->d_fq_nl, nl, recs, c->d_fq_res, c->d_fq_kept,
                                                  c->d_fq_len, (uint32_t)Lt, c->d_acc + c->acc_n * Lt, c->stream));
            }
            if ((rc = read_result(c))) return rc;
            if (R.first_odd == ~0ull && R.first_bad == ~0ull) {
                rows_done = true;
                R.kept = recs;
                R.first_kept = 0;
                R.kept_len = (uint32_t)Lt;
            } else {
                {
                    Timed t(c, "k_fastq", 24.0 * recs);
                    HIPCHK(c, launch_fq_rank(recs, c->d_fq_res, c->d_fq_kept, c->d_fq_len, c->d_fq_rank, c->d_fq_blk,
                                             fresh ? 0u : (uint32_t)c->L, c->stream));
                }
                if ((rc = read_result(c))) return rc;
            }
        } else {
            {
                // two newline positions and the sequence line per record (about a quarter of a
                // record), kept flag, length and offset written, each read twice more
                Timed t(c, "k_fastq", 0.25 * R.use + 48.0 * recs);
                HIPCHK(c, launch_fq_classify(d, size, c->d_fq_nl, nl, recs, c->d_fq_res, c->d_fq_kept, c->d_fq_len,
                                             c->d_fq_rank, c->d_fq_blk, fresh ? 0u : (uint32_t)c->L, c->stream));
            }
            if ((rc = read_result(c))) return rc;
        }
    }
    // The first offending record in file order decides the error, as the host reader meets them
    // in order; a block that cannot be framed comes after every record of the blocks before it.
    const uint32_t L = fresh ? R.kept_len : (uint32_t)c->L;
    if (R.first_bad != ~0ull && R.first_bad < R.first_len)
        return fail(c, EMP_ERR_BASE, "a kept read holds a byte other than ACGTacgt");
    if (R.first_len != ~0ull) {
        if (L > MAX_RD_LEN) return fail(c, EMP_ERR_READLEN, "read length %u outside [1, 255]", L);
        return fail(c, EMP_ERR_READLEN, "All reads must be of equal length (%d vs other)", (int)L);
    }
    if (R.err == FQ_ERR_FORMAT)
        return fail(c, EMP_ERR_FORMAT, "malformed FASTQ block %llu: it must start with '@' and hold a complete record",
                    (unsigned long long)R.blocks);
    const uint64_t kept = R.kept;
    if (kept) {
        if ((rc = set_len(c, (int)L))) return rc;
        if ((rc = ensure_dev_keep(c, &c->d_acc, c->acc_bytes, (c->acc_n + kept) * (uint64_t)L))) return rc;
        if (!rows_done) {
            Timed t(c, "k_fastq", 2.0 * kept * L + 24.0 * recs);
            HIPCHK(c, launch_fq_copy(d, size, c->d_fq_nl, nl, recs, c->d_fq_kept, c->d_fq_rank, L,
                                     c->d_acc + c->acc_n * (uint64_t)L, c->stream));
        }
        c->acc_n += kept;
        const uint64_t m = c->acc_n & ~15ull;
        if (m) {
            if ((rc = map_batch(c, c->d_acc, m))) return rc;
            const uint64_t rest = c->acc_n - m;  // fewer than 16 and m >= 16: no overlap
            if (rest)
                HIPCHK(c, hipMemcpyAsync(c->d_acc, c->d_acc + m * (uint64_t)L, rest * (uint64_t)L,
                                         hipMemcpyDeviceToDevice, c->stream));
            c->acc_n = rest;
        }
    }
    *consumed = (size_t)R.use;
    return EMP_OK;
}

// ---------------------------------------------------------------------------------------------
// A FASTQ file streamed into HBM: the CLI's block loop (main.cu:42-58 over read_block,
// preprocess.cu:195-237) with the file read ahead. FILE_READERS reader threads, alive for the whole
// file, each read their share of every chunk (pread) into one of FILE_SLOTS pinned chunks of
// FILE_CHUNK bytes; a copier thread sends each complete chunk to the device image of the file on a
// copy stream and frees its slot once that copy is done. Block j (map_mem/16 bytes from the
// position block j-1 left off) is parsed on the device as soon as the chunks it covers have
// arrived, while later chunks are read and copied. (Page-cache reads reach about 50 GB/s with 4
// threads and 80 with 8, the link about 50: with 8 readers and 3 slots the copies never wait.)
constexpr uint64_t FILE_CHUNK = 64ull << 20;
constexpr int FILE_READERS = 8;
constexpr int FILE_SLOTS = 3;

int add_fastq_file(emp_ctx *c, const char *path) {
    const uint64_t B = c->opts.map_mem / 16;  // main.cu:16
    if (B == 0) return fail(c, EMP_ERR_ARG, "map_mem / 16 is zero");
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(c, EMP_ERR_IO, "Could not open %s", path);
    struct stat stt;
    if (fstat(fd, &stt) != 0) {
        close(fd);
        return fail(c, EMP_ERR_IO, "Could not stat %s", path);
    }
    const uint64_t size = (uint64_t)stt.st_size;
    int rc = EMP_OK;
    if (size == 0) {
        close(fd);
        return EMP_OK;
    }
    if ((rc = ensure_dev(c, &c->d_file, c->file_cap, size))) {
        close(fd);
        return rc;
    }
    if (!c->cstream && hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking) != hipSuccess) {
        close(fd);
        return fail(c, EMP_ERR_HIP, "hipStreamCreate failed");
    }
    for (int i = 0; i < FILE_SLOTS; ++i)
        if (!c->h_chunk[i] && hipHostMalloc((void **)&c->h_chunk[i], FILE_CHUNK, hipHostMallocDefault) != hipSuccess) {
            close(fd);
            return fail(c, EMP_ERR_NOMEM, "hipHostMalloc of a %llu-byte chunk failed",
                        (unsigned long long)FILE_CHUNK);
        }
    const uint64_t nchunks = (size + FILE_CHUNK - 1) / FILE_CHUNK;
    std::vector<hipEvent_t> ev(nchunks, nullptr);
    for (auto &e : ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            rc = fail(c, EMP_ERR_HIP, "hipEventCreate failed");
            break;
        }
    std::mutex mu;
    std::condition_variable cv;
    std::vector<int> parts(nchunks, 0);  // readers done with each chunk
    uint64_t freed = 0;                  // chunk k's slot is free once k < freed + FILE_SLOTS
    uint64_t enqueued = 0;               // chunks whose copy (and event) is on the copy stream
    bool reader_failed = false;
    bool stop = false;
    std::vector<std::thread> readers;
    std::thread copier;
    if (!rc) {
        for (int p = 0; p < FILE_READERS; ++p) readers.emplace_back([&, p]() {
            for (uint64_t k = 0; k < nchunks; ++k) {
                {
                    std::unique_lock<std::mutex> g(mu);
                    cv.wait(g, [&]() { return stop || reader_failed || k < freed + FILE_SLOTS; });
                    if (stop || reader_failed) return;
                }
                const uint64_t off = k * FILE_CHUNK, len = std::min(FILE_CHUNK, size - off);
                const uint64_t per = (len + FILE_READERS - 1) / FILE_READERS;
                uint64_t a = std::min(len, per * (uint64_t)p);
                const uint64_t e = std::min(len, a + per);
                char *const buf = c->h_chunk[k % FILE_SLOTS];
                bool ok = true;
                while (a < e) {
                    const ssize_t got = pread(fd, buf + a, (size_t)(e - a), (off_t)(off + a));
                    if (got <= 0) {
                        ok = false;
                        break;
                    }
                    a += (uint64_t)got;
                }
                {
                    std::lock_guard<std::mutex> g(mu);
                    if (ok) ++parts[k];
                    else reader_failed = true;
                }
                cv.notify_all();
                if (!ok) return;
            }
        });
        copier = std::thread([&]() {
            (void)hipSetDevice(c->device);
            for (uint64_t k = 0; k < nchunks; ++k) {
                {
                    std::unique_lock<std::mutex> g(mu);
                    cv.wait(g, [&]() { return stop || reader_failed || parts[k] == FILE_READERS; });
                    if (stop || reader_failed) break;
                }
                const uint64_t off = k * FILE_CHUNK, len = std::min(FILE_CHUNK, size - off);
                bool ok = hipMemcpyAsync(c->d_file + off, c->h_chunk[k % FILE_SLOTS], len, hipMemcpyHostToDevice,
                                         c->cstream) == hipSuccess &&
                          hipEventRecord(ev[k], c->cstream) == hipSuccess;
                // slots whose copies are done are free; when the readers would otherwise wait for
                // one, wait for the oldest copy here
                uint64_t f = freed;
                while (ok && f < k + 1 && hipEventQuery(ev[f]) == hipSuccess) ++f;
                if (ok && k + 1 < nchunks && k + 1 >= f + FILE_SLOTS) {
                    ok = hipEventSynchronize(ev[f]) == hipSuccess;
                    ++f;
                }
                {
                    std::lock_guard<std::mutex> g(mu);
                    freed = f;
                    if (!ok) {
                        reader_failed = true;
                    } else {
                        enqueued = k + 1;
                    }
                }
                cv.notify_all();
                if (!ok) break;
            }
            cv.notify_all();
        });
    }
    // The arrived prefix is parsed in ranges of at least PARSE_STEP bytes (whole blocks of B each,
    // the short last one only with the file's end), while later chunks stream in.
    const uint64_t PARSE_STEP = std::max<uint64_t>(B, 8 * FILE_CHUNK);
    uint64_t pos = 0, waited = 0;  // waited: chunks the compute stream already waits for
    while (!rc && pos < size) {
        const uint64_t want = std::min(size, std::max(pos + PARSE_STEP, pos + B));
        const uint64_t need = (want + FILE_CHUNK - 1) / FILE_CHUNK;  // chunks [0, need) hold the range
        {
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&]() { return enqueued >= need || reader_failed; });
            if (enqueued < need) {
                rc = fail(c, EMP_ERR_IO, "reading %s failed", path);
                break;
            }
        }
        for (; waited < need; ++waited)
            if (hipStreamWaitEvent(c->stream, ev[waited], 0) != hipSuccess) rc = fail(c, EMP_ERR_HIP, "hipStreamWaitEvent");
        if (rc) break;
        const bool last = want == size;
        size_t used = 0;
        if ((rc = add_fastq_range(c, reinterpret_cast<const char *>(c->d_file + pos), (size_t)(want - pos), B, false,
                                  last, &used, true)))
            break;
        if (last) break;
        pos += used;
    }
    {
        std::lock_guard<std::mutex> g(mu);
        stop = true;
    }
    cv.notify_all();
    for (auto &t : readers) t.join();
    if (copier.joinable()) copier.join();
    close(fd);
    (void)hipStreamSynchronize(c->cstream);
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

// Kept reads in host memory (m of them, a multiple of 16, the context at a multiple of 16 with
// nothing staged), in batches of about PIN_BATCH bytes: copied into two device batches on the copy
// stream, each mapped on the compute stream while the next is copied. Pageable reads, and pinned
// ones of 2 PIN_BATCH bytes or more, go first into one of three pinned stages (the host fills batch
// b + 1 while batch b is on the link), packed by PACK_THREADS threads to 2-bit codes, a quarter of
// the bytes, which the device expands (k_unpack) before the map kernel: packing runs at the host's
// memory bandwidth, several times the link's rate. A batch holding a byte other than ACGTacgt
// crosses as it is (pageable: COPY_THREADS parallel copies into the stage first), and the map
// kernel's check reports it. Smaller pinned arrays are copied straight from the caller's buffer.
// The caller's buffer is free again once the copies are done; the map kernels may still run.
constexpr uint64_t PIN_BATCH = 256ull << 20;

int add_reads_pinned(emp_ctx *c, const char *reads, uint64_t m, bool pinned) {
    if (m == 0) return EMP_OK;
    const uint64_t L = (uint64_t)c->L;
    if (2 * (c->n + m) >= (1ull << 32)) return fail(c, EMP_ERR_TOO_MANY, "more than 2^31 - 1 reads");
    int rc;
    if ((rc = ensure_cols(c, c->n + m)) || (rc = ensure_entries(c, 2 * (c->n + m)))) return rc;
    const uint64_t per = std::max<uint64_t>(16, (PIN_BATCH / L) & ~15ull);
    const bool staged = !pinned || m * L >= 2 * PIN_BATCH;
    if (!c->cstream) HIPCHK(c, hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        if (c->pin_cap[k] < per * L) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            dfree(c, c->d_pin[k]);
            c->d_pin[k] = nullptr;
            void *p = nullptr;
            if ((rc = dalloc(c, &p, per * L))) return rc;
            c->d_pin[k] = (uint8_t *)p;
            c->pin_cap[k] = per * L;
        }
        if (staged && c->pack_cap[k] < per * L / 4 + 16) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            dfree(c, c->d_pack[k]);
            c->d_pack[k] = nullptr;
            void *p = nullptr;
            if ((rc = dalloc(c, &p, per * L / 4 + 16))) return rc;
            c->d_pack[k] = (uint8_t *)p;
            c->pack_cap[k] = per * L / 4 + 16;
        }
        if (!c->pin_ev[k]) {
            HIPCHK(c, hipEventCreateWithFlags(&c->pin_ev[k], hipEventDisableTiming));
            HIPCHK(c, hipEventCreateWithFlags(&c->pin_done[k], hipEventDisableTiming));
            HIPCHK(c, hipEventRecord(c->pin_done[k], c->stream));
        }
    }
    if (staged) {
        for (int k = 0; k < 3; ++k) {
            if (!c->h_big[k]) HIPCHK(c, hipHostMalloc((void **)&c->h_big[k], PIN_BATCH, hipHostMallocDefault));
            if (!c->big_done[k]) {
                HIPCHK(c, hipEventCreateWithFlags(&c->big_done[k], hipEventDisableTiming));
                HIPCHK(c, hipEventRecord(c->big_done[k], c->cstream));
            }
        }
    }
    for (uint64_t b = 0, off = 0; off < m; ++b, off += per) {
        const int slot = (int)(b & 1);
        const uint64_t cnt = std::min(per, m - off);
        const char *src = reads + off * L;
        bool packed = false;
        if (staged) {  // the stage's last copy to the device is done: refill it
            const int hs = (int)(b % 3);
            HIPCHK(c, hipEventSynchronize(c->big_done[hs]));
            packed = parallel_pack(reinterpret_cast<uint8_t *>(c->h_big[hs]), src, cnt * L);
            if (!packed && !pinned) parallel_copy(c->h_big[hs], src, cnt * L);
            if (packed || !pinned) src = c->h_big[hs];
        }
        HIPCHK(c, hipStreamWaitEvent(c->cstream, c->pin_done[slot], 0));  // the batch's last map has read the slot
        if (packed)
            HIPCHK(c, hipMemcpyAsync(c->d_pack[slot], src, (cnt * L + 3) / 4, hipMemcpyHostToDevice, c->cstream));
        else
            HIPCHK(c, hipMemcpyAsync(c->d_pin[slot], src, cnt * L, hipMemcpyHostToDevice, c->cstream));
        if (staged) HIPCHK(c, hipEventRecord(c->big_done[b % 3], c->cstream));
        HIPCHK(c, hipEventRecord(c->pin_ev[slot], c->cstream));
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->pin_ev[slot], 0));
        if (packed) {
            Timed t(c, "k_unpack", 1.25 * (double)(cnt * L));
            HIPCHK(c, launch_unpack(reinterpret_cast<const uint32_t *>(c->d_pack[slot]), c->d_pin[slot], cnt * L, c->stream));
        }
        if ((rc = map_batch(c, c->d_pin[slot], cnt))) return rc;
        HIPCHK(c, hipEventRecord(c->pin_done[slot], c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->cstream));
    return EMP_OK;
}

// Maps what the device accumulation still holds (the run's last batch).
int flush_acc(emp_ctx *c) {
    if (c->acc_n == 0) return EMP_OK;
    const uint64_t m = c->acc_n;
    c->acc_n = 0;
    return map_batch(c, c->d_acc, m);
}

// Partition plan of a list of x entries into 2^bb final buckets: passes of at most
// PART_MAX_BITS bits; intermediate buckets get capacity x / 2^bits * slack + 1024.
struct Level {
    int bits_in, dbits;
    uint32_t cap;
};

std::vector<Level> plan_levels(uint64_t x, int bb, double slack) {
    std::vector<Level> v;
    int done = 0;
    do {
        int d = std::min(PART_MAX_BITS, bb - done);
        Level l{done, d, 0};
        done += d;
        if (done == bb) {
            l.cap = BUCKET_CAP;
        } else {
            uint64_t per = (uint64_t)((double)x / (double)(1ull << done) * slack) + 1024;
            l.cap = (uint32_t)((per + 63) & ~63ull);
        }
        v.push_back(l);
    } while (done < bb);
    return v;
}

uint64_t inter_need(uint64_t x, int bb, double slack) {
    uint64_t need = 0;
    for (const Level &l : plan_levels(x, bb, slack))
        if (l.bits_in + l.dbits != bb) need = std::max<uint64_t>(need, (1ull << (l.bits_in + l.dbits)) * l.cap);
    return need;
}

int realloc_entries(emp_ctx *c, Entry **p, uint64_t n) {
    dfree(c, *p);
    *p = nullptr;
    void *q = nullptr;
    int rc = dalloc(c, &q, n * sizeof(Entry));
    if (rc) return rc;
    *p = (Entry *)q;
    return EMP_OK;
}

int realloc_spill(emp_ctx *c, SpillRec **p, uint32_t n) {
    dfree(c, *p);
    *p = nullptr;
    void *q = nullptr;
    int rc = dalloc(c, &q, (size_t)n * sizeof(SpillRec));
    if (rc) return rc;
    *p = (SpillRec *)q;
    return EMP_OK;
}

// Spill area of bucket set `set`, side `side`, with `sp` records. Its contents are dropped: call it
// only when the set holds no live round (before the kernel that fills it runs, or reruns).
int grow_bspill(emp_ctx *c, int set, int side, uint32_t sp) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int rc = realloc_spill(c, &c->d_bspill[set][side], sp);
    if (rc) return rc;
    c->bspill_cap[set][side] = sp;
    return EMP_OK;
}

// Buffers for lists of up to x entries in 2^bb buckets: the two bucket sets, and the dedup
// partition's intermediate buckets and spill area.
int ensure_run_buffers(emp_ctx *c, uint64_t x, int bb) {
    int rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const uint32_t nb = 1u << bb;
    if (nb > c->nb_alloc) {
        for (int set = 0; set < 2; ++set)
            for (int side = 0; side < 2; ++side) {
                if ((rc = realloc_entries(c, &c->d_bfine[set][side], (uint64_t)nb * BUCKET_CAP))) return rc;
                dfree(c, c->d_bcnt[set][side]);
                void *p = nullptr;
                if ((rc = dalloc(c, &p, (size_t)nb * sizeof(uint32_t)))) return rc;
                c->d_bcnt[set][side] = (uint32_t *)p;
            }
        dfree(c, c->d_dcnt);
        void *p = nullptr;
        if ((rc = dalloc(c, &p, (size_t)nb * sizeof(uint32_t)))) return rc;
        c->d_dcnt = (uint32_t *)p;
        c->nb_alloc = nb;
    }
    c->d_dfine = c->d_bfine[1][0];
    const uint64_t need = inter_need(x, bb, c->slack);
    if (need > c->inter_cap) {
        for (int k = 0; k < 2; ++k) {
            if ((rc = realloc_entries(c, &c->d_inter[k], need))) return rc;
            dfree(c, c->d_inter_pay[k]);
            c->d_inter_pay[k] = nullptr;
            void *p = nullptr;
            if ((rc = dalloc(c, &p, need * sizeof(uint32_t)))) return rc;
            c->d_inter_pay[k] = (uint32_t *)p;
        }
        c->inter_cap = need;
    }
    if ((uint64_t)nb * BUCKET_CAP > c->dpay_cap) {
        dfree(c, c->d_dpay);
        c->d_dpay = nullptr;
        void *p = nullptr;
        if ((rc = dalloc(c, &p, (size_t)nb * BUCKET_CAP * sizeof(uint32_t)))) return rc;
        c->d_dpay = (uint32_t *)p;
        c->dpay_cap = (uint64_t)nb * BUCKET_CAP;
    }
    if (!c->d_inter_cnt[0]) {
        for (int k = 0; k < 2; ++k) {
            void *p = nullptr;
            if ((rc = dalloc(c, &p, (size_t)(1u << 16) * sizeof(uint32_t)))) return rc;
            c->d_inter_cnt[k] = (uint32_t *)p;
        }
    }
    const uint32_t sp = (uint32_t)std::max<uint64_t>(1u << 16, x / 256);
    if (sp > c->dspill_cap) {
        if ((rc = realloc_spill(c, &c->d_dspill, sp))) return rc;
        c->dspill_cap = sp;
    }
    for (int set = 0; set < 2; ++set)
        for (int side = 0; side < 2; ++side)
            if (c->bspill_cap[set][side] < sp && (rc = grow_bspill(c, set, side, sp))) return rc;
    return EMP_OK;
}

// After an abort of the dedup partition (an intermediate bucket or the spill area overflowed):
// more slack, more spill.
int grow_after_abort(emp_ctx *c, uint64_t x, int bb) {
    c->slack *= 2.0;
    const uint32_t sp = c->dspill_cap * 4u;
    int rc = ensure_run_buffers(c, x, bb);
    if (rc) return rc;
    if (c->dspill_cap < sp) {
        if ((rc = realloc_spill(c, &c->d_dspill, sp))) return rc;
        c->dspill_cap = sp;
    }
    return EMP_OK;
}

// The dedup's sort (overlap.cu:211-212): the x map entries (linear, count in d_ctr->in) into
// 2^bb buckets by passes of at most PART_MAX_BITS bits.
// fin_nb != 0: the last pass puts the entries into fin_nb buckets (a multiple of 256, at most
// 2^bb) by bucket_of; needs a two-pass plan whose first pass takes the top 8 bits.
int partition(emp_ctx *c, uint64_t x, int bb, uint32_t fin_nb = 0) {
    const std::vector<Level> pl = plan_levels(x, bb, c->slack);
    if (fin_nb && (pl.size() != 2 || pl[0].dbits != 8 || fin_nb % 256 || fin_nb > (1u << bb)))
        return fail(c, EMP_ERR_ARG, "dedup partition into %u buckets", fin_nb);
    const Entry *in = c->d_entries;
    const uint32_t *in_cnt = nullptr;
    uint32_t in_cap = 0, groups = 1;
    int keyed = 0;
    for (size_t i = 0; i < pl.size(); ++i) {
        const Level &l = pl[i];
        const bool fin = i + 1 == pl.size();
        const int bits_out = l.bits_in + l.dbits;
        Entry *out = fin ? c->d_dfine : c->d_inter[i & 1];
        uint32_t *out_cnt = fin ? c->d_dcnt : c->d_inter_cnt[i & 1];
        PartArgs a{};
        a.in = in;
        a.in_cnt = in_cnt;
        a.in_n = &c->d_ctr->in;
        a.in_cap = in_cap;
        a.keyed = keyed;
        a.tiles_x = in_cnt ? std::max<uint32_t>(1, (in_cap + PART_TILE - 1) / PART_TILE)
                           : (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((x + PART_TILE - 1) / PART_TILE, 1u << 16));
        a.out = out;
        a.out_cnt = out_cnt;
        a.out_cap = l.cap;
        a.out_shift = 29 - bits_out;
        a.dbits = l.dbits;
        a.spill = fin ? c->d_dspill : nullptr;
        a.spill_n = &c->d_ctr->spill;
        a.spill_cap = c->dspill_cap;
        a.abort = &c->d_ctr->abort;
        // each entry's read's ends word goes along (round 0's regeneration in the dedup reads it)
        a.pay_in = i == 0 ? c->d_ends : c->d_inter_pay[(i - 1) & 1];
        a.pay_out = fin ? c->d_dpay : c->d_inter_pay[i & 1];
        if (!c->d_ends) a.pay_in = nullptr;
        a.fin_nb = fin ? fin_nb : 0u;
        {
            Timed t(c, "k_part", (a.pay_in ? 32.0 : 24.0) * (double)x);
            HIPCHK(c, launch_part(a, groups, c->stream));
        }
        in = out;
        in_cnt = out_cnt;
        in_cap = l.cap;
        groups = 1u << bits_out;
        keyed = 1;
    }
    return EMP_OK;
}

// Counter resets batched into one launch.
struct ZeroList {
    ZeroArgs a{};
    void add(uint32_t *p, uint64_t n, uint32_t v = 0) {
        a.p[a.count] = p;
        a.n[a.count] = (uint32_t)n;
        a.v[a.count] = v;
        ++a.count;
    }
};

// What the dedup partition of x entries into 2^bb buckets accumulates, its input size and flags.
void zero_partition(emp_ctx *c, ZeroList &z, uint64_t x, int bb) {
    const std::vector<Level> pl = plan_levels(x, bb, c->slack);
    for (size_t i = 0; i < pl.size(); ++i) {
        const int bits_out = pl[i].bits_in + pl[i].dbits;
        z.add(i + 1 == pl.size() ? c->d_dcnt : c->d_inter_cnt[i & 1], 1ull << bits_out);
    }
    z.add(&c->d_ctr->abort, 1);
    z.add(&c->d_ctr->spill, 1);
    z.add(&c->d_ctr->in, 1, (uint32_t)x);
    z.add(&c->d_ctr->too_many, 1);
}

// What a launch writing round t's bucket set (2^bb buckets per side) accumulates: its counts and
// the round's counter record.
void zero_next(emp_ctx *c, ZeroList &z, int t, int bb) {
    z.add(c->d_bcnt[t & 1][0], 1ull << bb);
    z.add(c->d_bcnt[t & 1][1], 1ull << bb);
    z.add(reinterpret_cast<uint32_t *>(&c->d_ctr->r[t]), sizeof(ListCtr) / 4);
}

int run_zero(emp_ctx *c, const ZeroList &z) {
    HIPCHK(c, launch_zero(z.a, c->stream));
    return EMP_OK;
}

// Edge pool: SUBS sub-ranges of `per_sub` edges, then an overflow region of `ovf` edges. The
// pool holds the edges of every round until the end-of-run compaction. With `keep` only the
// overflow region grows and everything already written stays at its offset.
int ensure_pool(emp_ctx *c, uint64_t per_sub, uint64_t ovf, bool keep) {
    if (keep) per_sub = c->pool_cap;
    if (per_sub <= c->pool_cap && ovf <= c->pool_ovf_cap) return EMP_OK;
    per_sub = std::max<uint64_t>(per_sub, c->pool_cap);
    ovf = std::max<uint64_t>(ovf, c->pool_ovf_cap);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const uint64_t old_total = c->pool_cap * SUBS + c->pool_ovf_cap;
    void *p = nullptr;
    if (int rc = dalloc(c, &p, (per_sub * SUBS + ovf) * sizeof(uint2))) return rc;
    if (keep && c->d_pool) {
        HIPCHK(c, hipMemcpyAsync(p, c->d_pool, old_total * sizeof(uint2), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    dfree(c, c->d_pool);
    c->d_pool = (uint2 *)p;
    c->pool_cap = per_sub;
    c->pool_ovf_cap = ovf;
    return EMP_OK;
}

// Edge slots for n (round, bucket) pairs, all counts zero.
int ensure_edge_slots(emp_ctx *c, uint64_t n) {
    if (n > c->eslots_cap) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dfree(c, c->d_edge_off);
        dfree(c, c->d_dst_off);
        dfree(c, c->d_edge_cnt);
        dfree(c, c->d_eblk);
        c->d_edge_off = c->d_dst_off = c->d_eblk = nullptr;
        c->d_edge_cnt = nullptr;
        c->eslots_cap = 0;
        void *p = nullptr;
        int rc;
        if ((rc = dalloc(c, &p, n * sizeof(unsigned long long)))) return rc;
        c->d_edge_off = (unsigned long long *)p;
        if ((rc = dalloc(c, &p, n * sizeof(unsigned long long)))) return rc;
        c->d_dst_off = (unsigned long long *)p;
        if ((rc = dalloc(c, &p, n * sizeof(uint32_t)))) return rc;
        c->d_edge_cnt = (uint32_t *)p;
        if ((rc = dalloc(c, &p, ((n + 4095) / 4096) * sizeof(unsigned long long)))) return rc;
        c->d_eblk = (unsigned long long *)p;
        c->eslots_cap = n;
    }
    HIPCHK(c, launch_fill_u32(c->d_edge_cnt, 0u, n, c->stream));
    return EMP_OK;
}

// Passes every pending record to the sink (output.bin order) and empties the arena.
int flush_records(emp_ctx *c, emp_record_sink sink, void *user) {
    if (c->records.empty()) return EMP_OK;
    if (c->out_used && !c->out_streamed) {
        HIPCHK(c, hipMemcpyAsync(c->h_from, c->d_out_from, c->out_used * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->h_to, c->d_out_to, c->out_used * 4, hipMemcpyDeviceToHost, c->stream));
    }
    int rc = sync(c);
    if (rc) return rc;
    for (size_t i = (size_t)std::min<uint64_t>(c->emitted, c->records.size()); i < c->records.size(); ++i) {
        const auto &r = c->records[i];
        if (!sink) continue;
        int src = sink(user, r.n, r.ovlp, c->h_from + r.off, c->h_to + r.off);
        if (src) return fail(c, src, "record sink returned %d", src);
    }
    c->emitted = 0;
    c->records.clear();
    c->out_used = 0;
    c->out_streamed = false;
    return EMP_OK;
}

// Output arena of at least `cap` edges (device + pinned host mirror); flushes first.
int ensure_arena(emp_ctx *c, uint64_t cap, emp_record_sink sink, void *user) {
    if (cap <= c->out_cap) return EMP_OK;
    int rc = flush_records(c, sink, user);
    if (rc) return rc;
    // the output file may still read records from the arena (chunks not written yet)
    if (c->out_file && c->out_file->settle() != EMP_OK) return fail(c, EMP_ERR_IO, "writing the output file failed");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, c->d_out_from);
    dfree(c, c->d_out_to);
    if (c->h_from) (void)hipHostFree(c->h_from);
    if (c->h_to) (void)hipHostFree(c->h_to);
    c->h_from = c->h_to = nullptr;
    c->out_cap = 0;
    void *p = nullptr;
    if ((rc = dalloc(c, &p, cap * 4))) return rc;
    c->d_out_from = (uint32_t *)p;
    if ((rc = dalloc(c, &p, cap * 4))) return rc;
    c->d_out_to = (uint32_t *)p;
    HIPCHK(c, hipHostMalloc((void **)&c->h_from, cap * 4, hipHostMallocDefault));
    HIPCHK(c, hipHostMalloc((void **)&c->h_to, cap * 4, hipHostMallocDefault));
    c->out_cap = cap;
    return EMP_OK;
}

RoundArgs round_args(const emp_ctx *c, int t, bool active) {
    RoundArgs r{};
    r.active = active ? 1 : 0;
    if (!active) return r;
    const int L = c->L;
    r.col_t = (uint32_t)t;
    r.col_back = (uint32_t)(L - 1 - t);
    r.m1 = modpow(R1, (uint64_t)(L - t - 1), P1);
    r.m2 = modpow(R2, (uint64_t)(L - t - 1), P2);
    r.G = col_groups8((uint32_t)L);
    return r;
}

int read_counters(emp_ctx *c) {
    HIPCHK(c, hipMemcpyAsync(c->h_ctr, c->d_ctr, sizeof(DevCounters), hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

int ensure_big_scratch(emp_ctx *c, uint64_t need) {
    if (need <= c->bscratch_cap) return EMP_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int rc;
    for (int s = 0; s < 2; ++s) {
        dfree(c, c->d_bkey[s]);
        dfree(c, c->d_brank[s]);
        dfree(c, c->d_baux[s]);
        void *p = nullptr;
        if ((rc = dalloc(c, &p, need * sizeof(uint64_t)))) return rc;
        c->d_bkey[s] = (uint64_t *)p;
        if ((rc = dalloc(c, &p, need * sizeof(uint32_t)))) return rc;
        c->d_brank[s] = (uint32_t *)p;
        if ((rc = dalloc(c, &p, need * sizeof(uint32_t)))) return rc;
        c->d_baux[s] = (uint32_t *)p;
    }
    c->bscratch_cap = need;
    return EMP_OK;
}

// One side of round t's lists as the input of a launch: buckets, counts, spill area with its count.
struct SideIn {
    const Entry *fine;
    uint32_t *cnt;
    const SpillRec *spill;
    const uint32_t *spill_n;
    uint32_t spill_cap;
};

SideIn side_in(emp_ctx *c, int t, int side) {
    const int set = t & 1;
    return SideIn{c->d_bfine[set][side], c->d_bcnt[set][side], c->d_bspill[set][side], &c->d_ctr->r[t].spill[side],
                  c->bspill_cap[set][side]};
}

// Round t's bucket set as the output of a launch over 2^src_bb source buckets: 2^bb buckets per
// side, regenerated for round rg (inactive: none, survivors only counted).
// Round t's lists in bucket set t % 2: nb buckets, written by a launch over src_nb source buckets
// (nb <= src_nb). The shift fields serve the out-of-core kernels, whose sets are powers of two.
NextOut next_out_nb(emp_ctx *c, int t, uint32_t nb, uint32_t src_nb, const RoundArgs &rg) {
    const int set = t & 1;
    NextOut o{};
    o.active = rg.active;
    o.nb = nb;
    o.src_nb = src_nb;
    o.shift = 29 - ceil_log2(nb);
    o.src_shift = 29 - ceil_log2(src_nb);
    o.rg = rg;
    for (int side = 0; side < 2; ++side) {
        o.fine[side] = c->d_bfine[set][side];
        o.cnt[side] = c->d_bcnt[set][side];
        o.spill[side] = c->d_bspill[set][side];
        o.spill_n[side] = &c->d_ctr->r[t].spill[side];
        o.spill_cap[side] = c->bspill_cap[set][side];
        o.total[side] = &c->d_ctr->r[t].total[side];
    }
    o.ovf = &c->d_ctr->r[t].nx_ovf;
    o.cols = c->d_cols;
    return o;
}

NextOut next_out(emp_ctx *c, int t, int bb, int src_bb, const RoundArgs &rg) {
    return next_out_nb(c, t, 1u << bb, 1u << src_bb, rg);
}

// ---------------------------------------------------------------------------------------------
// Multi-record framing of a round (cq_get_search_partitions, cqutils.cuh:107-137, used by
// emit_edges, overlap.cu:395-447). When the suffix or the prefix list reaches M/2 entries, the
// reference cuts both key-sorted lists into partitions of at most M/2 entries per side at a common
// key m (the smaller of the two sides' last keys in the window, then lower_bound on each side),
// and writes one record per partition with both sides non-empty. The lists here are in key order
// bucket by bucket, so every cut falls inside one bucket: positions and counts below a key come
// from the bucket counts plus that one bucket's entries, read back and sorted on the host. The
// round's edges are already in key order, so the records are consecutive slices of them.
struct FrameSide {
    const Entry *fine = nullptr;  // final buckets (BUCKET_CAP slots each)
    const SpillRec *spill = nullptr;
    uint32_t nspill = 0;
    std::vector<uint32_t> cnt;
    std::vector<uint64_t> pre;  // exclusive prefix sums of cnt, nb + 1 of them
    std::vector<SpillRec> hspill;
    bool spill_read = false;
    std::map<uint32_t, std::vector<std::pair<uint64_t, uint32_t>>> sorted;  // bucket -> (key, rid)
    int bb = 0;                 // bucket bits of the list's set
};

// Key of a wide entry of bucket b of a set of 2^bb buckets: h1's bits 21..28 are the bucket's top 8.
uint64_t entry_key(const Entry &e, uint32_t b, int bb) {
    const uint32_t h1 = (e.h1 & LOW_MASK) | ((b >> (bb - WIDE_MIN_BITS)) << WIDE_SHIFT);
    return ((uint64_t)h1 << 32) | (e.h2 & KEY_MASK);
}

int frame_bucket(emp_ctx *c, FrameSide &f, uint32_t b, const std::vector<std::pair<uint64_t, uint32_t>> **out) {
    auto it = f.sorted.find(b);
    if (it == f.sorted.end()) {
        const uint32_t n = f.cnt[b], in_fine = std::min<uint32_t>(n, BUCKET_CAP);
        std::vector<Entry> h(in_fine);
        if (in_fine)
            HIPCHK(c, hipMemcpyAsync(h.data(), f.fine + (size_t)b * BUCKET_CAP, in_fine * sizeof(Entry),
                                     hipMemcpyDeviceToHost, c->stream));
        if (n > in_fine && !f.spill_read) {  // the rest of a large bucket is in the spill area
            f.hspill.resize(f.nspill);
            if (f.nspill)
                HIPCHK(c, hipMemcpyAsync(f.hspill.data(), f.spill, f.nspill * sizeof(SpillRec), hipMemcpyDeviceToHost,
                                         c->stream));
            f.spill_read = true;
        }
        if (int rc = sync(c)) return rc;
        if (n > in_fine) {
            for (const SpillRec &r : f.hspill)
                if (r.bucket == b) h.push_back(r.e);
            if (h.size() != n)
                return fail(c, EMP_ERR_CAPACITY, "framing: bucket %u holds %zu entries, counted %u", b, h.size(), n);
        }
        std::vector<std::pair<uint64_t, uint32_t>> v(h.size());
        for (size_t i = 0; i < h.size(); ++i) v[i] = {entry_key(h[i], b, f.bb), h[i].rid};
        std::sort(v.begin(), v.end());
        it = f.sorted.emplace(b, std::move(v)).first;
    }
    *out = &it->second;
    return EMP_OK;
}

int frame_side(emp_ctx *c, FrameSide &f, const Entry *fine, const uint32_t *d_cnt, const SpillRec *spill,
               uint32_t nspill, uint32_t nb, uint64_t want) {
    f.fine = fine;
    f.spill = spill;
    f.nspill = nspill;
    f.bb = 0;
    while ((1u << f.bb) < nb) ++f.bb;
    f.cnt.resize(nb);
    HIPCHK(c, hipMemcpyAsync(f.cnt.data(), d_cnt, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    if (int rc = sync(c)) return rc;
    f.pre.assign(nb + 1, 0);
    for (uint32_t b = 0; b < nb; ++b) f.pre[b + 1] = f.pre[b] + f.cnt[b];
    if (f.pre[nb] != want)
        return fail(c, EMP_ERR_CAPACITY, "framing: %llu entries in buckets, %llu in the list",
                    (unsigned long long)f.pre[nb], (unsigned long long)want);
    return EMP_OK;
}

// Records of one round: the edge count of every partition with both sides non-empty, in order.
int frame_round(emp_ctx *c, int bb, FrameSide &S, FrameSide &P, const uint32_t *d_ecnt, uint64_t E, uint64_t M,
                std::vector<uint64_t> &recs) {
    const uint32_t nb = 1u << bb;
    const int shift = 29 - bb;
    std::vector<uint32_t> ecnt(nb);
    HIPCHK(c, hipMemcpyAsync(ecnt.data(), d_ecnt, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    if (int rc = sync(c)) return rc;
    std::vector<uint64_t> epre(nb + 1, 0);
    for (uint32_t b = 0; b < nb; ++b) epre[b + 1] = epre[b] + ecnt[b];
    if (epre[nb] != E) return fail(c, EMP_ERR_CAPACITY, "framing: %llu edges in buckets, %llu in the round",
                                   (unsigned long long)epre[nb], (unsigned long long)E);
    const uint64_t nA = S.pre[nb], nB = P.pre[nb], H = M / 2;
    auto bucket_of = [&](uint64_t m) { return (uint32_t)((m >> 32) >> shift); };
    auto key_at = [&](FrameSide &f, uint64_t r, uint64_t &key) -> int {
        const uint32_t b = (uint32_t)(std::upper_bound(f.pre.begin(), f.pre.end(), r) - f.pre.begin()) - 1;
        const std::vector<std::pair<uint64_t, uint32_t>> *v;
        int rc = frame_bucket(c, f, b, &v);
        if (rc) return rc;
        key = (*v)[r - f.pre[b]].first;
        return EMP_OK;
    };
    auto below = [&](FrameSide &f, uint64_t m, uint64_t &n) -> int {
        const uint32_t b = bucket_of(m);
        const std::vector<std::pair<uint64_t, uint32_t>> *v;
        int rc = frame_bucket(c, f, b, &v);
        if (rc) return rc;
        n = f.pre[b] + (uint64_t)(std::lower_bound(v->begin(), v->end(), std::make_pair(m, 0u)) - v->begin());
        return EMP_OK;
    };
    // edges whose key is below m: whole buckets below m's, then the valid pairs of m's bucket
    auto edges_below = [&](uint64_t m, uint64_t &n) -> int {
        const uint32_t b = bucket_of(m);
        const std::vector<std::pair<uint64_t, uint32_t>> *sv, *pv;
        int rc = frame_bucket(c, S, b, &sv);
        if (!rc) rc = frame_bucket(c, P, b, &pv);
        if (rc) return rc;
        n = epre[b];
        for (const auto &s : *sv) {
            if (s.first >= m) break;
            auto lo = std::lower_bound(pv->begin(), pv->end(), std::make_pair(s.first, 0u));
            for (; lo != pv->end() && lo->first == s.first; ++lo) n += (lo->second >> 1) != (s.second >> 1);
        }
        return EMP_OK;
    };
    uint64_t offA = 0, offB = 0, ebase = 0;
    int rc;
    while (offA < nA && offB < nB) {
        uint64_t nPa = std::min(nA - offA, H), nPb = std::min(nB - offB, H);
        if (nPa < H && nPb < H) {
            recs.push_back(E - ebase);  // the closing partition: everything left, both sides non-empty
            return EMP_OK;
        }
        uint64_t ka, kb, ca, cb, eb;
        if ((rc = key_at(S, offA + nPa - 1, ka)) || (rc = key_at(P, offB + nPb - 1, kb))) return rc;
        const uint64_t m = std::min(ka, kb);
        if ((rc = below(S, m, ca)) || (rc = below(P, m, cb)) || (rc = edges_below(m, eb))) return rc;
        nPa = ca - offA;
        nPb = cb - offB;
        if (nPa == 0 && nPb == 0)  // more than M/2 equal keys: the reference never leaves this loop
            return fail(c, EMP_ERR_PARTITION, "more than M/2 = %llu entries share one key", (unsigned long long)H);
        if (nPa && nPb) recs.push_back(eb - ebase);
        else if (eb != ebase) return fail(c, EMP_ERR_CAPACITY, "framing: edges in a one-sided partition");
        ebase = eb;
        offA += nPa;
        offB += nPb;
    }
    // one side is used up: what is left of the other forms a partition without a record
    if (ebase != E) return fail(c, EMP_ERR_CAPACITY, "framing: %llu edges outside the records",
                                (unsigned long long)(E - ebase));
    return EMP_OK;
}

// ---------------------------------------------------------------------------------------------
// Collision audit of one bucketed list (FPCHECK, overlap.cu:215-219,274-284,295-359): distinct
// keys and distinct strings [start, start + len) of the oriented reads, summed over keys. The LDS
// launch takes every bucket within capacity; the buckets over it follow in a launch over global
// scratch. Appends one record to c->audits.
int audit_list(emp_ctx *c, int list, int ovlp, uint64_t entries, const Entry *fine, const uint32_t *d_cnt,
               const SpillRec *spill, const uint32_t *d_spill_n, uint32_t spill_cap, uint32_t nb, int start, int len,
               bool wide) {
    int rc;
    if (!c->d_audit) {
        void *p = nullptr;
        if ((rc = dalloc(c, &p, 2 * sizeof(unsigned long long)))) return rc;
        c->d_audit = (unsigned long long *)p;
    }
    HIPCHK(c, hipMemsetAsync(c->d_audit, 0, 2 * sizeof(unsigned long long), c->stream));
    AuditArgs a{};
    a.fine = fine;
    a.cnt = d_cnt;
    a.cap = BUCKET_CAP;
    a.spill = spill;
    a.spill_n = d_spill_n;
    a.spill_cap = spill_cap;
    a.cols = c->d_cols;
    a.G = col_groups8((uint32_t)c->L);
    a.L = c->L;
    a.start = start;
    a.len = len;
    int bbl = 0;
    while ((1u << bbl) < nb) ++bbl;
    a.wshift = wide ? 29 - bbl : -1;  // a round's list holds wide entries; the dedup's, full keys
    a.out = c->d_audit;
    HIPCHK(c, launch_audit(a, nb, c->stream));
    std::vector<uint32_t> cnt(nb);
    HIPCHK(c, hipMemcpyAsync(cnt.data(), d_cnt, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    if ((rc = sync(c))) return rc;
    std::vector<uint32_t> ids;
    std::vector<unsigned long long> off;
    uint64_t tot = 0;
    for (uint32_t b = 0; b < nb; ++b)
        if (cnt[b] > BUCKET_CAP) {
            uint64_t p2 = 1;
            while (p2 < cnt[b]) p2 <<= 1;
            ids.push_back(b);
            off.push_back(tot);
            tot += p2;
        }
    if (!ids.empty()) {
        if (ids.size() > BIG_CAP) return fail(c, EMP_ERR_CAPACITY, "audit: %zu buckets over capacity", ids.size());
        if ((rc = ensure_big_scratch(c, tot))) return rc;
        if (!c->d_boff[0]) {
            for (int s2 = 0; s2 < 2; ++s2) {
                void *p = nullptr;
                if ((rc = dalloc(c, &p, BIG_CAP * sizeof(unsigned long long)))) return rc;
                c->d_boff[s2] = (unsigned long long *)p;
            }
        }
        HIPCHK(c, hipMemcpyAsync(c->d_big, ids.data(), ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_boff[0], off.data(), off.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                                 c->stream));
        a.ids = c->d_big;
        a.off = c->d_boff[0];
        a.skey = c->d_bkey[0];
        a.srid = c->d_brank[0];
        HIPCHK(c, launch_audit(a, (uint32_t)ids.size(), c->stream));
    }
    unsigned long long h[2];
    HIPCHK(c, hipMemcpyAsync(h, c->d_audit, sizeof h, hipMemcpyDeviceToHost, c->stream));
    if ((rc = sync(c))) return rc;
    emp_audit r{};
    r.list = list;
    r.ovlp = ovlp;
    r.entries = entries;
    r.keys = h[0];
    r.strings = h[1];
    c->audits.push_back(r);
    return EMP_OK;
}

// Buckets the LDS kernels handed over (more than BUCKET_CAP entries or MATCH_CAP matches):
// global-memory sort + join (or dedup), one work-group per bucket, over the inputs `sin` / `pin`
// (dedup: sin only); survivors (kept entries) go to nx.
int big_path(emp_ctx *c, uint32_t nbig, int bb, bool dedup, const RoundArgs &rg, uint64_t slot0, const SideIn &sin,
             const SideIn &pin, const NextOut &nx) {
    if (nbig > BIG_CAP) return fail(c, EMP_ERR_CAPACITY, "%u buckets over capacity (limit %u)", nbig, BIG_CAP);
    const uint32_t nb = 1u << bb;
    std::vector<uint32_t> ids(nbig), cs(nb), cp(nb);
    HIPCHK(c, hipMemcpyAsync(ids.data(), c->d_big, nbig * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(cs.data(), sin.cnt, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    if (!dedup) HIPCHK(c, hipMemcpyAsync(cp.data(), pin.cnt, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    int rc = sync(c);
    if (rc) return rc;
    std::vector<unsigned long long> so(nbig), po(nbig);
    uint64_t as = 0, ap = 0;
    auto p2 = [](uint64_t n) { uint64_t p = 1; while (p < n) p <<= 1; return p; };
    for (uint32_t i = 0; i < nbig; ++i) {
        so[i] = as;
        as += p2(cs[ids[i]]);
        po[i] = ap;
        ap += dedup ? 0 : p2(cp[ids[i]]);
    }
    if ((rc = ensure_big_scratch(c, std::max<uint64_t>(std::max(as, ap), 1)))) return rc;
    if (!c->d_boff[0]) {
        for (int s = 0; s < 2; ++s) {
            void *p = nullptr;
            if ((rc = dalloc(c, &p, BIG_CAP * sizeof(unsigned long long)))) return rc;
            c->d_boff[s] = (unsigned long long *)p;
        }
    }
    HIPCHK(c, hipMemcpyAsync(c->d_boff[0], so.data(), nbig * sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_boff[1], po.data(), nbig * sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
    BigArgs a{};
    a.ids = c->d_big;
    a.soff = c->d_boff[0];
    a.poff = c->d_boff[1];
    a.skey = c->d_bkey[0];
    a.pkey = c->d_bkey[1];
    a.srank = c->d_brank[0];
    a.prank = c->d_brank[1];
    a.saux = c->d_baux[0];
    a.paux = c->d_baux[1];
    a.s_in = sin.fine;
    a.s_cnt = sin.cnt;
    a.s_spill = sin.spill;
    a.s_spill_n = sin.spill_n;
    a.spill_cap = sin.spill_cap;
    if (!dedup) {
        a.p_in = pin.fine;
        a.p_cnt = pin.cnt;
        a.p_spill = pin.spill;
        a.p_spill_n = pin.spill_n;
        a.p_spill_cap = pin.spill_cap;
    }
    a.cap = BUCKET_CAP;
    a.rg = rg;
    a.cols = c->d_cols;
    a.nx = nx;
    a.pool = c->d_pool;
    a.pool_cursor = c->d_ctr->pool;
    a.pool_cap = c->pool_cap;
    a.ovf_cursor = &c->d_ctr->pool_ovf;
    a.ovf_cap = c->pool_ovf_cap;
    a.edge_off = c->d_edge_off + slot0;
    a.edge_cnt = c->d_edge_cnt + slot0;
    a.too_many = &c->d_ctr->too_many;
    {
        Timed t(c, dedup ? "k_dedup_big" : "k_join_big", 0.0);
        HIPCHK(c, dedup ? launch_dedup_big(a, nbig, c->stream) : launch_join_big(a, nbig, c->stream));
    }
    if (int rc = read_counters(c)) return rc;
    if (c->h_ctr->too_many)  // the reference's int edge offsets (overlap.cu:561,567)
        return fail(c, EMP_ERR_PARTITION, "a key group holds 2^31 edges or more");
    return EMP_OK;
}

// ---------------------------------------------------------------------------------------------
// The out-of-core order (2n > M). cq_sort_by_key (cqutils.cuh:179-247) orders equal keys of a
// list of more than M entries by its FIFO merge tree; every entry gets a tie key that reproduces
// that order (the default: in-order index of its chunk, order inside the chunk; explicit ranks
// for the keys the merge partitions reorder, found by the host planner of emp_ooc_plan.h). The
// dedup and the joins of such lists then sort each bucket by (key, tie) in global scratch.

int ensure_ooc(emp_ctx *c, uint64_t ids) {
    int rc;
    if (ids > c->ooc_ids) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        void *p = nullptr;
        for (int s = 0; s < 2; ++s) {
            dfree(c, c->d_tie[s]);
            c->d_tie[s] = nullptr;
            if ((rc = dalloc(c, &p, ids * sizeof(unsigned long long)))) return rc;
            c->d_tie[s] = (unsigned long long *)p;
        }
        dfree(c, c->d_chunk);
        c->d_chunk = nullptr;
        if ((rc = dalloc(c, &p, ids * sizeof(uint16_t)))) return rc;
        c->d_chunk = (uint16_t *)p;
        const uint64_t words = (ids + 31) / 32;
        dfree(c, c->d_obm);
        dfree(c, c->d_opc);
        dfree(c, c->d_obase);
        dfree(c, c->d_oblk);
        c->d_obm = c->d_opc = nullptr;
        c->d_obase = c->d_oblk = nullptr;
        if ((rc = dalloc(c, &p, words * sizeof(uint32_t)))) return rc;
        c->d_obm = (uint32_t *)p;
        if ((rc = dalloc(c, &p, words * sizeof(uint32_t)))) return rc;
        c->d_opc = (uint32_t *)p;
        if ((rc = dalloc(c, &p, (words + 1) * sizeof(unsigned long long)))) return rc;
        c->d_obase = (unsigned long long *)p;
        if ((rc = dalloc(c, &p, (fq_scan_blocks(words) + 1) * sizeof(unsigned long long)))) return rc;
        c->d_oblk = (unsigned long long *)p;
        c->ooc_ids = ids;
    }
    if (!c->d_obcnt) {
        void *p = nullptr;
        if ((rc = dalloc(c, &p, (PLAN_MAX_CELLS + PLAN_MAX_CHUNKS) * sizeof(uint32_t)))) return rc;
        c->d_obcnt = (uint32_t *)p;
        if ((rc = dalloc(c, &p, (PLAN_MAX_CHUNKS + 1) * sizeof(uint32_t)))) return rc;
        c->d_oio = (uint32_t *)p;
    }
    return EMP_OK;
}

template <class T>
int grow_dev(emp_ctx *c, T **p, uint64_t &cap, uint64_t need) {
    if (need <= cap) return EMP_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dfree(c, *p);
    *p = nullptr;
    const uint64_t n = std::max<uint64_t>(need, cap * 2);
    void *q = nullptr;
    int rc = dalloc(c, &q, n * sizeof(T));
    if (rc) return rc;
    *p = (T *)q;
    cap = n;
    return EMP_OK;
}

// The entries of one bucket group of a list, read from the device for the planner.
class DevSource : public PlanSource {
public:
    DevSource(emp_ctx *c, const OocList &l, int gshift, const std::vector<uint32_t> &cnt)
        : c_(c), l_(l), gshift_(gshift), cnt_(cnt) {}
    int fetch(uint32_t g, std::vector<PlanRec> &out) override {
        const uint32_t b0 = g << gshift_, b1 = std::min<uint32_t>(l_.nb, (g + 1) << gshift_);
        std::vector<uint32_t> ids;
        std::vector<unsigned long long> off;
        uint64_t tot = 0;
        for (uint32_t b = b0; b < b1; ++b)
            if (cnt_[b]) {
                ids.push_back(b);
                off.push_back(tot);
                tot += cnt_[b];
            }
        out.clear();
        if (ids.empty()) return EMP_OK;
        int rc;
        if ((rc = grow_dev(c_, &c_->d_ofids, c_->ofid_cap, ids.size())) ||
            (rc = grow_dev(c_, &c_->d_ofoff, c_->ofoff_cap, ids.size())) ||
            (rc = grow_dev(c_, &c_->d_orec, c_->orec_cap, tot)))
            return rc;
        std::vector<OocRec> h(tot);
        HIPCHK(c_, hipMemcpyAsync(c_->d_ofids, ids.data(), ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                  c_->stream));
        HIPCHK(c_, hipMemcpyAsync(c_->d_ofoff, off.data(), off.size() * sizeof(unsigned long long),
                                  hipMemcpyHostToDevice, c_->stream));
        HIPCHK(c_, launch_ooc_fetch(l_, c_->d_ofids, c_->d_ofoff, (uint32_t)ids.size(), c_->d_chunk, c_->d_orec,
                                    c_->stream));
        HIPCHK(c_, hipMemcpyAsync(h.data(), c_->d_orec, tot * sizeof(OocRec), hipMemcpyDeviceToHost, c_->stream));
        if ((rc = sync(c_))) return rc;
        out.resize(tot);
        for (uint64_t i = 0; i < tot; ++i) out[i] = PlanRec{h[i].key, h[i].rid, h[i].chunk};
        return EMP_OK;
    }

private:
    emp_ctx *c_;
    OocList l_;
    int gshift_;
    const std::vector<uint32_t> &cnt_;
};

// Tie keys of one list of n entries (d_tie[side]); with `unique`, also the check that the
// dedup's cq_get_partitions (cqutils.cuh:86-105) ends.
int ooc_ties(emp_ctx *c, const OocList &l, uint64_t n, bool rounds, int side, uint64_t M, bool unique) {
    const uint64_t F64 = n <= M ? 1 : (n + M - 1) / M;
    if (F64 > PLAN_MAX_CHUNKS)
        return fail(c, EMP_ERR_OUT_OF_CORE, "%llu entries need %llu sort chunks of M = %llu (at most %u): raise -r",
                    (unsigned long long)n, (unsigned long long)F64, (unsigned long long)M, PLAN_MAX_CHUNKS);
    const uint32_t F = (uint32_t)F64;
    const std::vector<uint32_t> io = plan_tree_io(F);
    int bb = 0;
    while ((1u << bb) < l.nb) ++bb;
    const int gshift = plan_group_shift(l.nb, F);
    const uint32_t G = std::max<uint32_t>(1u, l.nb >> gshift);
    OocTieArgs a{};
    a.l = l;
    a.rounds = rounds ? 1 : 0;
    a.M = M;
    a.F = F;
    a.io = c->d_oio;
    a.tie = c->d_tie[side];
    a.chunk = c->d_chunk;
    a.gshift = gshift;
    HIPCHK(c, hipMemcpyAsync(c->d_oio, io.data(), F * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    if (F > 1) {
        HIPCHK(c, hipMemsetAsync(c->d_obcnt, 0, (size_t)G * F * sizeof(uint32_t), c->stream));
        a.bcnt = c->d_obcnt;
        if (rounds) {  // rank of every id among the list's ids: bitmap, set bits per word, offsets
            const uint64_t words = (2 * c->n + 31) / 32;
            HIPCHK(c, hipMemsetAsync(c->d_obm, 0, words * sizeof(uint32_t), c->stream));
            HIPCHK(c, launch_ooc_mark(l, c->d_obm, c->stream));
            HIPCHK(c, launch_ooc_popc(c->d_obm, words, c->d_opc, c->stream));
            HIPCHK(c, launch_scan(c->d_opc, words, c->d_obase, c->d_oblk, c->stream));
            a.bitmap = c->d_obm;
            a.base = c->d_obase;
        }
    }
    {
        Timed t(c, "k_ooc_tie", 24.0 * (double)n);
        HIPCHK(c, launch_ooc_tie(a, c->stream));
    }
    if (F == 1) return sync(c);  // one chunk: the input order alone
    PlanList pl;
    pl.n = n;
    pl.M = M;
    pl.F = F;
    pl.G = G;
    pl.key_shift = 29 - bb + gshift;
    pl.rounds = a.rounds;
    pl.bcnt.resize((size_t)G * F);
    std::vector<uint32_t> cnt(l.nb);
    HIPCHK(c, hipMemcpyAsync(pl.bcnt.data(), c->d_obcnt, pl.bcnt.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(cnt.data(), l.cnt, l.nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    int rc = sync(c);
    if (rc) return rc;
    DevSource src(c, l, gshift, cnt);
    Planner p(pl, src);
    std::vector<uint32_t> ids;
    std::vector<unsigned long long> vals;
    if ((rc = p.overrides(ids, vals)))
        return rc == EMP_ERR_CAPACITY ? fail(c, rc, "out-of-core order: inconsistent counts of a list of %llu entries",
                                             (unsigned long long)n)
                                      : rc;
    if (unique && (rc = p.unique_check()))
        return rc == EMP_ERR_PARTITION
                   ? fail(c, rc, "dedup: a window of M = %llu entries holds one key (cq_get_partitions never ends)",
                          (unsigned long long)M)
                   : rc;
    if (!ids.empty()) {
        if ((rc = grow_dev(c, &c->d_ovid, c->ovid_cap, ids.size())) ||
            (rc = grow_dev(c, &c->d_ovval, c->ovval_cap, ids.size())))
            return rc;
        HIPCHK(c, hipMemcpyAsync(c->d_ovid, ids.data(), ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_ovval, vals.data(), vals.size() * sizeof(unsigned long long),
                                 hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, launch_ooc_override(c->d_ovid, c->d_ovval, (uint32_t)ids.size(), c->d_tie[side], c->stream));
    }
    return sync(c);
}

// Scratch offsets (powers of two of the counts) of every bucket of one or two lists; sizes the
// global-memory sort scratch.
int ooc_offsets(emp_ctx *c, uint32_t nb, const uint32_t *d_cnt_s, const uint32_t *d_cnt_p) {
    int rc;
    if (nb > c->ooff_nb) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int s = 0; s < 2; ++s) {
            dfree(c, c->d_ooff[s]);
            c->d_ooff[s] = nullptr;
            void *p = nullptr;
            if ((rc = dalloc(c, &p, (size_t)nb * sizeof(unsigned long long)))) return rc;
            c->d_ooff[s] = (unsigned long long *)p;
        }
        c->ooff_nb = nb;
    }
    std::vector<uint32_t> cs(nb), cp(d_cnt_p ? nb : 0);
    HIPCHK(c, hipMemcpyAsync(cs.data(), d_cnt_s, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    if (d_cnt_p) HIPCHK(c, hipMemcpyAsync(cp.data(), d_cnt_p, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    if ((rc = sync(c))) return rc;
    std::vector<unsigned long long> so(nb), po(nb);
    uint64_t as = 0, ap = 0;
    // as the kernels pad: an empty side of a bucket still takes one record
    auto p2 = [](uint64_t n) { uint64_t p = 1; while (p < n) p <<= 1; return p; };
    for (uint32_t b = 0; b < nb; ++b) {
        so[b] = as;
        as += p2(cs[b]);
        po[b] = ap;
        ap += d_cnt_p ? p2(cp[b]) : 0;
    }
    const uint64_t need = std::max<uint64_t>(std::max(as, ap), 1);
    if ((rc = ensure_big_scratch(c, need))) return rc;
    if (need > c->btie_cap) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int s = 0; s < 2; ++s) {
            dfree(c, c->d_btie[s]);
            c->d_btie[s] = nullptr;
            void *p = nullptr;
            if ((rc = dalloc(c, &p, need * sizeof(unsigned long long)))) return rc;
            c->d_btie[s] = (unsigned long long *)p;
        }
        c->btie_cap = need;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_ooff[0], so.data(), nb * sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_ooff[1], po.data(), nb * sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
    return sync(c);
}

BigArgs ooc_args(emp_ctx *c, const SideIn &sin, const SideIn *pin, const RoundArgs &rg, const NextOut &nx,
                 uint64_t slot0) {
    BigArgs a{};
    a.soff = c->d_ooff[0];
    a.poff = c->d_ooff[1];
    a.skey = c->d_bkey[0];
    a.pkey = c->d_bkey[1];
    a.srank = c->d_brank[0];
    a.prank = c->d_brank[1];
    a.saux = c->d_baux[0];
    a.paux = c->d_baux[1];
    a.stie = c->d_btie[0];
    a.ptie = c->d_btie[1];
    a.tie_s = c->d_tie[0];
    a.tie_p = c->d_tie[1];
    a.s_in = sin.fine;
    a.s_cnt = sin.cnt;
    a.s_spill = sin.spill;
    a.s_spill_n = sin.spill_n;
    a.spill_cap = sin.spill_cap;
    if (pin) {
        a.p_in = pin->fine;
        a.p_cnt = pin->cnt;
        a.p_spill = pin->spill;
        a.p_spill_n = pin->spill_n;
        a.p_spill_cap = pin->spill_cap;
    }
    a.cap = BUCKET_CAP;
    a.rg = rg;
    a.cols = c->d_cols;
    a.nx = nx;
    a.pool = c->d_pool;
    a.pool_cursor = c->d_ctr->pool;
    a.pool_cap = c->pool_cap;
    a.ovf_cursor = &c->d_ctr->pool_ovf;
    a.ovf_cap = c->pool_ovf_cap;
    a.edge_off = c->d_edge_off + slot0;
    a.edge_cnt = c->d_edge_cnt + slot0;
    a.too_many = &c->d_ctr->too_many;
    return a;
}

// ---------------------------------------------------------------------------------------------
// The overlap phase: dedup, then the rounds (synchronous or pipelined), then the output.

constexpr int RUN_FALLBACK = 1;  // pipelined rounds met an exceptional round: rerun synchronously

struct RunState {
    int rounds = 0, L = 0, bbD = 0;
    uint32_t nb0 = 0;     // buckets of round 0's lists (2^bbD unless the rounds are pipelined)
    uint32_t nbD = 0;     // buckets of the dedup partition (2^bbD, or nb0 for pipelined rounds)
    uint64_t M = 0, stride = 0, D = 0;
    int bbS[2] = {0, 0};  // bucket bits of the bucket set written for rounds of each parity
    int t = 0;            // rounds done
    bool streamed = false;  // the edges are already in the pinned arena, in output order
    std::vector<uint64_t> round_edges, rec_edges;  // edges per round; per record
    std::vector<int> rec_round;                      // round of each record
};

// Dedup (overlap.cu:212-225): partition all 2n entries, keep the smallest id per key. The kept
// entries (full-length keys), regenerated for round 0, go into bucket set 0 as round 0's suffix
// and prefix lists (overlap.cu:235-242). Round t's lists live in bucket set t % 2.
int dedup_phase(emp_ctx *c, RunState &st, uint64_t N0) {
    int rc;
    const int bbD = st.bbD, L = st.L;
    c->records.clear();
    c->out_used = 0;
    c->rounds.clear();
    c->audits.clear();
    st.round_edges.clear();
    st.rec_edges.clear();
    st.rec_round.clear();
    st.t = 0;
    st.streamed = false;
    st.bbS[0] = st.bbS[1] = bbD;
    if ((rc = ensure_pool(c, std::max<uint64_t>(4096, N0 / SUBS), std::max<uint64_t>(1u << 20, N0 / 8), false)))
        return rc;
    c->pool_prev.assign(SUBS + 2, 0ull);  // sub-range cursors, overflow cursor, abandoned edges
    if (st.rounds > 0 && (rc = ensure_edge_slots(c, (uint64_t)st.rounds * st.stride))) return rc;
    const RoundArgs none{};
    const RoundArgs rg0 = st.rounds > 0 ? round_args(c, 0, true) : none;
    const bool ooc = N0 > st.M;  // the out-of-core order decides which id of a key is kept
    bool need_part = true, ties = false;
    for (int attempt = 0;; ++attempt) {
        if (attempt > 8) return fail(c, EMP_ERR_CAPACITY, "dedup does not settle");
        {
            ZeroList z;
            if (need_part) zero_partition(c, z, N0, bbD);
            zero_next(c, z, 0, bbD);
            z.add((uint32_t *)c->d_ctr->pool, 2 * (SUBS + 2));  // edge-pool cursors
            if ((rc = run_zero(c, z))) return rc;
        }
        if (need_part && (rc = partition(c, N0, bbD, st.nbD == (1u << bbD) ? 0u : st.nbD))) return rc;
        if (ooc) {
            // cq_sort_by_key + cq_unique_by_key (overlap.cu:212-225, cqutils.cuh:179-247,271-306):
            // the first entry of every key in the merged order
            if ((rc = read_counters(c))) return rc;
            if (c->h_ctr->abort) {
                if ((rc = grow_after_abort(c, N0, bbD))) return rc;
                need_part = true;
                ties = false;
                continue;
            }
            need_part = false;
            const SideIn din{c->d_dfine, c->d_dcnt, c->d_dspill, &c->d_ctr->spill, c->dspill_cap};
            if (!ties) {
                const OocList l{din.fine, din.cnt, BUCKET_CAP, din.spill, din.spill_n, din.spill_cap, -1, 1u << bbD};
                if ((rc = ensure_ooc(c, N0)) || (rc = ooc_ties(c, l, N0, false, 0, st.M, true)) ||
                    (rc = ooc_offsets(c, 1u << bbD, din.cnt, nullptr)))
                    return rc;
                ties = true;
            }
            BigArgs a = ooc_args(c, din, nullptr, none, next_out_nb(c, 0, st.nb0, 1u << bbD, rg0), 0);
            {
                Timed t(c, "k_dedup_ooc", 0.0);
                HIPCHK(c, launch_dedup_ooc(a, 1u << bbD, c->stream));
            }
            if ((rc = read_counters(c))) return rc;
            if (c->h_ctr->r[0].nx_ovf) {
                for (int side = 0; side < 2; ++side)
                    if ((rc = grow_bspill(c, 0, side, c->bspill_cap[0][side] * 4))) return rc;
                continue;
            }
            break;
        }
        DedupArgs a{};
        a.in = c->d_dfine;
        a.cnt = c->d_dcnt;
        a.pay = c->d_ends ? c->d_dpay : nullptr;
        a.nx = next_out_nb(c, 0, st.nb0, st.nbD, rg0);
        a.big = c->d_big;
        a.n_big = &c->d_ctr->r[0].n_big;
        a.big_cap = BIG_CAP;
        a.abort = &c->d_ctr->abort;
        {
            Timed t(c, "k_dedup", (c->d_ends ? 16.0 : 12.0) * N0);  // entries and their payload
            HIPCHK(c, launch_dedup(a, st.nbD, c->stream));
        }
        if ((rc = read_counters(c))) return rc;
        if (c->h_ctr->abort) {
            if ((rc = grow_after_abort(c, N0, bbD))) return rc;
            need_part = true;
            continue;
        }
        need_part = false;
        const ListCtr &r0 = c->h_ctr->r[0];
        const SideIn din{c->d_dfine, c->d_dcnt, c->d_dspill, &c->d_ctr->spill, c->dspill_cap};
        if (!r0.nx_ovf && r0.n_big &&
            (rc = big_path(c, r0.n_big, bbD, true, none, 0, din, din, next_out_nb(c, 0, st.nb0, st.nbD, rg0))))
            return rc;
        if (c->h_ctr->r[0].nx_ovf) {
            for (int side = 0; side < 2; ++side)
                if ((rc = grow_bspill(c, 0, side, c->bspill_cap[0][side] * 4))) return rc;
            continue;
        }
        break;
    }
    // the 2n entries before dedup, in the dedup partition's buckets (overlap.cu:215-219)
    if (c->audit && (rc = audit_list(c, 0, L, N0, c->d_dfine, c->d_dcnt, c->d_dspill, &c->d_ctr->spill,
                                     c->dspill_cap, 1u << bbD, 0, L, false)))
        return rc;
    st.D = c->h_ctr->r[0].total[0];
    if (c->opts.profile && !ooc) {  // per kept entry: two 12-byte entries written (both sides)
        int id = kstat_id(c, "k_dedup");
        c->kstats[id].bytes += 24.0 * (double)st.D;
    }
    return EMP_OK;
}

// Rounds with the host between them: after every join the counters come back, so that a large
// bucket, a spill or pool overflow is handled (and the round rerun) at once, and framing and the
// audit can read the round's lists.
int rounds_sync(emp_ctx *c, RunState &st) {
    int rc;
    const int L = st.L, rounds = st.rounds;
    const uint64_t M = st.M, stride = st.stride;
    const RoundArgs none{};
    uint64_t St = c->h_ctr->r[0].total[0], Pt = c->h_ctr->r[0].total[1];
    int t = 0;
    for (; t < rounds && St > 0; ++t) {
        // No record can follow an empty prefix list. (The lists mirror each other under reverse
        // complement, so they empty together; the audit's records end here too.)
        if (Pt == 0) break;
        const bool framed = St >= M / 2 || Pt >= M / 2;  // several records (frame_round)
        const int sset = t & 1, nset = sset ^ 1;
        const int bb = st.bbS[sset];
        const uint32_t nbt = 1u << bb;
        const bool more = t + 1 < rounds;
        // the next round's buckets, chosen now because this round's join writes them
        const int bbn = more ? std::min(bb, round_bits(std::max(St, Pt))) : bb;
        const RoundArgs rg = round_args(c, t, true);
        const RoundArgs rgn = more ? round_args(c, t + 1, true) : none;
        emp_round_stats rs{};
        rs.s = St;
        rs.p = Pt;
        rs.ovlp = L - 1 - t;
        rs.bucket_bits = bb;
        rs.buckets = 1u << bb;
        const uint64_t slot0 = (uint64_t)t * stride;
        const SideIn sin = side_in(c, t, 0), pin = side_in(c, t, 1);
        // a list over M: cq_sort_by_key's merge order (overlap.cu:271-272) as tie keys per side
        const bool ooc = St > M || Pt > M;
        if (ooc) {
            const OocList ls{sin.fine, sin.cnt, BUCKET_CAP, sin.spill, sin.spill_n, sin.spill_cap, 29 - bb, nbt};
            const OocList lp{pin.fine, pin.cnt, BUCKET_CAP, pin.spill, pin.spill_n, pin.spill_cap, 29 - bb, nbt};
            if ((rc = ensure_ooc(c, 2 * c->n)) || (rc = ooc_ties(c, ls, St, true, 0, M, false)) ||
                (rc = ooc_ties(c, lp, Pt, true, 1, M, false)) || (rc = ooc_offsets(c, nbt, sin.cnt, pin.cnt)))
                return rc;
        }
        for (int attempt = 0;; ++attempt) {
            if (attempt > 8) return fail(c, EMP_ERR_CAPACITY, "round %d does not settle", t);
            if (attempt > 0)  // a rerun: the pool cursors go back to the end of the previous round
                HIPCHK(c, hipMemcpyAsync(c->d_ctr->pool, c->pool_prev.data(), (SUBS + 2) * sizeof(unsigned long long),
                                         hipMemcpyHostToDevice, c->stream));
            {
                ZeroList z;
                zero_next(c, z, t + 1, bbn);
                if ((rc = run_zero(c, z))) return rc;
            }
            if (ooc) {
                BigArgs a = ooc_args(c, sin, &pin, rg, next_out(c, t + 1, bbn, bb, rgn), slot0);
                {
                    Timed tm(c, "k_join_ooc", 0.0);
                    HIPCHK(c, launch_join_ooc(a, nbt, c->stream));
                }
                if ((rc = read_counters(c))) return rc;
                if (c->h_ctr->too_many)  // the reference's int edge offsets (overlap.cu:561,567)
                    return fail(c, EMP_ERR_PARTITION, "a key group holds 2^31 edges or more");
            } else {
                JoinArgs a{};
                a.s_in = sin.fine;
                a.p_in = pin.fine;
                a.s_cnt = sin.cnt;
                a.p_cnt = pin.cnt;
                a.zero_in = 0;
                a.nx = next_out(c, t + 1, bbn, bb, rgn);
                a.pool = c->d_pool;
                a.pool_cursor = c->d_ctr->pool;
                a.pool_cap = c->pool_cap;
                a.ovf_cursor = &c->d_ctr->pool_ovf;
                a.ovf_cap = c->pool_ovf_cap;
                a.edge_off = c->d_edge_off + slot0;
                a.edge_cnt = c->d_edge_cnt + slot0;
                a.big = c->d_big;
                a.n_big = &c->d_ctr->r[t + 1].n_big;
                a.big_cap = BIG_CAP;
                a.abort = &c->d_ctr->abort;
                {
                    Timed tm(c, "k_join", 0.0);
                    HIPCHK(c, launch_join(a, nbt, c->stream));
                }
                if ((rc = read_counters(c))) return rc;
                const ListCtr &rn = c->h_ctr->r[t + 1];
                rs.big_buckets = (int32_t)rn.n_big;
                if (!rn.nx_ovf && rn.n_big &&
                    (rc = big_path(c, rn.n_big, bb, false, rg, slot0, sin, pin, next_out(c, t + 1, bbn, bb, rgn))))
                    return rc;
            }
            if (c->h_ctr->r[t + 1].nx_ovf) {
                for (int side = 0; side < 2; ++side)
                    if ((rc = grow_bspill(c, nset, side, c->bspill_cap[nset][side] * 4))) return rc;
                continue;
            }
            const uint64_t ovf = c->h_ctr->pool_ovf;
            if (ovf > c->pool_ovf_cap) {
                if ((rc = ensure_pool(c, 0, std::max<uint64_t>(2 * c->pool_ovf_cap, ovf + ovf / 4 + 1024), true)))
                    return rc;
                continue;
            }
            break;
        }
        uint64_t E = 0;
        for (int i = 0; i <= SUBS + 1; ++i) {  // placed = reserved - abandoned
            const unsigned long long v =
                i < SUBS ? c->h_ctr->pool[i] : i == SUBS ? c->h_ctr->pool_ovf : c->h_ctr->pool_lost;
            if (i <= SUBS)
                E += v - c->pool_prev[i];
            else
                E -= v - c->pool_prev[i];
            c->pool_prev[i] = v;
        }
        const DevCounters &k = *c->h_ctr;
        const uint64_t Sn = k.r[t + 1].total[0], Pn = k.r[t + 1].total[1];
        rs.edges = E;
        rs.s_after = Sn;
        rs.p_after = Pn;
        if (c->opts.profile) {
            int id = kstat_id(c, "k_join");
            c->kstats[id].bytes += 12.0 * (St + Pt) + (more ? 12.0 * (Sn + Pn) : 0.0) + 8.0 * E;
        }
        // one record per partition with both sides non-empty (overlap.cu:398-422; written even when
        // E == 0): one per round unless a list reaches M/2
        size_t nrec0 = st.rec_edges.size();
        if (framed) {
            FrameSide fs, fp;
            const uint32_t nss = std::min(k.r[t].spill[0], sin.spill_cap);
            const uint32_t nps = std::min(k.r[t].spill[1], pin.spill_cap);
            if ((rc = frame_side(c, fs, sin.fine, sin.cnt, sin.spill, nss, nbt, St)) ||
                (rc = frame_side(c, fp, pin.fine, pin.cnt, pin.spill, nps, nbt, Pt)) ||
                (rc = frame_round(c, bb, fs, fp, c->d_edge_cnt + slot0, E, M, st.rec_edges)))
                return rc;
        } else {
            st.rec_edges.push_back(E);
        }
        if (c->audit) {  // the lists as the round joins them (overlap.cu:274-284)
            const int ell = L - 1 - t;
            if ((rc = audit_list(c, 1, ell, St, sin.fine, sin.cnt, sin.spill, sin.spill_n, sin.spill_cap, nbt, t + 1,
                                 ell, true)) ||
                (rc = audit_list(c, 2, ell, Pt, pin.fine, pin.cnt, pin.spill, pin.spill_n, pin.spill_cap, nbt, 0, ell,
                                 true)))
                return rc;
        }
        for (size_t i = nrec0; i < st.rec_edges.size(); ++i) {
            if (st.rec_edges[i] >= (1ull << 31))  // the reference's int edge offsets (overlap.cu:561,567)
                return fail(c, EMP_ERR_PARTITION, "round %d: %llu edges in one record reach 2^31", t,
                            (unsigned long long)st.rec_edges[i]);
            st.rec_round.push_back(t);
        }
        st.round_edges.push_back(E);
        rs.records = (int32_t)(st.rec_edges.size() - nrec0);
        c->rounds.push_back(rs);
        St = Sn;
        Pt = Pn;
        st.bbS[nset] = bbn;
    }
    st.t = t;
    return EMP_OK;
}

int file_sink(void *user, uint64_t n, int32_t ovlp, const uint32_t *from, const uint32_t *to);

// Writes records to the output file in order while the rounds run: each waits for the DMA of
// its round's edges into the pinned host arena (its event), then goes out as write_edges_binary
// lays it out (overlap.cu:148-161).
struct RecordWriter {
    OutFile *f;
    const uint32_t *from, *to;
    struct Item {
        uint64_t n, off;
        int32_t ovlp;
        hipEvent_t ev;
    };
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Item> q;
    size_t head = 0;
    bool done = false;
    int err = 0;
    std::thread th;
    int device;
    RecordWriter(OutFile *f_, const uint32_t *fr, const uint32_t *t, int dev) : f(f_), from(fr), to(t), device(dev) {
        th = std::thread([this]() {
            (void)hipSetDevice(device);
            for (;;) {
                Item it;
                {
                    std::unique_lock<std::mutex> g(mu);
                    cv.wait(g, [this]() { return head < q.size() || done; });
                    if (head == q.size()) return;
                    it = q[head++];
                }
                if (err) continue;
                if (hipEventSynchronize(it.ev) != hipSuccess) {
                    err = EMP_ERR_HIP;
                    continue;
                }
                err = f->append(it.n, it.ovlp, from + it.off, to + it.off);
                // every later record already in host memory joins this write (the direct side
                // takes a share of the file while the buffered writes lag behind)
                for (;;) {
                    {
                        std::lock_guard<std::mutex> g(mu);
                        if (err || head == q.size() || hipEventQuery(q[head].ev) != hipSuccess) break;
                        it = q[head++];
                    }
                    err = f->append(it.n, it.ovlp, from + it.off, to + it.off);
                }
                if (!err) err = f->flush();
            }
        });
    }
    void push(uint64_t n, uint64_t off, int32_t ovlp, hipEvent_t ev) {
        std::lock_guard<std::mutex> g(mu);
        q.push_back(Item{n, off, ovlp, ev});
        cv.notify_one();
    }
    int finish() {
        {
            std::lock_guard<std::mutex> g(mu);
            done = true;
            cv.notify_one();
        }
        if (th.joinable()) th.join();
        return err;
    }
    ~RecordWriter() { finish(); }
};

// Rounds back to back: no host synchronisation between joins. Each join clears the bucket counts
// it has read (the next writer of that set needs them zero) and fills round t+1's counter record;
// the host reads each record back one round late, only to size the buckets of the round after
// next (from an upper bound of the list sizes: lists only shrink) and to stop at an empty list.
// Any large bucket, spill or pool overflow returns RUN_FALLBACK: the rounds after it read
// incomplete lists, so the run restarts from the dedup on the synchronous path.
int rounds_pipelined(emp_ctx *c, RunState &st) {
    int rc;
    const int L = st.L, rounds = st.rounds;
    const uint64_t stride = st.stride;
    const RoundArgs none{};
    // Output while the rounds run, on a second stream: each round's edges are compacted into the
    // device arena in output order as soon as its join is done, and every finished prefix of the
    // arena goes on to the pinned host arena by DMA. The arena holds as many edges as the pool;
    // nothing is pending in it at the start of a run.
    if ((rc = ensure_arena(c, c->pool_cap * SUBS + c->pool_ovf_cap, nullptr, nullptr))) return rc;
    if (!c->ostream) HIPCHK(c, hipStreamCreateWithFlags(&c->ostream, hipStreamNonBlocking));
    if (!c->dstream) HIPCHK(c, hipStreamCreateWithFlags(&c->dstream, hipStreamNonBlocking));
    if (!c->h_orec) {
        void *p = nullptr, *dp = nullptr;
        HIPCHK(c, hipHostMalloc(&p, MAX_ROUNDS * 3 * sizeof(unsigned long long), hipHostMallocMapped));
        c->h_orec = (unsigned long long(*)[3])p;
        HIPCHK(c, hipHostGetDevicePointer(&dp, p, 0));
        c->d_orec_view = (unsigned long long(*)[3])dp;
    }
    if (c->orec_ev.empty()) {
        c->orec_ev.resize(MAX_ROUNDS);
        for (auto &e : c->orec_ev) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    // With an output file, every round whose record is final goes to the file while later rounds
    // run: its join handed no bucket to the large-bucket path and overflowed no spill area, the
    // lists it read were complete (the same for every earlier round, or the run has already
    // fallen back), and no edge of it or of an earlier round missed the pool (the overflow
    // cursor, read after its join, is within the pool). A run that falls back later repeats those
    // rounds with the same records; they are not written again.
    std::unique_ptr<RecordWriter> writer;
    if (c->out_file) {
        if (c->dma_ev.empty()) {
            c->dma_ev.resize(MAX_ROUNDS);
            for (auto &e : c->dma_ev) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        writer.reset(new RecordWriter(c->out_file, c->h_from, c->h_to, c->device));
    }
    bool emitting = writer != nullptr;
    auto emit_round = [&](int r) -> int {  // the edges of rounds <= r are in the device arena
        if (!emitting) return EMP_OK;
        const unsigned long long *o = c->h_orec[r];
        if ((uint64_t)r != c->emitted || o[2] > c->pool_ovf_cap || o[0] >= (1ull << 31)) {
            emitting = false;  // not final here: this round and the ones after it wait for the end
            return EMP_OK;
        }
        HIPCHK(c, hipEventRecord(c->dma_ev[r], c->dstream));
        writer->push(o[0], o[1], L - 1 - r, c->dma_ev[r]);
        ++c->emitted;
        return EMP_OK;
    };
    uint64_t copied = 0;  // arena edges already sent to the host
    auto send = [&](uint64_t upto) -> int {
        upto = std::min<uint64_t>(upto, c->out_cap);  // placed edges never exceed the arena (the pool's size)
        if (upto <= copied) return EMP_OK;
        const size_t bytes = (size_t)(upto - copied) * 4;
        // on a stream of its own: a copy of several MB queued on ostream would hold the scans of
        // later rounds, whose events the host waits on before it launches the next joins
        HIPCHK(c, hipMemcpyAsync(c->h_from + copied, c->d_out_from + copied, bytes, hipMemcpyDeviceToHost, c->dstream));
        HIPCHK(c, hipMemcpyAsync(c->h_to + copied, c->d_out_to + copied, bytes, hipMemcpyDeviceToHost, c->dstream));
        copied = upto;
        return EMP_OK;
    };
    // round 1's bucket set (set 1), every round's counter record and the output position start
    // at zero; set 0 was cleared before the dedup wrote it
    {
        ZeroList z;
        z.add(c->d_bcnt[1][0], 1ull << st.bbD);
        z.add(c->d_bcnt[1][1], 1ull << st.bbD);
        z.add(reinterpret_cast<uint32_t *>(&c->d_ctr->r[1]), (uint64_t)rounds * sizeof(ListCtr) / 4);
        z.add(reinterpret_cast<uint32_t *>(&c->d_ctr->obase), 2);
        if ((rc = run_zero(c, z))) return rc;
    }
    auto fallback = [&]() {
        (void)hipStreamSynchronize(c->ostream);
        (void)hipStreamSynchronize(c->dstream);
        if (writer && writer->finish()) return fail(c, EMP_ERR_IO, "writing the output file failed");
        // the synchronous rerun refills the arena: nothing may read it for the file any more
        if (c->out_file && c->out_file->settle() != EMP_OK) return fail(c, EMP_ERR_IO, "writing the output file failed");
        return RUN_FALLBACK;
    };
    if (!c->h_rec) {
        void *p = nullptr, *dp = nullptr;
        HIPCHK(c, hipHostMalloc(&p, (MAX_ROUNDS + 1) * sizeof(ListCtr), hipHostMallocMapped));
        c->h_rec = (ListCtr *)p;
        HIPCHK(c, hipHostGetDevicePointer(&dp, p, 0));
        c->d_rec_view = (ListCtr *)dp;
    }
    if (c->rec_ev.empty()) {
        c->rec_ev.resize(MAX_ROUNDS);
        for (auto &e : c->rec_ev) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    std::vector<uint32_t> nbs(rounds + 1, st.nb0);  // buckets of round t's lists
    c->h_rec[0] = c->h_ctr->r[0];
    uint64_t Sk = c->h_rec[0].total[0], Pk = c->h_rec[0].total[1];  // newest sizes known
    int launched = 0;
    for (int t = 0; t < rounds; ++t) {
        // Round t-3's edge scan has ended: the record of round t-2 (written by join t-3) is in
        // host memory, and so are round t-3's edges in the device arena. Check that record and
        // size round t+1 from it (lists only shrink). The scan of a round runs beside the next
        // join and often ends with it (the joins fill every CU), so the host looks three rounds
        // back: it then never holds the next join's launch.
        if (t >= 3) {
            HIPCHK(c, hipEventSynchronize(c->orec_ev[t - 3]));
            const ListCtr &r = c->h_rec[t - 2];
            if (r.n_big || r.nx_ovf) return fallback();
            Sk = r.total[0];
            Pk = r.total[1];
            if (Sk == 0 || Pk == 0) break;  // round t-2 writes no record, nor any later round
            if ((rc = send(c->h_orec[t - 3][1] + c->h_orec[t - 3][0]))) return rc;
            if ((rc = emit_round(t - 3))) return rc;
        }
        const bool more = t + 1 < rounds;
        const uint32_t nb = nbs[t];
        nbs[t + 1] = more ? std::min(nb, round_buckets(std::max(Sk, Pk))) : nb;
        const RoundArgs rgn = more ? round_args(c, t + 1, true) : none;
        const uint64_t slot0 = (uint64_t)t * stride;
        const SideIn sin = side_in(c, t, 0), pin = side_in(c, t, 1);
        JoinArgs a{};
        a.s_in = sin.fine;
        a.p_in = pin.fine;
        a.s_cnt = sin.cnt;
        a.p_cnt = pin.cnt;
        a.zero_in = 1;
        a.nx = next_out_nb(c, t + 1, nbs[t + 1], nb, rgn);
        a.pool = c->d_pool;
        a.pool_cursor = c->d_ctr->pool;
        a.pool_cap = c->pool_cap;
        a.ovf_cursor = &c->d_ctr->pool_ovf;
        a.ovf_cap = c->pool_ovf_cap;
        a.edge_off = c->d_edge_off + slot0;
        a.edge_cnt = c->d_edge_cnt + slot0;
        a.big = c->d_big;
        a.n_big = &c->d_ctr->r[t + 1].n_big;
        a.big_cap = BIG_CAP;
        a.abort = &c->d_ctr->abort;
        {
            Timed tm(c, "k_join", 0.0);
            HIPCHK(c, launch_join(a, nb, c->stream));
        }
        HIPCHK(c, hipEventRecord(c->rec_ev[t], c->stream));
        // the round's edges, in output order, into the device arena while later rounds run; the
        // scan also writes the round's output record and the next round's counter record into
        // pinned host memory, so the joins run back to back with no copy between them
        HIPCHK(c, hipStreamWaitEvent(c->ostream, c->rec_ev[t], 0));
        const RecMirror mirror{c->d_orec_view[t], reinterpret_cast<uint32_t *>(&c->d_rec_view[t + 1]),
                               reinterpret_cast<const uint32_t *>(&c->d_ctr->r[t + 1]), sizeof(ListCtr) / 4};
        HIPCHK(c, launch_round_edges(c->d_edge_cnt + slot0, c->d_edge_off + slot0, c->d_dst_off + slot0, nb,
                                     &c->d_ctr->obase, c->d_ctr->orec[t], &c->d_ctr->pool_ovf, c->d_pool,
                                     c->d_out_from, c->d_out_to, mirror, c->ostream));
        HIPCHK(c, hipEventRecord(c->orec_ev[t], c->ostream));
        launched = t + 1;
    }
    // every launched round's record, the pool cursors and the edges per round; the rest of the
    // arena to the host
    HIPCHK(c, hipStreamSynchronize(c->ostream));
    if ((rc = read_counters(c))) return rc;
    if (launched && (rc = send(c->h_ctr->obase))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->dstream));
    const DevCounters &k = *c->h_ctr;
    if (k.pool_ovf > c->pool_ovf_cap) return fallback();
    int t = 0;
    for (; t < launched; ++t) {
        const ListCtr &r = k.r[t + 1];
        if (r.n_big || r.nx_ovf) return fallback();
    }
    if (writer && writer->finish()) return fail(c, EMP_ERR_IO, "writing the output file failed");
    for (t = 0; t < launched; ++t) {
        const uint64_t St = k.r[t].total[0], Pt = k.r[t].total[1];
        if (St == 0 || Pt == 0) break;
        emp_round_stats rs{};
        rs.s = St;
        rs.p = Pt;
        rs.s_after = k.r[t + 1].total[0];
        rs.p_after = k.r[t + 1].total[1];
        rs.edges = k.orec[t][0];
        rs.ovlp = L - 1 - t;
        rs.records = 1;
        rs.buckets = nbs[t];
        rs.bucket_bits = ceil_log2(nbs[t]);
        if (c->opts.profile) {
            int id = kstat_id(c, "k_join");
            c->kstats[id].bytes += 12.0 * (St + Pt) + (t + 1 < rounds ? 12.0 * (rs.s_after + rs.p_after) : 0.0) +
                                   8.0 * rs.edges;
        }
        if (rs.edges >= (1ull << 31))  // the reference's int edge offsets (overlap.cu:561,567)
            return fail(c, EMP_ERR_PARTITION, "round %d: %llu edges in one record reach 2^31", t,
                        (unsigned long long)rs.edges);
        st.rec_edges.push_back(rs.edges);
        st.rec_round.push_back(t);
        st.round_edges.push_back(rs.edges);
        c->rounds.push_back(rs);
    }
    st.t = t;
    st.streamed = true;
    return EMP_OK;
}

int run_impl(emp_ctx *c, emp_record_sink sink, void *user) {
    int rc = flush_acc(c);
    if (rc) return rc;
    rc = stage_submit(c);
    if (rc) return rc;
    rc = sync(c);
    if (rc) return rc;
    for (int s = 0; s < 2; ++s) c->stage_busy[s] = false;
    uint32_t herr = 0;
    HIPCHK(c, hipMemcpy(&herr, c->d_err, sizeof herr, hipMemcpyDeviceToHost));
    if (herr) return fail(c, EMP_ERR_BASE, "a kept read holds a byte other than ACGTacgt");

    c->rounds.clear();
    c->audits.clear();
    emp_summary &S = c->summary;
    S.n_reads = c->n;
    S.read_len = c->L;
    S.n_rounds = 0;
    S.n_dedup = 0;
    S.n_records = 0;
    S.n_edges = 0;
    const uint64_t n = c->n;
    if (n == 0) return EMP_OK;  // overlap.cu:231,258: an empty file
    const int L = c->L;
    const int rounds = L - c->opts.min_ovlp;
    const uint64_t N0 = 2 * n;
    const uint64_t M = c->opts.reduce_mem / 16;
    if (M < 2) return fail(c, EMP_ERR_ARG, "reduce_mem / 16 must be at least 2");

    const int bbD = round_bits(N0);  // the dedup's buckets and round 0's
    if ((rc = ensure_run_buffers(c, N0, bbD))) return rc;
    const uint64_t stride = 1ull << bbD;  // edge slots per round
    RunState st{};
    st.rounds = rounds;
    st.L = L;
    st.M = M;
    st.bbD = bbD;
    st.stride = stride;
    // Rounds run back to back when nothing needs the host between them: no audit and no
    // framing (lists below M/2; they only shrink). An exceptional round (a bucket over capacity,
    // a spill or pool overflow) sends the run to the synchronous rounds, from the dedup on.
    bool pipelined = !c->audit && c->pipeline;
    for (int pass = 0;; ++pass) {
        // Round 0's lists get a bucket count sized to the list when the rounds will run pipelined
        // (the dedup keeps at most N0 entries, so N0 < M/2 settles it now); the synchronous rounds
        // (framing, the audit, the fallback) keep powers of two.
        const bool sized = pipelined && N0 < M / 2 && rounds > 0;
        st.nb0 = sized ? std::min<uint32_t>(1u << bbD, round_buckets(N0)) : 1u << bbD;
        // the dedup partition's buckets sized the same way where its plan allows (two passes, the
        // first on the top 8 key bits): otherwise half full for a list just above a power of two
        st.nbD = sized && bbD > 8 && bbD <= 8 + PART_MAX_BITS ? st.nb0 : 1u << bbD;
        if ((rc = dedup_phase(c, st, N0))) return rc;
        S.n_dedup = st.D;
        if (pipelined && st.D < M / 2 && st.D > 0 && rounds > 0) {
            rc = rounds_pipelined(c, st);
            if (rc == RUN_FALLBACK && pass == 0) {
                pipelined = false;
                continue;
            }
        } else {
            rc = rounds_sync(c, st);
        }
        if (rc) return rc;
        break;
    }
    for (const emp_round_stats &rs : c->rounds) {
        S.n_records += (uint64_t)rs.records;
        S.n_edges += rs.edges;
    }
    S.n_rounds = st.t;

    // Output: compact the edges of every round from the pool into the arena in (round, bucket)
    // order, then hand the records to the sink.
    uint64_t total = 0;
    for (uint64_t e : st.round_edges) total += e;
    if (!st.streamed && (rc = ensure_arena(c, std::max<uint64_t>(total, 1), sink, user))) return rc;
    if (total && !st.streamed) {
        EdgeCopyArgs ea{c->d_edge_cnt, c->d_edge_off, c->d_dst_off, c->d_eblk, (uint64_t)st.t * stride,
                        c->d_pool, c->d_out_from, c->d_out_to};
        Timed tm(c, "k_edge_compact", 16.0 * total + 20.0 * (double)st.t * stride);
        HIPCHK(c, launch_edge_compact(ea, c->stream));
    }
    uint64_t off = 0;  // the arena holds the rounds' edges back to back; records slice them in order
    for (size_t i = 0; i < st.rec_edges.size(); ++i) {
        c->records.push_back(emp_ctx::Record{st.rec_edges[i], off, L - 1 - st.rec_round[i]});
        off += st.rec_edges[i];
    }
    c->out_used = total;
    c->out_streamed = st.streamed;
    return flush_records(c, sink, user);
}

// One record as write_edges_binary lays it out (overlap.cu:148-161): n (8 bytes), the overlap
// (4 bytes), from[n], to[n], appended to the output file (emp_outfile.h).
int file_sink(void *user, uint64_t n, int32_t ovlp, const uint32_t *from, const uint32_t *to) {
    OutFile *f = static_cast<OutFile *>(user);
    const int rc = f->append(n, ovlp, from, to);
    return rc ? rc : f->flush();
}

}  // namespace

// ---------------------------------------------------------------------------------------------
extern "C" {

const char *emp_version(void) { return "empanada-mi355x 0.1 (gfx950)"; }

void emp_opts_default(emp_opts *o) {
    memset(o, 0, sizeof *o);
    o->min_ovlp = 31;
    o->map_mem = 1ull << 30;
    o->reduce_mem = 1ull << 30;
    o->base_blocksz = 32ull << 20;
    o->kv_chunksz = 64ull << 20;
}

int emp_ctx_create(const emp_opts *opts, emp_ctx **out) {
    if (!out) return EMP_ERR_ARG;
    *out = nullptr;
    emp_ctx *c = new emp_ctx();
    if (opts)
        c->opts = *opts;
    else
        emp_opts_default(&c->opts);
    if (c->opts.min_ovlp < 0) {
        delete c;
        return EMP_ERR_ARG;
    }
    c->device = c->opts.device;
    // EMPANADA_SYNC_ROUNDS=1: the host between every two rounds (the path that large buckets,
    // framing and the audit take anyway)
    const char *sync_env = getenv("EMPANADA_SYNC_ROUNDS");
    c->pipeline = !(sync_env && *sync_env && *sync_env != '0');
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) {
        delete c;
        return EMP_ERR_HIP;
    }
    if (c->opts.stream) {
        c->stream = (hipStream_t)c->opts.stream;
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return EMP_ERR_HIP;
        }
        c->own_stream = true;
    }
    void *p = nullptr;
    if (dalloc(c, &p, sizeof(uint32_t)) || hipMemset(p, 0, sizeof(uint32_t)) != hipSuccess) {
        emp_ctx_destroy(c);
        return EMP_ERR_NOMEM;
    }
    c->d_err = (uint32_t *)p;
    if (dalloc(c, &p, sizeof(DevCounters)) || dalloc(c, (void **)&c->d_big, BIG_CAP * sizeof(uint32_t))) {
        emp_ctx_destroy(c);
        return EMP_ERR_NOMEM;
    }
    c->d_ctr = (DevCounters *)p;
    if (hipHostMalloc((void **)&c->h_ctr, sizeof(DevCounters), hipHostMallocDefault) != hipSuccess) {
        emp_ctx_destroy(c);
        return EMP_ERR_HIP;
    }
    *out = c;
    return EMP_OK;
}

void emp_ctx_destroy(emp_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ostream) (void)hipStreamSynchronize(c->ostream);
    if (c->dstream) (void)hipStreamSynchronize(c->dstream);
    if (c->cstream) (void)hipStreamSynchronize(c->cstream);
    for (auto &a : c->allocs) (void)hipFree(a.first);
    c->allocs.clear();
    for (char *p : c->h_chunk)
        if (p) (void)hipHostFree(p);
    for (int i = 0; i < 2; ++i) {
        if (c->h_stage[i]) (void)hipHostFree(c->h_stage[i]);
        if (c->stage_ev[i]) (void)hipEventDestroy(c->stage_ev[i]);
        if (c->pin_ev[i]) (void)hipEventDestroy(c->pin_ev[i]);
        if (c->pin_done[i]) (void)hipEventDestroy(c->pin_done[i]);
    }
    for (int i = 0; i < 3; ++i) {
        if (c->h_big[i]) (void)hipHostFree(c->h_big[i]);
        if (c->big_done[i]) (void)hipEventDestroy(c->big_done[i]);
    }
    if (c->h_from) (void)hipHostFree(c->h_from);
    if (c->h_to) (void)hipHostFree(c->h_to);
    if (c->h_ctr) (void)hipHostFree(c->h_ctr);
    if (c->h_fq_res) (void)hipHostFree(c->h_fq_res);
    if (c->h_rec) (void)hipHostFree(c->h_rec);
    for (auto e : c->rec_ev) (void)hipEventDestroy(e);
    if (c->h_orec) (void)hipHostFree(c->h_orec);
    for (auto e : c->orec_ev) (void)hipEventDestroy(e);
    for (auto e : c->dma_ev) (void)hipEventDestroy(e);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : c->free_events) (void)hipEventDestroy(e);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    if (c->ostream) (void)hipStreamDestroy(c->ostream);
    if (c->dstream) (void)hipStreamDestroy(c->dstream);
    if (c->cstream) (void)hipStreamDestroy(c->cstream);
    delete c;
}

const char *emp_last_error(const emp_ctx *c) { return c ? c->err.c_str() : "no context"; }

int emp_add_fastq(emp_ctx *c, const char *data, size_t size, int final, size_t *consumed) {
    if (!c || (!data && size) || !consumed) return EMP_ERR_ARG;
    (void)hipSetDevice(c->device);
    return add_fastq_range(c, data, size, size, true, final != 0, consumed, false);
}

int emp_add_fastq_file(emp_ctx *c, const char *path) {
    if (!c || !path) return EMP_ERR_ARG;
    (void)hipSetDevice(c->device);
    return add_fastq_file(c, path);
}

int emp_add_fastq_device(emp_ctx *c, const void *d_data, size_t size, int final, size_t *consumed) {
    if (!c || (!d_data && size) || !consumed) return EMP_ERR_ARG;
    (void)hipSetDevice(c->device);
    const uint64_t B = c->opts.map_mem / 16;  // main.cu:16
    if (B == 0) return fail(c, EMP_ERR_ARG, "map_mem / 16 is zero");
    return add_fastq_range(c, static_cast<const char *>(d_data), size, B, false, final != 0, consumed, true);
}

int emp_add_reads(emp_ctx *c, const char *reads, uint64_t n, int32_t read_len) {
    if (!c || (!reads && n)) return EMP_ERR_ARG;
    if (n == 0) return EMP_OK;
    (void)hipSetDevice(c->device);
    int rc = set_len(c, read_len);
    if (rc) return rc;
    hipPointerAttribute_t attr;
    const bool pinned = hipPointerGetAttributes(&attr, reads) == hipSuccess && attr.type == hipMemoryTypeHost;
    (void)hipGetLastError();  // pageable memory is an error of the query, not of the context
    uint64_t i = 0;
    const bool large = n * (uint64_t)read_len >= 2 * PIN_BATCH;
    if ((pinned || large) && c->stage_n == 0 && c->acc_n == 0 && c->n % 16 == 0) {
        const uint64_t m = n & ~15ull;  // the rest (fewer than 16) is staged, so later batches still start at a multiple of 16
        if ((rc = add_reads_pinned(c, reads, m, pinned))) return rc;
        i = m;
    }
    // staged: whole spans of reads into the pinned stage (stage_read sets the stage up and moves
    // reads parsed on the device ahead of them)
    while (i < n) {
        if (c->acc_n || !c->h_stage[0] || c->stage_bytes < (size_t)STAGE_READS * c->L) {
            if ((rc = stage_read(c, reads + i * (uint64_t)read_len, (size_t)read_len))) return rc;
            ++i;
            continue;
        }
        const uint64_t k = std::min<uint64_t>(n - i, STAGE_READS - c->stage_n);
        memcpy(c->h_stage[c->stage_cur] + (size_t)c->stage_n * c->L, reads + i * (uint64_t)read_len,
               (size_t)(k * (uint64_t)read_len));
        c->stage_n += (uint32_t)k;
        i += k;
        if (c->stage_n == STAGE_READS && (rc = stage_submit(c))) return rc;
    }
    return EMP_OK;
}

int emp_add_reads_device(emp_ctx *c, const void *d_reads, uint64_t n, int32_t read_len) {
    if (!c || (!d_reads && n)) return EMP_ERR_ARG;
    if (n == 0) return EMP_OK;
    if (((uintptr_t)d_reads) & 3) return fail(c, EMP_ERR_ARG, "device reads must be 4-byte aligned");
    (void)hipSetDevice(c->device);
    int rc = set_len(c, read_len);
    if (rc) return rc;
    if ((rc = flush_acc(c))) return rc;
    if ((rc = stage_submit(c))) return rc;
    if ((rc = map_batch(c, (const uint8_t *)d_reads, n))) return rc;
    return sync(c);
}

int emp_run(emp_ctx *c, emp_record_sink sink, void *user) {
    if (!c) return EMP_ERR_ARG;
    (void)hipSetDevice(c->device);
    double t0 = now_ms();
    int rc = run_impl(c, sink, user);
    c->summary.overlap_ms = now_ms() - t0;
    if (rc) {
        (void)hipStreamSynchronize(c->stream);
        collect_timings(c);
    }
    return rc;
}

int emp_run_to_file(emp_ctx *c, const char *path) {
    if (!c || !path) return EMP_ERR_ARG;
    OutFile of;
    if (!c->dio) c->dio.reset(new DirectIO());
    if (of.open(path, c->dio.get()) != EMP_OK) return fail(c, EMP_ERR_IO, "Could not open %s", path);
    c->out_file = &of;  // the back-to-back rounds write finished records while later rounds run
    c->emitted = 0;
    int rc = emp_run(c, file_sink, &of);
    c->out_file = nullptr;
    c->emitted = 0;
    if (of.finish() != EMP_OK && rc == EMP_OK) rc = fail(c, EMP_ERR_IO, "writing %s failed", path);
    return rc;
}

int emp_reset(emp_ctx *c) {
    if (!c) return EMP_ERR_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    collect_timings(c);
    c->n = 0;
    c->stage_n = 0;
    c->acc_n = 0;
    c->stage_busy[0] = c->stage_busy[1] = false;
    c->rounds.clear();
    c->summary = emp_summary{};
    HIPCHK(c, hipMemset(c->d_err, 0, sizeof(uint32_t)));
    return EMP_OK;
}

int emp_get_summary(const emp_ctx *c, emp_summary *out) {
    if (!c || !out) return EMP_ERR_ARG;
    *out = c->summary;
    return EMP_OK;
}

int emp_get_round(const emp_ctx *c, int32_t round, emp_round_stats *out) {
    if (!c || !out || round < 0 || (size_t)round >= c->rounds.size()) return EMP_ERR_ARG;
    *out = c->rounds[(size_t)round];
    return EMP_OK;
}

int32_t emp_kernel_count(const emp_ctx *c) { return c ? (int32_t)c->kstats.size() : 0; }

int emp_kernel_info(const emp_ctx *c, int32_t i, const char **name, uint64_t *launches, double *ms, double *bytes) {
    if (!c || i < 0 || (size_t)i >= c->kstats.size()) return EMP_ERR_ARG;
    const KStat &k = c->kstats[(size_t)i];
    if (name) *name = k.name.c_str();
    if (launches) *launches = k.launches;
    if (ms) *ms = k.ms;
    if (bytes) *bytes = k.bytes;
    return EMP_OK;
}

void emp_kernel_clear(emp_ctx *c) {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    collect_timings(c);
    c->kstats.clear();
}

void *emp_stream(const emp_ctx *c) { return c ? (void *)c->stream : nullptr; }

void emp_set_audit(emp_ctx *c, int32_t on) {
    if (c) c->audit = on != 0;
}

int32_t emp_audit_count(const emp_ctx *c) { return c ? (int32_t)c->audits.size() : 0; }

int emp_get_audit(const emp_ctx *c, int32_t i, emp_audit *out) {
    if (!c || !out || i < 0 || (size_t)i >= c->audits.size()) return EMP_ERR_ARG;
    *out = c->audits[(size_t)i];
    return EMP_OK;
}

void emp_set_profile(emp_ctx *c, int32_t on) {
    if (!c) return;
    (void)hipStreamSynchronize(c->stream);
    collect_timings(c);
    c->opts.profile = on;
}

// ---------------------------------------------------------------------------------------------
// Stage entry points.

int emp_stage_fingerprints(int32_t device, const char *reads, uint64_t n, int32_t len, uint64_t *keys) {
    if ((!reads || !keys) && n) return EMP_ERR_ARG;
    if (n == 0) return EMP_OK;
    emp_opts o;
    emp_opts_default(&o);
    o.device = device;
    emp_ctx *c = nullptr;
    int rc = emp_ctx_create(&o, &c);
    if (rc) return rc;
    void *d = nullptr;
    rc = dalloc(c, &d, n * (uint64_t)len);
    if (!rc && hipMemcpy(d, reads, n * (uint64_t)len, hipMemcpyHostToDevice) != hipSuccess) rc = EMP_ERR_HIP;
    if (!rc) rc = emp_add_reads_device(c, d, n, len);
    uint32_t herr = 0;
    if (!rc && hipMemcpy(&herr, c->d_err, 4, hipMemcpyDeviceToHost) != hipSuccess) rc = EMP_ERR_HIP;
    if (!rc && herr) rc = EMP_ERR_BASE;
    if (!rc) {
        std::vector<Entry> e(2 * n);
        if (hipMemcpy(e.data(), c->d_entries, 2 * n * sizeof(Entry), hipMemcpyDeviceToHost) != hipSuccess)
            rc = EMP_ERR_HIP;
        for (uint64_t i = 0; !rc && i < 2 * n; ++i) keys[i] = ((uint64_t)e[i].h1 << 32) | e[i].h2;
    }
    emp_ctx_destroy(c);
    return rc;
}

int emp_stage_columns(int32_t device, const char *reads, uint64_t n, int32_t len, uint32_t *cols) {
    if ((!reads || !cols) && n) return EMP_ERR_ARG;
    if (len < 1 || len > MAX_RD_LEN) return EMP_ERR_ARG;
    if (n == 0) return EMP_OK;
    emp_opts o;
    emp_opts_default(&o);
    o.device = device;
    emp_ctx *c = nullptr;
    int rc = emp_ctx_create(&o, &c);
    if (rc) return rc;
    void *d = nullptr;
    rc = dalloc(c, &d, n * (uint64_t)len);
    if (!rc && hipMemcpy(d, reads, n * (uint64_t)len, hipMemcpyHostToDevice) != hipSuccess) rc = EMP_ERR_HIP;
    if (!rc) rc = emp_add_reads_device(c, d, n, len);
    // the store the rounds gather from, as k_map wrote it
    const uint64_t words = ((n + CHUNK_READS - 1) >> CHUNK_SHIFT) * col_groups8((uint32_t)len) * (uint64_t)CHUNK_WWORDS;
    if (!rc && hipMemcpy(cols, c->d_cols, words * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = EMP_ERR_HIP;
    emp_ctx_destroy(c);
    return rc;
}

int emp_stage_roll(int32_t device, const uint64_t *keys, const char *bases, uint64_t n, int32_t kind, int32_t len,
                   uint64_t *out) {
    if ((!keys || !bases || !out) && n) return EMP_ERR_ARG;
    if (kind != 0 && kind != 1) return EMP_ERR_ARG;
    if (kind == 0 && (len < 1 || len > MAX_RD_LEN)) return EMP_ERR_ARG;
    if (n == 0) return EMP_OK;
    if (hipSetDevice(device) != hipSuccess) return EMP_ERR_HIP;
    uint64_t *dk = nullptr, *dout = nullptr;
    uint8_t *db = nullptr;
    int rc = EMP_OK;
    if (hipMalloc(&dk, n * 8) != hipSuccess || hipMalloc(&dout, n * 8) != hipSuccess || hipMalloc(&db, n) != hipSuccess)
        rc = EMP_ERR_NOMEM;
    if (!rc && (hipMemcpy(dk, keys, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(db, bases, n, hipMemcpyHostToDevice) != hipSuccess))
        rc = EMP_ERR_HIP;
    uint32_t m1 = kind == 0 ? modpow(R1, (uint64_t)(len - 1), P1) : 0;
    uint32_t m2 = kind == 0 ? modpow(R2, (uint64_t)(len - 1), P2) : 0;
    if (!rc && launch_roll(dk, db, n, kind, m1, m2, dout, nullptr) != hipSuccess) rc = EMP_ERR_HIP;
    if (!rc && hipMemcpy(out, dout, n * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = EMP_ERR_HIP;
    (void)hipFree(dk);
    (void)hipFree(dout);
    (void)hipFree(db);
    return rc;
}

}  // extern "C"
