// Host side of the batched engine (include/empanada_batch.h): job admission, segment allocation,
// and the step loop that runs every resident job's round through one launch sequence.
//
// Per job the step does what run_impl (emp_lib.cpp) does for one context, and so what
// build_graph (overlap.cu:206-293) does for one read set: the map phase and dedup when the job is
// admitted, then one round per step (regenerate, partition, join, edges, used-read removal) while
// both of its lists are non-empty and its overlap length is at least its minimum overlap.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/empanada_batch.h"
#include "emp_batch.h"
#include "emp_fastq.h"
#include "emp_hostcopy.h"

using namespace emp;

namespace {

constexpr uint32_t BIG_MAX = 1u << 16;           // buckets the large-bucket path takes per step
constexpr uint64_t STAGE_BYTES = 64ull << 20;     // pinned staging buffer (x2) for host reads
// Estimated device bytes per entry of a resident job: its segment in both list arrays (24 B),
// and its share of a step's partition buffers (coarse ~1.4x and fine ~2x its live entries per
// side, spill, edge pool).
constexpr uint64_t BYTES_PER_ENTRY = 24 + 2 * (17 + 24) + 16;

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

uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Words per column group of a job of n reads (4 reads per word, 64-byte aligned).
uint32_t col_gs(uint64_t n) { return (uint32_t)((ceil_div(n, 4) + 15) & ~15ull); }

// First-fit allocator of [0, cap).
struct RangeAlloc {
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    uint64_t cap = 0;
    void grow(uint64_t ncap) {
        if (ncap <= cap) return;
        release(cap, ncap - cap);
        cap = ncap;
    }
    bool alloc(uint64_t len, uint64_t &off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                off = it->first;
                const uint64_t rest = it->second - len;
                free_.erase(it);
                if (rest) free_[off + len] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t len) {
        if (!len) return;
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && off + len == nx->first) {
            len += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {
                pv->second += len;
                return;
            }
        }
        free_[off] = len;
    }
    uint64_t free_total() const {
        uint64_t t = 0;
        for (const auto &f : free_) t += f.second;
        return t;
    }
    // End of the highest used range (0 when everything is free).
    uint64_t used_end() const {
        if (free_.empty()) return cap;
        auto last = std::prev(free_.end());
        return last->first + last->second == cap ? last->first : cap;
    }
};

struct HJob {
    const char *reads = nullptr;
    std::vector<char> owned;  // FASTQ jobs: the parsed reads
    bool dev = false;
    uint64_t n = 0;
    int L = 0, o = 0;
    int state = EMP_JOB_QUEUED, status = EMP_OK;
    uint64_t seg = 0, seglen = 0, cbase = 0, clen = 0;
    uint32_t gs = 0;
    uint32_t S = 0, P = 0;
    int t = 0, rounds = 0;
    uint64_t charge = 0, n_dedup = 0, n_edges = 0;
    double submit_ms = 0, admit_ms = 0, done_ms = 0;
    uint32_t fine_load = BTARGET, coarse_load = BCOARSE;  // from the job's tuning class
};

// The launches of a round step captured in a graph, replayed while the step fits: grids (tiles
// of each pass and side, fine buckets) and reset ranges are upper bounds; key holds every buffer
// address and capacity the captured arguments hold.
struct StepGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> key;
    uint32_t g1[2] = {0, 0}, g2[2] = {0, 0}, gf = 0;
    uint64_t gc = 0;
};

struct BCounters {
    unsigned long long pool[SUBS];
    unsigned long long pool_ovf;
    unsigned long long pool_lost;  // abandoned sub-range reservations (place_edges)
    uint32_t n_big, abort, spill[2];
    uint32_t too_many;  // a large bucket held 2^31 edges or more
    uint32_t pad;
};

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

struct emp_batch {
    emp_batch_opts opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    double t0 = 0;
    uint64_t bytes = 0;
    std::vector<std::pair<void *, size_t>> allocs;

    std::vector<HJob> jobs;
    std::deque<int> queued;
    std::vector<int> active;
    uint64_t charged = 0;

    // resident data: list arrays (job segments) and the column store (job regions)
    Entry *d_list[2] = {nullptr, nullptr};
    RangeAlloc lal;
    uint32_t *d_cols = nullptr;
    RangeAlloc cal;

    // host reads staging
    char *h_stage[2] = {nullptr, nullptr};
    uint8_t *d_stage[2] = {nullptr, nullptr};
    uint8_t *d_spack[2] = {nullptr, nullptr};  // a stage's reads as 2-bit codes, expanded into d_stage
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_busy[2] = {false, false};
    int stage_cur = 0;
    uint32_t *d_jerr = nullptr;  // per admitted job: a read holds a byte other than ACGTacgt

    // step tables (pinned host mirrors)
    BJob *d_jobs = nullptr, *h_jobs = nullptr;
    uint32_t *d_pre = nullptr, *h_pre = nullptr;  // t1[2], t2[2], fine prefix: 5 x (cap + 1)
    uint32_t *d_lcnt = nullptr, *h_lcnt = nullptr;  // per job, at a stride of tab_cap: [2] survivors, edges
    uint64_t tab_cap = 0;

    // step scratch
    Entry *d_coarse[2] = {nullptr, nullptr};
    uint64_t coarse_cap = 0;
    uint32_t *d_ccnt[2] = {nullptr, nullptr};
    uint64_t ccnt_cap = 0;
    Entry *d_fine[2] = {nullptr, nullptr};
    uint32_t *d_fcnt[2] = {nullptr, nullptr};
    uint64_t fine_cap = 0;
    SpillRec *d_spill[2] = {nullptr, nullptr};
    uint32_t spill_cap = 0, spill_min = 1u << 16;
    double slack = 1.3;
    uint2 *d_pool = nullptr;  // edge pool: (from, to) pairs
    uint64_t pool_cap = 0, pool_ovf_cap = 0;
    unsigned long long *d_edge_off = nullptr, *d_dst_off = nullptr, *d_eblk = nullptr;
    uint32_t *d_edge_cnt = nullptr;
    uint64_t slots_cap = 0;
    uint32_t *d_out_from = nullptr, *d_out_to = nullptr, *h_from = nullptr, *h_to = nullptr;
    uint64_t out_cap = 0;
    uint32_t *d_big = nullptr;
    unsigned long long *d_boff[2] = {nullptr, nullptr};
    uint64_t *d_bkey[2] = {nullptr, nullptr};
    uint32_t *d_brank[2] = {nullptr, nullptr}, *d_baux[2] = {nullptr, nullptr};
    uint64_t bscratch_cap = 0;
    BCounters *d_ctr = nullptr, *h_ctr = nullptr;

    std::vector<KStat> kstats;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> free_events;

    // tuning (emp_batch_set_tuning): the join/dedup variant and the per-class bucket loads
    uint32_t fcap = BCAP;
    std::vector<emp_batch_class> classes;
    // the round step's graph
    StepGraph graph;
    uint64_t graph_captures = 0, graph_replays = 0;
};

namespace {

int fail(emp_batch *b, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return code;
}

#define BCHK(b, expr)                                                                                     \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return fail((b), EMP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

int dalloc(emp_batch *b, void **p, size_t sz) {
    if (sz == 0) sz = 256;
    hipError_t e = hipMalloc(p, sz);
    if (e != hipSuccess) return fail(b, EMP_ERR_NOMEM, "hipMalloc(%zu) failed: %s", sz, hipGetErrorString(e));
    b->bytes += sz;
    b->allocs.push_back({*p, sz});
    return EMP_OK;
}

void dfree(emp_batch *b, void *p) {
    if (!p) return;
    for (size_t i = 0; i < b->allocs.size(); ++i)
        if (b->allocs[i].first == p) {
            b->bytes -= b->allocs[i].second;
            b->allocs.erase(b->allocs.begin() + (long)i);
            break;
        }
    (void)hipFree(p);
}

// Replaces *p by a fresh allocation of n elements of T (contents dropped).
template <class T>
int realloc_dev(emp_batch *b, T **p, uint64_t n) {
    dfree(b, *p);
    *p = nullptr;
    void *q = nullptr;
    int rc = dalloc(b, &q, n * sizeof(T));
    if (rc) return rc;
    *p = (T *)q;
    return EMP_OK;
}

int kstat_id(emp_batch *b, const char *name) {
    for (size_t i = 0; i < b->kstats.size(); ++i)
        if (b->kstats[i].name == name) return (int)i;
    b->kstats.push_back(KStat{name});
    return (int)b->kstats.size() - 1;
}

hipEvent_t get_event(emp_batch *b) {
    if (!b->free_events.empty()) {
        hipEvent_t e = b->free_events.back();
        b->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

struct Timed {
    emp_batch *b;
    int k = -1;
    hipEvent_t a = nullptr;
    double bytes;
    Timed(emp_batch *b_, const char *name, double bytes_) : b(b_), bytes(bytes_) {
        if (b->opts.profile) {
            k = kstat_id(b, name);
            a = get_event(b);
            (void)hipEventRecord(a, b->stream);
        }
    }
    ~Timed() {
        if (k >= 0) {
            hipEvent_t e = get_event(b);
            (void)hipEventRecord(e, b->stream);
            b->pending.push_back(Pending{k, a, e, bytes});
        }
    }
};

void collect_timings(emp_batch *b) {
    for (auto &p : b->pending) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, p.a, p.b);
        KStat &k = b->kstats[p.k];
        k.launches++;
        k.ms += ms;
        k.bytes += p.bytes;
        b->free_events.push_back(p.a);
        b->free_events.push_back(p.b);
    }
    b->pending.clear();
}

int sync(emp_batch *b) {
    BCHK(b, hipStreamSynchronize(b->stream));
    collect_timings(b);
    return EMP_OK;
}

int job_rounds(const HJob &h) { return h.L - h.o > 0 ? h.L - h.o : 0; }

uint64_t job_charge(uint64_t n, int L) {
    return 2 * n * BYTES_PER_ENTRY + (uint64_t)col_groups((uint32_t)L) * col_gs(n) * 4;
}

// Grows an arena of `per` bytes per element to at least `need` elements (contents kept).
template <class T>
int grow_arena(emp_batch *b, T **p, RangeAlloc &ra, uint64_t need, int copies) {
    uint64_t ncap = std::max<uint64_t>(need, ra.cap + ra.cap / 2);
    ncap = (ncap + 4095) & ~4095ull;
    BCHK(b, hipStreamSynchronize(b->stream));
    const uint64_t keep = ra.used_end();
    for (int k = 0; k < copies; ++k) {
        void *q = nullptr;
        int rc = dalloc(b, &q, ncap * sizeof(T));
        if (rc) return rc;
        if (p[k] && keep) BCHK(b, hipMemcpyAsync(q, p[k], keep * sizeof(T), hipMemcpyDeviceToDevice, b->stream));
        BCHK(b, hipStreamSynchronize(b->stream));
        dfree(b, p[k]);
        p[k] = (T *)q;
    }
    ra.grow(ncap);
    return EMP_OK;
}

// Segment of 2n entries in both list arrays and the column region of a job.
int place_job(emp_batch *b, HJob &h) {
    h.seglen = (2 * h.n + 15) & ~15ull;
    h.gs = col_gs(h.n);
    h.clen = (uint64_t)col_groups((uint32_t)h.L) * h.gs;
    int rc;
    if (!b->lal.alloc(h.seglen, h.seg)) {
        if ((rc = grow_arena(b, b->d_list, b->lal, b->lal.used_end() + h.seglen, 2))) return rc;
        if (!b->lal.alloc(h.seglen, h.seg)) return fail(b, EMP_ERR_NOMEM, "list arena");
    }
    if (!b->cal.alloc(h.clen, h.cbase)) {
        if ((rc = grow_arena(b, &b->d_cols, b->cal, b->cal.used_end() + h.clen, 1))) return rc;
        if (!b->cal.alloc(h.clen, h.cbase)) return fail(b, EMP_ERR_NOMEM, "column arena");
    }
    return EMP_OK;
}

void finish_job(emp_batch *b, int id, int status, emp_job_done done, void *user) {
    HJob &h = b->jobs[(size_t)id];
    if (h.state == EMP_JOB_ACTIVE || h.seglen) {
        b->lal.release(h.seg, h.seglen);
        b->cal.release(h.cbase, h.clen);
        h.seglen = h.clen = 0;
        b->charged -= h.charge;
    }
    h.state = EMP_JOB_DONE;
    h.status = status;
    h.done_ms = now_ms() - b->t0;
    h.owned.clear();
    h.owned.shrink_to_fit();
    h.reads = nullptr;
    if (done) done(user, id, status);
}

// Step tables of capacity n jobs (device + pinned host).
int ensure_tables(emp_batch *b, uint64_t n) {
    if (n <= b->tab_cap) return EMP_OK;
    BCHK(b, hipStreamSynchronize(b->stream));
    uint64_t cap = std::max<uint64_t>(n, 2 * b->tab_cap);
    int rc;
    if ((rc = realloc_dev(b, &b->d_jobs, cap))) return rc;
    if ((rc = realloc_dev(b, &b->d_pre, 5 * (cap + 1)))) return rc;
    if ((rc = realloc_dev(b, &b->d_lcnt, 3 * cap))) return rc;
    if ((rc = realloc_dev(b, &b->d_jerr, cap))) return rc;
    if (b->h_jobs) (void)hipHostFree(b->h_jobs);
    if (b->h_pre) (void)hipHostFree(b->h_pre);
    if (b->h_lcnt) (void)hipHostFree(b->h_lcnt);
    b->h_jobs = nullptr;
    b->h_pre = b->h_lcnt = nullptr;
    BCHK(b, hipHostMalloc((void **)&b->h_jobs, cap * sizeof(BJob), hipHostMallocDefault));
    BCHK(b, hipHostMalloc((void **)&b->h_pre, 5 * (cap + 1) * sizeof(uint32_t), hipHostMallocDefault));
    BCHK(b, hipHostMalloc((void **)&b->h_lcnt, 3 * cap * sizeof(uint32_t), hipHostMallocDefault));
    b->tab_cap = cap;
    return EMP_OK;
}

// Layout of one step over the jobs in h_jobs[0..nj): buckets and tiles from the live counts.
struct StepLayout {
    uint64_t ncoarse = 0, nfine = 0, coarse[2] = {0, 0};
    uint32_t tiles1[2] = {0, 0}, tiles2[2] = {0, 0};
    uint64_t live = 0;
};

int plan_step(emp_batch *b, uint32_t nj, bool dedup, const std::vector<int> &step_ids, StepLayout &sl) {
    uint32_t *t1 = b->h_pre, *t2 = b->h_pre + 2 * (b->tab_cap + 1), *fp = b->h_pre + 4 * (b->tab_cap + 1);
    const uint64_t stride = b->tab_cap + 1;
    sl = StepLayout{};
    for (uint32_t j = 0; j < nj; ++j) {
        BJob &J = b->h_jobs[j];
        const uint64_t x = std::max(J.cnt[0], J.cnt[1]);
        // the job's class loads; the fine load at least x / BSPLIT^2 (at most BSPLIT fine buckets
        // per coarse one)
        const HJob &h = b->jobs[(size_t)step_ids[j]];
        const uint64_t fl = std::max<uint64_t>(h.fine_load, ceil_div(x, (uint64_t)BSPLIT * BSPLIT));
        const uint32_t nb0 = (uint32_t)std::max<uint64_t>(1, ceil_div(x, fl));
        // coarse buckets from the class's load, at least enough that each fans out to at most
        // BSPLIT fine buckets
        J.nc = (uint32_t)std::min<uint64_t>(
            BSPLIT, std::max<uint64_t>(std::max<uint64_t>(1, ceil_div(x, h.coarse_load)), ceil_div(nb0, BSPLIT)));
        J.m = (uint32_t)ceil_div(nb0, J.nc);
        if (J.m > BSPLIT) return fail(b, EMP_ERR_CAPACITY, "a job list of %llu entries exceeds the batch limit",
                                      (unsigned long long)x);
        J.nb = J.nc * J.m;
        J.cb = (uint32_t)sl.ncoarse;
        J.fb = (uint32_t)sl.nfine;
        fp[j] = J.fb;
        for (int s = 0; s < 2; ++s) {
            t1[s * stride + j] = sl.tiles1[s];
            t2[s * stride + j] = sl.tiles2[s];
            const uint64_t c = J.cnt[s];
            uint64_t cap = c ? (uint64_t)((double)c / J.nc * b->slack) + 256 : 0;
            cap = (cap + 63) & ~63ull;
            if (cap > 0xFFFFFFFFull) return fail(b, EMP_ERR_CAPACITY, "coarse bucket too large");
            J.ccap[s] = (uint32_t)cap;
            J.coff[s] = sl.coarse[s];
            sl.coarse[s] += cap * J.nc;
            if (c) {
                sl.tiles1[s] += (uint32_t)ceil_div(c, PART_TILE);
                sl.tiles2[s] += J.nc * (uint32_t)ceil_div(cap, PART_TILE);
            }
            sl.live += c;
            if (dedup) break;
        }
        sl.ncoarse += J.nc;
        sl.nfine += J.nb;
    }
    for (uint64_t j = nj; j <= b->tab_cap; ++j) {  // the totals from the last job to the capacity
        for (int s = 0; s < 2; ++s) {
            t1[s * stride + j] = sl.tiles1[s];
            t2[s * stride + j] = sl.tiles2[s];
        }
        fp[j] = (uint32_t)sl.nfine;
    }
    if (sl.nfine >= (1ull << 32) / b->fcap) return fail(b, EMP_ERR_CAPACITY, "too many fine buckets in one step");
    return EMP_OK;
}

int upload_tables(emp_batch *b, uint32_t nj) {
    BCHK(b, hipMemcpyAsync(b->d_jobs, b->h_jobs, nj * sizeof(BJob), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipMemcpyAsync(b->d_pre, b->h_pre, 5 * (b->tab_cap + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                           b->stream));
    return EMP_OK;
}

int ensure_scratch(emp_batch *b, const StepLayout &sl) {
    int rc;
    const uint64_t cneed = std::max(sl.coarse[0], sl.coarse[1]);
    if (cneed > b->coarse_cap) {
        BCHK(b, hipStreamSynchronize(b->stream));
        const uint64_t cap = std::max<uint64_t>(cneed + cneed / 4, 1u << 16);
        for (int s = 0; s < 2; ++s)
            if ((rc = realloc_dev(b, &b->d_coarse[s], cap))) return rc;
        b->coarse_cap = cap;
    }
    if (sl.ncoarse > b->ccnt_cap) {
        BCHK(b, hipStreamSynchronize(b->stream));
        const uint64_t cap = std::max<uint64_t>(sl.ncoarse + sl.ncoarse / 4, 4096);
        for (int s = 0; s < 2; ++s)
            if ((rc = realloc_dev(b, &b->d_ccnt[s], cap))) return rc;
        b->ccnt_cap = cap;
    }
    if (sl.nfine > b->fine_cap) {
        BCHK(b, hipStreamSynchronize(b->stream));
        const uint64_t cap = std::max<uint64_t>(sl.nfine + sl.nfine / 4, 4096);
        for (int s = 0; s < 2; ++s) {
            if ((rc = realloc_dev(b, &b->d_fine[s], cap * b->fcap))) return rc;
            if ((rc = realloc_dev(b, &b->d_fcnt[s], cap))) return rc;
        }
        if ((rc = realloc_dev(b, &b->d_edge_off, cap))) return rc;
        if ((rc = realloc_dev(b, &b->d_dst_off, cap))) return rc;
        if ((rc = realloc_dev(b, &b->d_edge_cnt, cap))) return rc;
        if ((rc = realloc_dev(b, &b->d_eblk, ceil_div(cap, 4096)))) return rc;
        b->fine_cap = cap;
        b->slots_cap = cap;
    }
    const uint32_t sp = (uint32_t)std::min<uint64_t>(0x7FFFFFFF, std::max<uint64_t>(b->spill_min, sl.live / 128));
    if (sp > b->spill_cap) {
        BCHK(b, hipStreamSynchronize(b->stream));
        for (int s = 0; s < 2; ++s)
            if ((rc = realloc_dev(b, &b->d_spill[s], sp))) return rc;
        b->spill_cap = sp;
    }
    return EMP_OK;
}

// Edge pool with at least per_sub edges per sub-range and an overflow region of ovf edges.
int ensure_pool(emp_batch *b, uint64_t per_sub, uint64_t ovf) {
    if (per_sub <= b->pool_cap && ovf <= b->pool_ovf_cap) return EMP_OK;
    per_sub = std::max(per_sub, b->pool_cap);
    ovf = std::max(ovf, b->pool_ovf_cap);
    BCHK(b, hipStreamSynchronize(b->stream));
    int rc;
    if ((rc = realloc_dev(b, &b->d_pool, per_sub * SUBS + ovf))) return rc;
    b->pool_cap = per_sub;
    b->pool_ovf_cap = ovf;
    return EMP_OK;
}

int ensure_out(emp_batch *b, uint64_t n) {
    if (n <= b->out_cap) return EMP_OK;
    BCHK(b, hipStreamSynchronize(b->stream));
    const uint64_t cap = std::max<uint64_t>(n + n / 4, 1u << 20);
    int rc;
    if ((rc = realloc_dev(b, &b->d_out_from, cap))) return rc;
    if ((rc = realloc_dev(b, &b->d_out_to, cap))) return rc;
    if (b->h_from) (void)hipHostFree(b->h_from);
    if (b->h_to) (void)hipHostFree(b->h_to);
    b->h_from = b->h_to = nullptr;
    BCHK(b, hipHostMalloc((void **)&b->h_from, cap * 4, hipHostMallocDefault));
    BCHK(b, hipHostMalloc((void **)&b->h_to, cap * 4, hipHostMallocDefault));
    b->out_cap = cap;
    return EMP_OK;
}

int ensure_big_scratch(emp_batch *b, uint64_t need) {
    if (need <= b->bscratch_cap) return EMP_OK;
    BCHK(b, hipStreamSynchronize(b->stream));
    int rc;
    for (int s = 0; s < 2; ++s) {
        if ((rc = realloc_dev(b, &b->d_bkey[s], need))) return rc;
        if ((rc = realloc_dev(b, &b->d_brank[s], need))) return rc;
        if ((rc = realloc_dev(b, &b->d_baux[s], need))) return rc;
    }
    b->bscratch_cap = need;
    return EMP_OK;
}

// After an abort (a coarse bucket or the spill area overflowed): more slack, a larger spill area.
void after_abort(emp_batch *b) {
    b->slack *= 1.5;
    if (std::max(b->h_ctr->spill[0], b->h_ctr->spill[1]) >= b->spill_cap)
        b->spill_min = (uint32_t)std::min<uint64_t>(0x7FFFFFFF, 4ull * b->spill_cap);
}

int read_counters(emp_batch *b, uint32_t nj) {
    BCHK(b, hipMemcpyAsync(b->h_ctr, b->d_ctr, sizeof(BCounters), hipMemcpyDeviceToHost, b->stream));
    (void)nj;
    BCHK(b, hipMemcpyAsync(b->h_lcnt, b->d_lcnt, 3ull * b->tab_cap * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           b->stream));
    return sync(b);
}

// Counter resets of a step: coarse and fine counts (with `part`), survivor and edge counts,
// flags and pool cursors.
int zero_step(emp_batch *b, uint64_t ncoarse, uint64_t nfine, bool part, bool dedup) {
    ZeroArgs z{};
    auto add = [&](uint32_t *p, uint64_t n) {
        z.p[z.count] = p;
        z.n[z.count] = (uint32_t)n;
        z.v[z.count] = 0;
        ++z.count;
    };
    if (part) {
        for (int s = 0; s < (dedup ? 1 : 2); ++s) {
            add(b->d_ccnt[s], ncoarse);
            add(b->d_fcnt[s], nfine);
        }
        add(&b->d_ctr->abort, 1);
        add(b->d_ctr->spill, 2);
    }
    add(b->d_lcnt, 3 * b->tab_cap);
    add(&b->d_ctr->n_big, 1);
    add(&b->d_ctr->too_many, 1);
    add((uint32_t *)b->d_ctr->pool, 2 * (SUBS + 2));
    BCHK(b, launch_zero(z, b->stream));
    return EMP_OK;
}

// One partition (two passes) of side s of the step.
int partition_side(emp_batch *b, int s, bool regen, uint32_t groups1, uint32_t groups2, double bytes) {
    const uint64_t stride = b->tab_cap + 1;
    BPartArgs a{};
    a.jobs = b->d_jobs;
    a.njobs = (uint32_t)b->tab_cap;
    a.side = s;
    a.lists[0] = b->d_list[0];
    a.lists[1] = b->d_list[1];
    a.cols = b->d_cols;
    a.coarse = b->d_coarse[s];
    a.ccnt = b->d_ccnt[s];
    a.fine = b->d_fine[s];
    a.fcnt = b->d_fcnt[s];
    a.spill = b->d_spill[s];
    a.spill_n = &b->d_ctr->spill[s];
    a.spill_cap = b->spill_cap;
    a.abort = &b->d_ctr->abort;
    a.fcap = b->fcap;
    a.tpre = b->d_pre + s * stride;
    {
        Timed t(b, "k_bpart1", bytes);
        BCHK(b, launch_bpart(a, 1, regen, groups1, b->stream));
    }
    a.tpre = b->d_pre + (2 + s) * stride;
    {
        Timed t(b, "k_bpart2", bytes);
        BCHK(b, launch_bpart(a, 2, false, groups2, b->stream));
    }
    return EMP_OK;
}

BJoinArgs join_args(emp_batch *b, uint32_t nj) {
    BJoinArgs a{};
    a.jobs = b->d_jobs;
    a.njobs = (uint32_t)b->tab_cap;
    a.fpre = b->d_pre + 4 * (b->tab_cap + 1);
    a.s_in = b->d_fine[0];
    a.p_in = b->d_fine[1];
    a.s_cnt = b->d_fcnt[0];
    a.p_cnt = b->d_fcnt[1];
    a.lists[0] = b->d_list[0];
    a.lists[1] = b->d_list[1];
    a.lcnt = b->d_lcnt;
    a.ejob = b->d_lcnt + 2 * b->tab_cap;
    a.pool = b->d_pool;
    a.pool_cursor = b->d_ctr->pool;
    a.pool_cap = b->pool_cap;
    a.ovf_cursor = &b->d_ctr->pool_ovf;
    a.ovf_cap = b->pool_ovf_cap;
    a.edge_off = b->d_edge_off;
    a.edge_cnt = b->d_edge_cnt;
    a.big = b->d_big;
    a.n_big = &b->d_ctr->n_big;
    a.big_cap = BIG_MAX;
    a.abort = &b->d_ctr->abort;
    return a;
}

// Large-bucket path over the buckets the LDS kernels handed over.
int big_path(emp_batch *b, uint32_t nj, const StepLayout &sl, bool dedup) {
    const uint32_t nbig = b->h_ctr->n_big;
    if (nbig > BIG_MAX) return fail(b, EMP_ERR_CAPACITY, "%u buckets over capacity (limit %u)", nbig, BIG_MAX);
    std::vector<uint32_t> ids(nbig), cs(sl.nfine), cp(dedup ? 0 : sl.nfine);
    BCHK(b, hipMemcpy(ids.data(), b->d_big, nbig * sizeof(uint32_t), hipMemcpyDeviceToHost));
    BCHK(b, hipMemcpy(cs.data(), b->d_fcnt[0], sl.nfine * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (!dedup) BCHK(b, hipMemcpy(cp.data(), b->d_fcnt[1], sl.nfine * sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<unsigned long long> so(nbig), po(nbig);
    uint64_t as = 0, ap = 0;
    auto p2 = [](uint64_t n) { uint64_t p = 1; while (p < n) p <<= 1; return p; };
    for (uint32_t i = 0; i < nbig; ++i) {
        so[i] = as;
        as += p2(cs[ids[i]]);
        po[i] = ap;
        ap += dedup ? 0 : p2(cp[ids[i]]);
    }
    int rc = ensure_big_scratch(b, std::max<uint64_t>(std::max(as, ap), 1));
    if (rc) return rc;
    if (!b->d_boff[0])
        for (int s = 0; s < 2; ++s)
            if ((rc = realloc_dev(b, &b->d_boff[s], BIG_MAX))) return rc;
    BCHK(b, hipMemcpyAsync(b->d_boff[0], so.data(), nbig * sizeof(unsigned long long), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipMemcpyAsync(b->d_boff[1], po.data(), nbig * sizeof(unsigned long long), hipMemcpyHostToDevice, b->stream));
    BBigArgs a{};
    a.jobs = b->d_jobs;
    a.njobs = (uint32_t)b->tab_cap;
    a.fpre = b->d_pre + 4 * (b->tab_cap + 1);
    a.ids = b->d_big;
    a.soff = b->d_boff[0];
    a.poff = b->d_boff[1];
    a.skey = b->d_bkey[0];
    a.pkey = b->d_bkey[1];
    a.srank = b->d_brank[0];
    a.prank = b->d_brank[1];
    a.saux = b->d_baux[0];
    a.paux = b->d_baux[1];
    a.s_in = b->d_fine[0];
    a.p_in = b->d_fine[1];
    a.s_cnt = b->d_fcnt[0];
    a.p_cnt = b->d_fcnt[1];
    a.s_spill = b->d_spill[0];
    a.p_spill = b->d_spill[1];
    a.s_spill_n = &b->d_ctr->spill[0];
    a.p_spill_n = &b->d_ctr->spill[1];
    a.spill_cap = b->spill_cap;
    a.cols = b->d_cols;
    a.lists[0] = b->d_list[0];
    a.lists[1] = b->d_list[1];
    a.lcnt = b->d_lcnt;
    a.ejob = b->d_lcnt + 2 * b->tab_cap;
    a.pool = b->d_pool;
    a.pool_cursor = b->d_ctr->pool;
    a.pool_cap = b->pool_cap;
    a.ovf_cursor = &b->d_ctr->pool_ovf;
    a.ovf_cap = b->pool_ovf_cap;
    a.edge_off = b->d_edge_off;
    a.edge_cnt = b->d_edge_cnt;
    a.too_many = &b->d_ctr->too_many;
    a.fcap = b->fcap;
    {
        Timed t(b, dedup ? "k_bdedup_big" : "k_bjoin_big", 0.0);
        BCHK(b, dedup ? launch_bdedup_big(a, nbig, b->stream) : launch_bjoin_big(a, nbig, b->stream));
    }
    if ((rc = read_counters(b, nj))) return rc;
    if (b->h_ctr->too_many)  // the reference's int edge offsets (overlap.cu:561,567)
        return fail(b, EMP_ERR_PARTITION, "a key group of a batched job holds 2^31 edges or more");
    return EMP_OK;
}

// Map phase of job h: its reads (host, through the pinned double buffer, or device) become the
// job's column region and 2n entries in its segment of list array 1. `ji` indexes the job's
// error flag. Host reads cross PCIe as 2-bit codes (packed by host threads into the stage,
// expanded by k_unpack), or as they are when a stage holds a byte other than ACGTacgt, so that
// the map kernel's check reports it.
int map_job(emp_batch *b, HJob &h, uint32_t ji) {
    int Lpad = (h.L + 3) & ~3;
    if (((Lpad / 4) & 1) == 0) Lpad += 4;
    MapArgs a{};
    a.L = h.L;
    a.Lpad = Lpad;
    a.rows = 256u * (uint32_t)Lpad <= 65536u ? 256u : 128u;
    a.cols = b->d_cols;
    a.entries = b->d_list[1] + h.seg;
    a.err = b->d_jerr + ji;
    a.cbase = h.cbase;
    a.gs = h.gs;
    if (h.dev) {
        a.reads = (const uint8_t *)h.reads;
        a.n = (uint32_t)h.n;
        a.k0 = 0;
        Timed t(b, "k_map", (double)h.n * h.L * 1.25 + 24.0 * h.n);
        BCHK(b, launch_map(a, b->stream));
        return EMP_OK;
    }
    const uint64_t per = std::max<uint64_t>(16, (STAGE_BYTES / (uint64_t)h.L) & ~15ull);
    int &cur = b->stage_cur;
    for (uint64_t done = 0; done < h.n;) {
        const uint64_t m = std::min<uint64_t>(per, h.n - done);
        if (b->stage_busy[cur]) {
            BCHK(b, hipEventSynchronize(b->stage_ev[cur]));
            b->stage_busy[cur] = false;
        }
        const char *src = h.reads + done * (uint64_t)h.L;
        const uint64_t bytes = m * (uint64_t)h.L;
        if (parallel_pack(reinterpret_cast<uint8_t *>(b->h_stage[cur]), src, bytes)) {
            BCHK(b, hipMemcpyAsync(b->d_spack[cur], b->h_stage[cur], (bytes + 3) / 4, hipMemcpyHostToDevice, b->stream));
            Timed t(b, "k_unpack", 1.25 * (double)bytes);
            BCHK(b, launch_unpack(reinterpret_cast<const uint32_t *>(b->d_spack[cur]), b->d_stage[cur], bytes, b->stream));
        } else {
            parallel_copy(b->h_stage[cur], src, bytes);
            BCHK(b, hipMemcpyAsync(b->d_stage[cur], b->h_stage[cur], bytes, hipMemcpyHostToDevice, b->stream));
        }
        a.reads = b->d_stage[cur];
        a.n = (uint32_t)m;
        a.k0 = done;
        {
            Timed t(b, "k_map", (double)m * h.L * 1.25 + 24.0 * m);
            BCHK(b, launch_map(a, b->stream));
        }
        BCHK(b, hipEventRecord(b->stage_ev[cur], b->stream));
        b->stage_busy[cur] = true;
        cur ^= 1;
        done += m;
    }
    return EMP_OK;
}

int ensure_staging(emp_batch *b) {
    if (b->h_stage[0]) return EMP_OK;
    for (int i = 0; i < 2; ++i) {
        BCHK(b, hipHostMalloc((void **)&b->h_stage[i], STAGE_BYTES, hipHostMallocDefault));
        void *p = nullptr;
        int rc = dalloc(b, &p, STAGE_BYTES);
        if (rc) return rc;
        b->d_stage[i] = (uint8_t *)p;
        if ((rc = dalloc(b, &p, STAGE_BYTES / 4 + 16))) return rc;
        b->d_spack[i] = (uint8_t *)p;
        BCHK(b, hipEventCreateWithFlags(&b->stage_ev[i], hipEventDisableTiming));
    }
    return EMP_OK;
}

// Admission: queued jobs in FIFO order while the budget and max_active allow; their map phase and
// dedup (overlap.cu:212-225) run as one segmented step. Admitted jobs join `active` at round 0.
int admit(emp_batch *b, emp_job_done done, void *user) {
    std::vector<int> nw;
    while (!b->queued.empty()) {
        const int id = b->queued.front();
        HJob &h = b->jobs[(size_t)id];
        if (h.n == 0) {  // overlap.cu:231,258: no reads, no records
            b->queued.pop_front();
            h.admit_ms = now_ms() - b->t0;
            finish_job(b, id, EMP_OK, done, user);
            continue;
        }
        const uint64_t charge = job_charge(h.n, h.L);
        if (b->opts.max_active && b->active.size() + nw.size() >= b->opts.max_active) break;
        if (b->opts.hbm_budget && b->charged + charge > b->opts.hbm_budget) {
            if (b->active.empty() && nw.empty()) {  // can never fit
                b->queued.pop_front();
                finish_job(b, id, fail(b, EMP_ERR_NOMEM, "job %d needs %llu bytes, over the HBM budget", id,
                                       (unsigned long long)charge), done, user);
                continue;
            }
            break;
        }
        b->queued.pop_front();
        h.charge = charge;
        b->charged += charge;
        nw.push_back(id);
    }
    if (nw.empty()) return EMP_OK;
    int rc;
    const uint32_t nj = (uint32_t)nw.size();
    if ((rc = ensure_tables(b, std::max<uint64_t>(nj, b->active.size())))) return rc;
    BCHK(b, launch_fill_u32(b->d_jerr, 0u, nj, b->stream));
    {
        uint64_t segs = 0, cols = 0;
        for (int id : nw) {
            const HJob &h = b->jobs[(size_t)id];
            segs += (2 * h.n + 15) & ~15ull;
            cols += (uint64_t)col_groups((uint32_t)h.L) * col_gs(h.n);
        }
        if (b->lal.free_total() < segs && (rc = grow_arena(b, b->d_list, b->lal, b->lal.used_end() + segs, 2)))
            return rc;
        if (b->cal.free_total() < cols && (rc = grow_arena(b, &b->d_cols, b->cal, b->cal.used_end() + cols, 1)))
            return rc;
    }
    for (uint32_t i = 0; i < nj; ++i) {
        HJob &h = b->jobs[(size_t)nw[i]];
        h.admit_ms = now_ms() - b->t0;
        if ((rc = place_job(b, h))) return rc;
    }
    if ((rc = ensure_staging(b))) return rc;
    for (uint32_t i = 0; i < nj; ++i)
        if ((rc = map_job(b, b->jobs[(size_t)nw[i]], i))) return rc;

    // dedup step over the new jobs
    for (uint32_t i = 0; i < nj; ++i) {
        const HJob &h = b->jobs[(size_t)nw[i]];
        BJob &J = b->h_jobs[i];
        J = BJob{};
        J.seg = h.seg;
        J.cbase = h.cbase;
        J.gs = h.gs;
        J.cnt[0] = (uint32_t)(2 * h.n);
        J.src[0] = 1;
        J.keep = 1;
    }
    StepLayout sl;
    for (int attempt = 0;; ++attempt) {
        if (attempt > 8) return fail(b, EMP_ERR_CAPACITY, "dedup does not settle");
        if ((rc = plan_step(b, nj, true, nw, sl))) return rc;
        if ((rc = ensure_scratch(b, sl))) return rc;
        if ((rc = upload_tables(b, nj))) return rc;
        if ((rc = zero_step(b, sl.ncoarse, sl.nfine, true, true))) return rc;
        if ((rc = partition_side(b, 0, false, sl.tiles1[0], sl.tiles2[0], 24.0 * sl.live))) return rc;
        BJoinArgs a = join_args(b, nj);
        a.p_in = nullptr;
        a.p_cnt = nullptr;
        a.edge_cnt = nullptr;
        {
            Timed t(b, "k_bdedup", 12.0 * sl.live);
            BCHK(b, launch_bdedup(a, (uint32_t)sl.nfine, b->fcap, b->stream));
        }
        if ((rc = read_counters(b, nj))) return rc;
        if (b->h_ctr->abort) {
            after_abort(b);
            continue;
        }
        if (b->h_ctr->n_big && (rc = big_path(b, nj, sl, true))) return rc;
        break;
    }
    std::vector<uint32_t> jerr(nj);
    BCHK(b, hipMemcpy(jerr.data(), b->d_jerr, nj * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < nj; ++i) {
        const int id = nw[i];
        HJob &h = b->jobs[(size_t)id];
        h.state = EMP_JOB_ACTIVE;
        h.n_dedup = b->h_lcnt[i];
        h.S = h.P = b->h_lcnt[i];
        h.t = 0;
        h.rounds = job_rounds(h);
        if (jerr[i]) {
            finish_job(b, id, fail(b, EMP_ERR_BASE, "job %d: a kept read holds a byte other than ACGTacgt", id),
                       done, user);
        } else if (h.rounds == 0 || h.S == 0) {
            finish_job(b, id, EMP_OK, done, user);
        } else {
            b->active.push_back(id);
        }
    }
    return EMP_OK;
}

std::vector<uint64_t> graph_key(const emp_batch *b) {
    return {(uint64_t)(uintptr_t)b->d_jobs,     (uint64_t)(uintptr_t)b->d_pre,      (uint64_t)(uintptr_t)b->d_lcnt,
            b->tab_cap,                         (uint64_t)(uintptr_t)b->d_list[0],  (uint64_t)(uintptr_t)b->d_list[1],
            (uint64_t)(uintptr_t)b->d_cols,     (uint64_t)(uintptr_t)b->d_coarse[0], (uint64_t)(uintptr_t)b->d_coarse[1],
            (uint64_t)(uintptr_t)b->d_ccnt[0],  (uint64_t)(uintptr_t)b->d_ccnt[1],  b->ccnt_cap,
            (uint64_t)(uintptr_t)b->d_fine[0],  (uint64_t)(uintptr_t)b->d_fine[1],  (uint64_t)(uintptr_t)b->d_fcnt[0],
            (uint64_t)(uintptr_t)b->d_fcnt[1],  b->fine_cap,                        (uint64_t)(uintptr_t)b->d_spill[0],
            (uint64_t)(uintptr_t)b->d_spill[1], b->spill_cap,                       (uint64_t)(uintptr_t)b->d_pool,
            b->pool_cap,                        b->pool_ovf_cap,                    (uint64_t)(uintptr_t)b->d_edge_off,
            (uint64_t)(uintptr_t)b->d_edge_cnt, (uint64_t)(uintptr_t)b->d_big,      (uint64_t)(uintptr_t)b->d_ctr,
            b->fcap};
}

// The launches of a round step (counter resets, the two partition passes of both sides, the join)
// as one graph. The graph is captured with grids a quarter above the step's needs and replayed
// for later steps while they fit (lists only shrink; admitted jobs may need more) and no buffer
// it names has moved; far smaller steps capture anew, so that few empty work-groups launch.
int graph_step(emp_batch *b, uint32_t nj, const StepLayout &sl) {
    StepGraph &g = b->graph;
    const std::vector<uint64_t> key = graph_key(b);
    auto fits = [&](uint64_t need, uint64_t have) { return need <= have && have <= 2 * need + 1024; };
    const bool ok = g.exec && g.key == key && fits(sl.tiles1[0], g.g1[0]) && fits(sl.tiles1[1], g.g1[1]) &&
                    fits(sl.tiles2[0], g.g2[0]) && fits(sl.tiles2[1], g.g2[1]) && fits(sl.nfine, g.gf) &&
                    sl.ncoarse <= g.gc;
    if (!ok) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
        auto grow = [](uint64_t x) { return (uint32_t)std::min<uint64_t>(0x7FFFFFFF, x + x / 4 + 64); };
        for (int s = 0; s < 2; ++s) {
            g.g1[s] = grow(sl.tiles1[s]);
            g.g2[s] = grow(sl.tiles2[s]);
        }
        g.gf = (uint32_t)std::min<uint64_t>(grow(sl.nfine), b->fine_cap);
        g.gc = std::min<uint64_t>(grow(sl.ncoarse), b->ccnt_cap);
        g.key = key;
        hipGraph_t graph = nullptr;
        BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = zero_step(b, g.gc, g.gf, true, false);
        if (!rc) rc = partition_side(b, 0, true, g.g1[0], g.g2[0], 0.0);
        if (!rc) rc = partition_side(b, 1, true, g.g1[1], g.g2[1], 0.0);
        hipError_t e = hipSuccess;
        if (!rc) e = launch_bjoin(join_args(b, nj), g.gf, b->fcap, b->stream);
        const hipError_t ee = hipStreamEndCapture(b->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(b, EMP_ERR_HIP, "capturing the round step failed: %s",
                        hipGetErrorString(e != hipSuccess ? e : ee));
        }
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g.exec = nullptr;
            return fail(b, EMP_ERR_HIP, "instantiating the round step failed: %s", hipGetErrorString(e));
        }
        ++b->graph_captures;
    }
    BCHK(b, hipGraphLaunch(g.exec, b->stream));
    ++b->graph_replays;
    return EMP_OK;
}

// One round of every active job.
int round_step(emp_batch *b, emp_job_sink sink, emp_job_done done, void *user) {
    const uint32_t nj = (uint32_t)b->active.size();
    if (nj == 0) return EMP_OK;
    int rc;
    if ((rc = ensure_tables(b, nj))) return rc;
    for (uint32_t i = 0; i < nj; ++i) {
        const HJob &h = b->jobs[(size_t)b->active[i]];
        BJob &J = b->h_jobs[i];
        J = BJob{};
        J.seg = h.seg;
        J.cbase = h.cbase;
        J.gs = h.gs;
        J.cnt[0] = h.S;
        J.cnt[1] = h.P;
        J.src[0] = 0;
        J.src[1] = h.t == 0 ? 0 : 1;  // round 0: both sides come from the dedup list
        J.keep = h.t + 1 < h.rounds ? 1 : 0;
        J.rg.active = 1;
        J.rg.col_t = (uint32_t)h.t;
        J.rg.col_back = (uint32_t)(h.L - 1 - h.t);
        J.rg.m1 = modpow(R1, (uint64_t)(h.L - h.t - 1), P1);
        J.rg.m2 = modpow(R2, (uint64_t)(h.L - h.t - 1), P2);
        J.rg.G = col_groups((uint32_t)h.L);
    }
    StepLayout sl;
    bool need_part = true;
    for (int attempt = 0;; ++attempt) {
        if (attempt > 8) return fail(b, EMP_ERR_CAPACITY, "step does not settle");
        if (need_part) {
            if ((rc = plan_step(b, nj, false, b->active, sl))) return rc;
            if ((rc = ensure_scratch(b, sl))) return rc;
            if ((rc = ensure_pool(b, std::max<uint64_t>(4096, sl.live / SUBS / 2),
                                  std::max<uint64_t>(1u << 20, sl.live / 8))))
                return rc;
            if ((rc = upload_tables(b, nj))) return rc;
        }
        if (need_part && !b->opts.profile && !b->opts.no_graph) {
            if ((rc = graph_step(b, nj, sl))) return rc;
        } else {
            if ((rc = zero_step(b, sl.ncoarse, sl.nfine, need_part, false))) return rc;
            if (need_part) {
                uint64_t live[2] = {0, 0};
                for (uint32_t i = 0; i < nj; ++i) {
                    live[0] += b->h_jobs[i].cnt[0];
                    live[1] += b->h_jobs[i].cnt[1];
                }
                if ((rc = partition_side(b, 0, true, sl.tiles1[0], sl.tiles2[0], 24.0 * live[0]))) return rc;
                if ((rc = partition_side(b, 1, true, sl.tiles1[1], sl.tiles2[1], 24.0 * live[1]))) return rc;
            }
            Timed t(b, "k_bjoin", 12.0 * sl.live);
            BCHK(b, launch_bjoin(join_args(b, nj), (uint32_t)sl.nfine, b->fcap, b->stream));
        }
        if ((rc = read_counters(b, nj))) return rc;
        if (b->h_ctr->abort) {
            after_abort(b);
            need_part = true;
            continue;
        }
        need_part = false;
        if (b->h_ctr->n_big && (rc = big_path(b, nj, sl, false))) return rc;
        const uint64_t ovf = b->h_ctr->pool_ovf;
        if (ovf > b->pool_ovf_cap) {
            if ((rc = ensure_pool(b, b->pool_cap, std::max<uint64_t>(2 * b->pool_ovf_cap, ovf + ovf / 4 + 1024))))
                return rc;
            continue;
        }
        break;
    }
    const uint32_t *ejob = b->h_lcnt + 2 * b->tab_cap;
    uint64_t total = 0;
    for (uint32_t i = 0; i < nj; ++i) total += ejob[i];
    if ((rc = ensure_out(b, std::max<uint64_t>(total, 1)))) return rc;
    if (total) {
        EdgeCopyArgs ea{b->d_edge_cnt, b->d_edge_off, b->d_dst_off, b->d_eblk, sl.nfine,
                        b->d_pool, b->d_out_from, b->d_out_to};
        {
            Timed t(b, "k_edge_compact", 16.0 * total + 20.0 * (double)sl.nfine);
            BCHK(b, launch_edge_compact(ea, b->stream));
        }
        BCHK(b, hipMemcpyAsync(b->h_from, b->d_out_from, total * 4, hipMemcpyDeviceToHost, b->stream));
        BCHK(b, hipMemcpyAsync(b->h_to, b->d_out_to, total * 4, hipMemcpyDeviceToHost, b->stream));
        if ((rc = sync(b))) return rc;
    }
    // records in job order (each job's edges are contiguous: fine buckets are job-major)
    uint64_t off = 0;
    std::vector<int> still;
    still.reserve(nj);
    for (uint32_t i = 0; i < nj; ++i) {
        const int id = b->active[i];
        HJob &h = b->jobs[(size_t)id];
        const uint64_t n = ejob[i];
        if (sink) {
            const int src = sink(user, id, n, h.L - 1 - h.t, b->h_from + off, b->h_to + off);
            if (src) return fail(b, src, "record sink returned %d", src);
        }
        off += n;
        h.n_edges += n;
        h.S = b->h_lcnt[i];
        h.P = b->h_lcnt[b->tab_cap + i];
        ++h.t;
        if (h.t >= h.rounds || h.S == 0 || h.P == 0)
            finish_job(b, id, EMP_OK, done, user);
        else
            still.push_back(id);
    }
    b->active.swap(still);
    return EMP_OK;
}

// The job's bucket loads: the first tuning class it belongs to.
void classify(const emp_batch *b, HJob &h) {
    h.fine_load = BTARGET;
    h.coarse_load = BCOARSE;
    for (const emp_batch_class &c : b->classes)
        if (h.L <= c.read_len_max && h.n <= c.n_max) {
            h.fine_load = c.fine_load;
            h.coarse_load = c.coarse_load;
            return;
        }
}

int submit_job(emp_batch *b, HJob &&h, int32_t *job) {
    if (h.n > BJOB_MAX_READS)
        return fail(b, EMP_ERR_TOO_MANY, "a batched job holds at most %llu reads (got %llu)",
                    (unsigned long long)BJOB_MAX_READS, (unsigned long long)h.n);
    if (h.n && (h.L <= 0 || h.L > MAX_RD_LEN)) return fail(b, EMP_ERR_READLEN, "read length %d outside [1, 255]", h.L);
    if (h.o < 0) return fail(b, EMP_ERR_ARG, "negative minimum overlap");
    h.submit_ms = now_ms() - b->t0;
    classify(b, h);
    b->jobs.push_back(std::move(h));
    const int id = (int)b->jobs.size() - 1;
    b->queued.push_back(id);
    if (job) *job = id;
    return EMP_OK;
}

struct FastqJob {
    std::vector<char> *buf;
    int L;
};

int fastq_cb(void *user, const char *seq, size_t len) {
    FastqJob *f = (FastqJob *)user;
    if (f->L == 0) {
        if (len > (size_t)MAX_RD_LEN) return EMP_ERR_READLEN;
        f->L = (int)len;
    }
    if ((int)len != f->L) return EMP_ERR_READLEN;
    f->buf->insert(f->buf->end(), seq, seq + len);
    return EMP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
extern "C" {

void emp_batch_opts_default(emp_batch_opts *o) { memset(o, 0, sizeof *o); }

int emp_batch_create(const emp_batch_opts *opts, emp_batch **out) {
    if (!out) return EMP_ERR_ARG;
    *out = nullptr;
    emp_batch *b = new emp_batch();
    if (opts) b->opts = *opts;
    b->device = b->opts.device;
    b->t0 = now_ms();
    if (hipSetDevice(b->device) != hipSuccess) {
        delete b;
        return EMP_ERR_HIP;
    }
    if (b->opts.stream) {
        b->stream = (hipStream_t)b->opts.stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            return EMP_ERR_HIP;
        }
        b->own_stream = true;
    }
    void *p = nullptr;
    if (dalloc(b, &p, sizeof(BCounters)) || dalloc(b, (void **)&b->d_big, BIG_MAX * sizeof(uint32_t)) ||
        hipHostMalloc((void **)&b->h_ctr, sizeof(BCounters), hipHostMallocDefault) != hipSuccess) {
        emp_batch_destroy(b);
        return EMP_ERR_NOMEM;
    }
    b->d_ctr = (BCounters *)p;
    *out = b;
    return EMP_OK;
}

void emp_batch_destroy(emp_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &a : b->allocs) (void)hipFree(a.first);
    b->allocs.clear();
    for (int i = 0; i < 2; ++i) {
        if (b->h_stage[i]) (void)hipHostFree(b->h_stage[i]);
        if (b->stage_ev[i]) (void)hipEventDestroy(b->stage_ev[i]);
    }
    if (b->graph.exec) (void)hipGraphExecDestroy(b->graph.exec);
    void *host[] = {b->h_jobs, b->h_pre, b->h_lcnt, b->h_from, b->h_to, b->h_ctr};
    for (void *h : host)
        if (h) (void)hipHostFree(h);
    for (auto &p : b->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : b->free_events) (void)hipEventDestroy(e);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *emp_batch_last_error(const emp_batch *b) { return b ? b->err.c_str() : "no engine"; }

int emp_batch_set_tuning(emp_batch *b, const emp_batch_tuning *t) {
    if (!b) return EMP_ERR_ARG;
    if (!b->queued.empty() || !b->active.empty())
        return fail(b, EMP_ERR_ARG, "the tuning changes only while no job is queued or resident");
    const uint32_t fcap = t ? t->bucket_cap : (uint32_t)BCAP;
    if (fcap != (uint32_t)BCAP && fcap != (uint32_t)BCAP_L)
        return fail(b, EMP_ERR_ARG, "bucket_cap must be %d or %d", BCAP, BCAP_L);
    if (t && t->n_classes > EMP_BATCH_MAX_CLASSES) return fail(b, EMP_ERR_ARG, "more than %d classes", EMP_BATCH_MAX_CLASSES);
    std::vector<emp_batch_class> cls;
    for (uint32_t i = 0; t && i < t->n_classes; ++i) {
        const emp_batch_class &c = t->classes[i];
        if (c.fine_load < 32 || c.fine_load > fcap || c.coarse_load < c.fine_load || c.coarse_load > (1u << 20))
            return fail(b, EMP_ERR_ARG, "class %u: loads %u / %u outside 32..%u / fine..2^20", i, c.fine_load,
                        c.coarse_load, fcap);
        cls.push_back(c);
    }
    if (fcap != b->fcap) {  // the fine slots take the new stride at the next step
        BCHK(b, hipStreamSynchronize(b->stream));
        for (int s = 0; s < 2; ++s) {
            dfree(b, b->d_fine[s]);
            b->d_fine[s] = nullptr;
        }
        b->fine_cap = 0;
        b->fcap = fcap;
    }
    b->classes.swap(cls);
    return EMP_OK;
}

int emp_batch_graph_stats(const emp_batch *b, uint64_t *captures, uint64_t *replays) {
    if (!b) return EMP_ERR_ARG;
    if (captures) *captures = b->graph_captures;
    if (replays) *replays = b->graph_replays;
    return EMP_OK;
}

uint64_t emp_batch_job_bytes(uint64_t n, int32_t read_len) {
    return read_len > 0 ? job_charge(n, read_len) : 0;
}

int emp_batch_submit(emp_batch *b, const char *reads, uint64_t n, int32_t read_len, int32_t min_ovlp,
                     int32_t on_device, int32_t *job) {
    if (!b || (!reads && n)) return EMP_ERR_ARG;
    if (on_device && (((uintptr_t)reads) & 3)) return fail(b, EMP_ERR_ARG, "device reads must be 4-byte aligned");
    HJob h;
    h.reads = reads;
    h.dev = on_device != 0;
    h.n = n;
    h.L = read_len;
    h.o = min_ovlp;
    return submit_job(b, std::move(h), job);
}

int emp_batch_submit_fastq(emp_batch *b, const char *fastq, size_t size, int32_t min_ovlp, int32_t *job) {
    if (!b || (!fastq && size)) return EMP_ERR_ARG;
    HJob h;
    FastqJob f{&h.owned, 0};
    size_t used = 0;
    int rc = parse_fastq_block(fastq, size, true, &used, fastq_cb, &f);
    if (rc == EMP_ERR_FORMAT) return fail(b, rc, "malformed FASTQ: it must start with '@'");
    if (rc == EMP_ERR_BASE) return fail(b, rc, "a kept read holds a byte other than ACGTacgt");
    if (rc == EMP_ERR_READLEN) return fail(b, rc, "reads of a job must have one length, at most 255");
    if (rc) return fail(b, rc, "FASTQ parse failed");
    h.L = f.L;
    h.n = f.L ? h.owned.size() / (size_t)f.L : 0;
    h.reads = h.owned.data();
    h.o = min_ovlp;
    return submit_job(b, std::move(h), job);
}

int emp_batch_step(emp_batch *b, emp_job_sink sink, emp_job_done done, void *user, int32_t *active,
                   int32_t *queued) {
    if (!b) return EMP_ERR_ARG;
    (void)hipSetDevice(b->device);
    int rc = admit(b, done, user);
    if (!rc) rc = round_step(b, sink, done, user);
    if (rc) {
        (void)hipStreamSynchronize(b->stream);
        collect_timings(b);
        return rc;
    }
    if (active) *active = (int32_t)b->active.size();
    if (queued) *queued = (int32_t)b->queued.size();
    return EMP_OK;
}

int emp_batch_run(emp_batch *b, emp_job_sink sink, emp_job_done done, void *user) {
    if (!b) return EMP_ERR_ARG;
    while (!b->queued.empty() || !b->active.empty()) {
        int rc = emp_batch_step(b, sink, done, user, nullptr, nullptr);
        if (rc) return rc;
    }
    return EMP_OK;
}

int emp_batch_job_info(const emp_batch *b, int32_t job, emp_job_info *out) {
    if (!b || !out || job < 0 || (size_t)job >= b->jobs.size()) return EMP_ERR_ARG;
    const HJob &h = b->jobs[(size_t)job];
    out->state = h.state;
    out->status = h.status;
    out->read_len = h.L;
    out->min_ovlp = h.o;
    out->rounds = h.t;
    out->pad = 0;
    out->n_reads = h.n;
    out->n_dedup = h.n_dedup;
    out->n_edges = h.n_edges;
    out->submit_ms = h.submit_ms;
    out->admit_ms = h.admit_ms;
    out->done_ms = h.done_ms;
    return EMP_OK;
}

int emp_batch_reset(emp_batch *b) {
    if (!b) return EMP_ERR_ARG;
    if (!b->queued.empty() || !b->active.empty()) return fail(b, EMP_ERR_ARG, "jobs are queued or resident");
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    collect_timings(b);
    b->jobs.clear();
    b->charged = 0;
    return EMP_OK;
}

double emp_batch_clock_ms(const emp_batch *b) { return b ? now_ms() - b->t0 : 0.0; }

int32_t emp_batch_kernel_count(const emp_batch *b) { return b ? (int32_t)b->kstats.size() : 0; }

int emp_batch_kernel_info(const emp_batch *b, int32_t i, const char **name, uint64_t *launches, double *ms,
                          double *bytes) {
    if (!b || i < 0 || (size_t)i >= b->kstats.size()) return EMP_ERR_ARG;
    const KStat &k = b->kstats[(size_t)i];
    if (name) *name = k.name.c_str();
    if (launches) *launches = k.launches;
    if (ms) *ms = k.ms;
    if (bytes) *bytes = k.bytes;
    return EMP_OK;
}

void emp_batch_kernel_clear(emp_batch *b) {
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    collect_timings(b);
    b->kstats.clear();
}

void emp_batch_set_profile(emp_batch *b, int32_t on) {
    if (b) b->opts.profile = on;
}

void *emp_batch_stream(const emp_batch *b) { return b ? (void *)b->stream : nullptr; }

}  // extern "C"
