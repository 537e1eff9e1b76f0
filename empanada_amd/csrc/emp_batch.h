// Batched (segmented) form of the path: many independent jobs (read sets) share every launch.
//
// Every job owns a segment of the two list arrays and a region of the column store, and its
// entries carry job-local read ids. A step partitions the live entries of all jobs at once, job by
// job, into fine buckets: job j's keys go to fine buckets fb_j .. fb_j + nb_j - 1 by
// floor(h1 * nb_j / 2^29), which keeps key order and makes every bucket hold entries of one job
// only, so the join and the dedup compare keys exactly as the single-job kernels do. Each job
// carries its own round index, read length and minimum overlap: jobs that entered at different
// steps, with different read lengths, advance together.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emp_kernels.h"

namespace emp {

// A job's fine buckets hold BTARGET entries per side on average and at most BCAP (the rest
// spills to the large-bucket path); at that load the matched records of a bucket stay below BMCAP
// up to a match rate of about 100 %.
constexpr int BCAP = 512;             // fine-bucket capacity per side (the default kernel variant)
constexpr int BMCAP = 256;            // matched records per side a bucket may hold in LDS
constexpr int BJOIN_THREADS = 256;    // work-group size of the batched join and dedup
// The other variant of the join and dedup: buckets of 1,024 per side, 512 matched records,
// 512 threads (emp_batch_set_tuning chooses).
constexpr int BCAP_L = 1024, BMCAP_L = 512, BJOIN_THREADS_L = 512;
constexpr uint32_t BTARGET = 256;     // mean fine-bucket load of a job (default tuning)
constexpr uint32_t BCOARSE = 16384;   // mean coarse-bucket load of a job (default tuning)
constexpr uint32_t BSPLIT = 256;      // coarse buckets per job, and fine per coarse, at most
constexpr uint64_t BJOB_MAX_READS = 1ull << 23;  // 2^24 entries per side: nb_j <= BSPLIT^2

// A job as one step sees it (uploaded per step).
struct BJob {
    uint64_t seg;       // first entry of the job's segment in both list arrays
    uint64_t cbase;     // first word of the job's column region
    uint32_t gs;        // words per column group of the region
    uint32_t cnt[2];    // live entries per side entering the step
    uint32_t src[2];    // list array that side's first pass reads
    uint32_t nc, m, nb; // coarse buckets, fine buckets per coarse bucket (nb = nc * m)
    uint32_t cb, fb;    // first coarse / fine bucket of the job (global indices)
    uint32_t ccap[2];   // coarse-bucket capacity per side
    uint64_t coff[2];   // first coarse slot per side
    uint32_t keep;      // survivors are written (the job has a next round)
    uint32_t pad;
    RoundArgs rg;       // the job's round (rg.active = 0 in a dedup step)
};

// Every launch of a step takes the job tables' capacity as its job count: the prefix tables hold
// the step's totals from its last job on, job j's counters sit at a stride of that capacity, and a
// work-group past the step's total exits, so launches captured in a graph with larger grids can be
// replayed for later steps.
//
// One partition pass of one side over every job of a step.
//   pass 1: job segments -> coarse buckets (regenerating the entries for the job's round)
//   pass 2: coarse buckets -> fine buckets (overflow to the spill area)
struct BPartArgs {
    const BJob *jobs;
    uint32_t njobs;
    const uint32_t *tpre;    // [njobs + 1] first work-group of each job in this pass
    int32_t side;
    const Entry *lists[2];
    const uint32_t *cols;
    Entry *coarse;           // this side's coarse slots
    uint32_t *ccnt;          // this side's coarse counts
    Entry *fine;             // this side's fine slots, fcap per bucket
    uint32_t *fcnt;          // this side's fine counts
    uint32_t fcap;           // fine-bucket capacity (BCAP or BCAP_L)
    SpillRec *spill;
    uint32_t *spill_n;
    uint32_t spill_cap;
    uint32_t *abort;         // set when a coarse bucket or the spill area overflows
};

// Dedup (side 0) or join of every fine bucket of a step.
struct BJoinArgs {
    const BJob *jobs;
    uint32_t njobs;
    const uint32_t *fpre;    // [njobs + 1] first fine bucket of each job
    const Entry *s_in, *p_in;
    const uint32_t *s_cnt, *p_cnt;
    Entry *lists[2];         // survivors go to the job's segment of lists[side]
    uint32_t *lcnt;          // [2][njobs] survivor cursors
    uint32_t *ejob;          // [njobs] edges per job
    uint2 *pool;  // edge pool: (from, to) pairs
    unsigned long long *pool_cursor;
    unsigned long long pool_cap;
    unsigned long long *ovf_cursor;
    unsigned long long ovf_cap;
    unsigned long long *edge_off;
    uint32_t *edge_cnt;
    uint32_t *big, *n_big;
    uint32_t big_cap;
    const uint32_t *abort;
};

// Buckets over capacity: one work-group per bucket sorts it in global scratch memory.
struct BBigArgs {
    const BJob *jobs;
    uint32_t njobs;
    const uint32_t *fpre;
    const uint32_t *ids;
    const unsigned long long *soff, *poff;
    uint64_t *skey, *pkey;
    uint32_t *srank, *prank, *saux, *paux;
    const Entry *s_in, *p_in;
    const uint32_t *s_cnt, *p_cnt;
    const SpillRec *s_spill, *p_spill;
    const uint32_t *s_spill_n, *p_spill_n;
    uint32_t spill_cap;
    const uint32_t *cols;
    Entry *lists[2];
    uint32_t *lcnt, *ejob;
    uint2 *pool;  // edge pool: (from, to) pairs
    unsigned long long *pool_cursor;
    unsigned long long pool_cap;
    unsigned long long *ovf_cursor;
    unsigned long long ovf_cap;
    unsigned long long *edge_off;
    uint32_t *edge_cnt;
    uint32_t *too_many;  // set when a bucket holds 2^31 edges or more
    uint32_t fcap;       // fine-bucket capacity
};

hipError_t launch_bpart(const BPartArgs &a, int pass, bool regen, uint32_t groups, hipStream_t s);
// fcap selects the kernel variant (BCAP or BCAP_L); the fine slots have that stride.
hipError_t launch_bdedup(const BJoinArgs &a, uint32_t nb, uint32_t fcap, hipStream_t s);
hipError_t launch_bjoin(const BJoinArgs &a, uint32_t nb, uint32_t fcap, hipStream_t s);
hipError_t launch_bdedup_big(const BBigArgs &a, uint32_t n, hipStream_t s);
hipError_t launch_bjoin_big(const BBigArgs &a, uint32_t n, hipStream_t s);

}  // namespace emp
