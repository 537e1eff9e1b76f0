/*
 * empanada-mi355x — batched and served jobs (libempanada_hip.so).
 *
 * Many independent overlap jobs (read sets) share one engine and every kernel launch of the
 * path: a step partitions, dedups and joins the live entries of all resident jobs at once, each
 * job at its own round, read length and minimum overlap. Each job's records are exactly the
 * output.bin records (overlap.cu:148-161) that emp_run would write for that job alone with the
 * reference's defaults (in-core, one record per round).
 *
 * The reference has no counterpart: it runs one read set per process (main.cu:9-77). The engine
 * is for callers that hold thousands of small read sets, and for a request loop that admits jobs
 * while others run (continuous batching): emp_batch_step admits queued jobs whose estimated HBM
 * need fits the budget, then advances every resident job by one round.
 *
 * Threading: one engine per host thread; an engine owns its device memory and its HIP stream.
 */
#ifndef EMPANADA_BATCH_H
#define EMPANADA_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "empanada.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emp_batch emp_batch;

typedef struct {
    uint64_t hbm_budget;   /* admission budget: sum of the estimated device bytes of resident jobs
                              (emp_batch_job_bytes); 0 = admit every queued job at the next step */
    int32_t device;        /* HIP device ordinal */
    int32_t profile;       /* non-zero: time every kernel launch with HIP events */
    void *stream;          /* hipStream_t to run on; NULL = the engine creates one */
    uint32_t max_active;   /* resident jobs at most; 0 = no limit */
    int32_t no_graph;      /* non-zero: launch a round step's kernels one by one instead of replaying the
                              step's captured HIP graph (profile also launches them one by one) */
} emp_batch_opts;

/* A class of jobs and its bucket loads (emp_batch_set_tuning). */
typedef struct {
    int32_t read_len_max;  /* a job belongs to the first class with read_len <= read_len_max */
    uint32_t n_max;        /*   and n_reads <= n_max */
    uint32_t fine_load;    /* mean entries per fine bucket and side: 32 .. bucket_cap */
    uint32_t coarse_load;  /* mean entries per coarse bucket and side: fine_load .. 2^20 */
} emp_batch_class;

#define EMP_BATCH_MAX_CLASSES 16
typedef struct {
    uint32_t bucket_cap;   /* fine-bucket capacity, which selects the join and dedup kernels:
                              512 (256 threads) or 1024 (512 threads) */
    uint32_t n_classes;
    emp_batch_class classes[EMP_BATCH_MAX_CLASSES];  /* a job in no class: loads 256 / 16384 */
} emp_batch_tuning;

/* Job states reported by emp_batch_job_info. */
#define EMP_JOB_QUEUED 0
#define EMP_JOB_ACTIVE 1
#define EMP_JOB_DONE 2

typedef struct {
    int32_t state;         /* EMP_JOB_* */
    int32_t status;        /* EMP_OK or the EMP_ERR_* the job ended with */
    int32_t read_len, min_ovlp;
    int32_t rounds;        /* rounds run (= records written) */
    int32_t pad;
    uint64_t n_reads, n_dedup, n_edges;
    double submit_ms, admit_ms, done_ms;  /* engine clock (milliseconds since emp_batch_create) */
} emp_job_info;

/* Receives job `job`'s records in output.bin order (arrays valid during the call). A non-zero
 * return aborts the step with that value. A NULL sink drops the records (they are still copied
 * to host memory). */
typedef int (*emp_job_sink)(void *user, int32_t job, uint64_t n, int32_t ovlp, const uint32_t *from,
                            const uint32_t *to);
/* Called once per job when it ends (status EMP_OK or the job's error). */
typedef void (*emp_job_done)(void *user, int32_t job, int32_t status);

void emp_batch_opts_default(emp_batch_opts *opts);
int emp_batch_create(const emp_batch_opts *opts, emp_batch **b);
void emp_batch_destroy(emp_batch *b);
const char *emp_batch_last_error(const emp_batch *b);

/*
 * Queues a job of n kept reads (row-major n x read_len, ACGTacgt), minimum overlap min_ovlp.
 * The reads are borrowed until the job is admitted (emp_batch_job_info reports state != QUEUED);
 * on_device != 0: they are device memory (4-byte aligned). *job receives the job id (ids count
 * from 0 in submission order). A job holds at most 2^23 reads.
 */
int emp_batch_submit(emp_batch *b, const char *reads, uint64_t n, int32_t read_len, int32_t min_ovlp,
                     int32_t on_device, int32_t *job);

/* Queues a job given as FASTQ text, parsed with read_block's rules (preprocess.cu:195-237) into
 * a buffer the engine owns. */
int emp_batch_submit_fastq(emp_batch *b, const char *fastq, size_t size, int32_t min_ovlp, int32_t *job);

/* Estimated device bytes a job of n reads of length read_len holds while resident. */
uint64_t emp_batch_job_bytes(uint64_t n, int32_t read_len);

/*
 * One step: admits queued jobs (FIFO) while the budget and max_active allow, maps and dedups
 * them, then runs one round of every resident job and passes each job's record to `sink`.
 * Jobs that finish are reported to `done`. *active / *queued (may be NULL) receive the counts
 * after the step.
 */
int emp_batch_step(emp_batch *b, emp_job_sink sink, emp_job_done done, void *user, int32_t *active,
                   int32_t *queued);

/* Steps until no job is queued or resident. */
int emp_batch_run(emp_batch *b, emp_job_sink sink, emp_job_done done, void *user);

int emp_batch_job_info(const emp_batch *b, int32_t job, emp_job_info *out);

/* Forgets every finished job (ids restart at 0); allocations are kept. Fails while jobs are
 * queued or resident. */
int emp_batch_reset(emp_batch *b);

/*
 * Sets the engine's tuning, as the autotuner (empanada_amd/tune.py) picks it per class of jobs
 * from measurements and keeps it in a JSON cache; NULL restores the defaults (bucket_cap 512,
 * loads 256 / 16384 for every job). Outputs do not depend on it. Fails while jobs are queued or
 * resident.
 */
int emp_batch_set_tuning(emp_batch *b, const emp_batch_tuning *t);

/* Round steps launched from a captured graph so far, and graphs captured for them. */
int emp_batch_graph_stats(const emp_batch *b, uint64_t *captures, uint64_t *replays);

/* Milliseconds on the engine clock. */
double emp_batch_clock_ms(const emp_batch *b);

/* Kernel timing (profile != 0), as emp_kernel_info. */
int32_t emp_batch_kernel_count(const emp_batch *b);
int emp_batch_kernel_info(const emp_batch *b, int32_t i, const char **name, uint64_t *launches, double *ms,
                          double *bytes);
void emp_batch_kernel_clear(emp_batch *b);
void emp_batch_set_profile(emp_batch *b, int32_t on);
void *emp_batch_stream(const emp_batch *b);

#ifdef __cplusplus
}
#endif

#endif
