/*
 * empanada-mi355x — C ABI of the MI355X fingerprint -> sort -> suffix/prefix-overlap library
 * (libempanada_hip.so). Plain C types only; every entry point returns EMP_OK (0) or a negative
 * EMP_ERR_* code, never exits, and leaves a message in emp_last_error().
 *
 * The reference has no FFI: its boundary is the stage API of assembly.cuh:67-82 called from
 * main() (main.cu:27-71), plus the CLI (config.h:120-140) and output.bin (overlap.cu:148-161).
 * Each entry point below names the reference interface it replaces.
 *
 * Threading: one context per host thread; a context owns its device memory and its HIP stream.
 */
#ifndef EMPANADA_H
#define EMPANADA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMP_OK 0
#define EMP_ERR_ARG -1          /* bad argument or call order */
#define EMP_ERR_HIP -2          /* a HIP runtime call failed */
#define EMP_ERR_FORMAT -3       /* FASTQ block does not start with '@', or a record exceeds a block */
#define EMP_ERR_READLEN -4      /* kept reads of unequal length, or longer than 255 */
#define EMP_ERR_BASE -5         /* kept read holds a byte other than ACGTacgt */
#define EMP_ERR_OUT_OF_CORE -6  /* a list needs more than 65535 sort chunks of reduce_mem/16 */
#define EMP_ERR_NOMEM -7        /* device memory or the context's HBM budget exhausted */
#define EMP_ERR_TOO_MANY -8     /* 2n >= 2^32 oriented read ids */
#define EMP_ERR_PARTITION -9    /* a key-aligned search partition cannot be formed */
#define EMP_ERR_IO -10          /* output file */
#define EMP_ERR_CAPACITY -11    /* an internal fixed-capacity structure overflowed */

typedef struct emp_ctx emp_ctx;

/* Options: the reference's Config (config.h:120-140) plus context-level settings. */
typedef struct {
    int32_t min_ovlp;        /* -o, default 31 (config.h:124) */
    uint64_t map_mem;        /* -m, default 1 GiB; FASTQ block = map_mem / 16 bytes (main.cu:16) */
    uint64_t reduce_mem;     /* -r, default 1 GiB; M = reduce_mem / 16 (main.cu:69) */
    uint64_t base_blocksz;   /* -b, accepted; sizes host blocks in the reference, no effect here */
    uint64_t kv_chunksz;     /* -c, accepted; sizes host blocks in the reference, no effect here */
    uint64_t hbm_budget;     /* device bytes the context may allocate; 0 = no limit */
    int32_t device;          /* HIP device ordinal */
    int32_t profile;         /* non-zero: time every kernel launch with HIP events */
    void *stream;            /* hipStream_t to run on; NULL = the context creates one */
} emp_opts;

/* Fills the reference's defaults (config.h:124-128). */
void emp_opts_default(emp_opts *opts);

/* Replaces the Config/Kmp/Vmp singletons and init_match globals (main.cu:12-27, overlap.cu:117-127). */
int emp_ctx_create(const emp_opts *opts, emp_ctx **ctx);
void emp_ctx_destroy(emp_ctx *ctx);
const char *emp_last_error(const emp_ctx *ctx);

/*
 * Map phase, one FASTQ block at a time: read_block (preprocess.cu:195-237) followed by
 * transpose_reads, get_reverse_complements and generate_fingerprint_read_pairs
 * (fingerprint.cu:75-153, preprocess.cu:246-284) and the queue pushes of main.cu:50-55.
 * `data` is a block as read_block reads it (at most map_mem/16 bytes from the current file
 * position); `final` is non-zero for a short (last) read. *consumed receives the bytes up to the
 * last complete record; the caller continues from there (the reference's seekg, preprocess.cu:222).
 */
int emp_add_fastq(emp_ctx *ctx, const char *data, size_t size, int final, size_t *consumed);

/*
 * The CLI's map phase over a whole FASTQ file (main.cu:42-58: read_block in blocks of map_mem/16
 * bytes with its seek-back, preprocess.cu:195-237, then the map stages). The file is read ahead
 * into pinned memory by parallel preads and streamed into HBM on a copy stream, each block parsed
 * on the device once its bytes have arrived. Same reads, ids and errors as emp_add_fastq called
 * block by block.
 */
int emp_add_fastq_file(emp_ctx *ctx, const char *path);

/*
 * FASTQ bytes already in device memory (borrowed for the call), parsed as the CLI reads a file:
 * blocks of map_mem/16 bytes, each resumed after the previous one's complete records. `final`
 * non-zero: the bytes end the file (the last, shorter block is parsed too); otherwise parsing stops
 * before a block that would run past the end. *consumed receives the bytes used; the caller
 * continues from there.
 */
int emp_add_fastq_device(emp_ctx *ctx, const void *d_data, size_t size, int final, size_t *consumed);

/* Map phase over kept reads already parsed: host memory, row-major n x read_len, ACGTacgt. Reads in
 * pinned host memory (hipHostMalloc) are copied straight into HBM in double-buffered batches while
 * earlier batches are mapped; large pageable arrays go through pinned stages filled by parallel
 * copies, small ones through the context's staging buffers. */
int emp_add_reads(emp_ctx *ctx, const char *reads, uint64_t n, int32_t read_len);

/* Same, reads already in device memory (borrowed for the call). Unless this is the last batch,
 * the number of reads added so far must stay a multiple of 16. */
int emp_add_reads_device(emp_ctx *ctx, const void *d_reads, uint64_t n, int32_t read_len);

/*
 * Overlap phase: build_graph (overlap.cu:206-293) — dedup, then rounds of regenerate, sort,
 * suffix/prefix join, edge output and used-read removal. Each record of output.bin
 * (overlap.cu:148-161) is passed to `sink` in file order; the arrays are valid during the call.
 * A non-zero return from the sink aborts the run with that value. sink may be NULL: the records
 * are still produced and copied to host memory, then dropped.
 */
typedef int (*emp_record_sink)(void *user, uint64_t n, int32_t ovlp, const uint32_t *from, const uint32_t *to);
int emp_run(emp_ctx *ctx, emp_record_sink sink, void *user);

/* emp_run writing output.bin at `path` (overlap.cu:231-232,148-161). */
int emp_run_to_file(emp_ctx *ctx, const char *path);

/* Drops the reads of the context, keeping its allocations, for another run. */
int emp_reset(emp_ctx *ctx);

typedef struct {
    uint64_t s, p;            /* live suffix / prefix entries at the start of the round */
    uint64_t s_after, p_after;/* after used-read removal */
    uint64_t edges;           /* edges written (after self-edge removal) */
    int32_t ovlp;             /* overlap length L-1-t */
    int32_t records;          /* records written this round */
    int32_t bucket_bits;      /* ceil(log2(buckets)) */
    int32_t big_buckets;      /* buckets handled by the large-bucket path */
    uint32_t buckets;         /* buckets of the round's lists (key h in bucket floor(h buckets / 2^29)) */
    uint32_t pad;
} emp_round_stats;

typedef struct {
    uint64_t n_reads;
    int32_t read_len;
    int32_t n_rounds;
    uint64_t n_dedup;
    uint64_t n_records;
    uint64_t n_edges;
    double map_ms;            /* host wall time of the map phase calls */
    double overlap_ms;        /* host wall time of emp_run */
} emp_summary;

int emp_get_summary(const emp_ctx *ctx, emp_summary *out);
int emp_get_round(const emp_ctx *ctx, int32_t round, emp_round_stats *out);

/*
 * Collision audit: the FPCHECK diagnostic (overlap.cu:215-219,274-284,295-359), a compile-time
 * option of the reference, here switched per context. With it on, emp_run records, for the 2n
 * entries before dedup and then for every round's suffix and prefix list as they are joined, the
 * distinct fingerprints (n_uniq_fp) and the distinct strings summed over fingerprints (n_uniq_rd);
 * the reference logs 1 - keys/strings as "Collision rate". Bases compare by their 2-bit code. The
 * audit adds work and host synchronisation to every round; output.bin is unchanged.
 */
typedef struct {
    int32_t list;             /* 0: the 2n entries before dedup (full reads), 1: suffixes, 2: prefixes */
    int32_t ovlp;             /* string length: L for list 0, the round's overlap otherwise */
    uint64_t entries;         /* list size */
    uint64_t keys;            /* distinct fingerprints */
    uint64_t strings;         /* distinct strings, summed over fingerprints */
} emp_audit;

void emp_set_audit(emp_ctx *ctx, int32_t on);
int32_t emp_audit_count(const emp_ctx *ctx);
int emp_get_audit(const emp_ctx *ctx, int32_t i, emp_audit *out);

/* Kernel timing (profile != 0): launches and summed HIP-event milliseconds per kernel name. */
int32_t emp_kernel_count(const emp_ctx *ctx);
int emp_kernel_info(const emp_ctx *ctx, int32_t i, const char **name, uint64_t *launches, double *ms,
                    double *bytes);
void emp_kernel_clear(emp_ctx *ctx);

/* Turns per-launch event timing on or off for later calls. */
void emp_set_profile(emp_ctx *ctx, int32_t on);

/* The context's stream (hipStream_t) and device, for callers that record their own events. */
void *emp_stream(const emp_ctx *ctx);

/* Version string. */
const char *emp_version(void);

#ifdef __cplusplus
}
#endif

#endif
