/*
 * TEST INFRASTRUCTURE ONLY — the oracle.
 *
 * A CPU restatement, in C11 + OpenMP, of empanada's fingerprint -> sort -> suffix/prefix overlap
 * path. It follows the reference's steps (sort by read id, even/odd partition, gather, rolling
 * update, stable sort by key, lower/upper bound, edge expansion, self-edge removal, used-read
 * removal) so that its output.bin is the reference's. It is the parity checker for the HIP
 * library and the host-core baseline; the product library never links, loads or calls it.
 *
 * Semantics restated for every M = reduce_mem / 16: the key-aligned search partitions that split
 * a round into several records when a list reaches M/2, and the out-of-core order of lists over M
 * entries (the chunked sort and FIFO pairwise merge of cq_sort_by_key, the chunked regeneration,
 * the partitioned unique).
 */
#ifndef EMP_ORACLE_H
#define EMP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    EO_OK = 0,
    EO_ERR_IO = -1,
    EO_ERR_FORMAT = -2,      /* first byte of a block is not '@', or a record is longer than a block */
    EO_ERR_READLEN = -3,     /* kept reads of unequal length, or length > 255 */
    EO_ERR_BASE = -4,        /* kept read holds a byte other than ACGTacgt */
    EO_ERR_OUT_OF_CORE = -5, /* unused: every M is restated */
    EO_ERR_PARTITION = -6,   /* a partition with no cut (e.g. more than M/2 equal keys): the reference loops forever */
    EO_ERR_NOMEM = -7,
    EO_ERR_ARG = -8,
    EO_ERR_TOO_MANY = -9     /* 2n >= 2^32 read ids */
};

typedef struct {
    int min_ovlp;        /* -o */
    uint64_t map_mem;    /* -m; FASTQ block = map_mem / 16 bytes */
    uint64_t reduce_mem; /* -r; M = reduce_mem / 16 */
} eo_opts;

typedef struct {
    uint64_t n_reads;    /* kept reads */
    int read_len;
    uint64_t n_dedup;    /* entries after dedup */
    int n_rounds;        /* rounds executed */
    uint64_t n_records;
    uint64_t n_edges;
} eo_summary;

/* Per-round live counts (before and after removal) and edges written. */
typedef struct {
    uint64_t s, p, s_after, p_after, edges;
    int ovlp;
    int records;
} eo_round;

/* Output records are appended to a growable memory buffer. */
typedef struct {
    unsigned char *data;
    size_t size, cap;
} eo_buf;

void eo_buf_free(eo_buf *b);

/* Constants and rolling functions (unit level). */
uint32_t eo_modulo_exp(uint64_t a, unsigned n, unsigned p);
void eo_modulos(uint32_t *m1, uint32_t *m2); /* 256 entries each */
uint64_t eo_key_of(const char *s, int len);
uint64_t eo_key_of_rc(const char *s, int len);
uint64_t eo_next_sfx(uint64_t key, char base, int len);
uint64_t eo_prev_pfx(uint64_t key, char base);
char eo_revc(char c);

/*
 * FASTQ -> kept reads, with the reference's block / seek-back / N-filter rules. On success
 * *reads (malloc'd, row-major n x L) and *n, *L are set.
 */
int eo_parse_fastq_file(const char *path, uint64_t block_bytes, char **reads, uint64_t *n, int *L);
int eo_parse_fastq_mem(const char *data, uint64_t size, uint64_t block_bytes, char **reads, uint64_t *n,
                       int *L);

/*
 * The key-aligned search partitions of one round (cq_get_search_partitions, cqutils.cuh:107-137)
 * over key-sorted suffix keys a and prefix keys b: up to cap (na, nb) pairs into parts, *n_parts
 * of them. EO_ERR_PARTITION where the reference loops forever (a window with no cut).
 */
int eo_search_partitions(const uint64_t *a, size_t nA, const uint64_t *b, size_t nB, size_t M, size_t *parts,
                         size_t cap, size_t *n_parts);

/*
 * The out-of-core sort: cq_sort_by_key (cqutils.cuh:179-247) of n (key, value) pairs with device
 * buffer M (in place), and its merge partitions (cq_get_merge_partitions, cqutils.cuh:36-84) of
 * two key-sorted runs: up to cap (na, nb) pairs, *n_parts of them.
 */
int eo_cq_sort_by_key(uint64_t *keys, uint32_t *vals, size_t n, size_t M);
int eo_merge_partitions(const uint64_t *a, size_t nA, const uint64_t *b, size_t nB, size_t M, size_t *parts,
                        size_t cap, size_t *n_parts);

/* The dedup alone: the kept ids of the 2n entries of n reads in key order (ids: 2n of room). */
int eo_dedup(const char *reads, uint64_t n, int L, size_t M, uint32_t *ids, uint64_t *nd);

/* Kept reads -> output records (appended to out). rounds may be NULL (else >= 256 entries). */
int eo_overlap_reads(const char *reads, uint64_t n, int L, int min_ovlp, uint64_t reduce_mem, eo_buf *out,
                     eo_summary *sum, eo_round *rounds);

/*
 * Collision audit (the FPCHECK diagnostic, overlap.cu:215-219,274-284,295-359): for a key-sorted
 * list, the distinct fingerprints (n_uniq_fp) and the distinct strings summed over fingerprints
 * (n_uniq_rd), the strings being the oriented reads' substrings the keys stand for. One record for
 * the 2n entries before dedup (full reads), then per round the suffix list and the prefix list
 * as they are joined. Bases compare by their 2-bit code (case-insensitive).
 */
typedef struct {
    int32_t list;      /* 0: the 2n entries before dedup, 1: suffixes, 2: prefixes */
    int32_t ovlp;      /* string length: L for list 0, the round's overlap otherwise */
    uint64_t entries;  /* list size */
    uint64_t keys;     /* distinct fingerprints */
    uint64_t strings;  /* distinct strings, summed over fingerprints */
} eo_audit;

/* eo_overlap_reads that also fills up to audit_cap audit records (*n_audit of them). */
int eo_overlap_reads_audit(const char *reads, uint64_t n, int L, int min_ovlp, uint64_t reduce_mem, eo_buf *out,
                           eo_summary *sum, eo_round *rounds, eo_audit *audit, int audit_cap, int *n_audit);

/* Whole pipeline: FASTQ file -> output.bin file. */
int eo_run_file(const char *in_path, const char *out_path, const eo_opts *opts, eo_summary *sum,
                eo_round *rounds);

/* Same, with FASTQ bytes in memory and the records returned in memory. */
int eo_run_mem(const char *fastq, uint64_t size, const eo_opts *opts, eo_buf *out, eo_summary *sum,
               eo_round *rounds);

const char *eo_strerror(int code);

#ifdef __cplusplus
}
#endif

#endif
