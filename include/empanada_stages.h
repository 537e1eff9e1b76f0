/*
 * empanada-mi355x — stage-level C ABI: single kernels of the path exposed on host buffers, for
 * callers (and parity tests) that need one stage rather than the whole pipeline. Each call
 * allocates its device buffers, runs on the default stream of `device`, copies the result back
 * and frees; they are not the fast path.
 */
#ifndef EMPANADA_STAGES_H
#define EMPANADA_STAGES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/*
 * generate_fingerprint_read_pairs (fingerprint.cu:75-115) with get_reverse_complements
 * (preprocess.cu:246-284): for n kept reads (row-major n x len, ACGTacgt) writes 2n keys,
 * key(read_k) at 2k and key(revcomp(read_k)) at 2k+1, each HIGH32 = H(R=5,P=536870909),
 * LOW32 = H(R=7,P=536870879) (assembly.cuh:29-34).
 */
int emp_stage_fingerprints(int32_t device, const char *reads, uint64_t n, int32_t len, uint64_t *keys);

/*
 * transpose_reads + TBlockList::pack (fingerprint.cu:125-153, blocks.cuh:30-47): the 2-bit column
 * store as the map kernel writes it for the rounds, code (c>>1)&3 (A0 C1 T2 G3). Reads are grouped
 * in chunks of 2^20, columns in groups of 8; one 32-bit word holds 2 reads x 8 columns: base j of
 * read k is at bits 2*(8*(k%2) + j%8) of word (ch * G8 + j/8) * 2^19 + (k % 2^20) / 2, with
 * ch = k / 2^20 and G8 = ceil(len/8). cols holds ceil(n / 2^20) * G8 * 2^19 words; bits of reads
 * past n are unspecified, columns past len are zero.
 */
int emp_stage_columns(int32_t device, const char *reads, uint64_t n, int32_t len, uint32_t *cols);

/*
 * The rolling transforms of regenerate_fingerprints (overlap.cu:647,651): kind 0 = NextSfx with
 * length `len` (overlap.cu:57-75), kind 1 = PrevPfx (overlap.cu:77-91, including its uint64_t
 * wrap). bases are ASCII (ACGTacgt); out may alias keys.
 */
int emp_stage_roll(int32_t device, const uint64_t *keys, const char *bases, uint64_t n, int32_t kind, int32_t len,
                   uint64_t *out);

/*
 * The out-of-core order of cq_sort_by_key (cqutils.cuh:179-247: chunks of M sorted in input
 * order, the FIFO of sorted runs merged pairwise through cq_get_merge_partitions, 36-84, and
 * cq_merge_by_key, 139-177) as the overlap phase reproduces it once a list exceeds
 * M = reduce_mem / 16: writes a tie key per entry such that the entries sorted by (key, tie) are
 * in cq_sort_by_key's output order. The list is n entries (keys[i], rids[i]), keys of two 29-bit
 * halves (HIGH32 << 32 | LOW32), ids distinct; its input order is
 *   rounds = 0, the dedup (overlap.cu:212-213): id order, ids a permutation of 0..n-1;
 *   rounds = 1, a round (overlap.cu:266-272): the list sorted by id, cut into chunks of M, each
 *     chunk's even ids then its odd ids (regenerate_fingerprints, overlap.cu:613-659).
 * bucket_bits (1..29) is the bucketing the library would use for the list: bucket of a key =
 * HIGH32 >> (29 - bucket_bits). This is the host planner of that path on host arrays (no device
 * work). Returns EMP_ERR_OUT_OF_CORE when n needs more than 65535 chunks, and for rounds = 0
 * EMP_ERR_PARTITION (ties written) when cq_get_partitions (cqutils.cuh:86-105) would not end:
 * a window of M entries, not the last one, holding a single key.
 */
int emp_stage_ooc_ties(const uint64_t *keys, const uint32_t *rids, uint64_t n, uint64_t M, int32_t rounds,
                       int32_t bucket_bits, uint64_t *ties);

/*
 * Allocation-free form of emp_stage_fingerprints. It has no counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Input and output are the caller's device memory: d_reads (n x
 * len row-major, 4-byte aligned), d_keys (2n keys, as emp_stage_fingerprints writes them) and
 * d_status (one word: 0, or 1 when a read holds a byte other than ACGTacgt, the case where
 * emp_stage_fingerprints returns EMP_ERR_BASE). Everything else lives in a device array the
 * library holds on each GPU for the life of the process (about 24 MiB). It runs asynchronously on
 * `stream` (a hipStream_t of the current device; NULL = its null stream), calls no allocating
 * routine and never synchronises. At most EMP_NOALLOC_MAX_READS reads per call (EMP_ERR_TOO_MANY
 * above). Calls on one GPU share its array: calls that overlap on one GPU race.
 */
#define EMP_NOALLOC_MAX_READS 1048576
int emp_stage_fingerprints_noalloc(const void *d_reads, uint64_t n, int32_t len, uint64_t *d_keys, uint32_t *d_status,
                                   void *stream);

#ifdef __cplusplus
}
#endif

#endif
