// GPU FASTQ parser (emp_fastq_gpu.hip): launchers and the result block they fill.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emp {

constexpr uint32_t FQ_ERR_FORMAT = 1;  // a block without a leading '@', or holding no complete record
constexpr uint32_t FQ_ROW_MAX = 253;    // longest read the classify pass writes rows for (launch_fq_classify_copy)

struct FqResult {
    uint64_t newlines;              // '\n' bytes in the range
    uint64_t use;                   // bytes of the range in the blocks parsed (the caller resumes there)
    uint64_t records;               // sequence lines in those blocks
    uint64_t blocks;                // blocks parsed (the block after them failed when err is set)
    unsigned long long first_bad;   // first record whose kept-class read holds a byte outside ACGTacgt
    unsigned long long first_kept;  // first kept record
    unsigned long long first_len;   // first kept record of another length than L, or over 255
    uint64_t kept;                  // kept records
    uint32_t err;                   // FQ_ERR_FORMAT
    uint32_t kept_len;              // length of the first kept record
    unsigned long long first_odd;   // classify with copy: first record not kept or of another length
    uint64_t len0;                  // length of record 0's sequence line (the block chain)
};

// Exclusive scan of n counts into off[0..n] (off[n] = total); blk: fq_scan_blocks(n) scratch.
hipError_t launch_scan(const uint32_t *cnt, uint64_t n, unsigned long long *off, unsigned long long *blk,
                       hipStream_t s);

uint64_t fq_chunks(uint64_t size);
// Block totals a scan of n counts needs (the blk scratch of launch_fq_count / launch_fq_classify).
uint64_t fq_scan_blocks(uint64_t n);

// Newline counts per 4 KiB chunk of the range (cnt: fq_chunks(size) words) and their exclusive
// offsets (choff: fq_chunks(size) + 1, the last one the range's newline total).
// blk: fq_scan_blocks(fq_chunks(size)).
hipError_t launch_fq_count(const uint8_t *d, uint64_t size, uint32_t *cnt, unsigned long long *choff,
                           unsigned long long *blk, hipStream_t s);
// Every newline's position (nlpos: one per newline) and the block chain of the range into r: blocks
// of B bytes (the last, shorter one only when range_final), or with one_block the whole range as
// one block, final when one_final.
hipError_t launch_fq_frame(const uint8_t *d, uint64_t size, uint64_t B, int range_final, int one_block, int one_final,
                           const unsigned long long *choff, uint64_t *nlpos, FqResult *r, hipStream_t s);
// Per record of the parsed blocks: kept flag and length; rank = exclusive offsets of the kept
// records (records + 1); the first kept record and the first kept record of another length than L
// (L == 0: the first kept record's). blk: fq_scan_blocks(records).
hipError_t launch_fq_classify(const uint8_t *d, uint64_t size, const uint64_t *nlpos, uint64_t NL, uint64_t records,
                              FqResult *r, uint32_t *kept, uint32_t *len, unsigned long long *rank,
                              unsigned long long *blk, uint32_t L, hipStream_t s);
// The classify pass (kept flags, lengths, first_bad) with the rows of the records that are kept and
// L long (L <= FQ_ROW_MAX) copied to out + record * L; first_odd: the first record that is not
// (none: every record is a kept read of length L and the rows are final). Otherwise launch_fq_rank
// and launch_fq_copy finish as after launch_fq_classify.
hipError_t launch_fq_classify_copy(const uint8_t *d, uint64_t size, const uint64_t *nlpos, uint64_t NL,
                                   uint64_t records, FqResult *r, uint32_t *kept, uint32_t *len, uint32_t L,
                                   uint8_t *out, hipStream_t s);
// The rank pass of launch_fq_classify (after its flags and lengths).
hipError_t launch_fq_rank(uint64_t records, FqResult *r, const uint32_t *kept, const uint32_t *len,
                          unsigned long long *rank, unsigned long long *blk, uint32_t L, hipStream_t s);
// The kept reads, row-major, read i at out + i * L.
hipError_t launch_fq_copy(const uint8_t *d, uint64_t size, const uint64_t *nlpos, uint64_t NL, uint64_t records,
                          const uint32_t *kept, const unsigned long long *rank, uint32_t L, uint8_t *out,
                          hipStream_t s);

}  // namespace emp
