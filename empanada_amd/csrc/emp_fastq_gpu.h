// GPU FASTQ block parser (emp_fastq_gpu.hip): launchers and the result block they fill.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emp {

constexpr uint32_t FQ_ERR_FORMAT = 1;  // no leading '@', or a record longer than the block

struct FqResult {
    uint64_t newlines;    // '\n' bytes in the block
    uint64_t use;         // bytes of complete records used (the caller resumes there)
    uint64_t lines;       // lines in use
    uint64_t records;     // sequence lines in use
    unsigned long long first_bad;   // first record whose kept-class read holds a byte outside ACGTacgt
    unsigned long long first_kept;  // first kept record
    unsigned long long first_len;   // first kept record of another length than L, or over 255
    uint32_t err;
    uint32_t pad;
};

uint64_t fq_chunks(uint64_t size);

// Newline counts per chunk (cnt: fq_chunks(size) words) and their exclusive offsets (choff:
// fq_chunks(size) + 1, the last one the block's newline total).
hipError_t launch_fq_count(const uint8_t *d, uint64_t size, uint32_t *cnt, unsigned long long *choff,
                           hipStream_t s);
// Newline positions (nlpos: one per newline) and the block's framing into r.
hipError_t launch_fq_frame(const uint8_t *d, uint64_t size, int final, const unsigned long long *choff,
                           uint64_t *nlpos, FqResult *r, hipStream_t s);
// Per record: kept flag and length; rank = exclusive offsets of the kept records (records + 1).
hipError_t launch_fq_classify(const uint8_t *d, uint64_t size, const uint64_t *nlpos, uint64_t records, FqResult *r,
                              uint32_t *kept, uint32_t *len, unsigned long long *rank, hipStream_t s);
hipError_t launch_fq_lencheck(const uint32_t *kept, const uint32_t *len, uint64_t records, uint32_t L, FqResult *r,
                              hipStream_t s);
hipError_t launch_fq_copy(const uint8_t *d, const uint64_t *nlpos, uint64_t records, const uint32_t *kept,
                          const unsigned long long *rank, uint32_t L, uint8_t *out, hipStream_t s);

}  // namespace emp
