// FASTQ block parsing on the GPU with read_block's rules (preprocess.cu:195-237), the same rules
// emp_fastq.cpp applies on the host:
//
//   - the block must start with '@';
//   - unless it is the last block, only complete records are used: n_lines = newlines before the
//     last byte + 1; if n_lines % 4 != 0 or the block does not end in '\n', the first
//     ((n_lines - 1) / 4) * 4 lines are used (none: a record longer than the block);
//   - sequence lines (line % 4 == 1) that hold a byte in [64,127] with bit 3 set are dropped
//     (preprocess.cu:119-121,150-155), empty ones vanish, the others must be ACGTacgt;
//   - kept reads have one length, at most 255.
//
// Passes: newline counts per 4 KiB chunk -> chunk offsets (one work-group scan) -> newline
// positions -> framing (one thread) -> one wave per record classifies its sequence line ->
// kept-record offsets -> one wave per kept record copies its bases (row-major, read_len each).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emp_common.h"
#include "emp_fastq_gpu.h"

namespace emp {

namespace {

constexpr int FQ_THREADS = 256;
constexpr int FQ_PER = 16;
constexpr uint32_t FQ_CHUNK = FQ_THREADS * FQ_PER;  // bytes per work-group of the newline passes

__device__ __forceinline__ uint32_t newlines_of(const uint8_t *d, uint64_t p, uint64_t size, uint32_t &mask) {
    mask = 0;
    if (p + FQ_PER <= size) {
        const uint4 v = *reinterpret_cast<const uint4 *>(d + p);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (((w[i >> 2] >> (8 * (i & 3))) & 0xFFu) == '\n') mask |= 1u << i;
    } else {
        for (int i = 0; i < FQ_PER && p + i < size; ++i)
            if (d[p + i] == '\n') mask |= 1u << i;
    }
    return __popc(mask);
}

__global__ void __launch_bounds__(FQ_THREADS) k_fq_count(const uint8_t *d, uint64_t size, uint32_t *cnt) {
    __shared__ uint32_t sh[FQ_THREADS / 64];
    const uint64_t p = (uint64_t)blockIdx.x * FQ_CHUNK + (uint64_t)threadIdx.x * FQ_PER;
    uint32_t mask;
    const uint32_t c = p < size ? newlines_of(d, p, size, mask) : 0u;
    uint32_t tot;
    block_excl_scan<FQ_THREADS>(c, sh, tot);
    if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// Exclusive scan of n counts into off[0..n], off[n] = total (one work-group).
__global__ void __launch_bounds__(1024) k_fq_scan(const uint32_t *cnt, uint32_t n, unsigned long long *off) {
    __shared__ unsigned long long sh[16];
    unsigned long long carry = 0;
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const unsigned long long v = i < n ? cnt[i] : 0ull;
        unsigned long long tot;
        const unsigned long long ex = block_excl_scan64<1024>(v, sh, tot);
        if (i < n) off[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) off[n] = carry;
}

__global__ void __launch_bounds__(FQ_THREADS) k_fq_lines(const uint8_t *d, uint64_t size,
                                                       const unsigned long long *choff, uint64_t *nlpos) {
    __shared__ uint32_t sh[FQ_THREADS / 64];
    const uint64_t p = (uint64_t)blockIdx.x * FQ_CHUNK + (uint64_t)threadIdx.x * FQ_PER;
    uint32_t mask = 0;
    const uint32_t c = p < size ? newlines_of(d, p, size, mask) : 0u;
    uint32_t tot;
    uint64_t at = choff[blockIdx.x] + block_excl_scan<FQ_THREADS>(c, sh, tot);
    while (mask) {
        const int i = __ffs(mask) - 1;
        mask &= mask - 1;
        nlpos[at++] = p + (uint64_t)i;
    }
}

// Framing: which lines the block uses (read_block's resize, preprocess.cu:46-62,207-222).
__global__ void k_fq_frame(const uint8_t *d, uint64_t size, int final, const unsigned long long *choff,
                           uint32_t nchunks, const uint64_t *nlpos, FqResult *r) {
    if (threadIdx.x != 0) return;
    const uint64_t nl = choff[nchunks];
    r->newlines = nl;
    if (size == 0) return;
    if (d[0] != '@') {
        r->err = FQ_ERR_FORMAT;
        return;
    }
    const bool ends_nl = d[size - 1] == '\n';
    uint64_t lines;  // lines in use
    uint64_t use = size;
    if (!final) {
        const uint64_t n_lines = nl - (ends_nl ? 1 : 0) + 1;
        if (n_lines % 4 != 0 || !ends_nl) {
            const uint64_t keep = ((n_lines - 1) / 4) * 4;
            if (keep == 0) {
                r->err = FQ_ERR_FORMAT;
                return;
            }
            use = nlpos[keep - 1] + 1;
            lines = keep;
        } else {
            lines = nl;
        }
    } else {
        lines = nl + (ends_nl ? 0 : 1);
    }
    r->use = use;
    r->lines = lines;
    r->records = (lines + 2) / 4;  // sequence lines 1, 5, 9, ... below `lines`
}

// One wave per record: its sequence line's class, length and kept flag.
__global__ void __launch_bounds__(FQ_THREADS) k_fq_classify(const uint8_t *d, uint64_t size, const uint64_t *nlpos,
                                                          FqResult *r, uint32_t *kept, uint32_t *len) {
    const uint64_t rec = (uint64_t)blockIdx.x * (FQ_THREADS / 64) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (rec >= r->records) return;
    const uint64_t i = 4 * rec + 1;
    const uint64_t s = nlpos[i - 1] + 1;
    const uint64_t e = i < r->newlines ? nlpos[i] : size;
    bool drop = false, bad = false;
    for (uint64_t p = s + lane; p < e; p += 64) {
        const uint8_t c = d[p];
        const bool cls2 = c >= 64 && c < 128 && (c & 8);
        const uint8_t lc = c | 0x20;
        const bool ok = lc == 'a' || lc == 'c' || lc == 'g' || lc == 't';
        drop |= cls2;
        bad |= !ok && !cls2;
    }
    drop = __any(drop);
    bad = __any(bad);
    if (lane == 0) {
        const uint64_t n = e - s;
        const uint32_t k = (n > 0 && !drop && !bad) ? 1u : 0u;
        kept[rec] = k;
        len[rec] = (uint32_t)(n < 0xFFFFFFFFull ? n : 0xFFFFFFFFull);
        if (n > 0 && !drop && bad) atomicMin(&r->first_bad, (unsigned long long)rec);
        if (k) atomicMin(&r->first_kept, (unsigned long long)rec);
    }
}

// First kept record whose length is not L (or is over MAX_RD_LEN).
__global__ void __launch_bounds__(256) k_fq_lencheck(const uint32_t *kept, const uint32_t *len, uint64_t records,
                                                   uint32_t L, FqResult *r) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= records || !kept[i]) return;
    if (len[i] != L || len[i] > (uint32_t)MAX_RD_LEN) atomicMin(&r->first_len, (unsigned long long)i);
}

// One wave per record: kept ones copy their bases to out + (at + rank) * L.
__global__ void __launch_bounds__(FQ_THREADS) k_fq_copy(const uint8_t *d, const uint64_t *nlpos, uint64_t records,
                                                      const uint32_t *kept, const unsigned long long *rank,
                                                      uint32_t L, uint8_t *out) {
    const uint64_t rec = (uint64_t)blockIdx.x * (FQ_THREADS / 64) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (rec >= records || !kept[rec]) return;
    const uint64_t s = nlpos[4 * rec] + 1;
    uint8_t *o = out + rank[rec] * (uint64_t)L;
    for (uint32_t p = lane; p < L; p += 64) o[p] = d[s + p];
}

}  // namespace

uint64_t fq_chunks(uint64_t size) { return (size + FQ_CHUNK - 1) / FQ_CHUNK; }

hipError_t launch_fq_count(const uint8_t *d, uint64_t size, uint32_t *cnt, unsigned long long *choff,
                           hipStream_t s) {
    const uint64_t nch = fq_chunks(size);
    if (nch) {
        hipLaunchKernelGGL(k_fq_count, dim3((uint32_t)nch), dim3(FQ_THREADS), 0, s, d, size, cnt);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_fq_scan, dim3(1), dim3(1024), 0, s, cnt, (uint32_t)nch, choff);
    return hipGetLastError();
}

hipError_t launch_fq_frame(const uint8_t *d, uint64_t size, int final, const unsigned long long *choff,
                           uint64_t *nlpos, FqResult *r, hipStream_t s) {
    const uint64_t nch = fq_chunks(size);
    if (nch) {
        hipLaunchKernelGGL(k_fq_lines, dim3((uint32_t)nch), dim3(FQ_THREADS), 0, s, d, size, choff, nlpos);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_fq_frame, dim3(1), dim3(64), 0, s, d, size, final, choff, (uint32_t)nch, nlpos, r);
    return hipGetLastError();
}

hipError_t launch_fq_classify(const uint8_t *d, uint64_t size, const uint64_t *nlpos, uint64_t records, FqResult *r,
                              uint32_t *kept, uint32_t *len, unsigned long long *rank, hipStream_t s) {
    if (records == 0) return hipSuccess;
    const uint32_t per = FQ_THREADS / 64;
    hipLaunchKernelGGL(k_fq_classify, dim3((uint32_t)((records + per - 1) / per)), dim3(FQ_THREADS), 0, s, d, size,
                       nlpos, r, kept, len);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fq_scan, dim3(1), dim3(1024), 0, s, kept, (uint32_t)records, rank);
    return hipGetLastError();
}

hipError_t launch_fq_lencheck(const uint32_t *kept, const uint32_t *len, uint64_t records, uint32_t L, FqResult *r,
                              hipStream_t s) {
    if (records == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fq_lencheck, dim3((uint32_t)((records + 255) / 256)), dim3(256), 0, s, kept, len, records, L,
                       r);
    return hipGetLastError();
}

hipError_t launch_fq_copy(const uint8_t *d, const uint64_t *nlpos, uint64_t records, const uint32_t *kept,
                          const unsigned long long *rank, uint32_t L, uint8_t *out, hipStream_t s) {
    if (records == 0) return hipSuccess;
    const uint32_t per = FQ_THREADS / 64;
    hipLaunchKernelGGL(k_fq_copy, dim3((uint32_t)((records + per - 1) / per)), dim3(FQ_THREADS), 0, s, d, nlpos,
                       records, kept, rank, L, out);
    return hipGetLastError();
}

}  // namespace emp
