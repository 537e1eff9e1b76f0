// FASTQ parsing on the GPU with read_block's rules (preprocess.cu:195-237), the same rules
// emp_fastq.cpp applies on the host:
//
//   - each block must start with '@';
//   - unless it is the last block, only complete records are used: n_lines = newlines before the
//     last byte + 1; if n_lines % 4 != 0 or the block does not end in '\n', the first
//     ((n_lines - 1) / 4) * 4 lines are used (none: a record longer than the block);
//   - sequence lines (line % 4 == 1) that hold a byte in [64,127] with bit 3 set are dropped
//     (preprocess.cu:119-121,150-155), empty ones vanish, the others must be ACGTacgt;
//   - kept reads have one length, at most 255.
//
// A range of FASTQ bytes in HBM is parsed as the CLI reads it (main.cu:42-58): blocks of B bytes,
// each starting where the previous one's complete records ended. Passes over the range: newline
// counts per 4 KiB chunk -> chunk offsets (a three-launch scan) -> every newline's position ->
// the block chain (one wave: each block's framing from the newline counts) -> classify each
// record's sequence line -> kept-record offsets -> copy the kept bases (row-major, read_len each);
// classify and copy give each wave 64 records, 16 lanes per record.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emp_common.h"
#include "emp_fastq_gpu.h"

namespace emp {

namespace {

constexpr int FQ_THREADS = 256;
constexpr int FQ_PER = 16;
constexpr uint32_t FQ_CHUNK = FQ_THREADS * FQ_PER;  // bytes per chunk: a wave of the count and line passes

// Bit 7 of each byte of w set where that byte is '\n'.
__device__ __forceinline__ uint32_t nl_bytes(uint32_t w) {
    const uint32_t x = w ^ 0x0A0A0A0Au;
    return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}

// Bit i of the result: byte i of the 16 bytes w[0..3] is '\n'.
__device__ __forceinline__ uint32_t nl_mask16(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) {
    const uint32_t w[4] = {w0, w1, w2, w3};
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t b = nl_bytes(w[k]);
        m |= ((b >> 7) & 1u) << (4 * k) | ((b >> 15) & 1u) << (4 * k + 1) | ((b >> 23) & 1u) << (4 * k + 2) |
             ((b >> 31) & 1u) << (4 * k + 3);
    }
    return m;
}

// Newlines among bytes [p, p + 16) of the range at d (p a multiple of 16): mask bit i for byte
// p + i. The range may start anywhere; its offset from 16-byte alignment, the same for every lane,
// is s. Away from the range's ends two aligned 16-byte loads cover the 16 bytes (the 32 bytes
// from the aligned address below d + p lie inside the range once p >= 16 and p + 32 <= size);
// near the ends, byte loads.
__device__ __forceinline__ uint32_t newlines_of(const uint8_t *d, uint64_t p, uint64_t size, uint32_t s, uint32_t &mask) {
    if (p >= 16 && p + 32 <= size) {
        const uint4 *q = reinterpret_cast<const uint4 *>(d + p - s);
        const uint4 a = q[0];
        uint32_t m = nl_mask16(a.x, a.y, a.z, a.w);
        if (s) {
            const uint4 b = q[1];
            m = (m | (nl_mask16(b.x, b.y, b.z, b.w) << 16)) >> s;
        }
        mask = m & 0xFFFFu;
    } else {
        mask = 0;
        for (int i = 0; i < FQ_PER && p + i < size; ++i)
            if (d[p + i] == '\n') mask |= 1u << i;
    }
    return __popc(mask);
}

// The count and line passes give each wave one 4 KiB chunk, each lane 64 contiguous bytes of it
// (four 16-byte loads, five when the range is not 16-byte aligned, all in flight together): one
// wave scan orders the lanes' newlines, with no LDS and no barrier.
constexpr uint32_t FQ_LANE_BYTES = FQ_CHUNK / 64;
constexpr uint32_t FQ_CHUNKS_PER_GROUP = FQ_THREADS / 64;

// Newlines among bytes [p, p + 64) of the range (p a multiple of 64; s: the range's offset from
// 16-byte alignment), bit i for byte p + i.
__device__ __forceinline__ uint64_t newlines64(const uint8_t *d, uint64_t p, uint64_t size, uint32_t s) {
    if (p >= 16 && p + 80 <= size) {
        const uint4 *q = reinterpret_cast<const uint4 *>(d + p - s);
        const uint4 a0 = q[0], a1 = q[1], a2 = q[2], a3 = q[3];
        const uint4 a4 = s ? q[4] : make_uint4(0, 0, 0, 0);
        const uint64_t lo = (uint64_t)nl_mask16(a0.x, a0.y, a0.z, a0.w) | (uint64_t)nl_mask16(a1.x, a1.y, a1.z, a1.w) << 16 |
                            (uint64_t)nl_mask16(a2.x, a2.y, a2.z, a2.w) << 32 | (uint64_t)nl_mask16(a3.x, a3.y, a3.z, a3.w) << 48;
        if (!s) return lo;
        const uint64_t hi = nl_mask16(a4.x, a4.y, a4.z, a4.w);
        return (lo >> s) | (hi << (64 - s));
    }
    uint64_t m = 0;
    for (uint32_t i = 0; i < FQ_LANE_BYTES && p + i < size; ++i)
        if (d[p + i] == '\n') m |= 1ull << i;
    return m;
}

__device__ __forceinline__ uint32_t wave_excl_sum(uint32_t v, uint32_t &total) {
    const uint32_t lane = threadIdx.x & 63;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if ((int)lane >= o) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

__global__ void __launch_bounds__(FQ_THREADS) k_fq_count(const uint8_t *d, uint64_t size, uint32_t *cnt, uint64_t nch) {
    const uint32_t s = (uint32_t)(reinterpret_cast<uintptr_t>(d) & 15u);
    const uint64_t ch = (uint64_t)blockIdx.x * FQ_CHUNKS_PER_GROUP + (threadIdx.x >> 6);
    if (ch >= nch) return;  // the whole wave
    const uint64_t p = ch * FQ_CHUNK + (uint64_t)(threadIdx.x & 63) * FQ_LANE_BYTES;
    const uint32_t c = p < size ? (uint32_t)__popcll(newlines64(d, p, size, s)) : 0u;
    uint32_t tot;
    wave_excl_sum(c, tot);
    if ((threadIdx.x & 63) == 0) cnt[ch] = tot;
}

// Exclusive scan of n counts into off[0..n], off[n] = total, in three launches: offsets inside
// blocks of SCAN_BLOCK counts (and the block totals), an exclusive scan of the block totals in
// one work-group (which also writes off[n]), and the block offsets added.
constexpr uint32_t SCAN_BLOCK = 4096;

__global__ void __launch_bounds__(1024) k_fq_scan_local(const uint32_t *__restrict__ cnt, uint64_t n,
                                                      unsigned long long *__restrict__ off,
                                                      unsigned long long *__restrict__ blk) {
    __shared__ unsigned long long sh[16];
    const uint64_t i0 = (uint64_t)blockIdx.x * SCAN_BLOCK + 4ull * threadIdx.x;
    uint32_t v[4];
    unsigned long long sum = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        v[u] = i0 + u < n ? cnt[i0 + u] : 0u;
        sum += v[u];
    }
    unsigned long long tot;
    unsigned long long o = block_excl_scan64<1024>(sum, sh, tot);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (i0 + u < n) off[i0 + u] = o;
        o += v[u];
    }
    if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_fq_scan_top(unsigned long long *__restrict__ blk, uint32_t nblk,
                                                    unsigned long long *__restrict__ total) {
    __shared__ unsigned long long sh[16];
    unsigned long long carry = 0;
    for (uint32_t base = 0; base < nblk; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const unsigned long long v = i < nblk ? blk[i] : 0ull;
        unsigned long long tot;
        const unsigned long long ex = block_excl_scan64<1024>(v, sh, tot);
        if (i < nblk) blk[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256) k_fq_scan_add(unsigned long long *__restrict__ off, uint64_t n,
                                                   const unsigned long long *__restrict__ blk) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) off[i] += blk[i / SCAN_BLOCK];
}

}  // namespace

hipError_t launch_scan(const uint32_t *cnt, uint64_t n, unsigned long long *off, unsigned long long *blk,
                       hipStream_t s) {
    const uint64_t nblk = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    if (nblk) {
        hipLaunchKernelGGL(k_fq_scan_local, dim3((uint32_t)nblk), dim3(1024), 0, s, cnt, n, off, blk);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_fq_scan_top, dim3(1), dim3(1024), 0, s, blk, (uint32_t)nblk, off + n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nblk <= 1) return e;
    hipLaunchKernelGGL(k_fq_scan_add, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, off, n, blk);
    return hipGetLastError();
}

namespace {

__global__ void __launch_bounds__(FQ_THREADS) k_fq_lines(const uint8_t *d, uint64_t size,
                                                       const unsigned long long *choff, uint64_t *nlpos, uint64_t nch) {
    const uint32_t s = (uint32_t)(reinterpret_cast<uintptr_t>(d) & 15u);
    const uint64_t ch = (uint64_t)blockIdx.x * FQ_CHUNKS_PER_GROUP + (threadIdx.x >> 6);
    if (ch >= nch) return;  // the whole wave
    const uint64_t p = ch * FQ_CHUNK + (uint64_t)(threadIdx.x & 63) * FQ_LANE_BYTES;
    uint64_t mask = p < size ? newlines64(d, p, size, s) : 0ull;
    uint32_t tot;
    uint64_t at = choff[ch] + wave_excl_sum((uint32_t)__popcll(mask), tot);
    while (mask) {
        const int i = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        nlpos[at++] = p + (uint64_t)i;
    }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Newlines among bytes [c0, x) of the range, c0 the start of x's chunk (x - c0 < FQ_CHUNK): the
// wave's lanes take 16 bytes each in four steps, every step's loads issued before any is counted.
__device__ __forceinline__ uint32_t chunk_head_newlines(const uint8_t *d, uint64_t size, uint32_t s, uint64_t c0,
                                                        uint64_t x) {
    uint32_t mask[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const uint64_t p = c0 + (uint64_t)FQ_PER * (threadIdx.x + 64 * st);
        mask[st] = 0;
        if (p < x) newlines_of(d, p, size, s, mask[st]);
    }
    uint32_t k = 0;
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        const uint64_t p = c0 + (uint64_t)FQ_PER * (threadIdx.x + 64 * st);
        const uint64_t n = x > p ? x - p : 0;
        k += __popc(n >= 16 ? mask[st] : mask[st] & ((1u << n) - 1u));
    }
    return wave_sum(k);
}

// The block chain (read_block's framing, preprocess.cu:46-62,207-222), one wave. Block j starts at
// pos_j, a record boundary (4 * records-before-it newlines precede it), and holds min(B, rest)
// bytes; a block shorter than B is the last (when the range ends the file, `range_final`; else the
// rest waits for more data). With `one_block` the range is one block whose finality is `one_final`.
// Newlines before a position come from the chunk offsets plus the wave's count over that chunk.
// The blocks follow one another (each starts where the last one's records end), so the wave's
// time is a few dependent round trips per block: the leading '@', the chunk offset, the block's
// last byte and the partial chunk's bytes are loaded together, then the last kept newline.
__global__ void __launch_bounds__(64) k_fq_chain(const uint8_t *d, uint64_t size, uint64_t B, int range_final,
                                               int one_block, int one_final, const unsigned long long *choff,
                                               uint32_t nchunks, const uint64_t *nlpos, FqResult *r) {
    const uint32_t s = (uint32_t)(reinterpret_cast<uintptr_t>(d) & 15u);
    const uint64_t NL = choff[nchunks];
    uint64_t pos = 0, recs = 0, blocks = 0;
    uint32_t err = 0;
    while (pos < size) {
        const uint64_t got = one_block ? size : min(B, size - pos);
        bool final;
        if (one_block) {
            final = one_final != 0;
        } else {
            if (got != B && !range_final) break;  // the rest of the range waits for more data
            final = got != B;
        }
        const bool at = d[pos] == '@';  // checked once this block's other loads are out
        const uint64_t g0 = 4 * recs;  // newlines before pos
        uint64_t lines, use;
        if (final) {
            if (!at) {
                err = FQ_ERR_FORMAT;
                break;
            }
            lines = (NL - g0) + (d[size - 1] == '\n' ? 0 : 1);
            use = got;
        } else {
            const uint64_t x = pos + got, c = x / FQ_CHUNK;
            const unsigned long long cb = choff[c];
            const bool ends = d[x - 1] == '\n';
            const uint32_t head = chunk_head_newlines(d, size, s, c * FQ_CHUNK, x);
            if (!at) {
                err = FQ_ERR_FORMAT;
                break;
            }
            const uint64_t nlb = cb + head - g0;  // newlines in [pos, x)
            const uint64_t n_lines = nlb - (ends ? 1 : 0) + 1;
            if (n_lines % 4 == 0 && ends) {
                lines = nlb;
                use = got;
            } else {
                const uint64_t keep = ((n_lines - 1) / 4) * 4;
                if (keep == 0) {  // a record longer than the block
                    err = FQ_ERR_FORMAT;
                    break;
                }
                lines = keep;
                use = nlpos[g0 + keep - 1] + 1 - pos;
            }
        }
        recs += (lines + 2) / 4;  // sequence lines 1, 5, 9, ... below `lines`
        pos += use;
        ++blocks;
        if (final || one_block) break;
    }
    if (threadIdx.x == 0) {
        r->newlines = NL;
        r->use = pos;
        r->records = recs;
        r->blocks = blocks;
        r->err = err;
        r->len0 = recs ? (1 < NL ? nlpos[1] : size) - (nlpos[0] + 1) : 0;  // record 0's sequence line
    }
}

// Bit 7 of each byte of w set where that byte is not one of ACGTacgt.
__device__ __forceinline__ uint32_t fq_nz(uint32_t v) { return (((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u; }
__device__ __forceinline__ uint32_t non_bases(uint32_t w) {
    const uint32_t x = w | 0x20202020u;
    return fq_nz(x ^ 0x61616161u) & fq_nz(x ^ 0x63636363u) & fq_nz(x ^ 0x67676767u) & fq_nz(x ^ 0x74747474u);
}
// Bit 7 of each byte set where the byte is in [64, 127] with bit 3 set (drops the read): bit 7
// clear, bit 6 set (moved to bit 7 by the shift by 1), bit 3 set (by the shift by 4).
__device__ __forceinline__ uint32_t drop_bytes(uint32_t w) { return ~w & (w << 1) & (w << 4) & 0x80808080u; }

// Sequence line of record rec: [s, e).
__device__ __forceinline__ void seq_line(const uint64_t *nlpos, uint64_t NL, uint64_t size, uint64_t rec, uint64_t &s,
                                         uint64_t &e) {
    s = nlpos[4 * rec] + 1;
    e = 4 * rec + 1 < NL ? nlpos[4 * rec + 1] : size;
}

// Classify and copy take 64 records per wave: lane l first loads record l's line bounds (one round
// trip for the wave), then 16 lanes per record work on four records at a time, a batch of 16
// records whose loads are all issued before any is used.
constexpr uint32_t FQ_WAVE_RECS = 64;
constexpr int FQ_BATCH = 4;  // records per 16-lane group in one batch
constexpr int FQ_SLOTS = 4;  // aligned dwords per lane and record in classify: lines up to 253 bytes
static_assert(FQ_ROW_MAX == 64 * FQ_SLOTS - 3, "the classify pass holds lines of FQ_ROW_MAX bytes");

// Bit 7 of the bytes of the aligned dword at a that lie in [lo, hi) (a < hi, a + 4 > lo).
__device__ __forceinline__ uint32_t valid_bytes(uintptr_t a, uintptr_t lo, uintptr_t hi) {
    const uint32_t st = lo > a ? (uint32_t)(lo - a) : 0u;                 // 0..3
    const uint32_t en = hi - a < 4 ? (uint32_t)(hi - a) : 4u;             // 1..4
    return (0xFFFFFFFFu << (8 * st)) & (0xFFFFFFFFu >> (8 * (4 - en))) & 0x80808080u;
}

// The dword of the line at aligned address a, the bytes outside [d, d + size) zero (byte loads
// near the range's ends, where the aligned dword would leave it).
__device__ __forceinline__ uint32_t edge_dword(const uint8_t *d, uint64_t size, uintptr_t a) {
    const uintptr_t db = reinterpret_cast<uintptr_t>(d);
    uint32_t w = 0;
    for (int i = 0; i < 4; ++i)
        if (a + i >= db && a + i < db + size) w |= (uint32_t)d[a + i - db] << (8 * i);
    return w;
}

// Drop and bad flags of the bytes of dword w (valid: the bytes of the line).
__device__ __forceinline__ void flag_dword(uint32_t w, uint32_t valid, bool &drop, bool &bad) {
    const uint32_t dr = drop_bytes(w) & valid;
    drop |= dr != 0;
    bad |= (non_bases(w) & ~dr & valid) != 0;
}

__device__ __forceinline__ void copy_rows(const uint8_t *d, uint64_t size, uint64_t km, uint64_t src, uint64_t dst,
                                          uint32_t L, uint8_t *out, uint32_t g, uint32_t sub);

// Lane i of every 16-lane row gets the value of lane (i + 1) % 16 of its row (DPP row_ror:15), or
// of lane (i + 2) % 16 (row_ror:14). Every lane of the wave takes part.
__device__ __forceinline__ uint32_t row_next1(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x12F, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t row_next2(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x12E, 0xF, 0xF, false);
}

// Whether every aligned dword under the n bytes at d + s lies in the range (the classify loads
// then hold the whole line).
__device__ __forceinline__ bool line_in_range(uintptr_t db, uint64_t size, uintptr_t lo, uint64_t n) {
    const uintptr_t a = lo & ~(uintptr_t)3;
    return a >= db && a + ((lo - a + n + 3) & ~(uint64_t)3) <= db + size;
}

// The row of record rec (16-lane group: lane sub holds the line's aligned dwords sub + 16 j in
// w[j], the line starting at lo) written to out + rec * L from those registers: output dword m of
// the row's aligned part is the funnel shift of source dwords m + t and m + t + 1, which lanes
// fetch from their row neighbours; the few bytes before and after the aligned part are copied
// singly. row: the group writes (uniform in the group); every lane of the wave calls.
__device__ __forceinline__ void row_from_regs(const uint8_t *d, const uint32_t (&w)[FQ_SLOTS], uintptr_t lo, uint64_t rec,
                                              uint32_t L, uint8_t *out, uint32_t sub, bool row) {
    const uintptr_t dk = reinterpret_cast<uintptr_t>(out) + rec * L;
    const uint32_t h = (uint32_t)((4u - (dk & 3u)) & 3u);
    const uint32_t hb = h < L ? h : L, nd = (L - hb) >> 2, tl = L - hb - 4 * nd;
    const uint32_t o = (uint32_t)(lo & 3u), t = (o + hb) >> 2, rr = (o + hb) & 3u;
#pragma unroll
    for (int j = 0; j < FQ_SLOTS; ++j) {
        const uint32_t wn = j + 1 < FQ_SLOTS ? w[j + 1 < FQ_SLOTS ? j + 1 : j] : 0u;
        const uint32_t a1 = row_next1(w[j]), b1 = row_next1(wn), a2 = row_next2(w[j]), b2 = row_next2(wn);
        const uint32_t x1 = sub == 15 ? b1 : a1, x2 = sub >= 14 ? b2 : a2;
        const uint32_t q0 = t ? x1 : w[j], q1 = t ? x2 : x1;
        const uint32_t m = sub + 16 * j;
        if (row && m < nd) *reinterpret_cast<uint32_t *>(dk + hb + 4 * m) = __builtin_amdgcn_alignbyte(q1, q0, rr);
    }
    const uintptr_t db = reinterpret_cast<uintptr_t>(d);
    if (row && sub < hb) reinterpret_cast<uint8_t *>(dk)[sub] = d[lo - db + sub];
    if (row && sub >= 8 && sub - 8 < tl) {
        const uint32_t i = hb + 4 * nd + sub - 8;
        reinterpret_cast<uint8_t *>(dk)[i] = d[lo - db + i];
    }
}

// Per record of the range: its sequence line's length, kept flag (non-empty, no dropping byte, all
// ACGTacgt) and, for a kept-class line with another byte, the first such record. With COPY (L at
// most FQ_ROW_MAX), the records that are kept and L long also write their rows to out + record * L
// from the dwords the batch holds, and every other record is reported in first_odd: when there is
// none, the kept reads are the records and the rows are final; otherwise the rank pass and the
// copy run as without COPY.
template <bool COPY>
__global__ void __launch_bounds__(FQ_THREADS) k_fq_classify(const uint8_t *d, uint64_t size, const uint64_t *nlpos,
                                                          uint64_t NL, uint64_t records, uint32_t *kept, uint32_t *len,
                                                          FqResult *r, uint32_t L, uint8_t *out) {
    const uint32_t lane = threadIdx.x & 63, g = lane >> 4, sub = lane & 15;
    const uint64_t rec0 = (uint64_t)blockIdx.x * FQ_THREADS + (threadIdx.x & ~63u);
    if (rec0 >= records) return;  // the whole wave
    const uint64_t myrec = rec0 + lane;
    uint64_t s = 0, e = 0;
    if (myrec < records) seq_line(nlpos, NL, size, myrec, s, e);
    const uintptr_t db = reinterpret_cast<uintptr_t>(d);
    bool mydrop = false, mybad = false;
    constexpr int BATCH = COPY ? 2 : FQ_BATCH;  // the rows' registers leave room for two records per group
    for (uint32_t base = 0; base < FQ_WAVE_RECS; base += 4 * BATCH) {
        uint32_t w[BATCH][FQ_SLOTS];
        uintptr_t lo[BATCH], hi[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; ++k) {
            const uint32_t rl = base + 4 * k + g;
            lo[k] = db + (uintptr_t)__shfl((unsigned long long)s, (int)rl, 64);
            hi[k] = db + (uintptr_t)__shfl((unsigned long long)e, (int)rl, 64);
#pragma unroll
            for (int j = 0; j < FQ_SLOTS; ++j) {
                const uintptr_t a = (lo[k] & ~(uintptr_t)3) + 4 * (sub + 16 * j);
                const bool in = a < hi[k] && a >= db && a + 4 <= db + size;
                w[k][j] = in ? *reinterpret_cast<const uint32_t *>(a) : 0u;
            }
        }
        uint64_t bd[BATCH], bb[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; ++k) {
            bool drop = false, bad = false;
#pragma unroll
            for (int j = 0; j < FQ_SLOTS; ++j) {
                const uintptr_t a = (lo[k] & ~(uintptr_t)3) + 4 * (sub + 16 * j);
                if (a < hi[k]) {
                    const uint32_t x = (a >= db && a + 4 <= db + size) ? w[k][j] : edge_dword(d, size, a);
                    flag_dword(x, valid_bytes(a, lo[k], hi[k]), drop, bad);
                }
            }
            bd[k] = __ballot(drop);
            bb[k] = __ballot(bad);
        }
        // the batch's clean rows straight from the loaded dwords (lines the fast path covers)
        if (COPY) {
#pragma unroll
            for (int k = 0; k < BATCH; ++k) {
                const uint64_t rec = rec0 + base + 4 * k + g, n = hi[k] - lo[k];
                const bool flagged = ((bd[k] >> (16 * g)) & 0xFFFFull) != 0 || ((bb[k] >> (16 * g)) & 0xFFFFull) != 0;
                const bool row = rec < records && n == L && !flagged && line_in_range(db, size, lo[k], n);
                row_from_regs(d, w[k], lo[k], rec, L, out, sub, row);
            }
        }
        // lane base + 4k + g' takes its record's flags from group g' of batch k
        if (lane >= base && lane < base + 4 * BATCH) {
            const uint32_t k = (lane - base) >> 2, gg = (lane - base) & 3;
            uint64_t xd = bd[0], xb = bb[0];
#pragma unroll
            for (int q = 1; q < BATCH; ++q)
                if ((int)k == q) { xd = bd[q]; xb = bb[q]; }
            mydrop = ((xd >> (16 * gg)) & 0xFFFFull) != 0;
            mybad = ((xb >> (16 * gg)) & 0xFFFFull) != 0;
        }
    }
    // lines past the fast path's 253 bytes (no kept read is that long): the wave walks each
    // rest, 64 dwords at a time
    uint64_t lng = __ballot(myrec < records && e - s > (uint64_t)(64 * FQ_SLOTS - 3));
    while (lng) {
        const int rl = __ffsll((long long)lng) - 1;
        lng &= lng - 1;
        const uintptr_t lo = db + (uintptr_t)__shfl((unsigned long long)s, rl, 64);
        const uintptr_t hi = db + (uintptr_t)__shfl((unsigned long long)e, rl, 64);
        bool drop = false, bad = false;
        for (uintptr_t a = (lo & ~(uintptr_t)3) + 4 * (16 * FQ_SLOTS + lane); a < hi; a += 4 * 64) {
            const uint32_t x = (a >= db && a + 4 <= db + size) ? *reinterpret_cast<const uint32_t *>(a) : edge_dword(d, size, a);
            flag_dword(x, valid_bytes(a, lo, hi), drop, bad);
        }
        const bool gd = __ballot(drop) != 0, gb = __ballot(bad) != 0;
        if ((int)lane == rl) {
            mydrop |= gd;
            mybad |= gb;
        }
    }
    bool row = false;
    if (myrec < records) {
        const uint64_t n = e - s;
        const bool k = n > 0 && !mydrop && !mybad;
        kept[myrec] = k ? 1u : 0u;
        len[myrec] = (uint32_t)(n < 0xFFFFFFFFull ? n : 0xFFFFFFFFull);
        if (n > 0 && !mydrop && mybad) atomicMin(&r->first_bad, (unsigned long long)myrec);
        // a record that is not a kept read of length L, or whose row the batch did not write
        // (its line at the range's very end), sends the range to the rank pass and the copy
        row = k && n == L && line_in_range(db, size, db + s, n);
        if (COPY && !row) atomicMin(&r->first_odd, (unsigned long long)myrec);
    }
}

// The first kept record (the only kept one of rank 0) and its length; the kept total.
__global__ void __launch_bounds__(256) k_fq_first(const uint32_t *kept, const uint32_t *len,
                                                const unsigned long long *rank, uint64_t records, FqResult *r) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < records && kept[i] && rank[i] == 0) {
        r->first_kept = i;
        r->kept_len = len[i];
    }
    if (i == 0) r->kept = rank[records];
}

// First kept record whose length is not L (L == 0: the first kept record's length), or over
// MAX_RD_LEN.
__global__ void __launch_bounds__(256) k_fq_lencheck(const uint32_t *kept, const uint32_t *len, uint64_t records,
                                                   uint32_t L, FqResult *r) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= records || !kept[i]) return;
    const uint32_t want = L ? L : r->kept_len;
    if (len[i] != want || len[i] > (uint32_t)MAX_RD_LEN) atomicMin(&r->first_len, (unsigned long long)i);
}

// Kept records copy their bases to out + rank * L (row-major). A row's aligned output dwords come
// from two aligned source dwords each (a funnel shift); the bytes before and after them are copied
// singly. 64 records per wave, in batches of 16 whose loads are in flight together.
__device__ __forceinline__ uint32_t unaligned_dword(const uint8_t *d, uint64_t size, uintptr_t sa, uint32_t lo,
                                                    uint32_t hi, bool fast) {
    if (fast) return __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)(sa & 3u));
    const uintptr_t db = reinterpret_cast<uintptr_t>(d);
    uint32_t w = 0;
    for (int i = 0; i < 4; ++i) w |= (uint32_t)d[sa + i - db] << (8 * i);  // inside the line: in range
    (void)size;
    return w;
}

// The rows of the wave's records with their bit set in km (lane l: record l of the wave, its bases
// at d + src, its row at out + dst), L bytes each; g and sub: the lane's 16-lane group and place
// in it.
__device__ __forceinline__ void copy_rows(const uint8_t *d, uint64_t size, uint64_t km, uint64_t src, uint64_t dst,
                                          uint32_t L, uint8_t *out, uint32_t g, uint32_t sub) {
    const uintptr_t db = reinterpret_cast<uintptr_t>(d), ob = reinterpret_cast<uintptr_t>(out);
    for (uint32_t base = 0; base < FQ_WAVE_RECS; base += 4 * FQ_BATCH) {
        if (((km >> base) & 0xFFFFull) == 0) continue;
        bool act[FQ_BATCH];
        uintptr_t sk[FQ_BATCH], dk[FQ_BATCH];
        uint32_t hb[FQ_BATCH], nd[FQ_BATCH];
#pragma unroll
        for (int k = 0; k < FQ_BATCH; ++k) {
            const uint32_t rl = base + 4 * k + g;
            act[k] = ((km >> rl) & 1ull) != 0;
            sk[k] = db + (uintptr_t)__shfl((unsigned long long)src, (int)rl, 64);
            dk[k] = ob + (uintptr_t)__shfl((unsigned long long)dst, (int)rl, 64);
            const uint32_t h = (uint32_t)((4u - (dk[k] & 3u)) & 3u);
            hb[k] = h < L ? h : L;              // bytes before the row's first aligned dword
            nd[k] = (L - hb[k]) >> 2;           // aligned dwords
        }
        // the bytes before and after the aligned dwords: lanes 0..2 and 8..10 of the group
#pragma unroll
        for (int k = 0; k < FQ_BATCH; ++k) {
            if (!act[k]) continue;
            const uint32_t tl = L - hb[k] - 4 * nd[k], t0 = hb[k] + 4 * nd[k];
            if (sub < hb[k]) reinterpret_cast<uint8_t *>(dk[k])[sub] = reinterpret_cast<const uint8_t *>(sk[k])[sub];
            if (sub >= 8 && sub - 8 < tl)
                reinterpret_cast<uint8_t *>(dk[k])[t0 + sub - 8] = reinterpret_cast<const uint8_t *>(sk[k])[t0 + sub - 8];
        }
        for (uint32_t j0 = 0; 4 * j0 < L; j0 += 16) {
            uint32_t lo[FQ_BATCH], hi[FQ_BATCH];
            bool fast[FQ_BATCH];
#pragma unroll
            for (int k = 0; k < FQ_BATCH; ++k) {
                const uint32_t kd = j0 + sub;
                const uintptr_t sa = sk[k] + hb[k] + 4 * kd, a0 = sa & ~(uintptr_t)3;
                fast[k] = a0 >= db && a0 + 8 <= db + size;
                const bool ld = act[k] && kd < nd[k] && fast[k];
                lo[k] = ld ? *reinterpret_cast<const uint32_t *>(a0) : 0u;
                hi[k] = ld ? *reinterpret_cast<const uint32_t *>(a0 + 4) : 0u;
            }
#pragma unroll
            for (int k = 0; k < FQ_BATCH; ++k) {
                const uint32_t kd = j0 + sub;
                if (act[k] && kd < nd[k]) {
                    const uintptr_t sa = sk[k] + hb[k] + 4 * kd;
                    *reinterpret_cast<uint32_t *>(dk[k] + hb[k] + 4 * kd) = unaligned_dword(d, size, sa, lo[k], hi[k], fast[k]);
                }
            }
        }
    }
}

__global__ void __launch_bounds__(FQ_THREADS) k_fq_copy(const uint8_t *d, uint64_t size, const uint64_t *nlpos,
                                                      uint64_t records, const uint32_t *kept,
                                                      const unsigned long long *rank, uint32_t L, uint8_t *out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t rec0 = (uint64_t)blockIdx.x * FQ_THREADS + (threadIdx.x & ~63u);
    if (rec0 >= records) return;  // the whole wave
    const uint64_t myrec = rec0 + lane;
    const bool mine = myrec < records && kept[myrec] != 0;
    uint64_t src = 0, dst = 0;
    if (mine) {
        src = nlpos[4 * myrec] + 1;
        dst = rank[myrec] * (uint64_t)L;
    }
    copy_rows(d, size, __ballot(mine), src, dst, L, out, lane >> 4, lane & 15);
}

}  // namespace

uint64_t fq_chunks(uint64_t size) { return (size + FQ_CHUNK - 1) / FQ_CHUNK; }

uint64_t fq_scan_blocks(uint64_t n) { return (n + SCAN_BLOCK - 1) / SCAN_BLOCK + 1; }

hipError_t launch_fq_count(const uint8_t *d, uint64_t size, uint32_t *cnt, unsigned long long *choff,
                           unsigned long long *blk, hipStream_t s) {
    const uint64_t nch = fq_chunks(size);
    if (nch) {
        const uint64_t groups = (nch + FQ_CHUNKS_PER_GROUP - 1) / FQ_CHUNKS_PER_GROUP;
        hipLaunchKernelGGL(k_fq_count, dim3((uint32_t)groups), dim3(FQ_THREADS), 0, s, d, size, cnt, nch);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return launch_scan(cnt, nch, choff, blk, s);
}

hipError_t launch_fq_frame(const uint8_t *d, uint64_t size, uint64_t B, int range_final, int one_block, int one_final,
                           const unsigned long long *choff, uint64_t *nlpos, FqResult *r, hipStream_t s) {
    const uint64_t nch = fq_chunks(size);
    if (nch) {
        const uint64_t groups = (nch + FQ_CHUNKS_PER_GROUP - 1) / FQ_CHUNKS_PER_GROUP;
        hipLaunchKernelGGL(k_fq_lines, dim3((uint32_t)groups), dim3(FQ_THREADS), 0, s, d, size, choff, nlpos, nch);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_fq_chain, dim3(1), dim3(64), 0, s, d, size, B, range_final, one_block, one_final, choff,
                       (uint32_t)nch, nlpos, r);
    return hipGetLastError();
}

hipError_t launch_fq_classify(const uint8_t *d, uint64_t size, const uint64_t *nlpos, uint64_t NL, uint64_t records,
                              FqResult *r, uint32_t *kept, uint32_t *len, unsigned long long *rank,
                              unsigned long long *blk, uint32_t L, hipStream_t s) {
    if (records == 0) return hipSuccess;
    const uint64_t groups = (records + FQ_THREADS - 1) / FQ_THREADS;  // a wave per 64 records
    hipLaunchKernelGGL(k_fq_classify<false>, dim3((uint32_t)groups), dim3(FQ_THREADS), 0, s, d, size, nlpos, NL,
                       records, kept, len, r, 0u, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_fq_rank(records, r, kept, len, rank, blk, L, s);
}

hipError_t launch_fq_classify_copy(const uint8_t *d, uint64_t size, const uint64_t *nlpos, uint64_t NL,
                                   uint64_t records, FqResult *r, uint32_t *kept, uint32_t *len, uint32_t L,
                                   uint8_t *out, hipStream_t s) {
    if (records == 0) return hipSuccess;
    const uint64_t groups = (records + FQ_THREADS - 1) / FQ_THREADS;
    hipLaunchKernelGGL(k_fq_classify<true>, dim3((uint32_t)groups), dim3(FQ_THREADS), 0, s, d, size, nlpos, NL,
                       records, kept, len, r, L, out);
    return hipGetLastError();
}

hipError_t launch_fq_rank(uint64_t records, FqResult *r, const uint32_t *kept, const uint32_t *len,
                          unsigned long long *rank, unsigned long long *blk, uint32_t L, hipStream_t s) {
    if (records == 0) return hipSuccess;
    hipError_t e;
    if ((e = launch_scan(kept, records, rank, blk, s)) != hipSuccess) return e;
    const uint32_t g = (uint32_t)((records + 255) / 256);
    hipLaunchKernelGGL(k_fq_first, dim3(g), dim3(256), 0, s, kept, len, rank, records, r);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_fq_lencheck, dim3(g), dim3(256), 0, s, kept, len, records, L, r);
    return hipGetLastError();
}

hipError_t launch_fq_copy(const uint8_t *d, uint64_t size, const uint64_t *nlpos, uint64_t NL, uint64_t records,
                          const uint32_t *kept, const unsigned long long *rank, uint32_t L, uint8_t *out,
                          hipStream_t s) {
    if (records == 0) return hipSuccess;
    (void)NL;
    const uint64_t groups = (records + FQ_THREADS - 1) / FQ_THREADS;  // a wave per 64 records
    hipLaunchKernelGGL(k_fq_copy, dim3((uint32_t)groups), dim3(FQ_THREADS), 0, s, d, size, nlpos, records, kept, rank,
                       L, out);
    return hipGetLastError();
}

}  // namespace emp
