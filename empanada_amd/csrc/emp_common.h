// Device helpers shared by the single-job kernels (emp_kernels.hip) and the batched kernels
// (emp_batch.hip): regeneration of an entry for a round, block scans, sorts and searches.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emp_kernels.h"

namespace emp {

// Barrier for threads that share data only through LDS: each wave waits for its own LDS
// operations, not for its global loads, stores or atomics in flight (a result is waited for where
// it is used). Kernels whose threads exchange data through global memory use __syncthreads().
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Regeneration of a live entry for round t (keys of length L - t -> L - t - 1). Suffix list:
// even ids drop column t, odd ids the complement of column L-1-t; prefix list: the other way
// round (overlap.cu:266-267, 684-690: rid/2 picks the read, odd ids are complemented). An entry
// walks its columns forward (t = 0, 1, ...) or backward (L-1, L-2, ...); it gathers its read's
// word of a column group when it enters the group (or at round 0) and carries the codes of the
// group's remaining columns, in the order it uses them, in the spare key bits.
// Codes an entry carries after its regeneration at column `col` of group word `gw` (8 bits):
// forward, columns c+1 .. 3; backward, columns c-1 .. 0 (c = col % 4), first-used in bits 0-1.
__device__ __forceinline__ uint32_t carry_of(uint32_t gw, uint32_t col, bool fwd) {
    const uint32_t c = col & 3u;
    if (fwd) return gw >> (2 * (c + 1));
    uint32_t f = 0;
    for (uint32_t i = 0; i < c; ++i) f |= ((gw >> (2 * (c - 1 - i))) & 3u) << (2 * i);
    return f;
}

// Regenerates e for round r (suffix rule: NextSfx, else PrevPfx) from the gathered group word
// (0xFFFFFFFF: none, the carried codes supply the base) and returns the base's encoding (1..4).
// Column that round r consumes for an entry read forward (column t) or backward (column L-1-t).
// r is the same for every lane (kernel arguments, or one job's per work-group); taking both
// fields into scalar registers first makes the choice a select of values. A select of the two
// fields' addresses is what the compiler makes otherwise: a load from the argument per lane,
// waited for inside every regeneration.
__device__ __forceinline__ uint32_t round_col(const RoundArgs &r, bool fwd) {
    const uint32_t ct = __builtin_amdgcn_readfirstlane(r.col_t), cb = __builtin_amdgcn_readfirstlane(r.col_back);
    return fwd ? ct : cb;
}

__device__ __forceinline__ uint32_t regen_carry(const uint32_t *__restrict__ cols, const RoundArgs &r, Entry &e, bool sfx,
                                                uint32_t gathered) {
    const bool odd = e.rid & 1;
    const bool fwd = sfx != odd;
    const uint32_t col = round_col(r, fwd);
    uint32_t f = (e.h1 >> 29) | ((e.h2 >> 29) << 3), code;
    if (gathered != 0xFFFFFFFFu) {
        code = (gathered >> (2 * (col & 3u))) & 3u;
        f = carry_of(gathered, col, fwd);
    } else {
        code = f & 3u;
        f >>= 2;
    }
    uint32_t h1 = e.h1 & KEY_MASK, h2 = e.h2 & KEY_MASK;
    const uint32_t ev = enc_of_code(odd ? (code ^ 2u) : code);
    if (sfx)
        next_sfx(h1, h2, ev, r.m1, r.m2);
    else
        prev_pfx(h1, h2, ev);
    e.h1 = h1 | ((f & 7u) << 29);
    e.h2 = h2 | (((f >> 3) & 7u) << 29);
    return ev;
}

__device__ __forceinline__ void set_carry(Entry &e, uint32_t f) {
    e.h1 = (e.h1 & KEY_MASK) | ((f & 7u) << 29);
    e.h2 = (e.h2 & KEY_MASK) | (((f >> 3) & 7u) << 29);
}

// Wide carry (the single-job engine's round buckets, 14 bits: see WIDE_SHIFT). An entry gathers its
// read's 16-bit unit of 8 columns when it enters a group: every eighth round per direction.
// Forward: the unit's columns c+1 .. 7; backward: columns c-1 .. 0 in that order, which is the
// unit with its eight 2-bit fields reversed (a bit reversal, then the two bits of each field
// swapped back), shifted down to its last c fields. Both are computed and one is selected: the
// direction differs between lanes, and a branch per lane costs more than the other form.
__device__ __forceinline__ uint32_t wide_carry_of(uint32_t u, uint32_t col, bool fwd) {
    const uint32_t c = col & 7u, v = u & 0xFFFFu;
    const uint32_t fw = v >> (2 * (c + 1));
    uint32_t y = __builtin_bitreverse32(v) >> 16;
    y = ((y >> 1) & 0x5555u) | ((y & 0x5555u) << 1);
    const uint32_t bw = y >> (16 - 2 * c);
    return fwd ? fw : bw;
}

// Whether entry rid (side sfx) enters a column group this round (forward entries at a group's
// first column, backward ones at its last, every entry at round 0), and the group (g). The
// round's facts are uniform; per lane only the direction is picked, as integer selects.
__device__ __forceinline__ bool wide_needs_gather(const RoundArgs &r, uint32_t rid, bool sfx, uint32_t &g) {
    const uint32_t ct = __builtin_amdgcn_readfirstlane(r.col_t), cb = __builtin_amdgcn_readfirstlane(r.col_back);
    const uint32_t nf = (ct & 7u) == 0u ? 1u : 0u, nb = ((cb & 7u) == 7u || ct == 0u) ? 1u : 0u;
    const uint32_t fwd = (sfx ? 1u : 0u) ^ (rid & 1u);
    g = (fwd ? ct : cb) >> 3;
    return ((fwd & nf) | ((fwd ^ 1u) & nb)) != 0u;
}

// The 14-bit carry of a stored entry, and its h1 key with bits 21..28 restored from top8 (the top
// 8 bits of its bucket's index).
__device__ __forceinline__ uint32_t wide_carry(const Entry &e) {
    return (e.h1 >> 29) | ((e.h2 >> 29) << 3) | (((e.h1 >> WIDE_SHIFT) & 0xFFu) << 6);
}
__device__ __forceinline__ uint32_t wide_key1(uint32_t h1, uint32_t top8) { return (h1 & LOW_MASK) | (top8 << WIDE_SHIFT); }

// Regenerates keys h1, h2 (this round's, bits 0..28) of oriented read rid for round r and its carry
// f (in: this round's, out: the next round's), from the gathered 16-bit unit (0xFFFFFFFF: none, the
// carry supplies the base). Returns the base's encoding (1..4).
__device__ __forceinline__ uint32_t regen_wide(const RoundArgs &r, uint32_t &h1, uint32_t &h2, uint32_t rid, uint32_t &f,
                                               bool sfx, uint32_t gathered) {
    const bool odd = rid & 1;
    const bool fwd = sfx != odd;
    const uint32_t col = round_col(r, fwd);
    // the gathered unit's code and carry, or the carried ones, computed both and selected
    const bool has = gathered != 0xFFFFFFFFu;
    const uint32_t gcode = (gathered >> (2 * (col & 7u))) & 3u, gcarry = wide_carry_of(gathered, col, fwd);
    const uint32_t code = has ? gcode : (f & 3u);
    f = has ? gcarry : (f >> 2);
    const uint32_t ev = enc_of_code(code ^ (odd ? 2u : 0u));
    if (sfx)
        next_sfx(h1, h2, ev, r.m1, r.m2);
    else
        prev_pfx(h1, h2, ev);
    return ev;
}

// Stored form of an entry with keys h1, h2 (bits 0..28) and carry f.
__device__ __forceinline__ Entry wide_entry(uint32_t h1, uint32_t h2, uint32_t rid, uint32_t f) {
    return Entry{(h1 & LOW_MASK) | (((f >> 6) & 0xFFu) << WIDE_SHIFT) | ((f & 7u) << 29), h2 | (((f >> 3) & 7u) << 29), rid};
}

// Whether the entry must gather its group word this round, and which word (group index).
__device__ __forceinline__ bool regen_needs_gather(const RoundArgs &r, uint32_t rid, bool sfx, uint32_t &g) {
    const bool fwd = sfx != (bool)(rid & 1);
    const uint32_t col = round_col(r, fwd);
    g = col >> 2;
    return fwd ? (col & 3u) == 0 : ((col & 3u) == 3 || r.col_t == 0);
}

__device__ __forceinline__ Entry load_stream(const Entry *p) {
    const uint32_t *q = reinterpret_cast<const uint32_t *>(p);
    return Entry{__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1), __builtin_nontemporal_load(q + 2)};
}

// ---------------------------------------------------------------------------------------------
// Block helpers.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sh /* >= NT/64 */, uint32_t &total) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    lds_barrier();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        uint32_t s = sh[i];
        base += i < wv ? s : 0u;
        tot += s;
    }
    total = tot;
    lds_barrier();
    return base + x - v;
}

template <int NT>
__device__ __forceinline__ unsigned long long block_excl_scan64(unsigned long long v, unsigned long long *sh,
                                                               unsigned long long &total) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        unsigned long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    lds_barrier();
    unsigned long long base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        unsigned long long s = sh[i];
        base += i < wv ? s : 0ull;
        tot += s;
    }
    total = tot;
    lds_barrier();
    return base + x - v;
}

// Bitonic sort of n2 (power of two) records by (key, rank), ascending; `pay` (may be null) is
// permuted along. Works on LDS or on global memory private to the work-group.
__device__ inline void bitonic(uint64_t *key, uint32_t *rank, uint32_t *pay, uint32_t n2) {
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
                uint32_t ixj = i ^ j;
                if (ixj > i) {
                    uint64_t ka = key[i], kb = key[ixj];
                    uint32_t ra = rank[i], rb = rank[ixj];
                    bool a_gt = (ka > kb) || (ka == kb && ra > rb);
                    bool up = (i & k) == 0;
                    if (a_gt == up) {
                        key[i] = kb; key[ixj] = ka;
                        rank[i] = rb; rank[ixj] = ra;
                        if (pay) { uint32_t t = pay[i]; pay[i] = pay[ixj]; pay[ixj] = t; }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Edge-pool slot of n edges whose sub-range reservation returned c (atomicAdd on the cursor
// of sub-range `sub`): inside the sub-range when it fits, else the edges go to the overflow
// region; ~0 when that is full too (the host grows it and reruns). A reservation past the end is
// abandoned, not returned: the cursor only grows (moving it back could let a later reservation
// overlap one handed out in between); ovf_cursor[1] counts the abandoned edges, which the host
// subtracts when it counts edges from the cursors.
__device__ __forceinline__ unsigned long long place_edges(unsigned long long c, uint32_t n, uint32_t sub,
                                                          unsigned long long *cursor, unsigned long long cap,
                                                          unsigned long long *ovf_cursor, unsigned long long ovf_cap) {
    (void)cursor;
    if (c + n <= cap) return (unsigned long long)sub * cap + c;
    atomicAdd(ovf_cursor + 1, (unsigned long long)n);
    const unsigned long long o = atomicAdd(ovf_cursor, (unsigned long long)n);
    return o + n <= ovf_cap ? (unsigned long long)SUBS * cap + o : ~0ull;
}

__device__ __forceinline__ uint64_t key_of(const Entry &e) {
    return ((uint64_t)(e.h1 & KEY_MASK) << 32) | (e.h2 & KEY_MASK);
}

// Bucket sets: a set of nb buckets puts key h (< 2^29) into bucket floor(h nb / 2^29), the product
// taken as the high word of (8 h) nb; for nb = 2^k that is h >> (29 - k). Buckets are in key order,
// and when nb is a multiple of 256 every key of a bucket has the same top 8 bits (h >> 21 =
// floor(b / (nb / 256)) by floor(floor(x) / m) = floor(x / m)), the bits a round's entries carry
// codes in.
__device__ __forceinline__ uint32_t bucket_of(uint32_t h, uint32_t nb) { return __umulhi(h << 3, nb); }

// Smallest key of bucket b: ceil(b 2^29 / nb).
__device__ __forceinline__ uint32_t bucket_lo(uint32_t b, uint32_t nb) {
    return (uint32_t)((((uint64_t)b << 29) + nb - 1u) / nb);
}

// Next-round target slots (NextOut). The entries of one bucket of width w = 2^29 / src_nb land in
// a few buckets of the next round (width 2^29 / nb >= w), found from the entry's removed base e
// (encoding 1..4) and, for prefixes, k = (h1 - e) mod 5:
//   slots 0..11   suffixes: 3 (e - 1) + d, d = target - first target of base e (mod the bucket
//                 count; d <= 2, the third only when the shifted range wraps past P1);
//   slots 12..21  prefixes: 12 + 2 k + d, d <= 1;
//   slots 30, 31  suffixes / prefixes outside these (the reference's u64 wrap for h1 < e): one
//                 global atomic each.
// Work-groups rank their entries per slot with one LDS atomic per entry on a per-wave counter
// row, and reserve one run per slot.
constexpr int NSLOT = 32;
constexpr uint32_t SLOT_SFX_OTHER = 30, SLOT_PFX_OTHER = 31;

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// LDS of the join of one bucket: the prefix entries (12 B each), a used flag per prefix (then,
// for prefix survivors, the column-group byte they gathered), and one region that holds the hash
// table (chain heads + links) while probing and the matched records afterwards.
template <int CAP, int MCAP, int NT>
struct JoinShared {
    uint32_t pk1[CAP], pk2[CAP], prid[CAP];
    uint8_t used[CAP];
    union {
        struct {
            uint32_t head[CAP];
            uint16_t nxt[CAP];
        } h;
        struct {
            uint64_t msk[MCAP], mpk[MCAP];
            uint32_t msr[MCAP], mpr[MCAP], msc[MCAP];
            uint32_t wcnt[NT / 64][NSLOT];  // per wave and slot: entries, then their run start
            uint32_t hist[2][64];           // binned record sort: per side and key bin, count then start
        } m;
    } u;
    uint32_t ntot[2];  // survivors per side
    uint32_t sh[NT / 64];
    uint32_t ms, mp, ovf, etot;
    uint32_t rcnt[2];  // batched join: survivors per side
    unsigned long long base;
};

// Position of record i among n records (key, rank) ordered by (key, rank): the number of records
// smaller than it. (key, rank) pairs are distinct, so the positions of 0..n-1 are a permutation.
// Eight independent LDS reads per step (the same address for the whole wave: broadcasts), two
// 64-bit compares per record; ranks are compared only for a key that occurs more than once.
__device__ __forceinline__ uint32_t rank_of(const uint64_t *key, const uint32_t *rank, uint32_t n, uint32_t i,
                                            uint64_t &k, uint32_t &r) {
    k = key[i];
    r = rank[i];
    // records below in key order, and records with the key at most k (the record itself
    // included): two compares per record, no rank compare unless the key occurs twice
    uint32_t lt = 0, le = 0, j = 0;
    for (; j + 8 <= n; j += 8) {
        uint64_t kk[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) kk[u] = key[j + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            lt += kk[u] < k ? 1u : 0u;
            le += kk[u] <= k ? 1u : 0u;
        }
    }
    for (; j < n; ++j) {
        const uint64_t kk = key[j];
        lt += kk < k ? 1u : 0u;
        le += kk <= k ? 1u : 0u;
    }
    if (le - lt > 1u) {  // equal keys: the smaller ranks among them come first
        for (j = 0; j < n; ++j) lt += (key[j] == k && rank[j] < r) ? 1u : 0u;
    }
    return lt;
}

// Binary searches over sorted global-memory arrays.
__device__ __forceinline__ uint32_t lower_u64(const uint64_t *k, uint32_t lo, uint32_t hi, uint64_t x) {
    while (lo < hi) {
        uint32_t m = (lo + hi) >> 1;
        if (k[m] < x) lo = m + 1; else hi = m;
    }
    return lo;
}

__device__ __forceinline__ uint32_t upper_u64(const uint64_t *k, uint32_t lo, uint32_t hi, uint64_t x) {
    while (lo < hi) {
        uint32_t m = (lo + hi) >> 1;
        if (x < k[m]) hi = m; else lo = m + 1;
    }
    return lo;
}

__device__ __forceinline__ bool has_u32(const uint32_t *r, uint32_t lo, uint32_t hi, uint32_t x) {
    while (lo < hi) {
        uint32_t m = (lo + hi) >> 1;
        if (r[m] < x) lo = m + 1; else hi = m;
    }
    return lo < hi && r[lo] == x;
}

// Collects the entries of bucket b (slots + spilled records) into key/value scratch, padded with
// sentinels to n2. as_rank: store rid_rank(rid) (join) or rid (dedup).
// top8 >= 0: the entries are wide (round buckets of the single-job engine) and their keys' bits
// 21..28 are top8, the bucket's.
__device__ inline void gather_bucket(const Entry *slots, uint32_t cnt, uint32_t cap, const SpillRec *spill,
                              const uint32_t *spill_n, uint32_t spill_cap, uint32_t b, uint64_t *key, uint32_t *val,
                              uint32_t n2, bool as_rank, uint32_t *s_pos, int32_t top8 = -1) {
    const uint32_t nm = min(cnt, cap);
    auto kof = [&](const Entry &e) {
        return top8 >= 0 ? ((uint64_t)wide_key1(e.h1, (uint32_t)top8) << 32) | (e.h2 & KEY_MASK) : key_of(e);
    };
    for (uint32_t i = threadIdx.x; i < nm; i += blockDim.x) {
        Entry e = slots[(size_t)b * cap + i];
        key[i] = kof(e);
        val[i] = as_rank ? rid_rank(e.rid) : e.rid;
    }
    if (threadIdx.x == 0) *s_pos = nm;
    __syncthreads();
    if (cnt > cap) {
        const uint32_t sn = min(*spill_n, spill_cap);
        for (uint32_t j = threadIdx.x; j < sn; j += blockDim.x) {
            if (spill[j].bucket == b) {
                uint32_t k = atomicAdd(s_pos, 1u);
                if (k < cnt) {
                    Entry e = spill[j].e;
                    key[k] = kof(e);
                    val[k] = as_rank ? rid_rank(e.rid) : e.rid;
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t i = cnt + threadIdx.x; i < n2; i += blockDim.x) { key[i] = ~0ull; val[i] = ~0u; }
    __syncthreads();
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t n) {
    uint32_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

}  // namespace emp
