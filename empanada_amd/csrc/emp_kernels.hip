// HIP kernels of the fingerprint -> partition -> suffix/prefix join path, for gfx950 (wave64).
//
// Data in HBM:
//   reads    row-major ASCII kept reads (input of the map kernel only)
//   columns  2-bit codes per read position, chunked by 2^20 reads (emp_device.h)
//   entries  12-byte (h1, h2, rid) records. A list is linear after the kernel that produces it
//            (map, dedup, join), then one or more partition passes group it by the top bits of
//            h1 into buckets: bucket b of a pass with `bits` output bits holds the entries with
//            h1 >> (29 - bits) == b, in a fixed-capacity slot array [bucket][cap] plus a count.
//            Order inside a bucket is arbitrary: every order the output depends on is
//            re-established in LDS.
//
// Kernels:
//   k_map      transposition + forward/reverse-complement Rabin-Karp (fingerprint.cu:75-153)
//   k_part     tile-sorted partition pass (at most 8 more key bits per pass)
//   k_dedup    per bucket: keep the smallest read id of every key (overlap.cu:212-222), then
//              regenerate for round 0 as suffix and prefix entries (linear output)
//   k_join     per bucket of round t: LDS hash join of suffixes against prefixes, edges of the
//              matched key groups in (key, rid&1, rid) x (rid&1, rid) order, removal of used
//              entries, regeneration of the survivors for round t+1 (linear output)
//              (overlap.cu:258-290, 361-456, 557-710)
//   k_*_big    the same for buckets over capacity: global-memory sort + binary-search join
//   k_edge_*   bucket-ordered compaction of the edges of a round
#include "emp_kernels.h"

#include <hip/hip_runtime.h>

namespace emp {

// ---------------------------------------------------------------------------------------------
// Column gather: 2-bit code of read k at column `col`.
__device__ __forceinline__ uint32_t col_code(const uint32_t *const *__restrict__ chunks, uint32_t col, uint32_t k) {
    const uint32_t *ch = chunks[k >> CHUNK_SHIFT];
    uint32_t w = ch[(size_t)col * CHUNK_WORDS + ((k >> 4) & (CHUNK_WORDS - 1))];
    return (w >> (2 * (k & 15))) & 3u;
}

// Base code of a live entry for the round described by `r`. Suffix list: even ids read column
// t, odd ids the complement of column L-1-t; prefix list: the other way round (overlap.cu:266-267,
// 684-690: rid/2 picks the read, odd ids are complemented).
__device__ __forceinline__ uint32_t entry_code(const uint32_t *const *__restrict__ chunks, const RoundArgs &r, uint32_t rid,
                                               bool sfx) {
    uint32_t k = rid >> 1;
    bool odd = rid & 1;
    uint32_t col = (sfx != odd) ? r.col_t : r.col_back;
    uint32_t c = col_code(chunks, col, k);
    return odd ? (c ^ 2u) : c;
}

__device__ __forceinline__ void regen(const uint32_t *const *__restrict__ chunks, const RoundArgs &r, Entry &e, bool sfx) {
    uint32_t code = entry_code(chunks, r, e.rid, sfx);
    uint32_t ev = enc_of_code(code);
    if (sfx)
        next_sfx(e.h1, e.h2, ev, r.m1, r.m2);
    else
        prev_pfx(e.h1, e.h2, ev);
}

// ---------------------------------------------------------------------------------------------
// Map: one work-group of 256 threads per 256 reads.
__global__ void __launch_bounds__(256) k_map(MapArgs a) {
    extern __shared__ __align__(16) uint8_t tile[];
    const uint32_t first = blockIdx.x * a.rows;
    const uint32_t nt = min(a.rows, a.n - first);
    const int L = a.L, Lp = a.Lpad;
    const uint8_t *src = a.reads + (size_t)first * L;
    const uint32_t nbytes = nt * (uint32_t)L;
    const uint32_t ndw = nbytes >> 2;
    uint32_t bad = 0;
    for (uint32_t d = threadIdx.x; d < ndw; d += 256) {
        uint32_t w = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(src) + d);
        uint32_t g = d * 4;
        uint32_t r = g / (uint32_t)L, p = g - r * (uint32_t)L;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint8_t c = (uint8_t)(w >> (8 * i));
            uint8_t lc = c | 0x20;
            bad |= !(lc == 'a' || lc == 'c' || lc == 'g' || lc == 't');
            tile[r * Lp + p] = c;
            if (++p == (uint32_t)L) { p = 0; ++r; }
        }
    }
    for (uint32_t g = ndw * 4 + threadIdx.x; g < nbytes; g += 256) {
        uint8_t c = src[g];
        uint8_t lc = c | 0x20;
        bad |= !(lc == 'a' || lc == 'c' || lc == 'g' || lc == 't');
        uint32_t r = g / (uint32_t)L;
        tile[r * Lp + (g - r * (uint32_t)L)] = c;
    }
    if (bad) atomicOr(a.err, 1u);
    __syncthreads();

    const uint32_t t = threadIdx.x;
    if (t < nt) {
        const uint8_t *row = tile + t * Lp;
        uint32_t f1 = 0, f2 = 0, r1 = 0, r2 = 0;
        for (int p = 0; p < L; ++p) {
            uint32_t code = (row[p] >> 1) & 3u;
            uint32_t e = enc_of_code(code);
            f1 = step1(f1, e);
            f2 = step2(f2, e);
            uint32_t ce = enc_of_code(((row[L - 1 - p] >> 1) & 3u) ^ 2u);
            r1 = step1(r1, ce);
            r2 = step2(r2, ce);
        }
        uint64_t k = a.k0 + first + t;
        Entry ef = {f1, f2, (uint32_t)(2 * k)};
        Entry er = {r1, r2, (uint32_t)(2 * k + 1)};
        a.entries[2 * k] = ef;
        a.entries[2 * k + 1] = er;
    }

    // Columns: word w of the tile covers reads 16w..16w+15; task id = w * L + j keeps a wave on
    // consecutive columns of the same 16 rows (conflict-free LDS byte reads).
    const uint32_t nw = (nt + 15) / 16;
    const uint32_t ntask = nw * (uint32_t)L;
    const uint64_t w0 = (a.k0 + first) >> 4;
    for (uint32_t task = threadIdx.x; task < ntask; task += 256) {
        uint32_t w = task / (uint32_t)L, j = task - w * (uint32_t)L;
        uint32_t word = 0;
        uint32_t rmax = min(16u, nt - 16 * w);
        for (uint32_t i = 0; i < rmax; ++i) word |= ((uint32_t)(tile[(16 * w + i) * Lp + j] >> 1) & 3u) << (2 * i);
        uint64_t W = w0 + w;
        uint32_t *ch = a.chunks[W >> (CHUNK_SHIFT - 4)];
        ch[(size_t)j * CHUNK_WORDS + (W & (CHUNK_WORDS - 1))] = word;
    }
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
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        uint32_t s = sh[i];
        base += i < wv ? s : 0u;
        tot += s;
    }
    total = tot;
    __syncthreads();
    return base + x - v;
}

// Bitonic sort of n2 (power of two) records by (key, rank), ascending; `pay` (may be null) is
// permuted along. Works on LDS or on global memory private to the work-group.
__device__ void bitonic(uint64_t *key, uint32_t *rank, uint32_t *pay, uint32_t n2) {
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

__device__ __forceinline__ uint64_t key_of(const Entry &e) { return ((uint64_t)e.h1 << 32) | e.h2; }

// ---------------------------------------------------------------------------------------------
// Partition pass: each work-group takes tiles of PART_TILE entries of one input group, ranks them
// by their next `dbits` key bits in LDS, reserves one output range per (tile, bucket) with a
// single atomic, and writes each bucket's run contiguously.
__global__ void __launch_bounds__(256) k_part(PartArgs a) {
    constexpr int T = PART_TILE, PER = T / 256;
    __shared__ Entry buf[T];
    __shared__ uint8_t dig[T];
    __shared__ uint32_t hist[256], loc[256], base[256], sh[4];
    const uint32_t g = blockIdx.x / a.tiles_x;
    const uint32_t F = 1u << a.dbits;
    uint32_t n_g;
    const Entry *in;
    if (a.in_cnt) {
        n_g = min(a.in_cnt[g], a.in_cap);
        in = a.in + (size_t)g * a.in_cap;
    } else {
        n_g = *a.in_n;
        in = a.in;
    }
    for (uint32_t tile = blockIdx.x % a.tiles_x; (size_t)tile * T < n_g; tile += a.tiles_x) {
        const uint32_t t0 = tile * T, nt = min((uint32_t)T, n_g - t0);
        hist[threadIdx.x] = 0;
        __syncthreads();
        Entry e[PER];
        uint32_t r[PER], d[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            uint32_t i = threadIdx.x + 256 * q;
            if (i < nt) {
                e[q] = in[t0 + i];
                d[q] = (e[q].h1 >> a.out_shift) & (F - 1);
                r[q] = atomicAdd(&hist[d[q]], 1u);
            }
        }
        __syncthreads();
        const uint32_t h = hist[threadIdx.x];
        uint32_t tot;
        const uint32_t ex = block_excl_scan<256>(h, sh, tot);
        loc[threadIdx.x] = ex;
        const uint32_t gb = a.keyed ? (g << a.dbits) : 0u;
        base[threadIdx.x] = h ? atomicAdd(&a.out_cnt[gb + threadIdx.x], h) : 0u;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            uint32_t i = threadIdx.x + 256 * q;
            if (i < nt) {
                uint32_t p = loc[d[q]] + r[q];
                buf[p] = e[q];
                dig[p] = (uint8_t)d[q];
            }
        }
        __syncthreads();
        for (uint32_t p = threadIdx.x; p < nt; p += 256) {
            const uint32_t dd = dig[p];
            const uint32_t slot = base[dd] + (p - loc[dd]);
            const uint32_t bucket = gb + dd;
            if (slot < a.out_cap) {
                a.out[(size_t)bucket * a.out_cap + slot] = buf[p];
            } else if (a.spill) {
                uint32_t s = atomicAdd(a.spill_n, 1u);
                if (s < a.spill_cap) a.spill[s] = SpillRec{buf[p], bucket};
                else atomicOr(a.abort, 1u);
            } else {
                atomicOr(a.abort, 1u);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------
// Dedup: per bucket keep the entry with the smallest id of each key (stable sort by key of an
// id-ordered list followed by unique_by_key, overlap.cu:212-213,222 / cqutils.cuh:271-306), then
// regenerate it for round 0 as a suffix and as a prefix (overlap.cu:266-267).
template <int CAP, int NT>
__global__ void __launch_bounds__(NT) k_dedup(DedupArgs a) {
    constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    constexpr int PER = CAP / NT;
    __shared__ uint32_t h1s[CAP], h2s[CAP], rids[CAP];
    __shared__ uint32_t head[CAP];
    __shared__ uint16_t nxt[CAP];
    __shared__ uint32_t sh[NT / 64];
    __shared__ uint32_t s_base;
    if (*a.abort) return;
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.cnt[b];
    if (n > (uint32_t)CAP) {
        if (threadIdx.x == 0) {
            uint32_t i = atomicAdd(a.n_big, 1u);
            if (i < a.big_cap) a.big[i] = b;
        }
        return;
    }
    if (n == 0) return;
    const Entry *in = a.in + (size_t)b * CAP;
    for (uint32_t i = threadIdx.x; i < CAP; i += NT) head[i] = EMPTY;
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
        Entry e = in[i];
        h1s[i] = e.h1; h2s[i] = e.h2; rids[i] = e.rid;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += NT) nxt[i] = (uint16_t)atomicExch(head + (h2s[i] & (CAP - 1)), i);
    __syncthreads();
    uint32_t keepmask = 0, kept = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        uint32_t i = threadIdx.x + NT * q;
        if (i < n) {
            uint32_t h1 = h1s[i], h2 = h2s[i], rid = rids[i];
            bool keep = true;
            for (uint32_t j = head[h2 & (CAP - 1)]; j != EMPTY && keep; j = (nxt[j] == 0xFFFFu ? EMPTY : nxt[j]))
                if (h2s[j] == h2 && h1s[j] == h1 && rids[j] < rid) keep = false;
            if (keep) { keepmask |= 1u << q; ++kept; }
        }
    }
    uint32_t tot;
    const uint32_t off = block_excl_scan<NT>(kept, sh, tot);
    const uint32_t sub = b & (SUBS - 1);
    if (threadIdx.x == 0) {
        uint32_t c = tot ? atomicAdd(a.lin_n + sub, tot) : 0u;
        if (c + tot > a.lin_cap) { atomicOr(a.lin_ovf, 1u); c = 0xFFFFFFFFu; }
        s_base = c;
    }
    __syncthreads();
    if (a.r0.active && s_base != 0xFFFFFFFFu) {
        uint32_t pos = sub * a.lin_cap + s_base + off;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            if (keepmask & (1u << q)) {
                uint32_t i = threadIdx.x + NT * q;
                Entry s = {h1s[i], h2s[i], rids[i]}, p = s;
                regen(a.chunks, a.r0, s, true);
                regen(a.chunks, a.r0, p, false);
                a.s_lin[pos] = s;
                a.p_lin[pos] = p;
                ++pos;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Join of round t for one bucket.
//
// LDS: the prefix entries (12 B each), a used flag per prefix, and one region that holds the hash
// table (chain heads + links) while probing and the matched records afterwards. The suffix
// entries stay in registers. The base each entry needs for the next round is gathered right
// after the entries are loaded, so its latency overlaps the hash phase.
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
        } m;
    } u;
    uint32_t sh[NT / 64];
    uint32_t ms, mp, ovf, bs, bp;
    unsigned long long base;
};

// Base code for the next round of a live entry (rid), or 0 when there is no next round.
__device__ __forceinline__ uint32_t next_code(const uint32_t *const *__restrict__ chunks, const RoundArgs &r, uint32_t rid,
                                              bool sfx) {
    return r.active ? entry_code(chunks, r, rid, sfx) : 0u;
}

__device__ __forceinline__ void regen_code(const RoundArgs &r, Entry &e, uint32_t code, bool sfx) {
    uint32_t ev = enc_of_code(code);
    if (sfx)
        next_sfx(e.h1, e.h2, ev, r.m1, r.m2);
    else
        prev_pfx(e.h1, e.h2, ev);
}

// Sorts n <= NT records (key, rank[, payload]) in place by (key, rank): each thread counts the
// records smaller than its own, then stores its record at that position. (key, rank) pairs are
// distinct, so the positions are a permutation.
template <int NT>
__device__ __forceinline__ void rank_sort(uint64_t *key, uint32_t *rank, uint32_t *pay, uint32_t n, uint64_t &k,
                                          uint32_t &r, uint32_t &pv, uint32_t &pos) {
    const uint32_t i = threadIdx.x;
    pos = 0;
    if (i < n) {
        k = key[i];
        r = rank[i];
        pv = pay ? pay[i] : 0u;
        for (uint32_t j = 0; j < n; ++j) {
            const uint64_t kj = key[j];
            pos += (kj < k) || (kj == k && rank[j] < r);
        }
    }
}

template <int CAP, int MCAP, int NT>
__global__ void __launch_bounds__(NT) k_join(JoinArgs a) {
    constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    constexpr int PER = CAP / NT;
    static_assert(MCAP <= NT, "one matched record per thread in the rank sort");
    using SM = JoinShared<CAP, MCAP, NT>;
    static_assert(sizeof(((SM *)nullptr)->u.m) <= sizeof(((SM *)nullptr)->u.h), "matched records fit the hash table");
    __shared__ SM S;

    if (*a.abort) return;
    const uint32_t b = blockIdx.x;
    const uint32_t ns = a.s_cnt[b], np = a.p_cnt[b];
    if (ns > (uint32_t)CAP || np > (uint32_t)CAP) {
        if (threadIdx.x == 0) {
            uint32_t i = atomicAdd(a.n_big, 1u);
            if (i < a.big_cap) a.big[i] = b;
            a.edge_cnt[b] = 0;
        }
        return;
    }
    if (threadIdx.x == 0) { S.ms = 0; S.mp = 0; S.ovf = 0; }
    const Entry *sin = a.s_in + (size_t)b * CAP;
    const Entry *pin = a.p_in + (size_t)b * CAP;
    for (uint32_t i = threadIdx.x; i < CAP; i += NT) S.u.h.head[i] = EMPTY;
    Entry se[PER];
    uint32_t sc[PER], scode[PER], pcode[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (i < np) {
            Entry e = pin[i];
            S.pk1[i] = e.h1; S.pk2[i] = e.h2; S.prid[i] = e.rid; S.used[i] = 0;
            pcode[q] = next_code(a.chunks, a.rn, e.rid, false);
        }
        if (i < ns) {
            se[q] = sin[i];
            scode[q] = next_code(a.chunks, a.rn, se[q].rid, true);
        }
        sc[q] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (i < np) S.u.h.nxt[i] = (uint16_t)atomicExch(S.u.h.head + (S.pk2[i] & (CAP - 1)), i);
    }
    __syncthreads();

    // Probe: count valid partners (different read) of every suffix; flag the prefixes used.
    if (np) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const uint32_t i = threadIdx.x + NT * q;
            if (i < ns) {
                const uint32_t h1 = se[q].h1, h2 = se[q].h2, rd = se[q].rid >> 1;
                uint32_t c = 0;
                for (uint32_t j = S.u.h.head[h2 & (CAP - 1)]; j != EMPTY; j = (S.u.h.nxt[j] == 0xFFFFu ? EMPTY : S.u.h.nxt[j])) {
                    if (S.pk2[j] == h2 && S.pk1[j] == h1 && (S.prid[j] >> 1) != rd) {
                        ++c;
                        S.used[j] = 1;
                    }
                }
                sc[q] = c;
            }
        }
    }
    __syncthreads();  // the hash table is dead from here on; its space holds the matched records
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (i < ns && sc[q]) {
            uint32_t m = atomicAdd(&S.ms, 1u);
            if (m < (uint32_t)MCAP) {
                S.u.m.msk[m] = ((uint64_t)se[q].h1 << 32) | se[q].h2;
                S.u.m.msr[m] = rid_rank(se[q].rid);
                S.u.m.msc[m] = sc[q];
            } else {
                S.ovf = 1;
            }
        }
        if (i < np && S.used[i]) {
            uint32_t m = atomicAdd(&S.mp, 1u);
            if (m < (uint32_t)MCAP) {
                S.u.m.mpk[m] = ((uint64_t)S.pk1[i] << 32) | S.pk2[i];
                S.u.m.mpr[m] = rid_rank(S.prid[i]);
            } else {
                S.ovf = 1;
            }
        }
    }
    __syncthreads();
    if (S.ovf) {
        if (threadIdx.x == 0) {
            uint32_t i = atomicAdd(a.n_big, 1u);
            if (i < a.big_cap) a.big[i] = b;
            a.edge_cnt[b] = 0;
        }
        return;
    }
    const uint32_t nms = S.ms, nmp = S.mp;

    // Edges of the bucket: order matched suffixes and prefixes by (key, rank), then emit the
    // cross product of each key group skipping same-read pairs (overlap.cu:557-601).
    if (nms) {
        uint64_t k1 = 0, k2 = 0;
        uint32_t r1 = 0, r2 = 0, c1 = 0, c2 = 0, p1, p2;
        rank_sort<NT>(S.u.m.msk, S.u.m.msr, S.u.m.msc, nms, k1, r1, c1, p1);
        rank_sort<NT>(S.u.m.mpk, S.u.m.mpr, nullptr, nmp, k2, r2, c2, p2);
        __syncthreads();
        if (threadIdx.x < nms) { S.u.m.msk[p1] = k1; S.u.m.msr[p1] = r1; S.u.m.msc[p1] = c1; }
        if (threadIdx.x < nmp) { S.u.m.mpk[p2] = k2; S.u.m.mpr[p2] = r2; }
        __syncthreads();
        const uint32_t i0 = 2 * threadIdx.x, i1 = i0 + 1;
        const uint32_t e0 = i0 < nms ? S.u.m.msc[i0] : 0, e1 = i1 < nms ? S.u.m.msc[i1] : 0;
        uint32_t tot;
        const uint32_t off = block_excl_scan<NT>(e0 + e1, S.sh, tot);
        if (i0 < nms) S.u.m.msc[i0] = off;
        if (i1 < nms) S.u.m.msc[i1] = off + e0;
        if (threadIdx.x == 0) {
            const uint32_t sub = b & (SUBS - 1);
            unsigned long long c = atomicAdd(a.pool_cursor + sub, (unsigned long long)tot);
            S.base = c + tot <= a.pool_cap ? sub * a.pool_cap + c : ~0ull;
            a.edge_off[b] = sub * a.pool_cap + c;
            a.edge_cnt[b] = tot;
        }
        __syncthreads();
        const unsigned long long base = S.base;
        if (base != ~0ull) {
            for (uint32_t i = threadIdx.x; i < nms; i += NT) {
                const uint64_t k = S.u.m.msk[i];
                const uint32_t srid = (S.u.m.msr[i] << 1) | (S.u.m.msr[i] >> 31);
                uint32_t lo = 0, hi = nmp;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (S.u.m.mpk[mid] < k) lo = mid + 1; else hi = mid;
                }
                unsigned long long pos = base + S.u.m.msc[i];
                const unsigned long long end = base + (i + 1 < nms ? S.u.m.msc[i + 1] : tot);
                for (uint32_t j = lo; j < nmp && S.u.m.mpk[j] == k && pos < end; ++j) {
                    const uint32_t prd = (S.u.m.mpr[j] << 1) | (S.u.m.mpr[j] >> 31);
                    if ((prd >> 1) != (srid >> 1)) {
                        a.pool_from[pos] = srid;
                        a.pool_to[pos] = prd;
                        ++pos;
                    }
                }
            }
        }
    } else if (threadIdx.x == 0) {
        a.edge_cnt[b] = 0;
        a.edge_off[b] = 0;
    }

    // Survivors: suffixes with no edge, prefixes never a target (overlap.cu:424-429,694-710),
    // regenerated for the next round and appended to the linear lists (one reservation per list).
    uint32_t smask = 0, pmask = 0, ks = 0, kp = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (i < ns && sc[q] == 0) { smask |= 1u << q; ++ks; }
        if (i < np && !S.used[i]) { pmask |= 1u << q; ++kp; }
    }
    uint32_t tt;
    const uint32_t o2 = block_excl_scan<NT>(ks | (kp << 16), S.sh, tt);
    const uint32_t ts = tt & 0xFFFFu, tp = tt >> 16;
    const uint32_t sub = b & (SUBS - 1);
    if (threadIdx.x == 0) {
        uint32_t cs = ts ? atomicAdd(a.s_lin_n + sub, ts) : 0u;
        uint32_t cp = tp ? atomicAdd(a.p_lin_n + sub, tp) : 0u;
        if (cs + ts > a.lin_cap || cp + tp > a.lin_cap) atomicOr(a.lin_ovf, 1u);
        S.bs = cs;
        S.bp = cp;
    }
    __syncthreads();
    if (a.rn.active && S.bs + ts <= a.lin_cap && S.bp + tp <= a.lin_cap) {
        uint32_t pos = sub * a.lin_cap + S.bs + (o2 & 0xFFFFu);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            if (smask & (1u << q)) {
                Entry e = se[q];
                regen_code(a.rn, e, scode[q], true);
                a.s_lin[pos++] = e;
            }
        }
        pos = sub * a.lin_cap + S.bp + (o2 >> 16);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            if (pmask & (1u << q)) {
                const uint32_t i = threadIdx.x + NT * q;
                Entry e = {S.pk1[i], S.pk2[i], S.prid[i]};
                regen_code(a.rn, e, pcode[q], false);
                a.p_lin[pos++] = e;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Large buckets: global-memory sort + binary-search join, one work-group of 1024 per bucket.
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
__device__ void gather_bucket(const Entry *slots, uint32_t cnt, uint32_t cap, const SpillRec *spill,
                              const uint32_t *spill_n, uint32_t spill_cap, uint32_t b, uint64_t *key, uint32_t *val,
                              uint32_t n2, bool as_rank, uint32_t *s_pos) {
    const uint32_t nm = min(cnt, cap);
    for (uint32_t i = threadIdx.x; i < nm; i += blockDim.x) {
        Entry e = slots[(size_t)b * cap + i];
        key[i] = key_of(e);
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
                    key[k] = key_of(e);
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

__global__ void __launch_bounds__(1024) k_join_big(BigArgs a) {
    __shared__ uint32_t sh[16];
    __shared__ uint32_t s_pos;
    __shared__ unsigned long long s_base;
    const uint32_t b = a.ids[blockIdx.x];
    const uint32_t ns = a.s_cnt[b], np = a.p_cnt[b];
    const uint32_t ns2 = pow2_at_least(ns), np2 = pow2_at_least(np);
    uint64_t *SK = a.skey + a.soff[blockIdx.x], *PK = a.pkey + a.poff[blockIdx.x];
    uint32_t *SR = a.srank + a.soff[blockIdx.x], *PR = a.prank + a.poff[blockIdx.x];
    uint32_t *SA = a.saux + a.soff[blockIdx.x], *PA = a.paux + a.poff[blockIdx.x];
    gather_bucket(a.s_in, ns, a.cap, a.s_spill, a.s_spill_n, a.spill_cap, b, SK, SR, ns2, true, &s_pos);
    gather_bucket(a.p_in, np, a.cap, a.p_spill, a.p_spill_n, a.spill_cap, b, PK, PR, np2, true, &s_pos);
    bitonic(SK, SR, nullptr, ns2);
    bitonic(PK, PR, nullptr, np2);
    // valid partner counts: the key group on the other side minus the same-read members
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
        const uint64_t k = SK[i];
        const uint32_t lo = lower_u64(PK, 0, np, k), hi = upper_u64(PK, lo, np, k);
        const uint32_t rd = SR[i] & 0x7FFFFFFFu;  // read index of the id
        uint32_t same = has_u32(PR, lo, hi, rd) + has_u32(PR, lo, hi, rd | 0x80000000u);
        SA[i] = (hi - lo) - same;
    }
    for (uint32_t j = threadIdx.x; j < np; j += blockDim.x) {
        const uint64_t k = PK[j];
        const uint32_t lo = lower_u64(SK, 0, ns, k), hi = upper_u64(SK, lo, ns, k);
        const uint32_t rd = PR[j] & 0x7FFFFFFFu;
        uint32_t same = has_u32(SR, lo, hi, rd) + has_u32(SR, lo, hi, rd | 0x80000000u);
        PA[j] = ((hi - lo) - same) ? 1u : 0u;
    }
    __syncthreads();
    // exclusive scan of SA in place: contiguous chunk per thread
    const uint32_t per = (ns + blockDim.x - 1) / blockDim.x;
    const uint32_t lo = min(ns, threadIdx.x * per), hi = min(ns, lo + per);
    uint32_t run = 0;
    for (uint32_t i = lo; i < hi; ++i) run += SA[i];
    uint32_t tot;
    uint32_t off = block_excl_scan<1024>(run, sh, tot);
    for (uint32_t i = lo; i < hi; ++i) {
        uint32_t c = SA[i];
        SA[i] = off;
        off += c;
    }
    const uint32_t sub = b & (SUBS - 1);
    if (threadIdx.x == 0) {
        unsigned long long c = atomicAdd(a.pool_cursor + sub, (unsigned long long)tot);
        s_base = c + tot <= a.pool_cap ? sub * a.pool_cap + c : ~0ull;
        a.edge_off[b] = sub * a.pool_cap + c;
        a.edge_cnt[b] = tot;
    }
    __syncthreads();
    const unsigned long long base = s_base;
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
        const uint32_t o = SA[i], e = (i + 1 < ns) ? SA[i + 1] : tot;
        const uint32_t srid = (SR[i] << 1) | (SR[i] >> 31);
        if (e > o && base != ~0ull) {
            const uint64_t k = SK[i];
            const uint32_t plo = lower_u64(PK, 0, np, k);
            unsigned long long pos = base + o;
            for (uint32_t j = plo; j < np && PK[j] == k && pos < base + e; ++j) {
                const uint32_t prd = (PR[j] << 1) | (PR[j] >> 31);
                if ((prd >> 1) != (srid >> 1)) {
                    a.pool_from[pos] = srid;
                    a.pool_to[pos] = prd;
                    ++pos;
                }
            }
        }
        if (e == o) {
            Entry en = {(uint32_t)(SK[i] >> 32), (uint32_t)SK[i], srid};
            uint32_t pos = atomicAdd(a.s_lin_n + sub, 1u);
            if (pos >= a.lin_cap) atomicOr(a.lin_ovf, 1u);
            else if (a.rn.active) {
                regen(a.chunks, a.rn, en, true);
                a.s_lin[sub * a.lin_cap + pos] = en;
            }
        }
    }
    for (uint32_t j = threadIdx.x; j < np; j += blockDim.x) {
        if (!PA[j]) {
            const uint32_t prd = (PR[j] << 1) | (PR[j] >> 31);
            Entry en = {(uint32_t)(PK[j] >> 32), (uint32_t)PK[j], prd};
            uint32_t pos = atomicAdd(a.p_lin_n + sub, 1u);
            if (pos >= a.lin_cap) atomicOr(a.lin_ovf, 1u);
            else if (a.rn.active) {
                regen(a.chunks, a.rn, en, false);
                a.p_lin[sub * a.lin_cap + pos] = en;
            }
        }
    }
}

__global__ void __launch_bounds__(1024) k_dedup_big(BigArgs a) {
    __shared__ uint32_t s_pos;
    const uint32_t b = a.ids[blockIdx.x];
    const uint32_t n = a.s_cnt[b], n2 = pow2_at_least(n);
    uint64_t *K = a.skey + a.soff[blockIdx.x];
    uint32_t *R = a.srank + a.soff[blockIdx.x];
    gather_bucket(a.s_in, n, a.cap, a.s_spill, a.s_spill_n, a.spill_cap, b, K, R, n2, false, &s_pos);
    bitonic(K, R, nullptr, n2);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        if (i == 0 || K[i] != K[i - 1]) {
            Entry e = {(uint32_t)(K[i] >> 32), (uint32_t)K[i], R[i]};
            const uint32_t sub = b & (SUBS - 1);
            uint32_t pos = atomicAdd(a.s_lin_n + sub, 1u);
            if (pos >= a.lin_cap) atomicOr(a.lin_ovf, 1u);
            else if (a.rn.active) {
                Entry p = e;
                regen(a.chunks, a.rn, e, true);
                regen(a.chunks, a.rn, p, false);
                a.s_lin[sub * a.lin_cap + pos] = e;
                a.p_lin[sub * a.lin_cap + pos] = p;
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_zero(ZeroArgs a) {
    for (int r = 0; r < a.count; ++r)
        for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < a.n[r]; i += gridDim.x * 256u) a.p[r][i] = a.v[r];
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Edge compaction: exclusive scan of per-bucket edge counts in bucket (= key) order, then copy.
__global__ void __launch_bounds__(1024) k_edge_scan(const uint32_t *__restrict__ cnt, uint32_t nb,
                                                    unsigned long long *__restrict__ dst_off) {
    __shared__ unsigned long long part[1024];
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t lo = min(nb, threadIdx.x * per), hi = min(nb, lo + per);
    unsigned long long s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        unsigned long long v = threadIdx.x >= o ? part[threadIdx.x - o] : 0ull;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    unsigned long long run = part[threadIdx.x] - s;
    for (uint32_t i = lo; i < hi; ++i) {
        dst_off[i] = run;
        run += cnt[i];
    }
}

__global__ void __launch_bounds__(256) k_edge_copy(EdgeCopyArgs a) {
    const uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= a.nb) return;
    const uint32_t c = a.cnt[b];
    const unsigned long long src = a.src_off[b], dst = a.dst_base + a.dst_off[b];
    for (uint32_t i = threadIdx.x & 63; i < c; i += 64) {
        a.out_from[dst + i] = a.pool_from[src + i];
        a.out_to[dst + i] = a.pool_to[src + i];
    }
}

// ---------------------------------------------------------------------------------------------
// Stage kernel: rolling transforms on a plain key array.
__global__ void __launch_bounds__(256) k_roll(const uint64_t *__restrict__ keys, const uint8_t *__restrict__ bases,
                                              uint64_t n, int kind, uint32_t m1, uint32_t m2, uint64_t *__restrict__ out) {
    uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    if (i >= n) return;
    uint64_t k = keys[i];
    uint32_t h1 = (uint32_t)(k >> 32), h2 = (uint32_t)k;
    uint32_t e = enc_of_code((bases[i] >> 1) & 3u);
    if (kind == 0)
        next_sfx(h1, h2, e, m1, m2);
    else
        prev_pfx(h1, h2, e);
    out[i] = ((uint64_t)h1 << 32) | h2;
}

// ---------------------------------------------------------------------------------------------
// Launchers.
hipError_t launch_map(const MapArgs &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    uint32_t grid = (a.n + a.rows - 1) / a.rows;
    size_t lds = (size_t)a.rows * a.Lpad;
    hipLaunchKernelGGL(k_map, dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_part(const PartArgs &a, uint32_t groups, hipStream_t s) {
    uint64_t grid = (uint64_t)groups * a.tiles_x;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_part, dim3((uint32_t)grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_dedup(const DedupArgs &a, uint32_t nb, hipStream_t s) {
    hipLaunchKernelGGL((k_dedup<BUCKET_CAP, JOIN_THREADS>), dim3(nb), dim3(JOIN_THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_join(const JoinArgs &a, uint32_t nb, hipStream_t s) {
    hipLaunchKernelGGL((k_join<BUCKET_CAP, MATCH_CAP, JOIN_THREADS>), dim3(nb), dim3(JOIN_THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_dedup_big(const BigArgs &a, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_dedup_big, dim3(n), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_join_big(const BigArgs &a, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_join_big, dim3(n), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_edge_compact(const EdgeCopyArgs &a, unsigned long long *dst_off, hipStream_t s) {
    hipLaunchKernelGGL(k_edge_scan, dim3(1), dim3(1024), 0, s, a.cnt, a.nb, dst_off);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_edge_copy, dim3((a.nb + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t g = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)(g < 4096 ? g : 4096)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_zero(const ZeroArgs &a, hipStream_t s) {
    uint32_t mx = 0;
    for (int r = 0; r < a.count; ++r) mx = a.n[r] > mx ? a.n[r] : mx;
    if (mx == 0) return hipSuccess;
    uint32_t g = (mx + 255) / 256;
    hipLaunchKernelGGL(k_zero, dim3(g < 1024 ? g : 1024), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_roll(const uint64_t *keys, const uint8_t *bases, uint64_t n, int kind, uint32_t m1, uint32_t m2,
                       uint64_t *out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_roll, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, keys, bases, n, kind, m1, m2, out);
    return hipGetLastError();
}

}  // namespace emp
