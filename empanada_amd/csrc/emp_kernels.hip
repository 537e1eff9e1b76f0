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
//   k_part     tile-sorted partition pass (at most 8 more key bits per pass); the first pass
//              of a round regenerates its input for the round (overlap.cu:266-267, 684-690):
//              one launch per list, so its column gathers touch a single column
//   k_dedup    per bucket: keep the smallest read id of every key (overlap.cu:212-222)
//              (linear output; round 0 partitions it once as suffixes, once as prefixes)
//   k_join     per bucket of round t: LDS hash join of suffixes against prefixes, edges of the
//              matched key groups in (key, rid&1, rid) x (rid&1, rid) order, removal of used
//              entries (linear output, keys of round t) (overlap.cu:258-290, 361-456, 557-710)
//   k_*_big    the same for buckets over capacity: global-memory sort + binary-search join
//   k_edge_*   bucket-ordered compaction of the edges of a round
#include "emp_common.h"
#include "emp_kernels.h"

#include <hip/hip_runtime.h>

namespace emp {

// ---------------------------------------------------------------------------------------------
// The 4 codes (8 bits, column 4g + c at bits 2c) of read k in column group g.
__device__ __forceinline__ uint32_t group_codes(const uint32_t *__restrict__ cols, uint32_t G, uint32_t g, uint32_t k) {
    const size_t w = ((size_t)(k >> CHUNK_SHIFT) * G + g) * CHUNK_GWORDS + ((k >> 2) & (CHUNK_GWORDS - 1));
    return (cols[w] >> (8 * (k & 3))) & 0xFFu;
}

// The carried-code field of a live entry right after its regeneration for round r (recomputed
// from its read's group word; used where an entry is rebuilt from its key alone).
__device__ __forceinline__ uint32_t carry_after(const uint32_t *__restrict__ cols, const RoundArgs &r, uint32_t rid,
                                                bool sfx) {
    const bool fwd = sfx != (bool)(rid & 1);
    const uint32_t col = fwd ? r.col_t : r.col_back;
    return carry_of(group_codes(cols, r.G, col >> 2, rid >> 1), col, fwd);
}

// ---------------------------------------------------------------------------------------------
// Next round's suffix list written by the kernel that finds its entries (NextSfxOut).

// Work-group form, in three steps so that a kernel can overlap the reservation with other work:
//   next_targets  every thread: gathers, regenerates its entries e[q] (bit q of `mask`) for nx.rg
//                 and enters each target bucket in the LDS table;
//   (barrier)     then threads below NXT_SLOTS reserve one run per target (next_reserve) and one
//                 thread adds the entries to nx.total;
//   (barrier)     next_write: every thread rebuilds its entries and writes them.
// Between the steps an entry is kept as one word: its gathered codes (9 bits, 0x100 = none), its
// table slot (6 bits, NXT_SLOTS = table full) and its rank in the slot (12 bits).
__device__ __forceinline__ uint32_t tgt_word(uint32_t tg) { return (tg & 0x100u) ? 0xFFFFFFFFu : (tg & 0xFFu); }

template <int PER>
__device__ __forceinline__ void next_targets(const NextSfxOut &nx, NxtTable &T, const Entry (&e)[PER], uint32_t mask,
                                             uint32_t (&tgt)[PER]) {
    // An entry gathers when it enters a column group (one round in four): rounds where no lane
    // of the wave needs a word skip the loads; otherwise all of the wave's gathers are in flight
    // together (lanes without one read the store's first word, a cached line).
    bool need[PER];
    uint32_t grp[PER], any = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        grp[q] = 0;
        need[q] = ((mask >> q) & 1u) && regen_needs_gather(nx.rg, e[q].rid, true, grp[q]);
        any |= need[q];
        tgt[q] = 0x100u;
    }
    if (__any(any)) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const uint32_t w = group_codes(nx.cols, nx.rg.G, need[q] ? grp[q] : 0u, need[q] ? e[q].rid >> 1 : 0u);
            if (need[q]) tgt[q] = w;
        }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if ((mask >> q) & 1u) {
            Entry ne = e[q];
            regen_carry(nx.cols, nx.rg, ne, true, tgt_word(tgt[q]));
            uint32_t rk = 0;
            const uint32_t sl = nxt_slot(T, (ne.h1 & KEY_MASK) >> nx.shift, rk);
            tgt[q] |= (sl << 9) | (rk << 15);
        }
    }
    if (mask) atomicAdd(&T.total, (uint32_t)__popc(mask));
}

// Run base of table slot k (thread k < NXT_SLOTS, after a barrier that follows next_targets).
__device__ __forceinline__ uint32_t next_reserve(const NextSfxOut &nx, const NxtTable &T, uint32_t k) {
    const uint32_t key = T.key[k], c = T.cnt[k];
    return (key != 0xFFFFFFFFu && c) ? atomicAdd(nx.cnt + key, c) : 0u;
}

template <int PER>
__device__ __forceinline__ void next_write(const NextSfxOut &nx, const NxtTable &T, const Entry (&e)[PER], uint32_t mask,
                                           const uint32_t (&tgt)[PER]) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if ((mask >> q) & 1u) {
            Entry ne = e[q];
            regen_carry(nx.cols, nx.rg, ne, true, tgt_word(tgt[q]));
            const uint32_t nb = (ne.h1 & KEY_MASK) >> nx.shift, sl = (tgt[q] >> 9) & 63u;
            const uint32_t slot = sl < NXT_SLOTS ? T.base[sl] + (tgt[q] >> 15) : atomicAdd(nx.cnt + nb, 1u);
            if (slot < (uint32_t)BUCKET_CAP) {
                nx.fine[(size_t)nb * BUCKET_CAP + slot] = ne;
            } else {
                const uint32_t sp = atomicAdd(nx.spill_n, 1u);
                if (sp < nx.spill_cap) nx.spill[sp] = SpillRec{ne, nb};
                else atomicOr(nx.ovf, 1u);
            }
        }
    }
}

// The three steps back to back (every thread of the work-group calls it).
template <int PER>
__device__ __forceinline__ void scatter_next(const NextSfxOut &nx, NxtTable &T, const Entry (&e)[PER], uint32_t mask) {
    uint32_t tgt[PER];
    next_targets<PER>(nx, T, e, mask, tgt);
    lds_barrier();
    if (threadIdx.x < NXT_SLOTS) T.base[threadIdx.x] = next_reserve(nx, T, threadIdx.x);
    else if (threadIdx.x == NXT_SLOTS && T.total) atomicAdd(nx.total, T.total);
    lds_barrier();
    next_write<PER>(nx, T, e, mask, tgt);
}

// Single-entry form (large-bucket kernels): e is a suffix entry of this round with its carried
// codes in place.
__device__ __forceinline__ void put_next(const NextSfxOut &nx, Entry e) {
    uint32_t g, word = 0xFFFFFFFFu;
    if (regen_needs_gather(nx.rg, e.rid, true, g)) word = group_codes(nx.cols, nx.rg.G, g, e.rid >> 1);
    regen_carry(nx.cols, nx.rg, e, true, word);
    const uint32_t nb = (e.h1 & KEY_MASK) >> nx.shift;
    const uint32_t slot = atomicAdd(nx.cnt + nb, 1u);
    atomicAdd(nx.total, 1u);
    if (slot < (uint32_t)BUCKET_CAP) {
        nx.fine[(size_t)nb * BUCKET_CAP + slot] = e;
    } else {
        const uint32_t sp = atomicAdd(nx.spill_n, 1u);
        if (sp < nx.spill_cap) nx.spill[sp] = SpillRec{e, nb};
        else atomicOr(nx.ovf, 1u);
    }
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
    lds_barrier();

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

    // Columns: word w of the tile covers reads 4w..4w+3; task id = g * nw + w keeps a wave on
    // consecutive words of one column group (contiguous stores).
    const uint32_t G = col_groups((uint32_t)L);
    const uint32_t nw = (nt + 3) / 4;
    const uint32_t ntask = nw * G;
    const uint64_t w0 = (a.k0 + first) >> 2;
    for (uint32_t task = threadIdx.x; task < ntask; task += 256) {
        const uint32_t g = task / nw, w = task - g * nw;
        uint32_t word = 0;
        const uint32_t rmax = min(4u, nt - 4 * w);
        for (uint32_t i = 0; i < rmax; ++i) {
            const uint8_t *row = tile + (4 * w + i) * Lp;
#pragma unroll
            for (uint32_t c = 0; c < 4; ++c) {
                const uint32_t j = 4 * g + c;
                if (j < (uint32_t)L) word |= ((uint32_t)(row[j] >> 1) & 3u) << (2 * (4 * i + c));
            }
        }
        const uint64_t W = w0 + w;
        if (a.gs)
            a.cols[a.cbase + (uint64_t)g * a.gs + W] = word;
        else
            a.cols[((W >> (CHUNK_SHIFT - 2)) * G + g) * CHUNK_GWORDS + (W & (CHUNK_GWORDS - 1))] = word;
    }
}

// ---------------------------------------------------------------------------------------------
// Partition pass: each work-group takes tiles of PART_TILE entries of one input group, ranks them
// by their next `dbits` key bits in LDS, reserves one output range per (tile, bucket) with a
// single atomic, and writes each bucket's run contiguously. With a.rg.active (first pass of a
// round) the entries are streamed in with non-temporal loads, so that L2 keeps the one column
// the launch gathers from, and regenerated before they are bucketed (REGEN = a.rg.active).
template <bool REGEN>
__global__ void __launch_bounds__(PART_THREADS) __attribute__((amdgpu_waves_per_eu(EMP_PART_WAVES))) k_part(PartArgs a) {
    constexpr int NT = PART_THREADS, T = PART_TILE, PER = T / NT, NB = 1 << PART_MAX_BITS;
    static_assert(NT >= NB && T % NT == 0, "one bin per thread");
    __shared__ Entry buf[T];
    __shared__ uint32_t hist[NB], loc[NB], base[NB], sh[NT / 64];
    // Regenerating pass: the input groups are the SUBS sub-ranges of a survivor list; the
    // work-groups of read range r (sub-ranges r * SUBS_PER_RANGE ...) all have id = r mod 8.
    uint32_t g, tile0;
    if (REGEN) {
        const uint32_t x = blockIdx.x & (NRANGE - 1), j = blockIdx.x / NRANGE;
        g = x * SUBS_PER_RANGE + j / a.tiles_x;
        tile0 = j % a.tiles_x;
    } else {
        g = blockIdx.x / a.tiles_x;
        tile0 = blockIdx.x % a.tiles_x;
    }
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
    const uint32_t gb = a.keyed ? (g << a.dbits) : 0u;
    for (uint32_t tile = tile0; (size_t)tile * T < n_g; tile += a.tiles_x) {
        const uint32_t t0 = tile * T, nt = min((uint32_t)T, n_g - t0);
        if (threadIdx.x < NB) hist[threadIdx.x] = 0;
        lds_barrier();
        // Without regeneration every lane loads (lanes past the tile re-read its last entry), so the
        // loads of all PER entries are in flight together.
        Entry e[PER];
        uint32_t r[PER], d[PER];
        const uint32_t last = nt - 1;
        if (REGEN) {
            // entry and gather per slot: within the 80 registers of 6 waves per SIMD, loading all
            // 8 entries first spills
            uint32_t word[PER];
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                uint32_t i = threadIdx.x + NT * q;
                word[q] = 0xFFFFFFFFu;
                if (i < nt) {
                    e[q] = load_stream(in + t0 + i);
                    uint32_t g;
                    if (regen_needs_gather(a.rg, e[q].rid, a.sfx, g)) word[q] = group_codes(a.cols, a.rg.G, g, e[q].rid >> 1);
                }
            }
#pragma unroll
            for (int q = 0; q < PER; ++q)
                if (threadIdx.x + NT * q < nt) regen_carry(a.cols, a.rg, e[q], a.sfx, word[q]);
        } else {
#pragma unroll
            for (int q = 0; q < PER; ++q) e[q] = in[t0 + min(threadIdx.x + NT * q, last)];
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            uint32_t i = threadIdx.x + NT * q;
            if (i < nt) {
                d[q] = (e[q].h1 >> a.out_shift) & (F - 1);
                r[q] = atomicAdd(&hist[d[q]], 1u);
            }
        }
        lds_barrier();
        const uint32_t h = threadIdx.x < NB ? hist[threadIdx.x] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<NT>(h, sh, tot);
        if (threadIdx.x < NB) {
            loc[threadIdx.x] = ex;
            base[threadIdx.x] = h ? atomicAdd(&a.out_cnt[gb + threadIdx.x], h) : 0u;
        }
        lds_barrier();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            uint32_t i = threadIdx.x + NT * q;
            if (i < nt) buf[loc[d[q]] + r[q]] = e[q];
        }
        lds_barrier();
        for (uint32_t p = threadIdx.x; p < nt; p += NT) {
            const Entry v = buf[p];
            const uint32_t dd = (v.h1 >> a.out_shift) & (F - 1);
            const uint32_t slot = base[dd] + (p - loc[dd]);
            const uint32_t bucket = gb + dd;
            if (slot < a.out_cap) {
                a.out[(size_t)bucket * a.out_cap + slot] = v;
            } else if (a.spill) {
                uint32_t s = atomicAdd(a.spill_n, 1u);
                if (s < a.spill_cap) a.spill[s] = SpillRec{v, bucket};
                else atomicOr(a.abort, 1u);
            } else {
                atomicOr(a.abort, 1u);
            }
        }
        lds_barrier();
    }
}

// ---------------------------------------------------------------------------------------------
// Dedup: per bucket keep the entry with the smallest id of each key (stable sort by key of an
// id-ordered list followed by unique_by_key, overlap.cu:212-213,222 / cqutils.cuh:271-306).
template <int CAP, int NT>
__global__ void __launch_bounds__(NT) k_dedup(DedupArgs a) {
    constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    constexpr int PER = CAP / NT;
    __shared__ uint32_t h1s[CAP], h2s[CAP], rids[CAP];
    __shared__ uint32_t head[CAP];
    __shared__ uint16_t nxt[CAP];
    __shared__ uint32_t rcnt[NRANGE], rbase[NRANGE];
    __shared__ NxtTable T;
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
    if (threadIdx.x < NRANGE) rcnt[threadIdx.x] = 0;
    nxt_init(T);
    {
        Entry e[PER];  // all loads in flight together (lanes past n re-read the last entry)
#pragma unroll
        for (int q = 0; q < PER; ++q) e[q] = in[min(threadIdx.x + NT * q, n - 1)];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const uint32_t i = threadIdx.x + NT * q;
            if (i < n) { h1s[i] = e[q].h1; h2s[i] = e[q].h2; rids[i] = e[q].rid; }
        }
    }
    lds_barrier();
    for (uint32_t i = threadIdx.x; i < n; i += NT) nxt[i] = (uint16_t)atomicExch(head + (h2s[i] & (CAP - 1)), i);
    lds_barrier();
    // kept entries: a slot in their read range's run of this work-group (one reservation per range)
    uint32_t keepmask = 0, slot[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        uint32_t i = threadIdx.x + NT * q;
        if (i < n) {
            uint32_t h1 = h1s[i], h2 = h2s[i], rid = rids[i];
            bool keep = true;
            for (uint32_t j = head[h2 & (CAP - 1)]; j != EMPTY && keep; j = (nxt[j] == 0xFFFFu ? EMPTY : nxt[j]))
                if (((h2s[j] ^ h2) & KEY_MASK) == 0 && ((h1s[j] ^ h1) & KEY_MASK) == 0 && rids[j] < rid) keep = false;
            if (keep) {
                keepmask |= 1u << q;
                slot[q] = atomicAdd(&rcnt[read_range(rid)], 1u);
            }
        }
    }
    lds_barrier();
    if (threadIdx.x < NRANGE) {
        const uint32_t r = threadIdx.x, cnt = rcnt[r], sub = range_sub(r, b);
        uint32_t c = cnt ? atomicAdd(a.lin_n + sub, cnt) : 0u;
        if (c + cnt > a.lin_cap) { atomicOr(a.lin_ovf, 1u); c = 0xFFFFFFFFu; }
        rbase[r] = c == 0xFFFFFFFFu ? c : sub * a.lin_cap + c;
    }
    lds_barrier();
    Entry kept[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if (keepmask & (1u << q)) {
            const uint32_t i = threadIdx.x + NT * q, base = rbase[read_range(rids[i])];
            kept[q] = Entry{h1s[i], h2s[i], rids[i]};
            if (base != 0xFFFFFFFFu) a.lin[base + slot[q]] = kept[q];
        }
    }
    // the kept entries are also round 0's suffix list
    if (a.nx.active) scatter_next<PER>(a.nx, T, kept, keepmask);
}

// ---------------------------------------------------------------------------------------------
// Join of round t for one bucket.
//
// LDS: the prefix entries (12 B each), a used flag per prefix, and one region that holds the hash
// table (chain heads + links) while probing and the matched records afterwards. The suffix
// entries stay in registers.

// Loads the suffix and prefix entries of bucket b (nothing for a bucket over capacity).
template <int CAP, int NT>
__device__ __forceinline__ void load_bucket(const JoinArgs &a, uint32_t b, uint32_t ns, uint32_t np, Entry (&se)[CAP / NT],
                                            Entry (&pe)[CAP / NT]) {
    const Entry *sin = a.s_in + (size_t)b * CAP;
    const Entry *pin = a.p_in + (size_t)b * CAP;
    const bool fit = ns <= (uint32_t)CAP && np <= (uint32_t)CAP;
#pragma unroll
    for (int q = 0; q < CAP / NT; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (fit && i < ns) se[q] = sin[i];
        if (fit && i < np) pe[q] = pin[i];
    }
}

// Join of one bucket: suffixes in registers se[] (slot threadIdx.x + NT * q), prefixes already
// in S.pk1/pk2/prid with S.used cleared. Buckets over capacity, or with more matched records
// than MCAP, are handed to the large-bucket path.
template <int CAP, int MCAP, int NT>
__device__ __forceinline__ void join_bucket(const JoinArgs &a, JoinShared<CAP, MCAP, NT> &S, uint32_t b, uint32_t ns,
                                            uint32_t np, const Entry (&se)[CAP / NT]) {
    constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    constexpr int PER = CAP / NT;
    if (ns > (uint32_t)CAP || np > (uint32_t)CAP) {
        if (threadIdx.x == 0) {
            uint32_t i = atomicAdd(a.n_big, 1u);
            if (i < a.big_cap) a.big[i] = b;
            a.edge_cnt[b] = 0;
        }
        return;
    }
    if (threadIdx.x == 0) { S.ms = 0; S.mp = 0; S.ovf = 0; S.etot = 0; }
    if (threadIdx.x < 2 * NRANGE) S.rcnt[threadIdx.x] = 0;
    nxt_init(S.nt);
    for (uint32_t i = threadIdx.x; i < CAP; i += NT) S.u.h.head[i] = EMPTY;
    uint32_t sc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) sc[q] = 0;
    lds_barrier();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (i < np) S.u.h.nxt[i] = (uint16_t)atomicExch(S.u.h.head + (S.pk2[i] & (CAP - 1)), i);
    }
    lds_barrier();

    // Probe: count valid partners (different read) of every suffix; flag the prefixes used.
    if (np) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const uint32_t i = threadIdx.x + NT * q;
            if (i < ns) {
                const uint32_t h1 = se[q].h1 & KEY_MASK, h2 = se[q].h2 & KEY_MASK, rd = se[q].rid >> 1;
                uint32_t c = 0;
                for (uint32_t j = S.u.h.head[h2 & (CAP - 1)]; j != EMPTY; j = (S.u.h.nxt[j] == 0xFFFFu ? EMPTY : S.u.h.nxt[j])) {
                    if ((S.pk2[j] & KEY_MASK) == h2 && (S.pk1[j] & KEY_MASK) == h1 && (S.prid[j] >> 1) != rd) {
                        ++c;
                        S.used[j] = 1;
                    }
                }
                sc[q] = c;
            }
        }
    }
    lds_barrier();  // the hash table is dead from here on; its space holds the matched records

    // Matched records (key, rank) of both sides, the bucket's edge total, and a slot for every
    // survivor (suffixes with no edge, prefixes never a target: overlap.cu:424-429,694-710) in its
    // run of (side, read range).
    uint32_t smask = 0, pmask = 0, slot[PER], ecount = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        slot[q] = 0;
        if (i < ns) {
            if (sc[q]) {
                ecount += sc[q];
                uint32_t m = atomicAdd(&S.ms, 1u);
                if (m < (uint32_t)MCAP) {
                    S.u.m.msk[m] = ((uint64_t)(se[q].h1 & KEY_MASK) << 32) | (se[q].h2 & KEY_MASK);
                    S.u.m.msr[m] = rid_rank(se[q].rid);
                    S.u.m.msc[m] = sc[q];
                } else {
                    S.ovf = 1;
                }
            } else {
                smask |= 1u << q;  // a survivor: into the next round's buckets below
            }
        }
        if (i < np) {
            const uint32_t rid = S.prid[i];
            if (S.used[i]) {
                uint32_t m = atomicAdd(&S.mp, 1u);
                if (m < (uint32_t)MCAP) {
                    S.u.m.mpk[m] = ((uint64_t)(S.pk1[i] & KEY_MASK) << 32) | (S.pk2[i] & KEY_MASK);
                    S.u.m.mpr[m] = rid_rank(rid);
                } else {
                    S.ovf = 1;
                }
            } else {
                pmask |= 1u << q;
                slot[q] |= atomicAdd(&S.rcnt[NRANGE + read_range(rid)], 1u) << 16;
            }
        }
    }
    // Suffix survivors: their next-round targets (the entries are rebuilt when written, below).
    uint32_t tgt[PER];
    if (a.nx.active) next_targets<PER>(a.nx, S.nt, se, smask, tgt);
    else if (smask) atomicAdd(&S.nt.total, (uint32_t)__popc(smask));  // last round: only counted
    if (ecount) atomicAdd(&S.etot, ecount);
    lds_barrier();
    if (S.ovf) {
        if (threadIdx.x == 0) {
            uint32_t i = atomicAdd(a.n_big, 1u);
            if (i < a.big_cap) a.big[i] = b;
            a.edge_cnt[b] = 0;
        }
        return;
    }
    const uint32_t nms = S.ms, nmp = S.mp, tot = S.etot;

    // Reservations, issued now and read back after the sort so that their latency overlaps it:
    // thread k in [8, 16) reserves the prefix survivor run of read range k % 8, thread 16 the edges,
    // threads 64 + k the next-round suffix run of table slot k.
    constexpr uint32_t NXT0 = 64;
    static_assert(NT >= NXT0 + NXT_SLOTS + 1, "reservation threads");
    unsigned long long resv = 0;
    if (threadIdx.x >= NRANGE && threadIdx.x < 2 * NRANGE) {
        const uint32_t k = threadIdx.x, cnt = S.rcnt[k], sub = range_sub(k & (NRANGE - 1), b);
        resv = cnt ? atomicAdd(a.p_lin_n + sub, cnt) : 0u;
    } else if (threadIdx.x == 2 * NRANGE && tot) {
        resv = atomicAdd(a.pool_cursor + (b & (SUBS - 1)), (unsigned long long)tot);
    } else if (threadIdx.x >= NXT0 && threadIdx.x < NXT0 + NXT_SLOTS && a.nx.active) {
        resv = next_reserve(a.nx, S.nt, threadIdx.x - NXT0);
    } else if (threadIdx.x == NXT0 + NXT_SLOTS && S.nt.total) {
        atomicAdd(a.nx.total, S.nt.total);
    }

    // Edges of the bucket: order matched suffixes and prefixes by (key, rank), then emit the
    // cross product of each key group skipping same-read pairs (overlap.cu:557-601).
    if (nms) {
        // one record per thread: suffix t, or prefix t - nms (both sides at once when they fit)
        const bool both = nms + nmp <= (uint32_t)NT;
        uint64_t k1 = 0, k2 = 0;
        uint32_t r1 = 0, r2 = 0, c1 = 0, p1 = 0, p2 = 0;
        const uint32_t t = threadIdx.x;
        if (t < nms) {
            p1 = rank_of(S.u.m.msk, S.u.m.msr, nms, t, k1, r1);
            c1 = S.u.m.msc[t];
        } else if (both && t - nms < nmp) {
            p2 = rank_of(S.u.m.mpk, S.u.m.mpr, nmp, t - nms, k2, r2);
        }
        if (!both && t < nmp) p2 = rank_of(S.u.m.mpk, S.u.m.mpr, nmp, t, k2, r2);
        const bool hasp = both ? (t >= nms && t - nms < nmp) : t < nmp;
        lds_barrier();
        if (t < nms) { S.u.m.msk[p1] = k1; S.u.m.msr[p1] = r1; S.u.m.msc[p1] = c1; }
        if (hasp) { S.u.m.mpk[p2] = k2; S.u.m.mpr[p2] = r2; }
    }
    if (threadIdx.x >= NRANGE && threadIdx.x < 2 * NRANGE) {
        const uint32_t k = threadIdx.x, cnt = S.rcnt[k], sub = range_sub(k & (NRANGE - 1), b);
        uint32_t c = (uint32_t)resv;
        if (c + cnt > a.lin_cap) { atomicOr(a.lin_ovf, 1u); c = 0xFFFFFFFFu; }
        S.rbase[k] = c == 0xFFFFFFFFu ? c : sub * a.lin_cap + c;
    } else if (threadIdx.x == 2 * NRANGE) {
        const uint32_t sub = b & (SUBS - 1);
        const unsigned long long base =
            tot ? place_edges(resv, tot, sub, a.pool_cursor + sub, a.pool_cap, a.ovf_cursor, a.ovf_cap) : ~0ull;
        S.base = base;
        a.edge_off[b] = base;
        a.edge_cnt[b] = tot;
    } else if (threadIdx.x >= NXT0 && threadIdx.x < NXT0 + NXT_SLOTS) {
        S.nt.base[threadIdx.x - NXT0] = (uint32_t)resv;
    }
    lds_barrier();
    if (nms) {
        const uint32_t i0 = 2 * threadIdx.x, i1 = i0 + 1;
        const uint32_t e0 = i0 < nms ? S.u.m.msc[i0] : 0, e1 = i1 < nms ? S.u.m.msc[i1] : 0;
        uint32_t stot;
        const uint32_t off = block_excl_scan<NT>(e0 + e1, S.sh, stot);
        if (i0 < nms) S.u.m.msc[i0] = off;
        if (i1 < nms) S.u.m.msc[i1] = off + e0;
        lds_barrier();
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
    }

    // Survivors: prefixes into their runs (the next round's first partition pass regenerates
    // them), suffixes regenerated here into the next round's suffix buckets.
    if (a.keep) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            if (pmask & (1u << q)) {
                const uint32_t i = threadIdx.x + NT * q, rid = S.prid[i];
                const uint32_t base = S.rbase[NRANGE + read_range(rid)];
                if (base != 0xFFFFFFFFu) a.p_lin[base + (slot[q] >> 16)] = Entry{S.pk1[i], S.pk2[i], rid};
            }
        }
    }
    if (a.nx.active) next_write<PER>(a.nx, S.nt, se, smask, tgt);
}

// Prefix entries of a bucket (registers) into LDS, used flags cleared.
template <int CAP, int MCAP, int NT>
__device__ __forceinline__ void stage_prefixes(JoinShared<CAP, MCAP, NT> &S, uint32_t np, const Entry (&pe)[CAP / NT]) {
    if (np > (uint32_t)CAP) return;
#pragma unroll
    for (int q = 0; q < CAP / NT; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (i < np) { S.pk1[i] = pe[q].h1; S.pk2[i] = pe[q].h2; S.prid[i] = pe[q].rid; S.used[i] = 0; }
    }
}

// Join kernel. PERSIST: work-group g takes buckets g, g + G, g + 2G, ... (G = grid size); while
// it joins bucket b, the entries of bucket b + G and the counts of bucket b + 2G are in flight.
// Otherwise one work-group per bucket.
template <int CAP, int MCAP, int NT, bool PERSIST>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) k_join(JoinArgs a) {
    constexpr int PER = CAP / NT;
    static_assert(MCAP <= NT, "one matched record per thread in the rank sort");
    using SM = JoinShared<CAP, MCAP, NT>;
    static_assert(sizeof(((SM *)nullptr)->u.m) <= sizeof(((SM *)nullptr)->u.h), "matched records fit the hash table");
    __shared__ SM S;

    if (*a.abort) return;
    const uint32_t G = gridDim.x, nb = a.nb;
    uint32_t b = blockIdx.x;
    if (b >= nb) return;
    if (!PERSIST) {
        const uint32_t ns = a.s_cnt[b], np = a.p_cnt[b];
        Entry se[PER], pe[PER];
        load_bucket<CAP, NT>(a, b, ns, np, se, pe);
        stage_prefixes<CAP, MCAP, NT>(S, np, pe);
        join_bucket<CAP, MCAP, NT>(a, S, b, ns, np, se);
        return;
    }
    uint32_t ns = a.s_cnt[b], np = a.p_cnt[b];
    uint32_t b1 = b + G, ns1 = 0, np1 = 0;
    if (b1 < nb) { ns1 = a.s_cnt[b1]; np1 = a.p_cnt[b1]; }
    Entry se[PER], pe[PER];
    load_bucket<CAP, NT>(a, b, ns, np, se, pe);
    stage_prefixes<CAP, MCAP, NT>(S, np, pe);
    for (;;) {
        Entry sn[PER], pn[PER];
        const uint32_t b2 = b1 + G;
        uint32_t ns2 = 0, np2 = 0;
        if (b1 < nb) {
            load_bucket<CAP, NT>(a, b1, ns1, np1, sn, pn);
            if (b2 < nb) { ns2 = a.s_cnt[b2]; np2 = a.p_cnt[b2]; }
        }
        join_bucket<CAP, MCAP, NT>(a, S, b, ns, np, se);
        if (b1 >= nb) break;
        lds_barrier();  // LDS of bucket b is dead
        stage_prefixes<CAP, MCAP, NT>(S, np1, pn);
        b = b1; ns = ns1; np = np1;
        b1 = b2; ns1 = ns2; np1 = np2;
#pragma unroll
        for (int q = 0; q < PER; ++q) se[q] = sn[q];
    }
}

// ---------------------------------------------------------------------------------------------
// Large buckets: global-memory sort + binary-search join, one work-group of 1024 per bucket.
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
    gather_bucket(a.p_in, np, a.cap, a.p_spill, a.p_spill_n, a.p_spill_cap, b, PK, PR, np2, true, &s_pos);
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
        unsigned long long base = ~0ull;
        if (tot) {
            const unsigned long long c = atomicAdd(a.pool_cursor + sub, (unsigned long long)tot);
            base = place_edges(c, tot, sub, a.pool_cursor + sub, a.pool_cap, a.ovf_cursor, a.ovf_cap);
        }
        s_base = base;
        a.edge_off[b] = base;
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
        if (e == o) {  // a survivor
            if (a.nx.active) {
                Entry en = {(uint32_t)(SK[i] >> 32), (uint32_t)SK[i], srid};
                set_carry(en, carry_after(a.cols, a.rg, srid, true));
                put_next(a.nx, en);
            } else {
                atomicAdd(a.nx.total, 1u);
            }
        }
    }
    for (uint32_t j = threadIdx.x; j < np; j += blockDim.x) {
        if (!PA[j]) {
            const uint32_t prd = (PR[j] << 1) | (PR[j] >> 31);
            Entry en = {(uint32_t)(PK[j] >> 32), (uint32_t)PK[j], prd};
            const uint32_t rs = range_sub(read_range(prd), b);
            uint32_t pos = atomicAdd(a.p_lin_n + rs, 1u);
            if (pos >= a.lin_cap) atomicOr(a.lin_ovf, 1u);
            else if (a.keep) {
                set_carry(en, carry_after(a.cols, a.rg, prd, false));
                a.p_lin[rs * a.lin_cap + pos] = en;
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
            const uint32_t sub = range_sub(read_range(R[i]), b);
            uint32_t pos = atomicAdd(a.s_lin_n + sub, 1u);
            if (pos >= a.lin_cap) atomicOr(a.lin_ovf, 1u);
            else {
                a.s_lin[sub * a.lin_cap + pos] = e;
            }
            if (a.nx.active) put_next(a.nx, e);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Collision audit (FPCHECK: fpcheck and print_false_positive_rate, overlap.cu:320-359). One
// work-group per bucket sorts its entries by key (LDS, or global scratch for a bucket over
// capacity); every entry then counts as a new key when it opens its key group, and as a new
// string when no earlier entry of its group spells the same string.
constexpr int AUDIT_THREADS = 256;

// 2-bit code of base `pos` of oriented read `rid` (odd ids: reverse complement, code ^ 2).
__device__ __forceinline__ uint32_t oriented_code(const AuditArgs &a, uint32_t rid, int pos) {
    const int c = (rid & 1) ? a.L - 1 - pos : pos;
    const uint32_t v = (group_codes(a.cols, a.G, (uint32_t)c >> 2, rid >> 1) >> (2 * (c & 3))) & 3u;
    return (rid & 1) ? v ^ 2u : v;
}

__device__ bool same_string(const AuditArgs &a, uint32_t r1, uint32_t r2) {
    if (((r1 ^ r2) & 1u) == 0) {
        // one orientation: the same columns of both reads ([start, start+len) forward,
        // [L-start-len, L-start) reverse-complemented), four at a time
        const int c0 = (r1 & 1) ? a.L - a.start - a.len : a.start, c1 = c0 + a.len;
        for (int g = c0 >> 2; g <= (c1 - 1) >> 2; ++g) {
            const int lo = max(c0 - 4 * g, 0), hi = min(c1 - 4 * g, 4);
            const uint32_t m = ((1u << (2 * hi)) - 1u) & ~((1u << (2 * lo)) - 1u);
            if ((group_codes(a.cols, a.G, (uint32_t)g, r1 >> 1) ^ group_codes(a.cols, a.G, (uint32_t)g, r2 >> 1)) & m)
                return false;
        }
        return true;
    }
    for (int p = 0; p < a.len; ++p)
        if (oriented_code(a, r1, a.start + p) != oriented_code(a, r2, a.start + p)) return false;
    return true;
}

__global__ void __launch_bounds__(AUDIT_THREADS) k_audit(AuditArgs a) {
    __shared__ uint64_t lk[BUCKET_CAP];
    __shared__ uint32_t lr[BUCKET_CAP];
    __shared__ uint32_t s_pos;
    __shared__ unsigned long long red[2];
    const bool big = a.ids != nullptr;
    const uint32_t b = big ? a.ids[blockIdx.x] : blockIdx.x;
    const uint32_t n = a.cnt[b];
    if (n == 0 || (!big && n > a.cap)) return;  // a bucket over capacity comes in the large-bucket launch
    uint64_t *K = big ? a.skey + a.off[blockIdx.x] : lk;
    uint32_t *R = big ? a.srid + a.off[blockIdx.x] : lr;
    const uint32_t n2 = pow2_at_least(n);
    if (threadIdx.x == 0) { red[0] = 0; red[1] = 0; }
    gather_bucket(a.fine, n, a.cap, a.spill, a.spill_n, a.spill_cap, b, K, R, n2, false, &s_pos);
    bitonic(K, R, nullptr, n2);
    unsigned long long keys = 0, strs = 0;
    for (uint32_t i = threadIdx.x; i < n; i += AUDIT_THREADS) {
        const uint64_t k = K[i];
        keys += (i == 0 || K[i - 1] != k) ? 1u : 0u;
        bool fresh = true;
        for (uint32_t j = i; j-- > 0 && K[j] == k;)
            if (same_string(a, R[i], R[j])) { fresh = false; break; }
        strs += fresh ? 1u : 0u;
    }
    if (keys) atomicAdd(&red[0], keys);
    if (strs) atomicAdd(&red[1], strs);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(a.out, red[0]);
        atomicAdd(a.out + 1, red[1]);
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
// Edge compaction at the end of a run: the per-bucket edge counts of all rounds, in (round,
// bucket) order (= output.bin order), get exclusive offsets in two levels (blocks of
// ESCAN_BLOCK counts, then the block totals), and every bucket's edges are copied from the pool
// to their offset in the output arena.
constexpr int ESCAN_BLOCK = 4096;

__global__ void __launch_bounds__(1024) k_escan_local(const uint32_t *__restrict__ cnt, uint64_t n,
                                                     unsigned long long *__restrict__ dst,
                                                     unsigned long long *__restrict__ blk) {
    __shared__ unsigned long long sh[16];
    const uint64_t i0 = (uint64_t)blockIdx.x * ESCAN_BLOCK + 4ull * threadIdx.x;
    uint32_t v[4];
    unsigned long long sum = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        v[u] = i0 + u < n ? cnt[i0 + u] : 0u;
        sum += v[u];
    }
    unsigned long long tot;
    unsigned long long off = block_excl_scan64<1024>(sum, sh, tot);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (i0 + u < n) dst[i0 + u] = off;
        off += v[u];
    }
    if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_escan_top(unsigned long long *__restrict__ blk, uint32_t nblk) {
    __shared__ unsigned long long part[1024];
    unsigned long long carry = 0;
    for (uint32_t base = 0; base < nblk; base += 1024) {
        const uint32_t i = base + threadIdx.x;
        const unsigned long long v = i < nblk ? blk[i] : 0ull;
        part[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            const unsigned long long w = threadIdx.x >= o ? part[threadIdx.x - o] : 0ull;
            __syncthreads();
            part[threadIdx.x] += w;
            __syncthreads();
        }
        if (i < nblk) blk[i] = carry + part[threadIdx.x] - v;
        carry += part[1023];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_edge_copy(EdgeCopyArgs a) {
    const uint64_t b = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= a.n) return;
    const uint32_t c = a.cnt[b];
    if (c == 0) return;
    const unsigned long long src = a.src_off[b], dst = a.dst_off[b] + a.blk[b / ESCAN_BLOCK];
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
    if (a.rg.active)
        hipLaunchKernelGGL(k_part<true>, dim3((uint32_t)grid), dim3(PART_THREADS), 0, s, a);
    else
        hipLaunchKernelGGL(k_part<false>, dim3((uint32_t)grid), dim3(PART_THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_dedup(const DedupArgs &a, uint32_t nb, hipStream_t s) {
    hipLaunchKernelGGL((k_dedup<BUCKET_CAP, JOIN_THREADS>), dim3(nb), dim3(JOIN_THREADS), 0, s, a);
    return hipGetLastError();
}

// Work-groups of `kernel` that fit on the current device at once (CUs x resident per CU).
static uint32_t resident_groups(const void *kernel, int threads) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1024;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
    return (uint32_t)(cus * per);
}

hipError_t launch_join(const JoinArgs &a, uint32_t nb, hipStream_t s) {
    constexpr auto kern = k_join<BUCKET_CAP, MATCH_CAP, JOIN_THREADS, JOIN_PERSIST>;
    static uint32_t resident[64];
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) dev = 0;
    if (!resident[dev]) resident[dev] = resident_groups(reinterpret_cast<const void *>(kern), JOIN_THREADS);
    JoinArgs b = a;
    b.nb = nb;
    const uint32_t grid = JOIN_PERSIST ? (nb < resident[dev] ? nb : resident[dev]) : nb;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(JOIN_THREADS), 0, s, b);
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

hipError_t launch_edge_compact(const EdgeCopyArgs &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    const uint64_t nblk = (a.n + ESCAN_BLOCK - 1) / ESCAN_BLOCK;
    hipLaunchKernelGGL(k_escan_local, dim3((uint32_t)nblk), dim3(1024), 0, s, a.cnt, a.n, a.dst_off, a.blk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_escan_top, dim3(1), dim3(1024), 0, s, a.blk, (uint32_t)nblk);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_edge_copy, dim3((uint32_t)((a.n + 3) / 4)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t g = (n + 255) / 256;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)(g < 4096 ? g : 4096)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_audit(const AuditArgs &a, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_audit, dim3(n), dim3(AUDIT_THREADS), 0, s, a);
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
