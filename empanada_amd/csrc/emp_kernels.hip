// HIP kernels of the fingerprint -> partition -> suffix/prefix join path, for gfx950 (wave64).
//
// Data in HBM:
//   reads    row-major ASCII kept reads (input of the map kernel only)
//   columns  2-bit codes per read position, chunked by 2^20 reads (emp_device.h)
//   entries  12-byte (h1, h2, rid) records, in buckets in key order: bucket b of a list of nb
//            buckets (a multiple of 256 for a round's lists) holds the entries with
//            floor(h1 nb / 2^29) == b (bucket_of; h1 >> (29 - bits) for nb = 2^bits), in a
//            fixed-capacity slot array [bucket][cap] plus a count (spilled records past it). A
//            round's entries keep carried codes in h1's bits 21..28 (implied by the bucket:
//            wide_key1) and 29..31 and h2's bits 29..31. Order inside a bucket is arbitrary: every
//            order the output depends on is re-established in LDS.
//
// Kernels:
//   k_map      transposition + forward/reverse-complement Rabin-Karp as byte dot products
//              (fingerprint.cu:75-153)
//   k_part     tile-sorted partition pass (at most 8 more key bits per pass): the dedup's sort of
//              the 2n map entries (overlap.cu:211-212)
//   k_dedup    per bucket: keep the smallest read id of every key (overlap.cu:212-222); the kept
//              entries, regenerated, go straight into round 0's suffix and prefix buckets
//   k_join     per bucket of round t: LDS hash join of suffixes against prefixes, edges of the
//              matched key groups in (key, rid&1, rid) x (rid&1, rid) order, removal of used
//              entries (overlap.cu:258-290, 361-456, 557-710); the survivors, regenerated for
//              round t+1 (overlap.cu:266-267, 605-692), go straight into that round's buckets
//   k_*_big    the same for buckets over capacity: global-memory sort + binary-search join
//   k_edge_*   bucket-ordered compaction of the edges of a round
//   k_audit    collision audit of one list (FPCHECK)
#include "emp_common.h"
#include "emp_kernels.h"
#include "emp_ooc.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

namespace emp {

// ---------------------------------------------------------------------------------------------
// The word of the wide column store holding read k's 8 codes of column group g (2 reads per word;
// G8 groups per chunk).
__device__ __forceinline__ size_t wide_word(uint32_t G8, uint32_t g, uint32_t k) {
    return ((size_t)(k >> CHUNK_SHIFT) * G8 + g) * CHUNK_WWORDS + ((k >> 1) & (CHUNK_WWORDS - 1));
}

// The 8 codes (16 bits, column 8g + c at bits 2c) of read k in column group g.
__device__ __forceinline__ uint32_t wide_codes(const uint32_t *__restrict__ cols, uint32_t G8, uint32_t g, uint32_t k) {
    return (cols[wide_word(G8, g, k)] >> (16 * (k & 1))) & 0xFFFFu;
}

// The carry of a live entry of round r (the codes it holds after that round's regeneration),
// recomputed from its read's unit: used where an entry is rebuilt from its key alone.
__device__ __forceinline__ uint32_t carry_after(const uint32_t *__restrict__ cols, const RoundArgs &r, uint32_t rid,
                                                bool sfx) {
    const bool fwd = sfx != (bool)(rid & 1);
    const uint32_t col = round_col(r, fwd);
    return wide_carry_of(wide_codes(cols, r.G, col >> 3, rid >> 1), col, fwd);
}

// ---------------------------------------------------------------------------------------------
// The next round's lists written by the kernel that finds their entries (NextOut), by slot (see
// NSLOT): count (every thread: gathers, regenerates, finds its entries' slots and ranks them per
// wave), reserve (one thread per slot: one run per target bucket), write.

// The functions below take the source bucket as its smallest key lo (bucket_lo, once per
// work-group).

// First target bucket of the suffixes of the source bucket whose removed base has encoding ev: the
// bucket of its smallest key minus NextSfx's subtrahend e * R1^(l-1) mod P1.
__device__ __forceinline__ uint32_t sfx_first(const NextOut &nx, uint32_t lo, uint32_t ev) {
    const uint32_t c = red1(ev * nx.rg.m1);
    return bucket_of(lo >= c ? lo - c : lo + P1 - c, nx.nb);
}

// First target bucket of the prefixes of the source bucket with (h1 - e) mod 5 == k: PrevPfx gives
// (h1 - e + k P1) / 5, smallest for the smallest h1 - e (h1 >= lo, e <= 4).
__device__ __forceinline__ uint32_t pfx_first(const NextOut &nx, uint32_t lo, uint32_t k) {
    const uint32_t u = lo >= 4u ? lo - 4u : 0u;
    return bucket_of((u + k * P1) / 5u, nx.nb);
}

// d: the target's distance from the first target of base ev, modulo the bucket count (the shifted
// range can wrap past P1 to bucket 0).
__device__ __forceinline__ uint32_t sfx_slot(const NextOut &nx, uint32_t lo, uint32_t ev, uint32_t h1n) {
    const uint32_t t = bucket_of(h1n, nx.nb), f = sfx_first(nx, lo, ev);
    const uint32_t d = t >= f ? t - f : t + nx.nb - f;
    return d < 3u ? 3u * (ev - 1u) + d : SLOT_SFX_OTHER;
}

// h1o: the prefix's key before PrevPfx (the u64 wrap for h1o < e lands anywhere: the fallback
// slot; k and d are formed for it too and not used).
__device__ __forceinline__ uint32_t pfx_slot(const NextOut &nx, uint32_t lo, uint32_t h1o, uint32_t ev, uint32_t h1n) {
    const uint32_t k = (h1o - ev) % 5u;
    const uint32_t d = bucket_of(h1n, nx.nb) - pfx_first(nx, lo, k);
    return (h1o >= ev && d < 2u) ? 12u + 2u * k + d : SLOT_PFX_OTHER;
}

// Target bucket of direct slot s (< 22) of the source bucket (called only for a slot that holds
// entries, so the bucket exists).
__device__ __forceinline__ uint32_t slot_bucket(const NextOut &nx, uint32_t lo, uint32_t s) {
    if (s < 12u) {
        const uint32_t t = sfx_first(nx, lo, s / 3u + 1u) + s % 3u;
        return t >= nx.nb ? t - nx.nb : t;
    }
    return pfx_first(nx, lo, (s - 12u) / 2u) + (s - 12u) % 2u;
}

// Stores entry e of side `side` at position pos of bucket nb (past the capacity: the spill area).
__device__ __forceinline__ void next_put(const NextOut &nx, uint32_t side, uint32_t nb, uint32_t pos, const Entry &e) {
    if (pos < (uint32_t)BUCKET_CAP) {
        nx.fine[side][(size_t)nb * BUCKET_CAP + pos] = e;
    } else {
        const uint32_t sp = atomicAdd(nx.spill_n[side], 1u);
        if (sp < nx.spill_cap[side]) nx.spill[side][sp] = SpillRec{e, nb};
        else atomicOr(nx.ovf, 1u);
    }
}

// Between gather and rank an entry's word is its gathered unit (16 bits; bit 16 set: none). After
// the rank it holds the entry's carry bits 6..13 (bits 0-7, stored in h1 bits 21..28 when written),
// its slot (bits 9-13) and, from bit 14, its rank among the wave's entries of the slot (for a
// fallback slot: its position in the target bucket).
__device__ __forceinline__ uint32_t tgt_word(uint32_t tg) { return (tg & 0x10000u) ? 0xFFFFFFFFu : (tg & 0xFFFFu); }

// Gather step for register-resident entries e[q] (bit q of mask) of side `side`, in two halves.
// The issue half loads, into tgt[q], the wide-store word of each entry that enters a column group
// this round (one round in eight per direction) and returns those entries as a mask; the finish
// half turns the words into tgt form (the read's 16 bits, or bit 16 set: no word). Between the two the caller
// issues the other side's gathers, so that all of them are in flight together. Rounds where no
// lane of the wave needs a word skip the loads; otherwise lanes without one read the store's first
// word (a cached line). The finish half uses every loaded word whether it is needed or not: a
// word used only under its condition would let the compiler move each load into a branch of its
// own and wait for each in turn.
template <int PER>
__device__ __forceinline__ uint32_t next_gather(const NextOut &nx, const Entry (&e)[PER], uint32_t mask, uint32_t (&tgt)[PER],
                                                uint32_t side) {
    const bool sfx = side == 0;
    uint32_t nmask = 0, grp[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t act = (mask >> q) & 1u;  // slots past the bucket hold no loaded entry
        const bool need = wide_needs_gather(nx.rg, act ? e[q].rid : 0u, sfx, grp[q]);
        nmask |= (act & (need ? 1u : 0u)) << q;
        tgt[q] = 0u;
    }
    if (__any(nmask != 0)) {
        const uint32_t G8 = __builtin_amdgcn_readfirstlane(nx.rg.G);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const size_t w = wide_word(G8, grp[q], e[q].rid >> 1);  // formed for every slot, selected
            tgt[q] = nx.cols[(nmask >> q) & 1u ? w : 0];
        }
    }
    return nmask;
}

// The read's 16 bits of a loaded word (need), else the word with bit 16 set (both formed, one
// selected).
__device__ __forceinline__ uint32_t gathered(uint32_t w, uint32_t rid, bool need) {
    const uint32_t u = (w >> (16 * ((rid >> 1) & 1u))) & 0xFFFFu, none = w | 0x10000u;
    return need ? u : none;
}

template <int PER>
__device__ __forceinline__ void next_gather_finish(const Entry (&e)[PER], uint32_t nmask, uint32_t (&tgt)[PER]) {
#pragma unroll
    for (int q = 0; q < PER; ++q) tgt[q] = gathered(tgt[q], e[q].rid, (nmask >> q) & 1u);
}

// An entry's rank in its slot (table slots: LDS counter of the wave), or for the two fallback slots
// its position in the target bucket, taken here with a global atomic (clamped to BUCKET_CAP: any
// position from there on spills). Taking the fallback positions in the rank step keeps atomics
// out of the write step, whose stores would otherwise all wait for them.
__device__ __forceinline__ uint32_t slot_rank(const NextOut &nx, uint32_t *row, uint32_t slot, uint32_t h1n, uint32_t side) {
    if (slot < 22u) return atomicAdd(row + slot, 1u);
    const uint32_t pos = atomicAdd((side ? nx.cnt[1] : nx.cnt[0]) + bucket_of(h1n, nx.nb), 1u);
    return pos < (uint32_t)BUCKET_CAP ? pos : (uint32_t)BUCKET_CAP;
}

// Regenerates entry e of this round in place: in, its stored form (wide: key bits 21..28 restored
// from top8, the top 8 bits of its bucket's index) or, from the dedup (wide false), its full keys
// with no carry; out, the next round's keys in bits 0..28 with carry bits 0..5 above them. Returns
// the removed base's encoding; x and h1o get the carry bits 6..13 and the old h1 key.
__device__ __forceinline__ uint32_t regen_entry(const RoundArgs &r, Entry &e, bool wide, uint32_t top8, bool sfx,
                                                uint32_t gathered_unit, uint32_t &x, uint32_t &h1o) {
    uint32_t h1 = wide ? wide_key1(e.h1, top8) : e.h1 & KEY_MASK, h2 = e.h2 & KEY_MASK;
    uint32_t f = wide ? wide_carry(e) : 0u;
    h1o = h1;
    const uint32_t ev = regen_wide(r, h1, h2, e.rid, f, sfx, gathered_unit);
    e.h1 = h1 | ((f & 7u) << 29);
    e.h2 = h2 | (((f >> 3) & 7u) << 29);
    x = (f >> 6) & 0xFFu;
    return ev;
}

// Rank step (after next_gather): regenerates each entry in place (e[q] becomes its next-round
// entry, regen_entry), finds its slot for the source bucket (smallest key lo) and takes a rank
// among the wave's entries of that slot from the wave's own counters (row: NSLOT of them, zeroed;
// no other wave touches them). ntot: the work-group's per-side totals. tgt[q] becomes the carry's top bits, the slot and
// the rank (tgt_word).
template <int PER>
__device__ __forceinline__ void next_rank(const NextOut &nx, uint32_t lo, uint32_t *row, uint32_t *ntot, Entry (&e)[PER],
                                          uint32_t mask, uint32_t (&tgt)[PER], uint32_t side, bool wide, uint32_t top8) {
    const bool sfx = side == 0;
    uint32_t wn = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const bool act = (mask >> q) & 1u;
        if (act) {
            uint32_t x, h1o;
            const uint32_t ev = regen_entry(nx.rg, e[q], wide, top8, sfx, tgt_word(tgt[q]), x, h1o);
            const uint32_t h1n = e[q].h1 & KEY_MASK;
            const uint32_t slot = sfx ? sfx_slot(nx, lo, ev, h1n) : pfx_slot(nx, lo, h1o, ev, h1n);
            tgt[q] = x | (slot << 9) | (slot_rank(nx, row, slot, h1n, side) << 14);
        }
        wn += (uint32_t)__popcll(__ballot(act));
    }
    if (lane_id() == 0 && wn) atomicAdd(ntot + side, wn);
}

// Stored form of a regenerated entry (keys in bits 0..28): h1's bits 21..28, implied by the
// bucket, hold the carry bits 6..13 from tgt.
__device__ __forceinline__ Entry stored(const Entry &e, uint32_t tg) {
    return Entry{(e.h1 & ~(0xFFu << WIDE_SHIFT)) | ((tg & 0xFFu) << WIDE_SHIFT), e.h2, e.rid};
}

// Write step: `row` now holds this wave's run start per slot; e[q] are the next-round entries.
template <int PER>
__device__ __forceinline__ void next_write(const NextOut &nx, const uint32_t *row, const Entry (&e)[PER], uint32_t mask,
                                           const uint32_t (&tgt)[PER], uint32_t side) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if ((mask >> q) & 1u) {
            const uint32_t tb = bucket_of(e[q].h1 & KEY_MASK, nx.nb), sl = (tgt[q] >> 9) & 31u;
            const uint32_t pos = sl < 22u ? row[sl] + (tgt[q] >> 14) : tgt[q] >> 14;
            next_put(nx, side, tb, pos, stored(e[q], tgt[q]));
        }
    }
}

// Reserve step, thread for slot s of an NW-wave work-group: the slot's entries over all waves,
// one run in its target bucket (returned: the run start), to be turned into per-wave starts by
// slot_starts once the reservation is back.
// SEL: pick the side's count array with a select of the two pointers instead of an index into
// the argument (which the compiler loads per lane); the select holds more scalar registers, which
// pays in the dedup and not in the join.
template <int NW, bool SEL = false>
__device__ __forceinline__ uint32_t slot_reserve(const NextOut &nx, uint32_t lo, const uint32_t (*wcnt)[NSLOT], uint32_t s) {
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) tot += wcnt[w][s];
    if (s >= 22u || tot == 0) return 0u;
    uint32_t *const cnt = SEL ? (s < 12u ? nx.cnt[0] : nx.cnt[1]) : nx.cnt[s < 12u ? 0 : 1];
    return atomicAdd(cnt + slot_bucket(nx, lo, s), tot);
}

template <int NW>
__device__ __forceinline__ void slot_starts(uint32_t (*wcnt)[NSLOT], uint32_t s, uint32_t base) {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const uint32_t c = wcnt[w][s];
        wcnt[w][s] = base;
        base += c;
    }
}

// Single-entry form (large-bucket kernels): keys h1, h2 (bits 0..28) and carry f of an entry of
// this round (side `side`).
__device__ __forceinline__ void put_next(const NextOut &nx, uint32_t h1, uint32_t h2, uint32_t rid, uint32_t f,
                                         uint32_t side) {
    const bool sfx = side == 0;
    uint32_t g, unit = 0xFFFFFFFFu;
    if (wide_needs_gather(nx.rg, rid, sfx, g)) unit = wide_codes(nx.cols, nx.rg.G, g, rid >> 1);
    regen_wide(nx.rg, h1, h2, rid, f, sfx, unit);
    const uint32_t tb = bucket_of(h1, nx.nb);
    const uint32_t pos = atomicAdd(nx.cnt[side] + tb, 1u);
    atomicAdd(nx.total[side], 1u);
    next_put(nx, side, tb, pos, wide_entry(h1, h2, rid, f));
}

// ---------------------------------------------------------------------------------------------
// Map: one work-group of 256 threads per 256 reads.
// Bytes of w outside ACGTacgt (non-zero when any), four bytes at a time: after setting the
// lower-case bit, a byte is a base when it equals one of a, c, g, t; nz(v) has bit 7 of each
// byte set where that byte of v is not zero.
__device__ __forceinline__ uint32_t nz_bytes(uint32_t v) { return (((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u; }
__device__ __forceinline__ uint32_t bad_bases(uint32_t w) {
    const uint32_t x = w | 0x20202020u;
    return nz_bytes(x ^ 0x61616161u) & nz_bytes(x ^ 0x63636363u) & nz_bytes(x ^ 0x67676767u) &
           nz_bytes(x ^ 0x74747474u);
}

// The 2-bit codes of the 4 bases in dword v of a row (column 4g + c at bits 2c), columns past the
// read cleared by keep.
__device__ __forceinline__ uint32_t pack_codes(uint32_t v, uint32_t keep) {
    const uint32_t c = (v >> 1) & 0x03030303u;
    return (c | (c >> 6) | (c >> 12) | (c >> 18)) & keep;
}
__device__ __forceinline__ uint32_t group_keep(uint32_t g, int L) {
    const int rest = L - 4 * (int)g;
    return rest >= 4 ? 0xFFu : (1u << (2 * rest)) - 1u;
}
// The 16-bit unit of 8-column group g of a row of ng dwords (no dword past the row is read).
__device__ __forceinline__ uint32_t wide_unit(const uint32_t *row4, uint32_t g, int ng, int L) {
    const uint32_t lo = pack_codes(row4[2 * g], group_keep(2 * g, L));
    const uint32_t hi = 2 * (int)g + 1 < ng ? pack_codes(row4[2 * g + 1], group_keep(2 * g + 1, L)) : 0u;
    return lo | (hi << 8);
}

constexpr int MAP_Q = 7;  // 16-byte loads per thread in flight (a 256 x 100 bp block is 1,600)

__global__ void __launch_bounds__(256) k_map(MapArgs a) {
    extern __shared__ __align__(16) uint8_t tile[];
    const uint32_t first = blockIdx.x * a.rows;
    const uint32_t nt = min(a.rows, a.n - first);
    const int L = a.L, Lp = a.Lpad;
    const uint8_t *src = a.reads + (size_t)first * L;
    const uint32_t nbytes = nt * (uint32_t)L;
    uint32_t bad = 0;
    // Reads -> LDS tile (row stride Lp). A 16-byte aligned block is read with 16-byte loads, up to
    // MAP_Q per thread in flight; when Lp == L the tile is the block itself.
    uint32_t done = 0;
    if ((reinterpret_cast<uintptr_t>(src) & 15u) == 0) {
        const uint32_t nq = nbytes >> 4;
        for (uint32_t q0 = 0; q0 < nq; q0 += 256 * MAP_Q) {
            // every lane loads (past the block: its last 16 bytes again), so that no branch
            // separates the loads and all of them are in flight together
            uint4 v[MAP_Q];
#pragma unroll
            for (int u = 0; u < MAP_Q; ++u)
                v[u] = reinterpret_cast<const uint4 *>(src)[min(q0 + (uint32_t)u * 256 + threadIdx.x, nq - 1)];
#pragma unroll
            for (int u = 0; u < MAP_Q; ++u) {
                const uint32_t q = q0 + (uint32_t)u * 256 + threadIdx.x;
                if (q < nq) {
                    bad |= bad_bases(v[u].x) | bad_bases(v[u].y) | bad_bases(v[u].z) | bad_bases(v[u].w);
                    const uint32_t g = q * 16;
                    if (Lp == L) {
                        *reinterpret_cast<uint4 *>(tile + g) = v[u];
                    } else {
                        const uint32_t w4[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
                        uint32_t r = g / (uint32_t)L, p = g - r * (uint32_t)L;
#pragma unroll
                        for (int i = 0; i < 16; ++i) {
                            tile[r * Lp + p] = (uint8_t)(w4[i >> 2] >> (8 * (i & 3)));
                            if (++p == (uint32_t)L) { p = 0; ++r; }
                        }
                    }
                }
            }
        }
        done = nq * 16;
    }
    for (uint32_t g = done + threadIdx.x; g < nbytes; g += 256) {
        const uint8_t c = src[g];
        bad |= bad_bases(c | 0x61616100u);
        const uint32_t r = g / (uint32_t)L;
        tile[r * Lp + (g - r * (uint32_t)L)] = c;
    }
    if (bad) atomicOr(a.err, 1u);
    lds_barrier();

    // Fingerprints of the read and of its reverse complement as byte dot products, four bases per
    // instruction: the 2-bit codes of a dword of the row give the encodings e (the read) and e'
    // (their complements, the reverse complement's bases in reverse order); a.wt holds the bytes
    // of the position weights, 16 dwords per 4 positions (map_weights), read as scalars.
    const uint32_t t = threadIdx.x;
    const int ng = (L + 3) >> 2;
    if (t < nt) {
        const uint32_t *row4 = reinterpret_cast<const uint32_t *>(tile + t * Lp);
        const __attribute__((address_space(4))) uint32_t *wt =
            (const __attribute__((address_space(4))) uint32_t *)a.wt;
        uint32_t sd[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) sd[i] = 0;
        for (int d = 0; d < ng; ++d) {
            const uint32_t c = (row4[d] >> 1) & 0x03030303u;  // bytes past L: weight 0
            const uint32_t e = c + 0x01010101u, ec = (c ^ 0x02020202u) + 0x01010101u;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                sd[j] = __builtin_amdgcn_udot4(e, wt[16 * d + j], sd[j], false);
                sd[4 + j] = __builtin_amdgcn_udot4(e, wt[16 * d + 4 + j], sd[4 + j], false);
                sd[8 + j] = __builtin_amdgcn_udot4(ec, wt[16 * d + 8 + j], sd[8 + j], false);
                sd[12 + j] = __builtin_amdgcn_udot4(ec, wt[16 * d + 12 + j], sd[12 + j], false);
            }
        }
        uint64_t k = a.k0 + first + t;
        Entry ef = {fold1(sd[0], sd[1], sd[2], sd[3]), fold2(sd[4], sd[5], sd[6], sd[7]), (uint32_t)(2 * k)};
        Entry er = {fold1(sd[8], sd[9], sd[10], sd[11]), fold2(sd[12], sd[13], sd[14], sd[15]), (uint32_t)(2 * k + 1)};
        a.entries[2 * k] = ef;
        a.entries[2 * k + 1] = er;
        if (a.ends) {
            const uint32_t gl = col_groups8((uint32_t)L) - 1u;
            a.ends[k] = wide_unit(row4, 0, ng, L) | (wide_unit(row4, gl, ng, L) << 16);
        }
    }

    // Columns. A batched job's region (gs != 0): 4-column groups, word w of the tile covers reads
    // 4w..4w+3. The wide store: 8-column groups, word w covers reads 2w, 2w+1. Task id = g * nw + w
    // keeps a wave on consecutive words of one column group (contiguous stores). No store (cols
    // null): fingerprints only.
    if (!a.cols) {
        return;
    } else if (a.gs) {
        const uint32_t G = col_groups((uint32_t)L);
        const uint32_t nw = (nt + 3) / 4;
        const uint32_t ntask = nw * G;
        const uint64_t w0 = (a.k0 + first) >> 2;
        for (uint32_t task = threadIdx.x; task < ntask; task += 256) {
            const uint32_t g = task / nw, w = task - g * nw;
            uint32_t word = 0;
            const uint32_t rmax = min(4u, nt - 4 * w), keep = group_keep(g, L);
            for (uint32_t i = 0; i < rmax; ++i)
                word |= pack_codes(reinterpret_cast<const uint32_t *>(tile + (4 * w + i) * Lp)[g], keep) << (8 * i);
            a.cols[a.cbase + (uint64_t)g * a.gs + w0 + w] = word;
        }
    } else {
        const uint32_t G8 = col_groups8((uint32_t)L);
        const uint32_t nw = (nt + 1) / 2;
        const uint32_t ntask = nw * G8;
        const uint64_t w0 = (a.k0 + first) >> 1;
        for (uint32_t task = threadIdx.x; task < ntask; task += 256) {
            const uint32_t g = task / nw, w = task - g * nw;
            uint32_t word = 0;
            const uint32_t rmax = min(2u, nt - 2 * w);
            for (uint32_t i = 0; i < rmax; ++i)
                word |= wide_unit(reinterpret_cast<const uint32_t *>(tile + (2 * w + i) * Lp), g, ng, L) << (16 * i);
            const uint64_t W = w0 + w;
            a.cols[((W >> (CHUNK_SHIFT - 1)) * G8 + g) * CHUNK_WWORDS + (W & (CHUNK_WWORDS - 1))] = word;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Partition pass (the dedup's sort, overlap.cu:211-212): each work-group takes tiles of PART_TILE
// entries of one input group, ranks them by their next `dbits` key bits in LDS, reserves one output
// range per (tile, bucket) with a single atomic, and writes each bucket's run contiguously.
__global__ void __launch_bounds__(PART_THREADS) __attribute__((amdgpu_waves_per_eu(EMP_PART_WAVES))) k_part(PartArgs a) {
    constexpr int NT = PART_THREADS, T = PART_TILE, PER = T / NT, NB = 1 << PART_MAX_BITS;
    static_assert(NT >= NB && T % NT == 0, "one bin per thread");
    __shared__ Entry buf[T];
    __shared__ uint32_t hist[NB], loc[NB], base[NB], sh[NT / 64];
    const uint32_t g = blockIdx.x / a.tiles_x, tile0 = blockIdx.x % a.tiles_x;
    const uint32_t F = a.fin_nb ? a.fin_nb >> 8 : 1u << a.dbits;
    uint32_t n_g;
    const Entry *in;
    if (a.in_cnt) {
        n_g = min(a.in_cnt[g], a.in_cap);
        in = a.in + (size_t)g * a.in_cap;
    } else {
        n_g = *a.in_n;
        in = a.in;
    }
    const uint32_t gb = a.keyed ? g * F : 0u;
    // an entry's bin: its output bucket less the group's first
    auto digit = [&](uint32_t h1) { return a.fin_nb ? bucket_of(h1 & KEY_MASK, a.fin_nb) - gb : (h1 >> a.out_shift) & (F - 1); };
    for (uint32_t tile = tile0; (size_t)tile * T < n_g; tile += a.tiles_x) {
        const uint32_t t0 = tile * T, nt = min((uint32_t)T, n_g - t0);
        if (threadIdx.x < NB) hist[threadIdx.x] = 0;
        lds_barrier();
        // every lane loads (lanes past the tile re-read its last entry), so the loads of all PER
        // entries are in flight together
        Entry e[PER];
        uint32_t r[PER], d[PER];
        const uint32_t last = nt - 1;
#pragma unroll
        for (int q = 0; q < PER; ++q) e[q] = in[t0 + min(threadIdx.x + NT * q, last)];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            uint32_t i = threadIdx.x + NT * q;
            if (i < nt) {
                d[q] = digit(e[q].h1);
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
        uint32_t at[PER];  // each entry's position in the tile's bin order (the payload's too)
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            uint32_t i = threadIdx.x + NT * q;
            at[q] = i < nt ? loc[d[q]] + r[q] : 0u;
            if (i < nt) buf[at[q]] = e[q];
        }
        lds_barrier();
        static_assert(PART_MAX_BITS <= 8 && PER % 4 == 0, "bin digits packed four to a word");
        uint32_t dig[PER / 4];  // the bin of each position this thread writes (read again for the payload)
#pragma unroll
        for (int k = 0; k < PER / 4; ++k) dig[k] = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t p = threadIdx.x + NT * k;
            if (p < nt) {
                const Entry v = buf[p];
                const uint32_t dd = digit(v.h1);
                dig[k / 4] |= dd << (8 * (k % 4));
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
        }
        lds_barrier();
        if (a.pay_in) {
            // the payloads take the same path through the tile's LDS, loaded now that the entries'
            // registers are free; each thread writes the payloads of the positions it wrote
            uint32_t pv[PER];
            const uint32_t *pin = a.in_cnt ? a.pay_in + (size_t)g * a.in_cap : a.pay_in;
            const uint32_t psh = a.in_cnt ? 0u : 1u;
#pragma unroll
            for (int q = 0; q < PER; ++q) pv[q] = pin[(t0 + min(threadIdx.x + NT * q, last)) >> psh];
            uint32_t *pb = reinterpret_cast<uint32_t *>(buf);
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const uint32_t i = threadIdx.x + NT * q;
                if (i < nt) pb[at[q]] = pv[q];
            }
            lds_barrier();
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const uint32_t p = threadIdx.x + NT * k;
                if (p < nt) {
                    const uint32_t dd = (dig[k / 4] >> (8 * (k % 4))) & 0xFFu;
                    const uint32_t slot = base[dd] + (p - loc[dd]);
                    if (slot < a.out_cap) a.pay_out[(size_t)(gb + dd) * a.out_cap + slot] = pb[p];
                }
            }
            lds_barrier();
        }
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
    __shared__ uint32_t wcnt[NT / 64][NSLOT];
    __shared__ uint32_t ntot[2];
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
    if (threadIdx.x < 2) ntot[threadIdx.x] = 0;
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
    uint32_t keepmask = 0;
    Entry kept[PER];
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
                kept[q] = Entry{h1, h2, rid};
            }
        }
    }
    // the kept entries, regenerated, are round 0's suffix list and prefix list (overlap.cu:235-242)
    if (a.nx.active) {
        uint32_t *row = wcnt[threadIdx.x >> 6];
        if ((threadIdx.x & 63) < NSLOT) row[threadIdx.x & 63] = 0;
        uint32_t ts[PER], tp[PER];
        if (a.pay && a.nx.rg.col_t == 0) {
            // round 0 gathers column group 0 or the last one (suffix and prefix in turn): both
            // are in the read's ends word, which the partition carried to the entry's slot
            uint32_t ew[PER];
#pragma unroll
            for (int q = 0; q < PER; ++q) ew[q] = a.pay[(size_t)b * CAP + min(threadIdx.x + NT * q, n - 1)];
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                ts[q] = tp[q] = 0x10000u;
                if ((keepmask >> q) & 1u) {
                    uint32_t g;
                    if (wide_needs_gather(a.nx.rg, kept[q].rid, true, g)) ts[q] = (g ? ew[q] >> 16 : ew[q]) & 0xFFFFu;
                    if (wide_needs_gather(a.nx.rg, kept[q].rid, false, g)) tp[q] = (g ? ew[q] >> 16 : ew[q]) & 0xFFFFu;
                }
            }
        } else {
            const uint32_t gs = next_gather<PER>(a.nx, kept, keepmask, ts, 0);
            const uint32_t gp = next_gather<PER>(a.nx, kept, keepmask, tp, 1);
            next_gather_finish<PER>(kept, gs, ts);
            next_gather_finish<PER>(kept, gp, tp);
        }
        Entry kp[PER];  // the prefix side's copy (each side is regenerated in place)
#pragma unroll
        for (int q = 0; q < PER; ++q) kp[q] = kept[q];
        const uint32_t lo = __builtin_amdgcn_readfirstlane(bucket_lo(b, a.nx.src_nb));
        next_rank<PER>(a.nx, lo, row, ntot, kept, keepmask, ts, 0, false, 0u);
        next_rank<PER>(a.nx, lo, row, ntot, kp, keepmask, tp, 1, false, 0u);
        lds_barrier();
        if (threadIdx.x < NSLOT) {
            slot_starts<NT / 64>(wcnt, threadIdx.x, slot_reserve<NT / 64, true>(a.nx, lo, wcnt, threadIdx.x));
        } else if (threadIdx.x < NSLOT + 2) {
            const uint32_t sd = threadIdx.x - NSLOT;
            if (ntot[sd]) atomicAdd(sd ? a.nx.total[1] : a.nx.total[0], ntot[sd]);
        }
        lds_barrier();
        next_write<PER>(a.nx, row, kept, keepmask, ts, 0);
        next_write<PER>(a.nx, row, kp, keepmask, tp, 1);
    } else if (keepmask) {  // no round follows: the kept entries are only counted
        atomicAdd(a.nx.total[0], (uint32_t)__popc(keepmask));
        atomicAdd(a.nx.total[1], (uint32_t)__popc(keepmask));
    }
}

// ---------------------------------------------------------------------------------------------
// Join of round t for one bucket.
//
// LDS: the prefix entries (12 B each), a used flag per prefix, and one region that holds the hash
// table (chain heads + links) while probing and the matched records afterwards. The suffix
// entries stay in registers.

// Loads the suffix and prefix entries of bucket b (nothing past the first PRE * NT slots for a
// bucket over capacity). The first PRE * NT slots of each side are loaded whatever the counts, so
// that these loads do not wait for the counts' (buckets are sized for about 1,800 entries per
// side; slots past a count hold stale entries that nothing reads).
template <int CAP, int NT>
__device__ __forceinline__ void load_bucket(const JoinArgs &a, uint32_t b, uint32_t ns, uint32_t np, Entry (&se)[CAP / NT],
                                            Entry (&pe)[CAP / NT]) {
    constexpr int PRE = CAP / NT >= 2 ? 2 : 0;
    const Entry *sin = a.s_in + (size_t)b * CAP;
    const Entry *pin = a.p_in + (size_t)b * CAP;
    const bool fit = ns <= (uint32_t)CAP && np <= (uint32_t)CAP;
#pragma unroll
    for (int q = 0; q < CAP / NT; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (q < PRE) {
            se[q] = sin[i];
            pe[q] = pin[i];
        } else {
            if (fit && i < ns) se[q] = sin[i];
            if (fit && i < np) pe[q] = pin[i];
        }
    }
}

// Prefix survivors of the join (in LDS, slot threadIdx.x + NT * q for bit q of pmask): the gather
// and rank steps of next_gather / next_rank with the entries read from LDS.
template <int CAP, int MCAP, int NT>
__device__ __forceinline__ uint32_t pfx_gather(const NextOut &nx, const JoinShared<CAP, MCAP, NT> &S, uint32_t pmask,
                                               uint32_t (&tgt)[CAP / NT]) {
    constexpr int PER = CAP / NT;
    uint32_t nmask = 0, grp[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const bool need = wide_needs_gather(nx.rg, S.prid[threadIdx.x + NT * q], false, grp[q]);
        nmask |= (((pmask >> q) & 1u) & (need ? 1u : 0u)) << q;
        tgt[q] = 0u;
    }
    if (__any(nmask != 0)) {
        const uint32_t G8 = __builtin_amdgcn_readfirstlane(nx.rg.G);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const size_t w = wide_word(G8, grp[q], S.prid[threadIdx.x + NT * q] >> 1);
            tgt[q] = nx.cols[(nmask >> q) & 1u ? w : 0];
        }
    }
    return nmask;
}

template <int CAP, int MCAP, int NT>
__device__ __forceinline__ void pfx_gather_finish(const JoinShared<CAP, MCAP, NT> &S, uint32_t nmask, uint32_t (&tgt)[CAP / NT]) {
#pragma unroll
    for (int q = 0; q < CAP / NT; ++q) tgt[q] = gathered(tgt[q], S.prid[threadIdx.x + NT * q], (nmask >> q) & 1u);
}

template <int CAP, int MCAP, int NT>
__device__ __forceinline__ void pfx_rank(const NextOut &nx, JoinShared<CAP, MCAP, NT> &S, uint32_t lo, uint32_t *row,
                                         uint32_t pmask, uint32_t (&tgt)[CAP / NT], uint32_t top8) {
    constexpr int PER = CAP / NT;
    uint32_t wn = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const bool act = (pmask >> q) & 1u;
        if (act) {
            // the regenerated entry replaces the old key in LDS (the matched records hold copies
            // of the keys the edges need)
            const uint32_t i = threadIdx.x + NT * q;
            Entry e{S.pk1[i], S.pk2[i], S.prid[i]};
            uint32_t x, h1o;
            const uint32_t ev = regen_entry(nx.rg, e, true, top8, false, tgt_word(tgt[q]), x, h1o);
            S.pk1[i] = e.h1;
            S.pk2[i] = e.h2;
            const uint32_t slot = pfx_slot(nx, lo, h1o, ev, e.h1 & KEY_MASK);
            tgt[q] = x | (slot << 9) | (slot_rank(nx, row, slot, e.h1 & KEY_MASK, 1) << 14);
        }
        wn += (uint32_t)__popcll(__ballot(act));
    }
    if (lane_id() == 0 && wn) atomicAdd(&S.ntot[1], wn);
}

template <int CAP, int MCAP, int NT>
__device__ __forceinline__ void pfx_write(const NextOut &nx, const JoinShared<CAP, MCAP, NT> &S, const uint32_t *row,
                                          uint32_t pmask, const uint32_t (&tgt)[CAP / NT]) {
    constexpr int PER = CAP / NT;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if ((pmask >> q) & 1u) {
            const uint32_t i = threadIdx.x + NT * q;
            const Entry e{S.pk1[i], S.pk2[i], S.prid[i]};
            const uint32_t tb = bucket_of(e.h1 & KEY_MASK, nx.nb), sl = (tgt[q] >> 9) & 31u;
            const uint32_t pos = sl < 22u ? row[sl] + (tgt[q] >> 14) : tgt[q] >> 14;
            next_put(nx, 1, tb, pos, stored(e, tgt[q]));
        }
    }
}

// Join of one bucket: suffixes in registers se[] (slot threadIdx.x + NT * q), prefixes already
// in S.pk1/pk2/prid with S.used cleared. Buckets over capacity, or with more matched records
// than MCAP, are handed to the large-bucket path.
template <int CAP, int MCAP, int NT>
__device__ __forceinline__ void join_bucket(const JoinArgs &a, JoinShared<CAP, MCAP, NT> &S, uint32_t b, uint32_t ns,
                                            uint32_t np, Entry (&se)[CAP / NT]) {
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
    if (threadIdx.x == 0) { S.ms = 0; S.mp = 0; S.ovf = 0; S.etot = 0; S.ntot[0] = 0; S.ntot[1] = 0; }
    for (uint32_t i = threadIdx.x; i < CAP; i += NT) S.u.h.head[i] = EMPTY;
    // the bucket's smallest key and the top 8 key bits all its keys share
    const uint32_t lo = __builtin_amdgcn_readfirstlane(bucket_lo(b, a.nx.src_nb)), top8 = lo >> WIDE_SHIFT;
    uint32_t sc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) sc[q] = 0;
    lds_barrier();
    // Every thread has read the bucket's counts (before this barrier) and nothing else reads
    // them: with zero_in they are cleared for the launch that next writes this bucket set.
    if (a.zero_in && threadIdx.x == 0) { a.s_cnt[b] = 0; a.p_cnt[b] = 0; }
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
                // keys of one bucket differ only in h1's low bits (its bits 21..28 carry codes)
                const uint32_t h1 = se[q].h1 & LOW_MASK, h2 = se[q].h2 & KEY_MASK, rd = se[q].rid >> 1;
                uint32_t c = 0;
                for (uint32_t j = S.u.h.head[h2 & (CAP - 1)]; j != EMPTY; j = (S.u.h.nxt[j] == 0xFFFFu ? EMPTY : S.u.h.nxt[j])) {
                    if ((S.pk2[j] & KEY_MASK) == h2 && (S.pk1[j] & LOW_MASK) == h1 && (S.prid[j] >> 1) != rd) {
                        ++c;
                        S.used[j] = 1;
                    }
                }
                sc[q] = c;
            }
        }
    }
    // The bucket's edge total. It bounds both sides' matched-record counts (every matched record
    // has an edge), so a bucket whose total is at most MCAP cannot overflow the records.
    {
        uint32_t ecount = 0;
#pragma unroll
        for (int q = 0; q < PER; ++q) ecount += sc[q];
        if (ecount) atomicAdd(&S.etot, ecount);
    }
    lds_barrier();  // the hash table is dead from here on; its space holds the matched records
    if (threadIdx.x < 128) S.u.m.hist[threadIdx.x >> 6][threadIdx.x & 63] = 0;
    const bool may_overflow = __builtin_amdgcn_readfirstlane(S.etot) > (uint32_t)MCAP;

    // Matched records (key, rank) of both sides and the survivors (suffixes with no edge, prefixes
    // never a target: overlap.cu:424-429,694-710).
    uint32_t smask = 0, pmask = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (i < ns) {
            if (sc[q]) {
                uint32_t m = atomicAdd(&S.ms, 1u);
                if (m < (uint32_t)MCAP) {
                    S.u.m.msk[m] = ((uint64_t)(se[q].h1 & LOW_MASK) << 32) | (se[q].h2 & KEY_MASK);
                    S.u.m.msr[m] = rid_rank(se[q].rid);
                    S.u.m.msc[m] = sc[q];
                } else {
                    S.ovf = 1;
                }
            } else {
                smask |= 1u << q;
            }
        }
        if (i < np) {
            const uint32_t rid = S.prid[i];
            if (S.used[i]) {
                uint32_t m = atomicAdd(&S.mp, 1u);
                if (m < (uint32_t)MCAP) {
                    S.u.m.mpk[m] = ((uint64_t)(S.pk1[i] & LOW_MASK) << 32) | (S.pk2[i] & KEY_MASK);
                    S.u.m.mpr[m] = rid_rank(rid);
                } else {
                    S.ovf = 1;
                }
            } else {
                pmask |= 1u << q;
            }
        }
    }
    // A bucket with more matched records than MCAP goes to the large-bucket path, which writes
    // its survivors itself: decided before the rank step, whose fallback-slot positions are
    // global reservations in the next round's buckets that this launch would then never fill.
    if (may_overflow) {
        lds_barrier();
        if (S.ovf) {
            if (threadIdx.x == 0) {
                uint32_t i = atomicAdd(a.n_big, 1u);
                if (i < a.big_cap) a.big[i] = b;
                a.edge_cnt[b] = 0;
            }
            return;
        }
    }
    // Survivors: their next-round slots and ranks (the entries are rebuilt when written, below).
    // The wave's slot counters live past the matched records (the hash table is dead).
    uint32_t tgt[PER], ptg[PER];
    uint32_t *row = S.u.m.wcnt[threadIdx.x >> 6];
    if (a.nx.active) {
        if ((threadIdx.x & 63) < NSLOT) row[threadIdx.x & 63] = 0;
        const uint32_t gs = next_gather<PER>(a.nx, se, smask, tgt, 0);  // both sides' gathers in flight
        const uint32_t gp = pfx_gather<CAP, MCAP, NT>(a.nx, S, pmask, ptg);
        next_gather_finish<PER>(se, gs, tgt);
        pfx_gather_finish<CAP, MCAP, NT>(S, gp, ptg);
        next_rank<PER>(a.nx, lo, row, S.ntot, se, smask, tgt, 0, true, top8);
        pfx_rank<CAP, MCAP, NT>(a.nx, S, lo, row, pmask, ptg, top8);
    } else {  // last round: only counted
        if (smask) atomicAdd(&S.ntot[0], (uint32_t)__popc(smask));
        if (pmask) atomicAdd(&S.ntot[1], (uint32_t)__popc(pmask));
    }
    lds_barrier();
    const uint32_t nms = S.ms, nmp = S.mp, tot = S.etot;

    // Reservations, issued now and read back after the sort so that their latency overlaps it:
    // thread 16 reserves the edges, threads 64 + k the next-round run of table slot k, and threads
    // 128 and 129 (another wave: nothing in it waits for the returns) add the side totals.
    // The edge reservation's address is made lane-dependent (lane 16: offset 0) so that it stays
    // one lane's atomic; the compiler's wave-wide form of an atomic at a uniform address reads
    // the result back at once.
    constexpr uint32_t NXT0 = 64, TOT0 = 128;
    static_assert(NT >= TOT0 + 2, "reservation threads");
    unsigned long long resv = 0;
    if (threadIdx.x == 16 && tot) {
        resv = atomicAdd(a.pool_cursor + (b & (SUBS - 1)) + (lane_id() - 16u), (unsigned long long)tot);
    } else if (threadIdx.x >= NXT0 && threadIdx.x < NXT0 + NSLOT && a.nx.active) {
        resv = slot_reserve<NT / 64>(a.nx, lo, S.u.m.wcnt, threadIdx.x - NXT0);
    }
    if (threadIdx.x >= TOT0 && threadIdx.x < TOT0 + 2) {
        const uint32_t sd = threadIdx.x - TOT0;
        if (S.ntot[sd]) atomicAdd(a.nx.total[sd], S.ntot[sd]);
    }

    // Edges of the bucket: order matched suffixes and prefixes by (key, rank), then emit the
    // cross product of each key group skipping same-read pairs (overlap.cu:557-601).
    if (nms) {
        // one record per thread: suffix t, or prefix t - nms (both sides at once when they fit)
        const bool both = nms + nmp <= (uint32_t)NT;
        uint64_t k1 = 0, k2 = 0;
        uint32_t r1 = 0, r2 = 0, c1 = 0, p1 = 0, p2 = 0;
        const uint32_t t = threadIdx.x, j2 = both ? t - nms : t;
        const bool has1 = t < nms, has2 = both ? (t >= nms && t - nms < nmp) : t < nmp;
        if (has1) { k1 = S.u.m.msk[t]; r1 = S.u.m.msr[t]; c1 = S.u.m.msc[t]; }
        if (has2) { k2 = S.u.m.mpk[j2]; r2 = S.u.m.mpr[j2]; }
        if (nms + nmp > 48u) {
            // Binned: 64 bins per side, the bucket's key range cut in 64 (the bin of h1 in the set
            // of 64 times as many buckets, less 64 b: the bins are in key order), counted, scanned
            // by wave 0, record indices scattered into bin order (in the dead used flags), then
            // each record ranked only against its own bin.
            static_assert(4 * MCAP <= CAP, "record indices in the used flags");
            const uint32_t nb64 = a.nx.src_nb << 6, hi8 = top8 << WIDE_SHIFT, b64 = b << 6;
            uint32_t b1 = 0, b2 = 0, l1 = 0, l2 = 0;
            if (has1) {
                b1 = bucket_of((uint32_t)(k1 >> 32) | hi8, nb64) - b64;
                l1 = atomicAdd(&S.u.m.hist[0][b1], 1u);
            }
            if (has2) {
                b2 = bucket_of((uint32_t)(k2 >> 32) | hi8, nb64) - b64;
                l2 = atomicAdd(&S.u.m.hist[1][b2], 1u);
            }
            lds_barrier();
            if (t < 64) {
                const uint32_t x0 = S.u.m.hist[0][t], x1 = S.u.m.hist[1][t];
                uint32_t s0 = x0, s1 = x1;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const uint32_t y0 = __shfl_up(s0, o, 64), y1 = __shfl_up(s1, o, 64);
                    if ((int)t >= o) { s0 += y0; s1 += y1; }
                }
                S.u.m.hist[0][t] = s0 - x0;
                S.u.m.hist[1][t] = s1 - x1;
            }
            lds_barrier();
            uint16_t *const idx = reinterpret_cast<uint16_t *>(S.used);  // suffixes, then prefixes from MCAP
            uint32_t st1 = 0, st2 = 0;
            if (has1) { st1 = S.u.m.hist[0][b1]; idx[st1 + l1] = (uint16_t)t; }
            if (has2) { st2 = S.u.m.hist[1][b2]; idx[MCAP + st2 + l2] = (uint16_t)j2; }
            lds_barrier();
            if (has1) {
                const uint32_t e = b1 == 63u ? nms : S.u.m.hist[0][b1 + 1];
                uint32_t lt = 0;
                for (uint32_t q = st1; q < e; ++q) {
                    const uint32_t o = idx[q];
                    const uint64_t kk = S.u.m.msk[o];
                    lt += (kk < k1 || (kk == k1 && S.u.m.msr[o] < r1)) ? 1u : 0u;
                }
                p1 = st1 + lt;
            }
            if (has2) {
                const uint32_t e = b2 == 63u ? nmp : S.u.m.hist[1][b2 + 1];
                uint32_t lt = 0;
                for (uint32_t q = st2; q < e; ++q) {
                    const uint32_t o = idx[MCAP + q];
                    const uint64_t kk = S.u.m.mpk[o];
                    lt += (kk < k2 || (kk == k2 && S.u.m.mpr[o] < r2)) ? 1u : 0u;
                }
                p2 = st2 + lt;
            }
        } else {
            if (has1) p1 = rank_of(S.u.m.msk, S.u.m.msr, nms, t, k1, r1);
            if (has2) p2 = rank_of(S.u.m.mpk, S.u.m.mpr, nmp, j2, k2, r2);
        }
        lds_barrier();
        if (has1) { S.u.m.msk[p1] = k1; S.u.m.msr[p1] = r1; S.u.m.msc[p1] = c1; }
        if (has2) { S.u.m.mpk[p2] = k2; S.u.m.mpr[p2] = r2; }
    }
    if (threadIdx.x == 16) {
        const uint32_t sub = b & (SUBS - 1);
        const unsigned long long base =
            tot ? place_edges(resv, tot, sub, a.pool_cursor + sub, a.pool_cap, a.ovf_cursor, a.ovf_cap) : ~0ull;
        S.base = base;
        a.edge_off[b] = base;
        // edges that found no room are not written: the slot counts none, so that nothing copies
        // from the pool for it (the host sees the overflow cursor past the end and reruns)
        a.edge_cnt[b] = base == ~0ull ? 0u : tot;
    } else if (threadIdx.x >= NXT0 && threadIdx.x < NXT0 + NSLOT && a.nx.active) {
        slot_starts<NT / 64>(S.u.m.wcnt, threadIdx.x - NXT0, (uint32_t)resv);
    }
    lds_barrier();
    // Edge offsets (an exclusive scan of the partner counts, two records per thread) and the
    // edges. Up to 128 matched suffixes (nearly every bucket) wave 0 does both alone: a wave scan,
    // its own LDS writes read back in order, no work-group barrier; the other waves go on to the
    // survivors.
    const bool one_wave = nms <= 128u;
    if (nms && (!one_wave || threadIdx.x < 64)) {
        const uint32_t i0 = 2 * threadIdx.x, i1 = i0 + 1;
        const uint32_t e0 = i0 < nms ? S.u.m.msc[i0] : 0, e1 = i1 < nms ? S.u.m.msc[i1] : 0;
        uint32_t off;
        if (one_wave) {
            uint32_t x = e0 + e1;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, 64);
                if ((int)lane_id() >= o) x += y;
            }
            off = x - (e0 + e1);
        } else {
            uint32_t stot;
            off = block_excl_scan<NT>(e0 + e1, S.sh, stot);
        }
        if (i0 < nms) S.u.m.msc[i0] = off;
        if (i1 < nms) S.u.m.msc[i1] = off + e0;
        if (one_wave)
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront", "local");
        else
            lds_barrier();
        const unsigned long long base = S.base;
        const uint32_t nthr = one_wave ? 64u : (uint32_t)NT;
        if (base != ~0ull) {
            for (uint32_t i = threadIdx.x; i < nms; i += nthr) {
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
                        a.pool[pos] = make_uint2(srid, prd);
                        ++pos;
                    }
                }
            }
        }
    }

    // Survivors into the next round's buckets.
    if (a.nx.active) {
        pfx_write<CAP, MCAP, NT>(a.nx, S, row, pmask, ptg);
        next_write<PER>(a.nx, row, se, smask, tgt, 0);
    }
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

    const uint32_t G = gridDim.x, nb = a.nb;
    uint32_t b = blockIdx.x;
    if (b >= nb) return;
    if (!PERSIST) {
        // the abort flag is read beside the counts and the first loads, not before them
        const uint32_t abort_set = *a.abort;
        const uint32_t ns = a.s_cnt[b], np = a.p_cnt[b];
        Entry se[PER], pe[PER];
        load_bucket<CAP, NT>(a, b, ns, np, se, pe);
        if (abort_set) return;
        stage_prefixes<CAP, MCAP, NT>(S, np, pe);
        join_bucket<CAP, MCAP, NT>(a, S, b, ns, np, se);
        return;
    }
    if (*a.abort) return;
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
    __shared__ unsigned long long sh64[16];
    __shared__ uint32_t s_pos;
    __shared__ unsigned long long s_base;
    const uint32_t b = a.ids[blockIdx.x];
    const uint32_t ns = a.s_cnt[b], np = a.p_cnt[b];
    const uint32_t ns2 = pow2_at_least(ns), np2 = pow2_at_least(np);
    uint64_t *SK = a.skey + a.soff[blockIdx.x], *PK = a.pkey + a.poff[blockIdx.x];
    uint32_t *SR = a.srank + a.soff[blockIdx.x], *PR = a.prank + a.poff[blockIdx.x];
    uint32_t *SA = a.saux + a.soff[blockIdx.x], *PA = a.paux + a.poff[blockIdx.x];
    // round buckets hold wide entries: their keys are restored from the bucket index
    const int32_t top8 = (int32_t)(bucket_lo(b, a.nx.src_nb) >> WIDE_SHIFT);
    gather_bucket(a.s_in, ns, a.cap, a.s_spill, a.s_spill_n, a.spill_cap, b, SK, SR, ns2, true, &s_pos, top8);
    gather_bucket(a.p_in, np, a.cap, a.p_spill, a.p_spill_n, a.p_spill_cap, b, PK, PR, np2, true, &s_pos, top8);
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
    // exclusive scan of SA in place: contiguous chunk per thread, in 64 bits; a bucket of 2^31
    // edges or more (the reference's int edge offsets, overlap.cu:561,567) writes none and sets
    // the run's error flag
    const uint32_t per = (ns + blockDim.x - 1) / blockDim.x;
    const uint32_t lo = min(ns, threadIdx.x * per), hi = min(ns, lo + per);
    unsigned long long run = 0;
    for (uint32_t i = lo; i < hi; ++i) run += SA[i];
    unsigned long long tot64;
    unsigned long long off64 = block_excl_scan64<1024>(run, sh64, tot64);
    const bool too_many = tot64 >= (1ull << 31);
    const uint32_t tot = too_many ? 0u : (uint32_t)tot64;
    if (!too_many) {
        uint32_t off = (uint32_t)off64;
        for (uint32_t i = lo; i < hi; ++i) {
            uint32_t c = SA[i];
            SA[i] = off;
            off += c;
        }
    }
    const uint32_t sub = b & (SUBS - 1);
    if (threadIdx.x == 0) {
        unsigned long long base = ~0ull;
        if (too_many) atomicOr(a.too_many, 1u);
        if (tot) {
            const unsigned long long c = atomicAdd(a.pool_cursor + sub, (unsigned long long)tot);
            base = place_edges(c, tot, sub, a.pool_cursor + sub, a.pool_cap, a.ovf_cursor, a.ovf_cap);
        }
        s_base = base;
        a.edge_off[b] = base;
        a.edge_cnt[b] = base == ~0ull ? 0u : tot;  // unplaced edges are not written (see k_join)
    }
    __syncthreads();
    if (too_many) return;
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
                    a.pool[pos] = make_uint2(srid, prd);
                    ++pos;
                }
            }
        }
        if (e == o) {  // a survivor
            if (a.nx.active) {
                put_next(a.nx, (uint32_t)(SK[i] >> 32), (uint32_t)SK[i], srid, carry_after(a.cols, a.rg, srid, true), 0);
            } else {
                atomicAdd(a.nx.total[0], 1u);
            }
        }
    }
    for (uint32_t j = threadIdx.x; j < np; j += blockDim.x) {
        if (!PA[j]) {
            const uint32_t prd = (PR[j] << 1) | (PR[j] >> 31);
            if (a.nx.active) {
                put_next(a.nx, (uint32_t)(PK[j] >> 32), (uint32_t)PK[j], prd, carry_after(a.cols, a.rg, prd, false), 1);
            } else {
                atomicAdd(a.nx.total[1], 1u);
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
        if (i == 0 || K[i] != K[i - 1]) {  // the smallest id of its key: round 0's suffix and prefix
            if (a.nx.active) {  // the dedup's buckets hold full keys and no carry
                put_next(a.nx, (uint32_t)(K[i] >> 32), (uint32_t)K[i], R[i], 0u, 0);
                put_next(a.nx, (uint32_t)(K[i] >> 32), (uint32_t)K[i], R[i], 0u, 1);
            } else {
                atomicAdd(a.nx.total[0], 1u);
                atomicAdd(a.nx.total[1], 1u);
            }
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
    const uint32_t v = (wide_codes(a.cols, a.G, (uint32_t)c >> 3, rid >> 1) >> (2 * (c & 7))) & 3u;
    return (rid & 1) ? v ^ 2u : v;
}

__device__ bool same_string(const AuditArgs &a, uint32_t r1, uint32_t r2) {
    if (((r1 ^ r2) & 1u) == 0) {
        // one orientation: the same columns of both reads ([start, start+len) forward,
        // [L-start-len, L-start) reverse-complemented), eight at a time
        const int c0 = (r1 & 1) ? a.L - a.start - a.len : a.start, c1 = c0 + a.len;
        for (int g = c0 >> 3; g <= (c1 - 1) >> 3; ++g) {
            const int lo = max(c0 - 8 * g, 0), hi = min(c1 - 8 * g, 8);
            const uint32_t m = ((1u << (2 * hi)) - 1u) & ~((1u << (2 * lo)) - 1u);
            if ((wide_codes(a.cols, a.G, (uint32_t)g, r1 >> 1) ^ wide_codes(a.cols, a.G, (uint32_t)g, r2 >> 1)) & m)
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
    gather_bucket(a.fine, n, a.cap, a.spill, a.spill_n, a.spill_cap, b, K, R, n2, false, &s_pos,
                  a.wshift >= 0 ? (int32_t)(b >> (WIDE_SHIFT - a.wshift)) : -1);
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

// Kept reads packed on the host (2-bit codes, four to a byte, base i at bits 2 (i % 4) of byte
// i / 4) back to one byte per base, the code's base in capitals (A C T G for codes 0-3: the map
// kernel reads (c >> 1) & 3 from them). A thread expands one packed dword into 16 bases.
__global__ void __launch_bounds__(256) k_unpack(const uint32_t *__restrict__ packed, uint8_t *__restrict__ out, uint64_t n) {
    const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    if (16 * i >= n) return;
    const uint32_t w = packed[i];
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t b = w >> (8 * k);  // four codes of this byte
        // 'A' 0x41, 'C' 0x43, 'T' 0x54, 'G' 0x47 per code, one byte each
        uint32_t v = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) v |= ((0x47544341u >> (8 * ((b >> (2 * j)) & 3u))) & 0xFFu) << (8 * j);
        o[k] = v;
    }
    if (16 * i + 16 <= n) {
        *reinterpret_cast<uint4 *>(out + 16 * i) = make_uint4(o[0], o[1], o[2], o[3]);
    } else {
        for (uint64_t j = 16 * i; j < n; ++j) out[j] = (uint8_t)(o[(j - 16 * i) / 4] >> (8 * ((j - 16 * i) % 4)));
    }
}

hipError_t launch_unpack(const uint32_t *packed, uint8_t *out, uint64_t n, hipStream_t s) {
    const uint64_t threads = (n + 15) / 16;
    if (threads == 0) return hipSuccess;
    hipLaunchKernelGGL(k_unpack, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, packed, out, n);
    return hipGetLastError();
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

// One round's edge offsets in the output, in bucket order: dst[i] = base + the edges of slots
// before i, with base = *run (the edges of the rounds before) ; then *run += the round's edges and
// rec = {edges, base, the edge pool's overflow cursor as read now}. One work-group; rounds of one
// run are scanned in order on one stream, after their join and while later joins may run (the
// cursor read is then at least its value after this round's join).
__global__ void __launch_bounds__(1024) k_round_scan(const uint32_t *__restrict__ cnt, uint32_t n,
                                                    unsigned long long *__restrict__ dst,
                                                    unsigned long long *__restrict__ run,
                                                    unsigned long long *__restrict__ rec,
                                                    const unsigned long long *ovf_cursor, RecMirror m) {
    __shared__ unsigned long long sh[16];
    const unsigned long long base = *run;
    unsigned long long carry = base;
    for (uint32_t c0 = 0; c0 < n; c0 += 4096) {
        const uint32_t i0 = c0 + 4 * threadIdx.x;
        uint32_t v[4];
        unsigned long long sum = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v[u] = i0 + u < n ? cnt[i0 + u] : 0u;
            sum += v[u];
        }
        unsigned long long tot;
        unsigned long long off = carry + block_excl_scan64<1024>(sum, sh, tot);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (i0 + u < n) dst[i0 + u] = off;
            off += v[u];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        const unsigned long long o = __hip_atomic_load(ovf_cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        rec[0] = carry - base;
        rec[1] = base;
        rec[2] = o;
        *run = carry;
        if (m.rec) {
            m.rec[0] = carry - base;
            m.rec[1] = base;
            m.rec[2] = o;
        }
    }
    if (m.list && threadIdx.x < m.list_words) m.list[threadIdx.x] = m.list_src[threadIdx.x];
}

__global__ void __launch_bounds__(256) k_edge_copy(EdgeCopyArgs a) {
    const uint64_t b = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= a.n) return;
    const uint32_t c = a.cnt[b];
    if (c == 0) return;
    const unsigned long long src = a.src_off[b], dst = a.dst_off[b] + (a.blk ? a.blk[b / ESCAN_BLOCK] : 0ull);
    if (src == ~0ull) return;  // edges that found no room in the pool (their count is zero too)
    for (uint32_t i = threadIdx.x & 63; i < c; i += 64) {
        const uint2 v = a.pool[src + i];
        a.out_from[dst + i] = v.x;
        a.out_to[dst + i] = v.y;
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
// Weights of the map kernel's dot products for read length L (1..255): dword 16 d + 4 h + j holds,
// in byte u, byte j of the weight of position k = 4d + u (0 past the read), for h = 0: R1^(L-1-k)
// mod P1 and h = 1: R2^(L-1-k) mod P2 (the read), h = 2: R1^k mod P1 and h = 3: R2^k mod P2 (its
// reverse complement: its base L-1-k is the complement of the read's base k). One table per
// device and length, built at its first use and kept until the process ends.
static hipError_t map_weights(int L, const uint32_t **out) {
    static std::mutex mu;
    static const uint32_t *tab[64][256];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64 || L < 1 || L > 255) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(mu);
    if (!tab[dev][L]) {
        const int ng = (L + 3) / 4;
        std::vector<uint32_t> h((size_t)ng * 16, 0u);
        std::vector<uint64_t> p1(L), p2(L);
        p1[0] = p2[0] = 1;
        for (int k = 1; k < L; ++k) {
            p1[k] = p1[k - 1] * R1 % P1;
            p2[k] = p2[k - 1] * R2 % P2;
        }
        for (int k = 0; k < L; ++k) {
            const uint64_t w[4] = {p1[L - 1 - k], p2[L - 1 - k], p1[k], p2[k]};
            for (int q = 0; q < 4; ++q)
                for (int j = 0; j < 4; ++j)
                    h[16 * (k / 4) + 4 * q + j] |= (uint32_t)((w[q] >> (8 * j)) & 0xFFu) << (8 * (k % 4));
        }
        void *d = nullptr;
        if ((e = hipMalloc(&d, h.size() * 4)) != hipSuccess) return e;
        if ((e = hipMemcpy(d, h.data(), h.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) {
            (void)hipFree(d);
            return e;
        }
        tab[dev][L] = static_cast<const uint32_t *>(d);
    }
    *out = tab[dev][L];
    return hipSuccess;
}

hipError_t launch_map_wt(const MapArgs &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    uint32_t grid = (a.n + a.rows - 1) / a.rows;
    size_t lds = (size_t)a.rows * a.Lpad;
    hipLaunchKernelGGL(k_map, dim3(grid), dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_map(const MapArgs &a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    MapArgs b = a;
    hipError_t e = map_weights(a.L, &b.wt);
    if (e != hipSuccess) return e;
    return launch_map_wt(b, s);
}

hipError_t launch_part(const PartArgs &a, uint32_t groups, hipStream_t s) {
    uint64_t grid = (uint64_t)groups * a.tiles_x;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_part, dim3((uint32_t)grid), dim3(PART_THREADS), 0, s, a);
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

hipError_t launch_round_edges(const uint32_t *cnt, const unsigned long long *src_off, unsigned long long *dst_off,
                              uint32_t n, unsigned long long *run, unsigned long long *rec,
                              const unsigned long long *ovf_cursor, const uint2 *pool, uint32_t *out_from,
                              uint32_t *out_to, const RecMirror &m, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_round_scan, dim3(1), dim3(1024), 0, s, cnt, n, dst_off, run, rec, ovf_cursor, m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    EdgeCopyArgs a{cnt, src_off, dst_off, nullptr, n, pool, out_from, out_to};
    hipLaunchKernelGGL(k_edge_copy, dim3((n + 3) / 4), dim3(256), 0, s, a);
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

// ---------------------------------------------------------------------------------------------
// The out-of-core order (emp_ooc.h).
namespace emp {

// Key of entry e of bucket b of list l (a round's wide entry has h1's bits 21..28 in its bucket).
__device__ __forceinline__ uint64_t ooc_key(const OocList &l, const Entry &e, uint32_t b) {
    if (l.wshift < 0) return key_of(e);
    return ((uint64_t)wide_key1(e.h1, b >> (WIDE_SHIFT - l.wshift)) << 32) | (e.h2 & KEY_MASK);
}

// Visits every entry of the list: fine slots by work-group per bucket (grid-stride), spilled
// records by the whole grid.
template <class Fn>
__device__ __forceinline__ void ooc_each(const OocList &l, Fn fn) {
    for (uint32_t b = blockIdx.x; b < l.nb; b += gridDim.x) {
        const uint32_t n = min(l.cnt[b], l.cap);
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) fn(l.fine[(size_t)b * l.cap + i], b);
    }
    if (l.spill) {
        const uint32_t sn = min(*l.spill_n, l.spill_cap);
        for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < sn; j += gridDim.x * blockDim.x)
            fn(l.spill[j].e, l.spill[j].bucket);
    }
}

__global__ void __launch_bounds__(256) k_ooc_mark(OocList l, uint32_t *bitmap) {
    ooc_each(l, [&](const Entry &e, uint32_t) { atomicOr(bitmap + (e.rid >> 5), 1u << (e.rid & 31)); });
}

__global__ void __launch_bounds__(256) k_ooc_popc(const uint32_t *bitmap, uint64_t n, uint32_t *cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) cnt[i] = (uint32_t)__popc(bitmap[i]);
}

// Chunk and tie key of every entry: tie = io(chunk) << 32 | order inside the chunk, the id for
// the dedup (its input is id-ordered), (rid & 1) << 31 | rid >> 1 for a round (each chunk holds its
// even ids, then its odd ids, each in id order).
__global__ void __launch_bounds__(256) k_ooc_tie(OocTieArgs a) {
    ooc_each(a.l, [&](const Entry &e, uint32_t b) {
        const uint32_t rid = e.rid;
        uint32_t c = 0;
        if (a.F > 1) {
            uint64_t pos = rid;
            if (a.rounds) {
                const uint32_t w = rid >> 5;
                pos = a.base[w] + (uint64_t)__popc(a.bitmap[w] & ((1u << (rid & 31)) - 1u));
            }
            c = (uint32_t)(pos / a.M);
            atomicAdd(a.bcnt + (size_t)(b >> a.gshift) * a.F + c, 1u);
        }
        const uint64_t inner = a.rounds ? (((uint64_t)(rid & 1) << 31) | (rid >> 1)) : rid;
        a.tie[rid] = ((unsigned long long)a.io[c] << 32) | inner;
        a.chunk[rid] = (uint16_t)c;
    });
}

__global__ void __launch_bounds__(256) k_ooc_fetch(OocList l, const uint32_t *ids, const unsigned long long *off,
                                                   const uint16_t *chunk, OocRec *out) {
    __shared__ uint32_t s_pos;
    const uint32_t b = ids[blockIdx.x];
    if (threadIdx.x == 0) s_pos = 0;
    __syncthreads();
    OocRec *o = out + off[blockIdx.x];
    const uint32_t n = min(l.cnt[b], l.cap);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const Entry e = l.fine[(size_t)b * l.cap + i];
        o[i] = OocRec{ooc_key(l, e, b), e.rid, chunk[e.rid]};
    }
    if (l.cnt[b] > l.cap && l.spill) {
        const uint32_t sn = min(*l.spill_n, l.spill_cap);
        for (uint32_t j = threadIdx.x; j < sn; j += blockDim.x)
            if (l.spill[j].bucket == b) {
                const uint32_t k = n + atomicAdd(&s_pos, 1u);
                if (k < l.cnt[b]) o[k] = OocRec{ooc_key(l, l.spill[j].e, b), l.spill[j].e.rid, chunk[l.spill[j].e.rid]};
            }
    }
}

__global__ void __launch_bounds__(256) k_ooc_override(const uint32_t *ids, const unsigned long long *vals, uint32_t n,
                                                      unsigned long long *tie) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) tie[ids[i]] = vals[i];
}

// Bitonic sort of n2 (power of two) records by (key, tie), ascending, with the id along; global
// memory private to the work-group.
__device__ void bitonic_kt(uint64_t *key, unsigned long long *tie, uint32_t *rid, uint32_t n2) {
    for (uint32_t k = 2; k <= n2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t ka = key[i], kb = key[ixj];
                    const unsigned long long ta = tie[i], tb = tie[ixj];
                    const bool a_gt = ka > kb || (ka == kb && ta > tb);
                    if (a_gt == ((i & k) == 0)) {
                        key[i] = kb; key[ixj] = ka;
                        tie[i] = tb; tie[ixj] = ta;
                        const uint32_t r = rid[i]; rid[i] = rid[ixj]; rid[ixj] = r;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Bucket b's entries (slots + spilled records) as (key, tie, id) at K/T/R, padded to n2 with
// maximal records.
__device__ void ooc_gather(const Entry *slots, uint32_t cnt, uint32_t cap, const SpillRec *spill, const uint32_t *spill_n,
                           uint32_t spill_cap, uint32_t b, int32_t wshift, const unsigned long long *ties, uint64_t *K,
                           unsigned long long *T, uint32_t *R, uint32_t n2, uint32_t *s_pos) {
    OocList l{slots, nullptr, cap, spill, spill_n, spill_cap, wshift, 0};
    const uint32_t nm = min(cnt, cap);
    for (uint32_t i = threadIdx.x; i < nm; i += blockDim.x) {
        const Entry e = slots[(size_t)b * cap + i];
        K[i] = ooc_key(l, e, b);
        T[i] = ties[e.rid];
        R[i] = e.rid;
    }
    if (threadIdx.x == 0) *s_pos = nm;
    __syncthreads();
    if (cnt > cap) {
        const uint32_t sn = min(*spill_n, spill_cap);
        for (uint32_t j = threadIdx.x; j < sn; j += blockDim.x)
            if (spill[j].bucket == b) {
                const uint32_t k = atomicAdd(s_pos, 1u);
                if (k < cnt) {
                    const Entry e = spill[j].e;
                    K[k] = ooc_key(l, e, b);
                    T[k] = ties[e.rid];
                    R[k] = e.rid;
                }
            }
    }
    __syncthreads();
    for (uint32_t i = cnt + threadIdx.x; i < n2; i += blockDim.x) { K[i] = ~0ull; T[i] = ~0ull; R[i] = ~0u; }
    __syncthreads();
}

__global__ void __launch_bounds__(1024) k_dedup_ooc(BigArgs a) {
    __shared__ uint32_t s_pos;
    const uint32_t b = blockIdx.x;
    const uint32_t n = a.s_cnt[b];
    if (n == 0) return;
    const uint32_t n2 = pow2_at_least(n);
    uint64_t *K = a.skey + a.soff[b];
    unsigned long long *T = a.stie + a.soff[b];
    uint32_t *R = a.srank + a.soff[b];
    ooc_gather(a.s_in, n, a.cap, a.s_spill, a.s_spill_n, a.spill_cap, b, -1, a.tie_s, K, T, R, n2, &s_pos);
    bitonic_kt(K, T, R, n2);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        if (i == 0 || K[i] != K[i - 1]) {  // the first of its key in the merged order
            if (a.nx.active) {
                put_next(a.nx, (uint32_t)(K[i] >> 32), (uint32_t)K[i], R[i], 0u, 0);
                put_next(a.nx, (uint32_t)(K[i] >> 32), (uint32_t)K[i], R[i], 0u, 1);
            } else {
                atomicAdd(a.nx.total[0], 1u);
                atomicAdd(a.nx.total[1], 1u);
            }
        }
    }
}

// Join of every bucket in (key, tie) order. A suffix's partners are the prefix key group minus
// the same read's members (counted over the group), a prefix is used when a suffix of another
// read shares its key (overlap.cu:413-415,424-429,557-603,694-710).
__global__ void __launch_bounds__(1024) k_join_ooc(BigArgs a) {
    __shared__ unsigned long long sh64[16];
    __shared__ uint32_t s_pos;
    __shared__ unsigned long long s_base;
    const uint32_t b = blockIdx.x;
    const uint32_t ns = a.s_cnt[b], np = a.p_cnt[b];
    if (ns == 0 && np == 0) {
        if (threadIdx.x == 0) a.edge_cnt[b] = 0;
        return;
    }
    const uint32_t ns2 = pow2_at_least(ns), np2 = pow2_at_least(np);
    uint64_t *SK = a.skey + a.soff[b], *PK = a.pkey + a.poff[b];
    unsigned long long *ST = a.stie + a.soff[b], *PT = a.ptie + a.poff[b];
    uint32_t *SR = a.srank + a.soff[b], *PR = a.prank + a.poff[b];
    uint32_t *SA = a.saux + a.soff[b], *PA = a.paux + a.poff[b];
    const int32_t ws = a.nx.src_shift;
    ooc_gather(a.s_in, ns, a.cap, a.s_spill, a.s_spill_n, a.spill_cap, b, ws, a.tie_s, SK, ST, SR, ns2, &s_pos);
    ooc_gather(a.p_in, np, a.cap, a.p_spill, a.p_spill_n, a.p_spill_cap, b, ws, a.tie_p, PK, PT, PR, np2, &s_pos);
    bitonic_kt(SK, ST, SR, ns2);
    bitonic_kt(PK, PT, PR, np2);
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
        const uint64_t k = SK[i];
        const uint32_t lo = lower_u64(PK, 0, np, k), hi = upper_u64(PK, lo, np, k), rd = SR[i] >> 1;
        uint32_t c = 0;
        for (uint32_t j = lo; j < hi; ++j) c += (PR[j] >> 1) != rd ? 1u : 0u;
        SA[i] = c;
    }
    for (uint32_t j = threadIdx.x; j < np; j += blockDim.x) {
        const uint64_t k = PK[j];
        const uint32_t lo = lower_u64(SK, 0, ns, k), hi = upper_u64(SK, lo, ns, k), rd = PR[j] >> 1;
        uint32_t used = 0;
        for (uint32_t i = lo; i < hi && !used; ++i) used = (SR[i] >> 1) != rd ? 1u : 0u;
        PA[j] = used;
    }
    __syncthreads();
    const uint32_t per = (ns + blockDim.x - 1) / blockDim.x;
    const uint32_t lo = min(ns, threadIdx.x * per), hi = min(ns, lo + per);
    unsigned long long run = 0;
    for (uint32_t i = lo; i < hi; ++i) run += SA[i];
    unsigned long long tot64;
    const unsigned long long off64 = block_excl_scan64<1024>(run, sh64, tot64);
    const bool too_many = tot64 >= (1ull << 31);
    const uint32_t tot = too_many ? 0u : (uint32_t)tot64;
    if (!too_many) {
        uint32_t off = (uint32_t)off64;
        for (uint32_t i = lo; i < hi; ++i) {
            const uint32_t c = SA[i];
            SA[i] = off;
            off += c;
        }
    }
    const uint32_t sub = b & (SUBS - 1);
    if (threadIdx.x == 0) {
        unsigned long long base = ~0ull;
        if (too_many) atomicOr(a.too_many, 1u);
        if (tot) {
            const unsigned long long c = atomicAdd(a.pool_cursor + sub, (unsigned long long)tot);
            base = place_edges(c, tot, sub, a.pool_cursor + sub, a.pool_cap, a.ovf_cursor, a.ovf_cap);
        }
        s_base = base;
        a.edge_off[b] = base;
        a.edge_cnt[b] = base == ~0ull ? 0u : tot;
    }
    __syncthreads();
    if (too_many) return;
    const unsigned long long base = s_base;
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
        const uint32_t o = SA[i], e = (i + 1 < ns) ? SA[i + 1] : tot;
        const uint32_t srid = SR[i];
        if (e > o && base != ~0ull) {
            const uint64_t k = SK[i];
            unsigned long long pos = base + o;
            for (uint32_t j = lower_u64(PK, 0, np, k); j < np && PK[j] == k && pos < base + e; ++j)
                if ((PR[j] >> 1) != (srid >> 1)) {
                    a.pool[pos] = make_uint2(srid, PR[j]);
                    ++pos;
                }
        }
        if (e == o) {  // a survivor
            if (a.nx.active)
                put_next(a.nx, (uint32_t)(SK[i] >> 32), (uint32_t)SK[i], srid, carry_after(a.cols, a.rg, srid, true), 0);
            else
                atomicAdd(a.nx.total[0], 1u);
        }
    }
    for (uint32_t j = threadIdx.x; j < np; j += blockDim.x) {
        if (!PA[j]) {
            if (a.nx.active)
                put_next(a.nx, (uint32_t)(PK[j] >> 32), (uint32_t)PK[j], PR[j], carry_after(a.cols, a.rg, PR[j], false), 1);
            else
                atomicAdd(a.nx.total[1], 1u);
        }
    }
}

hipError_t launch_ooc_mark(const OocList &l, uint32_t *bitmap, hipStream_t s) {
    const uint32_t g = l.nb < 65536 ? (l.nb ? l.nb : 1) : 65536;
    hipLaunchKernelGGL(k_ooc_mark, dim3(g), dim3(256), 0, s, l, bitmap);
    return hipGetLastError();
}

hipError_t launch_ooc_popc(const uint32_t *bitmap, uint64_t n, uint32_t *cnt, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ooc_popc, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, bitmap, n, cnt);
    return hipGetLastError();
}

hipError_t launch_ooc_tie(const OocTieArgs &a, hipStream_t s) {
    const uint32_t g = a.l.nb < 65536 ? (a.l.nb ? a.l.nb : 1) : 65536;
    hipLaunchKernelGGL(k_ooc_tie, dim3(g), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ooc_fetch(const OocList &l, const uint32_t *ids, const unsigned long long *off, uint32_t n,
                            const uint16_t *chunk, OocRec *out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ooc_fetch, dim3(n), dim3(256), 0, s, l, ids, off, chunk, out);
    return hipGetLastError();
}

hipError_t launch_ooc_override(const uint32_t *ids, const unsigned long long *vals, uint32_t n,
                               unsigned long long *tie, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ooc_override, dim3((n + 255) / 256), dim3(256), 0, s, ids, vals, n, tie);
    return hipGetLastError();
}

hipError_t launch_dedup_ooc(const BigArgs &a, uint32_t nb, hipStream_t s) {
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_dedup_ooc, dim3(nb), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_join_ooc(const BigArgs &a, uint32_t nb, hipStream_t s) {
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_join_ooc, dim3(nb), dim3(1024), 0, s, a);
    return hipGetLastError();
}

}  // namespace emp
