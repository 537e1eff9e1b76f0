// Kernel argument blocks and launchers (host + device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emp_device.h"

namespace emp {

// Bucket capacity (entries per bucket per list), matched-record capacity and work-group size of
// the LDS join.
// Build-time tuning knobs: EMP_BUCKET_CAP, EMP_MATCH_CAP, EMP_JOIN_THREADS.
#ifndef EMP_BUCKET_CAP
#define EMP_BUCKET_CAP 2048
#endif
#ifndef EMP_MATCH_CAP
#define EMP_MATCH_CAP 384
#endif
#ifndef EMP_JOIN_THREADS
#define EMP_JOIN_THREADS 512
#endif
constexpr int BUCKET_CAP = EMP_BUCKET_CAP;
constexpr int MATCH_CAP = EMP_MATCH_CAP;
constexpr int JOIN_THREADS = EMP_JOIN_THREADS;
// Join work-groups loop over buckets with the next bucket's loads in flight (true), or take one
// bucket each (false). Build-time tuning knob (EMP_JOIN_PERSIST).
#ifndef EMP_JOIN_PERSIST
#define EMP_JOIN_PERSIST 0
#endif
constexpr bool JOIN_PERSIST = EMP_JOIN_PERSIST != 0;
// Target mean bucket load: the partition fan-out is chosen so that lists average at most this
// many entries per bucket. 7/8 of BUCKET_CAP keeps a hash-uniform bucket 6 standard deviations
// (sqrt of the mean, about 42) below the capacity; fuller buckets spread the join's fixed
// per-work-group costs over more entries (measured on the 15 M-read join: 1024 -> 1664 took it
// from 12.7 to 10.6 ms, 1664 -> 1792 from 9.65 to 9.44 ms). A bucket that still overflows takes
// the large-bucket path. Build-time tuning knob (EMP_BUCKET_TARGET).
#ifndef EMP_BUCKET_TARGET
#define EMP_BUCKET_TARGET (EMP_BUCKET_CAP / 8 * 7)
#endif
constexpr uint32_t BUCKET_TARGET = EMP_BUCKET_TARGET;
// Entries per work-group tile of a partition pass, its work-group size, and the widest digit of
// one pass. A 4096-entry tile gives runs of 16 entries (192 bytes) per bucket at 256 buckets.
// Build-time tuning knobs: EMP_PART_THREADS (work-group size) and EMP_PART_WAVES (waves per SIMD
// the kernel is compiled for, which caps its registers).
#ifndef EMP_PART_THREADS
#define EMP_PART_THREADS 512
#endif
#ifndef EMP_PART_WAVES
#define EMP_PART_WAVES 6
#endif
constexpr int PART_TILE = 4096;
constexpr int PART_THREADS = EMP_PART_THREADS;
constexpr int PART_MAX_BITS = 8;
// The edge pool is split into SUBS sub-ranges, each with its own cursor: work-group b reserves in
// sub-range b % SUBS, so no single counter serialises a launch and, with work-groups dealt
// round-robin to the 8 XCDs, each cursor stays on one XCD.
constexpr int SUBS = 256;

// An entry that did not fit its final bucket, with the bucket it belongs to.
struct SpillRec {
    Entry e;
    uint32_t bucket;
};

// Round description for regeneration: keys of length l become those of length l-1 for round t.
struct RoundArgs {
    int32_t active;    // 0: no regeneration
    uint32_t col_t;    // column t
    uint32_t col_back; // column L-1-t
    uint32_t m1, m2;   // R1^(l-1) mod P1, R2^(l-1) mod P2, l = L - t
    uint32_t G;        // column groups of the column store: ceil(L / 8) (wide store), ceil(L / 4) (batched)
};

// The next round's two lists (side 0 suffixes, side 1 prefixes), written by the kernel that finds
// their entries, already regenerated and in their final buckets (bucket_of(h1, nb)). A bucket of
// width w feeds few buckets of the next round (as many buckets or fewer):
//   - NextSfx subtracts e * R1^(l-1) mod P1 from h1 (e in 1..4): at most 2 buckets per base;
//   - PrevPfx multiplies h1 - e by R1^-1 = (P1 + 1) / 5, which is y = (h1 - e + k P1) / 5 with
//     k = (h1 - e) mod 5: 5 ranges of width w / 5, at most 2 buckets each (6 in practice; the
//     reference's u64 wrap for h1 < e sends an entry anywhere).
// A work-group therefore collects its entries per (side, target bucket) in a small LDS table and
// reserves one run per target; an entry whose target does not fit the table takes one global
// atomic of its own.
struct NextOut {
    int32_t active;          // 0: no next round (survivors are only counted)
    uint32_t nb;             // its bucket count: target bucket of key h = bucket_of(h, nb)
    uint32_t src_nb;         // bucket count of the writing launch's source buckets
    int32_t shift;           // 29 - log2(nb) when nb is a power of two (the out-of-core kernels)
    int32_t src_shift;       // the same for src_nb
    RoundArgs rg;            // the next round
    Entry *fine[2];          // its buckets per side, BUCKET_CAP slots each
    uint32_t *cnt[2];        // their counts
    SpillRec *spill[2];      // entries past a bucket's capacity
    uint32_t *spill_n[2];
    uint32_t spill_cap[2];
    uint32_t *total[2];      // entries of the next round per side
    uint32_t *ovf;           // set when a spill area overflows
    const uint32_t *cols;
};

struct MapArgs {
    const uint8_t *reads;    // n x L row-major, 4-byte aligned
    uint32_t n;              // reads in this batch
    int32_t L, Lpad;         // Lpad: LDS row stride (multiple of 4, odd number of dwords)
    uint32_t rows;           // reads per work-group (256, or 128 when 256 rows exceed 64 KiB of LDS)
    uint64_t k0;             // global index of the first read (multiple of 16)
    uint32_t *cols;          // column store (grouped layout of emp_device.h); null: no columns written
    Entry *entries;          // linear entries, index 2k and 2k+1
    uint32_t *err;           // set to 1 when a byte is not ACGTacgt
    // gs != 0: a batched job's column region instead of the chunked store: word of read k,
    // group g at cbase + g * gs + k / 4 (k and entry ids local to the job)
    uint64_t cbase;
    uint32_t gs;
    // non-null: per read k, the codes of its first and last 8-column groups (group 0 in bits 0-15,
    // group ceil(L/8) - 1 in bits 16-31): everything round 0's regeneration reads from the columns
    uint32_t *ends;
    const uint32_t *wt;      // set by launch_map: the dot-product weights for L
};

// One partition pass: input groups (keys sharing their top bits) -> output buckets refined by
// `dbits` more key bits. Output bucket id = h1 >> out_shift.
struct PartArgs {
    const Entry *in;
    const uint32_t *in_cnt;  // per-group counts; null for one linear group of *in_n entries
    const uint32_t *in_n;
    uint32_t in_cap;         // group stride in entries (ignored for a linear group)
    int32_t keyed;           // groups share their top key bits (bucket = group << dbits | digit),
                             // or are plain segments of one list (bucket = digit)
    uint32_t tiles_x;        // work-groups per input group
    Entry *out;
    uint32_t *out_cnt;
    uint32_t out_cap;
    int32_t out_shift;
    int32_t dbits;
    SpillRec *spill;         // final pass: overflowing entries go here; null for intermediate passes
    uint32_t *spill_n;
    uint32_t spill_cap;
    uint32_t *abort;         // set when an intermediate bucket or the spill area overflows
    // a 4-byte payload that moves with each entry (null: none): a linear input's entry j has
    // pay_in[j / 2] (the map's per-read word, shared by a read's two entries), a grouped input's
    // the word at its own slot; pay_out is written at the output slot (spilled entries drop it)
    const uint32_t *pay_in;
    uint32_t *pay_out;
    // fin_nb != 0: the last pass of a partition into fin_nb buckets (a multiple of 256; key h in
    // bucket floor(h fin_nb / 2^29)) over groups keyed by the top 8 key bits: group g holds output
    // buckets g F .. g F + F - 1, F = fin_nb / 256. 0: output bucket h1 >> out_shift.
    uint32_t fin_nb;
};

struct DedupArgs {
    const Entry *in;
    const uint32_t *cnt;
    const uint32_t *pay;     // per slot (b * CAP + i): the entry's read's ends word (MapArgs::ends),
                             // carried through the partition; null: gathered from the columns
    NextOut nx;              // the kept entries, regenerated: round 0's suffix and prefix lists
    uint32_t *big, *n_big;
    uint32_t big_cap;
    const uint32_t *abort;
};

struct JoinArgs {
    const Entry *s_in, *p_in;
    uint32_t *s_cnt, *p_cnt;
    int32_t zero_in;         // clear the input counts once read (no later launch reads them)
    uint32_t nb;             // buckets
    NextOut nx;              // survivors -> the next round's buckets (nx.active: a next round exists)
    uint2 *pool;                     // edge pool, (from, to) pairs: SUBS sub-ranges of pool_cap edges, then the
    unsigned long long *pool_cursor; // overflow region [SUBS * pool_cap, + ovf_cap) with its
    unsigned long long pool_cap;     // own cursor for reservations a sub-range cannot take
    unsigned long long *ovf_cursor;
    unsigned long long ovf_cap;
    unsigned long long *edge_off;
    uint32_t *edge_cnt;
    uint32_t *big, *n_big;
    uint32_t big_cap;
    const uint32_t *abort;
};

// Large buckets (more than BUCKET_CAP entries, or more matched records than MATCH_CAP): one
// work-group per bucket sorts the bucket in global scratch memory and joins by binary search.
struct BigArgs {
    const uint32_t *ids;           // bucket ids
    const unsigned long long *soff;// scratch offset (records) of each bucket's suffix side
    const unsigned long long *poff;// and prefix side (dedup: unused)
    uint64_t *skey, *pkey;         // scratch keys
    uint32_t *srank, *prank;       // scratch ranks (join) or read ids (dedup)
    uint32_t *saux, *paux;         // per-record counts / flags
    const Entry *s_in, *p_in;
    const uint32_t *s_cnt, *p_cnt;
    const SpillRec *s_spill, *p_spill;
    const uint32_t *s_spill_n, *p_spill_n;
    uint32_t spill_cap, p_spill_cap;  // capacities of s_spill / p_spill
    uint32_t cap;
    RoundArgs rg;                  // the round (join): survivors get their carried codes back
    const uint32_t *cols;
    NextOut nx;                    // survivors (join) / kept entries (dedup) for the next round
    uint2 *pool;                     // edge pool, (from, to) pairs
    unsigned long long *pool_cursor;
    unsigned long long pool_cap;
    unsigned long long *ovf_cursor;
    unsigned long long ovf_cap;
    unsigned long long *edge_off;
    uint32_t *edge_cnt;
    uint32_t *too_many;            // set when a bucket holds 2^31 edges or more (join)
    // the out-of-core kernels (emp_ooc.h): tie-key scratch and the tie keys per id
    unsigned long long *stie, *ptie;
    const unsigned long long *tie_s, *tie_p;
};

// Up to 16 ranges of 32-bit words set to a value in one launch (per-round counter resets).
constexpr int ZERO_RANGES = 16;
struct ZeroArgs {
    uint32_t *p[ZERO_RANGES];
    uint32_t n[ZERO_RANGES];
    uint32_t v[ZERO_RANGES];
    int32_t count;
};

// Edge compaction over n (round, bucket) slots: counts, pool offsets, and scratch for the
// exclusive offsets (dst_off: n, blk: ceil(n / 4096)).
struct EdgeCopyArgs {
    const uint32_t *cnt;
    const unsigned long long *src_off;
    unsigned long long *dst_off;
    unsigned long long *blk;  // block offsets added to dst_off (null: dst_off is final)
    uint64_t n;
    const uint2 *pool;               // edge pool: (from, to) pairs
    uint32_t *out_from, *out_to;
};

// Collision audit of one key-bucketed list (FPCHECK, overlap.cu:295-359): the distinct keys and,
// summed over keys, the distinct strings [start, start + len) of the oriented reads.
struct AuditArgs {
    const Entry *fine;              // final buckets of cap slots, with counts and spilled records
    const uint32_t *cnt;
    uint32_t cap;
    const SpillRec *spill;
    const uint32_t *spill_n;
    uint32_t spill_cap;
    const uint32_t *cols;           // the wide column store, G = ceil(L / 8) groups
    uint32_t G;
    int32_t L, start, len;
    int32_t wshift;                 // >= 0: wide entries of buckets h1 >> wshift (a round's list)
    const uint32_t *ids;            // large-bucket launch: bucket of each work-group; else null
    const unsigned long long *off;  // large-bucket launch: scratch offset of each work-group
    uint64_t *skey;
    uint32_t *srid;
    unsigned long long *out;        // [0] += distinct keys, [1] += distinct strings
};

hipError_t launch_map(const MapArgs &a, hipStream_t s);
// The map kernel with the dot-product weights a.wt supplied by the caller (launch_map builds and
// keeps one table per device and read length): ceil(L / 4) * 16 dwords, dword 16 d + 4 h + j
// holding in byte u byte j of the weight of position 4 d + u (h = 0: R1^(L-1-k) mod P1, 1: R2^(L-1-k)
// mod P2, 2: R1^k mod P1, 3: R2^k mod P2; 0 past the read).
hipError_t launch_map_wt(const MapArgs &a, hipStream_t s);
// LDS row stride and reads per work-group of the map kernel for read length L.
inline void map_shape(int L, int32_t *Lpad, uint32_t *rows) {
    int lp = (L + 3) & ~3;
    if (((lp / 4) & 1) == 0) lp += 4;
    *Lpad = lp;
    *rows = 256u * (uint32_t)lp <= 65536u ? 256u : 128u;
}
hipError_t launch_part(const PartArgs &a, uint32_t groups, hipStream_t s);
hipError_t launch_dedup(const DedupArgs &a, uint32_t nb, hipStream_t s);
hipError_t launch_join(const JoinArgs &a, uint32_t nb, hipStream_t s);
hipError_t launch_dedup_big(const BigArgs &a, uint32_t n, hipStream_t s);
hipError_t launch_join_big(const BigArgs &a, uint32_t n, hipStream_t s);
hipError_t launch_edge_compact(const EdgeCopyArgs &a, hipStream_t s);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
// n bases packed four to a byte (ceil(n / 16) dwords read) -> n bytes at out (16-byte aligned)
hipError_t launch_unpack(const uint32_t *packed, uint8_t *out, uint64_t n, hipStream_t s);
hipError_t launch_zero(const ZeroArgs &a, hipStream_t s);
// Copies of a round's records in host memory (device views of pinned buffers), written by the
// round's edge scan so that no copy runs between joins: rec gets {edges, first position, overflow
// cursor}; list gets list_words words of list_src (the next round's counter record). Null: none.
struct RecMirror {
    unsigned long long *rec;
    uint32_t *list;
    const uint32_t *list_src;
    uint32_t list_words;
};
// One round's edges in output order: the n bucket slots (counts cnt, pool offsets src_off) scanned
// into dst_off from *run (the output position after the earlier rounds; advanced by the round's
// edges; rec = {edges, first position, *ovf_cursor as read after the round's join}), then copied
// from the pool to out_from / out_to (which may be pinned host memory). Two launches on stream s.
hipError_t launch_round_edges(const uint32_t *cnt, const unsigned long long *src_off, unsigned long long *dst_off,
                              uint32_t n, unsigned long long *run, unsigned long long *rec,
                              const unsigned long long *ovf_cursor, const uint2 *pool, uint32_t *out_from,
                              uint32_t *out_to, const RecMirror &m, hipStream_t s);
// n work-groups: all buckets (a.ids null; buckets over a.cap are skipped) or the listed ones.
hipError_t launch_audit(const AuditArgs &a, uint32_t n, hipStream_t s);
hipError_t launch_roll(const uint64_t *keys, const uint8_t *bases, uint64_t n, int kind, uint32_t m1, uint32_t m2,
                       uint64_t *out, hipStream_t s);

}  // namespace emp
