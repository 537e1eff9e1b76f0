// The out-of-core order on the device (emp_ooc.hip): the reference sorts a list of more than M
// entries (M = reduce_mem / 16) by sorting chunks of M in input order and merging the sorted runs
// pairwise in FIFO order (cq_sort_by_key, cqutils.cuh:179-247), which orders equal keys by the
// merge tree rather than by id. Every entry of such a list gets a 64-bit tie key: by default the
// in-order index of its chunk in the merge tree, then its order inside the chunk; keys that a
// merge partition splits get explicit ties from the host. Kernels over all buckets of a list then
// sort each bucket by (key, tie) and dedup or join it.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emp_kernels.h"

namespace emp {

// One bucketed list as a launch reads it.
struct OocList {
    const Entry *fine;      // buckets of cap slots
    const uint32_t *cnt;    // entries per bucket
    uint32_t cap;
    const SpillRec *spill;  // entries past a bucket's capacity
    const uint32_t *spill_n;
    uint32_t spill_cap;
    int32_t wshift;         // >= 0: a round's wide entries, bucket = h1 >> wshift; -1: the dedup's full keys
    uint32_t nb;            // buckets
};

// Tie keys of a list's entries. An entry's position in the sorted list's input order gives its
// chunk c = position / M: the dedup's input is in id order (position = id); a round's input is
// the id-sorted list cut in chunks of M, evens then odds in each (position from the rank of the id
// among the list's ids: rank = base[w] + set bits of bitmap word w below it).
// A list of at most M entries is sorted in one piece: F = 1, every entry in chunk 0, the tie its
// order in the input alone.
struct OocTieArgs {
    OocList l;
    int32_t rounds;                 // a round's list (inner order: evens, then odds) or the dedup's (ids)
    const uint32_t *bitmap;         // rounds with F > 1: one bit per live id
    const unsigned long long *base; // rounds with F > 1: set bits before each 32-bit word
    unsigned long long M;
    uint32_t F;                     // chunks
    const uint32_t *io;             // in-order index of each chunk in the merge tree
    unsigned long long *tie;        // out, per id
    uint16_t *chunk;                // out, per id
    uint32_t *bcnt;                 // F > 1: entries per bucket group and chunk, bcnt[(b >> gshift) * F + c]
    int32_t gshift;                 // (zeroed by the caller)
};

struct OocRec {
    unsigned long long key;
    uint32_t rid, chunk;
};

// Sets bit rid of `bitmap` for every entry of the list.
hipError_t launch_ooc_mark(const OocList &l, uint32_t *bitmap, hipStream_t s);
// Set bits per 32-bit word (n words) into cnt.
hipError_t launch_ooc_popc(const uint32_t *bitmap, uint64_t n, uint32_t *cnt, hipStream_t s);
hipError_t launch_ooc_tie(const OocTieArgs &a, hipStream_t s);
// The entries of buckets ids[0..n) as (key, id, chunk) records at out + off[i] (any order).
hipError_t launch_ooc_fetch(const OocList &l, const uint32_t *ids, const unsigned long long *off, uint32_t n,
                            const uint16_t *chunk, OocRec *out, hipStream_t s);
// tie[ids[i]] = vals[i].
hipError_t launch_ooc_override(const uint32_t *ids, const unsigned long long *vals, uint32_t n,
                               unsigned long long *tie, hipStream_t s);

// Dedup / join of every bucket (one work-group each) in (key, tie) order, in global scratch:
// BigArgs with ids null, soff / poff the scratch offsets per bucket (powers of two of the
// counts), skey / pkey, stie / ptie, srank / prank (ids), saux / paux; tie_s / tie_p the tie keys per
// id. The dedup keeps the first entry of every key; the join writes the cross product of every
// key group in (key, tie) x tie order, survivors to the next round.
hipError_t launch_dedup_ooc(const BigArgs &a, uint32_t nb, hipStream_t s);
hipError_t launch_join_ooc(const BigArgs &a, uint32_t nb, hipStream_t s);

}  // namespace emp
