// Host side of the out-of-core order (emp_ooc_plan.cpp). No device code: the planner reads a
// list's entries through a source (the device kernels of emp_ooc.h, or host arrays).
//
// cq_sort_by_key (cqutils.cuh:179-247) sorts a list of n > M entries in chunks of M taken in input
// order, then merges the FIFO of sorted runs pairwise: the front two runs A, B are merged and the
// result goes to the back. A merge cuts A and B into partitions (cq_get_merge_partitions,
// cqutils.cuh:36-84) and stably sorts each mixed one as [A part; B part] (cq_merge_by_key,
// cqutils.cuh:139-177). The merged runs are key-sorted, so only the order of equal keys depends
// on the tree:
//   - the runs under a node are the chunks of an interval of the tree's in-order leaf sequence;
//   - a key whose A entries all come before its B entries at every merge ends in the order
//     (in-order index of its chunk, order inside the chunk): the default tie;
//   - at a merge, B entries of key k come before A entries of k exactly when some cut between two
//     partitions has b[offB - 1] == a[offA] == k (consumed entries never exceed unconsumed
//     ones). Such keys are replayed through the tree, partition by partition, and every entry of
//     them gets its rank in the final order as an explicit tie.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <map>
#include <utility>
#include <vector>

namespace emp {

// An entry as the planner sees it: key, id, chunk.
struct PlanRec {
    uint64_t key;
    uint32_t rid, chunk;
};

// The entries of one bucket group (any order).
class PlanSource {
public:
    virtual ~PlanSource() {}
    virtual int fetch(uint32_t group, std::vector<PlanRec> &out) = 0;
};

constexpr uint32_t PLAN_MAX_CHUNKS = 65535;      // chunk ids are 16-bit on the device
constexpr uint64_t PLAN_MAX_CELLS = 1ull << 22;  // groups x chunks of the count table

// The FIFO merge tree of F sorted runs: in-order index of every chunk (io[c]).
std::vector<uint32_t> plan_tree_io(uint32_t F);

// Group bits for nb buckets and F chunks: (nb >> shift) * F <= PLAN_MAX_CELLS.
int plan_group_shift(uint32_t nb, uint32_t F);

struct PlanList {
    uint64_t n = 0, M = 0;
    uint32_t F = 0, G = 0;    // chunks, bucket groups
    int key_shift = 0;        // group of a key = (key >> 32) >> key_shift
    int rounds = 0;           // inner order: (rid & 1) << 31 | rid >> 1 (a round) or rid (the dedup)
    std::vector<uint32_t> bcnt;  // entries per group and chunk, bcnt[g * F + c]
};

class Planner {
public:
    Planner(const PlanList &l, PlanSource &src);
    // Explicit ties (id, value) for every entry of a key the merges order otherwise than the
    // default tie; an error code from the source, or EMP_ERR_CAPACITY on inconsistent counts.
    int overrides(std::vector<uint32_t> &ids, std::vector<unsigned long long> &vals);
    // cq_get_partitions (cqutils.cuh:86-105) over the merged list: EMP_ERR_PARTITION when a
    // window of M entries that is not the last holds one key (the reference never leaves it).
    int unique_check();
    const std::vector<uint32_t> &io() const { return io_; }
    int error() const { return err_; }

private:
    struct Rec {
        uint64_t key;
        uint32_t io;
        uint32_t rid;
        uint64_t inner;
    };
    struct Node {
        int a = -1, b = -1;  // children (-1: a leaf)
        uint32_t lo = 0, hi = 0;  // in-order leaf interval
    };
    const PlanList &l_;
    PlanSource &src_;
    int err_ = 0;
    std::vector<uint32_t> io_;
    std::vector<Node> nodes_;
    int root_ = 0;
    std::vector<uint64_t> T_;  // T_[g * (F + 1) + i]: entries of groups < g with in-order index < i
    std::map<uint32_t, std::vector<Rec>> groups_;  // sorted by (key, io, inner)
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> cuts_;  // per node: (offA, offB) after each partition

    uint64_t below(uint32_t lo, uint32_t hi, uint32_t g) const {
        const size_t W = (size_t)l_.F + 1;
        return T_[(size_t)g * W + hi] - T_[(size_t)g * W + lo];
    }
    uint32_t group_of(uint64_t key) const { return (uint32_t)((key >> 32) >> l_.key_shift); }
    const std::vector<Rec> *group(uint32_t g);
    bool key_at(const Node &r, uint64_t pos, uint64_t &key);
    bool count_below(const Node &r, uint64_t key, bool inclusive, uint64_t &n);
    bool merge_cuts(int v);
    bool replay(uint64_t key, std::vector<uint32_t> &order);
    void seq(int v, uint64_t key, const std::vector<Rec> &recs, std::vector<const Rec *> &out);
};

}  // namespace emp
