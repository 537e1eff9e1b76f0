// Planner of the out-of-core order (emp_ooc_plan.h): the FIFO merge tree of cq_sort_by_key
// (cqutils.cuh:204-245), the merge partitions of every node (cq_get_merge_partitions,
// cqutils.cuh:36-84) from order statistics of the runs, the keys whose equal entries the
// partitions reorder, and their replay through the tree.
//
// A run is the set of entries whose chunk lies in a node's in-order interval. Positions in a run
// come from a count table (entries per bucket group and chunk, summed over groups below and
// in-order indices below: T) and, inside one group, from the group's entries read through the
// source, sorted by (key, in-order index, inner order), which is also their default tie order.
#include "emp_ooc_plan.h"

#include <algorithm>
#include <deque>
#include <functional>
#include <set>

#include "../../include/empanada.h"

namespace emp {

namespace {

struct TreeNode {
    int a = -1, b = -1;
    uint32_t lo = 0, hi = 0;
};

// Leaves 0..F-1 are the chunks; every merge takes the front two runs (A first) and appends the
// result (cq_sort_by_key, cqutils.cuh:230-245).
int build_tree(uint32_t F, std::vector<TreeNode> &nodes, std::vector<uint32_t> &io) {
    nodes.assign(F, TreeNode{});
    std::deque<int> q;
    for (uint32_t c = 0; c < F; ++c) q.push_back((int)c);
    while (q.size() > 1) {
        TreeNode n;
        n.a = q.front();
        q.pop_front();
        n.b = q.front();
        q.pop_front();
        nodes.push_back(n);
        q.push_back((int)nodes.size() - 1);
    }
    const int root = F ? q.front() : -1;
    io.assign(F, 0);
    // in-order numbering of the leaves; a node spans its children's intervals
    uint32_t next = 0;
    std::function<void(int)> walk = [&](int v) {
        TreeNode &n = nodes[v];
        if (n.a < 0) {
            io[v] = next;
            n.lo = next;
            n.hi = ++next;
            return;
        }
        walk(n.a);
        walk(n.b);
        n.lo = nodes[n.a].lo;
        n.hi = nodes[n.b].hi;
    };
    if (root >= 0) walk(root);
    return root;
}

}  // namespace

std::vector<uint32_t> plan_tree_io(uint32_t F) {
    std::vector<TreeNode> nodes;
    std::vector<uint32_t> io;
    build_tree(F, nodes, io);
    return io;
}

int plan_group_shift(uint32_t nb, uint32_t F) {
    int s = 0;
    while ((uint64_t)std::max<uint32_t>(1u, nb >> s) * F > PLAN_MAX_CELLS && (nb >> s) > 1) ++s;
    return s;
}

Planner::Planner(const PlanList &l, PlanSource &src) : l_(l), src_(src) {
    std::vector<TreeNode> tn;
    root_ = build_tree(l.F, tn, io_);
    nodes_.resize(tn.size());
    for (size_t i = 0; i < tn.size(); ++i) nodes_[i] = Node{tn[i].a, tn[i].b, tn[i].lo, tn[i].hi};
    cuts_.resize(nodes_.size());
    std::vector<uint32_t> chunk_at(l.F);
    for (uint32_t c = 0; c < l.F; ++c) chunk_at[io_[c]] = c;
    const size_t W = (size_t)l.F + 1;
    T_.assign((size_t)(l.G + 1) * W, 0);
    for (uint32_t g = 0; g < l.G; ++g) {
        uint64_t acc = 0;
        for (uint32_t i = 0; i < l.F; ++i) {
            acc += l.bcnt[(size_t)g * l.F + chunk_at[i]];
            T_[(size_t)(g + 1) * W + i + 1] = T_[(size_t)g * W + i + 1] + acc;
        }
    }
    if (T_[(size_t)l.G * W + l.F] != l.n) err_ = EMP_ERR_CAPACITY;
}

const std::vector<Planner::Rec> *Planner::group(uint32_t g) {
    auto it = groups_.find(g);
    if (it != groups_.end()) return &it->second;
    if (g >= l_.G) {
        err_ = EMP_ERR_CAPACITY;
        return nullptr;
    }
    std::vector<PlanRec> raw;
    if (int rc = src_.fetch(g, raw)) {
        err_ = rc;
        return nullptr;
    }
    if (raw.size() != below(0, l_.F, g + 1) - below(0, l_.F, g)) {
        err_ = EMP_ERR_CAPACITY;
        return nullptr;
    }
    std::vector<Rec> v(raw.size());
    for (size_t i = 0; i < raw.size(); ++i) {
        const PlanRec &r = raw[i];
        if (r.chunk >= l_.F) {
            err_ = EMP_ERR_CAPACITY;
            return nullptr;
        }
        const uint64_t inner = l_.rounds ? (((uint64_t)(r.rid & 1) << 31) | (r.rid >> 1)) : r.rid;
        v[i] = Rec{r.key, io_[r.chunk], r.rid, inner};
    }
    std::sort(v.begin(), v.end(), [](const Rec &x, const Rec &y) {
        if (x.key != y.key) return x.key < y.key;
        if (x.io != y.io) return x.io < y.io;
        return x.inner < y.inner;
    });
    return &groups_.emplace(g, std::move(v)).first->second;
}

// Key of the entry at `pos` of run r (0-based, key order).
bool Planner::key_at(const Node &r, uint64_t pos, uint64_t &key) {
    uint32_t lo = 0, hi = l_.G;
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (below(r.lo, r.hi, mid) <= pos) lo = mid;
        else hi = mid;
    }
    const std::vector<Rec> *v = group(lo);
    if (!v) return false;
    uint64_t k = pos - below(r.lo, r.hi, lo);
    for (const Rec &e : *v)
        if (e.io >= r.lo && e.io < r.hi && k-- == 0) {
            key = e.key;
            return true;
        }
    err_ = EMP_ERR_CAPACITY;
    return false;
}

// Entries of run r below `key` (inclusive: at most key).
bool Planner::count_below(const Node &r, uint64_t key, bool inclusive, uint64_t &n) {
    const uint32_t g = group_of(key);
    if (g >= l_.G) {
        n = below(r.lo, r.hi, l_.G);
        return true;
    }
    const std::vector<Rec> *v = group(g);
    if (!v) return false;
    n = below(r.lo, r.hi, g);
    for (const Rec &e : *v) {
        if (e.key > key || (!inclusive && e.key == key)) break;
        n += e.io >= r.lo && e.io < r.hi;
    }
    return true;
}

// cq_get_merge_partitions (cqutils.cuh:36-84) of node v's runs: (offA, offB) after every
// partition, from (0, 0) to (|A|, |B|).
bool Planner::merge_cuts(int v) {
    std::vector<std::pair<uint64_t, uint64_t>> &cuts = cuts_[v];
    if (!cuts.empty()) return true;
    const Node &A = nodes_[nodes_[v].a], &B = nodes_[nodes_[v].b];
    const uint64_t nA = below(A.lo, A.hi, l_.G), nB = below(B.lo, B.hi, l_.G), H = l_.M / 2;
    uint64_t offA = 0, offB = 0;
    cuts.push_back({0, 0});
    while (offA < nA && offB < nB) {
        uint64_t nPa = std::min(nA - offA, H), nPb = std::min(nB - offB, H);
        if (nPa < H && nPb < H) {
            offA += nPa;
            offB += nPb;
            cuts.push_back({offA, offB});
            break;
        }
        uint64_t la, fb, lb, fa;
        if (!key_at(A, offA + nPa - 1, la) || !key_at(B, offB, fb)) return false;
        if (la <= fb) {  // case 1: A's window alone
            offA += nPa;
        } else {
            if (!key_at(B, offB + nPb - 1, lb) || !key_at(A, offA, fa)) return false;
            if (lb <= fa) {  // case 2: B's window alone
                offB += nPb;
            } else {  // case 3: the window with the larger last key cut at upper_bound of the other's
                uint64_t u;
                if (la < lb) {
                    if (!count_below(B, la, true, u)) return false;
                    nPb = std::min(std::max(u, offB), offB + nPb) - offB;
                } else if (lb < la) {
                    if (!count_below(A, lb, true, u)) return false;
                    nPa = std::min(std::max(u, offA), offA + nPa) - offA;
                }
                offA += nPa;
                offB += nPb;
            }
        }
        cuts.push_back({offA, offB});
    }
    if (offA == nA && offB < nB) cuts.push_back({nA, nB});
    else if (offB == nB && offA < nA) cuts.push_back({nA, nB});
    return true;
}

// The entries of `key` under node v in the order node v's merge leaves them.
void Planner::seq(int v, uint64_t key, const std::vector<Rec> &recs, std::vector<const Rec *> &out) {
    out.clear();
    const Node &n = nodes_[v];
    auto first = std::lower_bound(recs.begin(), recs.end(), n.lo, [](const Rec &e, uint32_t io) { return e.io < io; });
    auto last = std::lower_bound(first, recs.end(), n.hi, [](const Rec &e, uint32_t io) { return e.io < io; });
    if (first == last || err_) return;
    if (n.a < 0) {
        for (auto it = first; it != last; ++it) out.push_back(&*it);
        return;
    }
    std::vector<const Rec *> sa, sb;
    seq(n.a, key, recs, sa);
    seq(n.b, key, recs, sb);
    if (err_) return;
    if (sa.empty() || sb.empty()) {
        out = sa.empty() ? sb : sa;
        return;
    }
    const Node &A = nodes_[n.a], &B = nodes_[n.b];
    uint64_t lowA, lowB, hiA, hiB;
    if (!count_below(A, key, false, lowA) || !count_below(B, key, false, lowB) || !count_below(A, key, true, hiA) ||
        !count_below(B, key, true, hiB) || !merge_cuts(v))
        return;
    if (hiA - lowA != sa.size() || hiB - lowB != sb.size()) {
        err_ = EMP_ERR_CAPACITY;
        return;
    }
    // per partition in order: its A entries of the key, then its B entries (the stable sort of
    // [A part; B part], cqutils.cuh:160-171)
    const std::vector<std::pair<uint64_t, uint64_t>> &cuts = cuts_[v];
    // the first partition that ends past lowA or lowB (cuts grow in both coordinates)
    size_t j = (size_t)(std::partition_point(cuts.begin() + 1, cuts.end(),
                                             [&](const std::pair<uint64_t, uint64_t> &c) {
                                                 return c.first <= lowA && c.second <= lowB;
                                             }) -
                        (cuts.begin() + 1));
    size_t ia = 0, ib = 0;
    for (; j + 1 < cuts.size() && (ia < sa.size() || ib < sb.size()); ++j) {
        while (ia < sa.size() && lowA + ia < cuts[j + 1].first) out.push_back(sa[ia++]);
        while (ib < sb.size() && lowB + ib < cuts[j + 1].second) out.push_back(sb[ib++]);
    }
    if (ia != sa.size() || ib != sb.size()) err_ = EMP_ERR_CAPACITY;
}

bool Planner::replay(uint64_t key, std::vector<uint32_t> &order) {
    const std::vector<Rec> *v = group(group_of(key));
    if (!v) return false;
    std::vector<Rec> recs;
    for (const Rec &e : *v)
        if (e.key == key) recs.push_back(e);  // in (in-order index, inner order)
    std::vector<const Rec *> out;
    seq(root_, key, recs, out);
    if (err_) return false;
    if (out.size() != recs.size()) {
        err_ = EMP_ERR_CAPACITY;
        return false;
    }
    order.clear();
    for (const Rec *e : out) order.push_back(e->rid);
    return true;
}

int Planner::overrides(std::vector<uint32_t> &ids, std::vector<unsigned long long> &vals) {
    ids.clear();
    vals.clear();
    if (err_) return err_;
    std::set<uint64_t> special;
    for (int v = 0; v < (int)nodes_.size(); ++v) {
        if (nodes_[v].a < 0) continue;
        if (!merge_cuts(v)) return err_;
        const Node &A = nodes_[nodes_[v].a], &B = nodes_[nodes_[v].b];
        const uint64_t nA = below(A.lo, A.hi, l_.G);
        for (const auto &cut : cuts_[v]) {
            if (cut.second == 0 || cut.first >= nA) continue;
            uint64_t kb, ka;
            if (!key_at(B, cut.second - 1, kb) || !key_at(A, cut.first, ka)) return err_;
            if (ka == kb) special.insert(ka);  // B's entries of ka already out, A's still to come
        }
    }
    std::vector<uint32_t> order;
    for (uint64_t k : special) {
        if (!replay(k, order)) return err_;
        for (size_t i = 0; i < order.size(); ++i) {
            ids.push_back(order[i]);
            vals.push_back((unsigned long long)i);
        }
    }
    return EMP_OK;
}

int Planner::unique_check() {
    if (err_) return err_;
    const Node &r = nodes_[root_];
    const uint64_t n = below(r.lo, r.hi, l_.G), M = l_.M;
    uint64_t off = 0;
    while (n - off > M) {
        uint64_t k, c;
        if (!key_at(r, off + M - 1, k) || !count_below(r, k, false, c)) return err_;
        if (c <= off) return EMP_ERR_PARTITION;  // the window is one key
        off = c;
    }
    return EMP_OK;
}

}  // namespace emp

// ---------------------------------------------------------------------------------------------
// emp_stage_ooc_ties: the planner over host arrays (include/empanada_stages.h).
namespace {

class HostSource : public emp::PlanSource {
public:
    std::vector<std::vector<emp::PlanRec>> g;
    int fetch(uint32_t group, std::vector<emp::PlanRec> &out) override {
        out = g[group];
        return EMP_OK;
    }
};

}  // namespace

extern "C" int emp_stage_ooc_ties(const uint64_t *keys, const uint32_t *rids, uint64_t n, uint64_t M, int32_t rounds,
                                  int32_t bucket_bits, uint64_t *ties) {
    using namespace emp;
    if (n && (!keys || !rids || !ties)) return EMP_ERR_ARG;
    if (M < 2 || bucket_bits < 1 || bucket_bits > 29 || n >= (1ull << 32)) return EMP_ERR_ARG;
    for (uint64_t i = 0; i < n; ++i)
        if ((keys[i] >> 32) >= (1u << 29) || (uint32_t)keys[i] >= (1u << 29) || (!rounds && rids[i] >= n))
            return EMP_ERR_ARG;
    const uint64_t F64 = n <= M ? 1 : (n + M - 1) / M;
    if (F64 > PLAN_MAX_CHUNKS) return EMP_ERR_OUT_OF_CORE;
    const uint32_t F = (uint32_t)F64;
    auto inner = [&](uint32_t rid) { return rounds ? (((uint64_t)(rid & 1) << 31) | (rid >> 1)) : (uint64_t)rid; };
    if (F == 1) {
        for (uint64_t i = 0; i < n; ++i) ties[i] = inner(rids[i]);
        return EMP_OK;
    }
    // chunk of every entry: its position in the input order / M
    std::vector<uint32_t> chunk(n);
    if (rounds) {
        std::vector<std::pair<uint32_t, uint64_t>> by_id(n);
        for (uint64_t i = 0; i < n; ++i) by_id[i] = {rids[i], i};
        std::sort(by_id.begin(), by_id.end());
        for (uint64_t r = 0; r < n; ++r) {
            if (r && by_id[r].first == by_id[r - 1].first) return EMP_ERR_ARG;
            chunk[by_id[r].second] = (uint32_t)(r / M);
        }
    } else {
        for (uint64_t i = 0; i < n; ++i) chunk[i] = (uint32_t)(rids[i] / M);
    }
    const uint32_t nb = 1u << bucket_bits;
    const int gshift = plan_group_shift(nb, F);
    PlanList l;
    l.n = n;
    l.M = M;
    l.F = F;
    l.G = std::max<uint32_t>(1u, nb >> gshift);
    l.key_shift = 29 - bucket_bits + gshift;
    l.rounds = rounds ? 1 : 0;
    l.bcnt.assign((size_t)l.G * F, 0);
    HostSource src;
    src.g.resize(l.G);
    for (uint64_t i = 0; i < n; ++i) {
        const uint32_t g = (uint32_t)((keys[i] >> 32) >> l.key_shift);
        ++l.bcnt[(size_t)g * F + chunk[i]];
        src.g[g].push_back(PlanRec{keys[i], rids[i], chunk[i]});
    }
    Planner p(l, src);
    const std::vector<uint32_t> &io = p.io();
    std::vector<uint32_t> ids;
    std::vector<unsigned long long> vals;
    int rc = p.overrides(ids, vals);
    if (rc) return rc;
    std::vector<std::pair<uint32_t, uint64_t>> idx(n);  // (id, entry) by id: ids are distinct
    for (uint64_t i = 0; i < n; ++i) {
        ties[i] = ((uint64_t)io[chunk[i]] << 32) | inner(rids[i]);
        idx[i] = {rids[i], i};
    }
    std::sort(idx.begin(), idx.end());
    for (size_t k = 0; k < ids.size(); ++k) {
        auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(ids[k], (uint64_t)0));
        if (it == idx.end() || it->first != ids[k]) return EMP_ERR_CAPACITY;
        ties[it->second] = vals[k];
    }
    return rounds ? EMP_OK : p.unique_check();
}
