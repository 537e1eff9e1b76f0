// Batched kernels: the partition, dedup and join of emp_kernels.hip over the live entries of many
// jobs in one launch each (emp_batch.h describes the layout).
//
//   k_bpart1   job segments -> coarse buckets, floor(h1 * nc_j / 2^29); in a round step the
//              entries are regenerated for the job's own round first (overlap.cu:266-267, 684-690)
//   k_bpart2   coarse bucket c of job j -> fine buckets fb_j + floor(h1 * nb_j / 2^29), which lie
//              in [c * m_j, (c + 1) * m_j) because nb_j = nc_j * m_j
//   k_bdedup   per fine bucket: smallest read id per key (overlap.cu:212-222)
//   k_bjoin    per fine bucket: suffix/prefix join, edges in reference order, used-entry removal
//              (overlap.cu:258-290, 361-456, 557-710)
//   k_b*_big   buckets over capacity: global-memory sort + binary-search join / dedup
//
// A work-group finds its job by a binary search of the step's prefix arrays (tiles per job, fine
// buckets per job); every other index is job-local.
#include "emp_batch.h"
#include "emp_common.h"

#include <hip/hip_runtime.h>

namespace emp {

// Index j < n with pre[j] <= x < pre[j + 1]; pre ascending from pre[0] = 0 and x < pre[n].
__device__ __forceinline__ uint32_t job_of(const uint32_t *__restrict__ pre, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

// The 4 codes of job-local read k in column group g of the job's column region.
__device__ __forceinline__ uint32_t bgroup_codes(const uint32_t *__restrict__ cols, const BJob &J, uint32_t g,
                                                 uint32_t k) {
    return (cols[J.cbase + (uint64_t)g * J.gs + (k >> 2)] >> (8 * (k & 3))) & 0xFFu;
}

__device__ __forceinline__ uint32_t bcarry_after(const uint32_t *__restrict__ cols, const BJob &J, uint32_t rid,
                                                 bool sfx) {
    const bool fwd = sfx != (bool)(rid & 1);
    const uint32_t col = round_col(J.rg, fwd);
    return carry_of(bgroup_codes(cols, J, col >> 2, rid >> 1), col, fwd);
}

// Bucket of key h1 among `parts` equal key ranges (order-preserving).
__device__ __forceinline__ uint32_t range_of(uint32_t h1, uint32_t parts) {
    return (uint32_t)(((uint64_t)(h1 & KEY_MASK) * parts) >> 29);
}

// ---------------------------------------------------------------------------------------------
// Partition passes: each work-group ranks a tile of PART_TILE entries of one job by bucket in
// LDS, reserves one output run per (tile, bucket) and writes every run contiguously.
template <bool REGEN>
__global__ void __launch_bounds__(PART_THREADS) k_bpart1(BPartArgs a) {
    constexpr int NT = PART_THREADS, T = PART_TILE, PER = T / NT, NB = BSPLIT;
    static_assert(NT >= NB && T % NT == 0, "one bin per thread");
    __shared__ Entry buf[T];
    __shared__ uint32_t hist[NB], loc[NB], base[NB], sh[NT / 64];
    if (blockIdx.x >= a.tpre[a.njobs]) return;  // a grid sized for a larger step (a replayed graph)
    const uint32_t j = job_of(a.tpre, a.njobs, blockIdx.x);
    const BJob &J = a.jobs[j];
    const int side = a.side;
    const uint32_t n = J.cnt[side], t0 = (blockIdx.x - a.tpre[j]) * (uint32_t)T;
    if (t0 >= n) return;
    const uint32_t nt = min((uint32_t)T, n - t0), nc = J.nc, ccap = J.ccap[side];
    const Entry *in = a.lists[J.src[side]] + J.seg + t0;
    if (threadIdx.x < NB) hist[threadIdx.x] = 0;
    lds_barrier();
    Entry e[PER];
    uint32_t r[PER], d[PER];
    // every lane loads (lanes past the tile re-read its last entry): all loads in flight together
    const uint32_t last = nt - 1;
#pragma unroll
    for (int q = 0; q < PER; ++q) e[q] = load_stream(in + min(threadIdx.x + NT * q, last));
    if (REGEN) {
        const RoundArgs rg = J.rg;
        uint32_t word[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            uint32_t g;
            const bool need = regen_needs_gather(rg, e[q].rid, side == 0, g);
            // entries without a gather this round read the job's first column word (cached)
            const uint32_t w = bgroup_codes(a.cols, J, need ? g : 0u, need ? e[q].rid >> 1 : 0u);
            word[q] = need ? w : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int q = 0; q < PER; ++q)
            if (threadIdx.x + NT * q < nt) regen_carry(a.cols, rg, e[q], side == 0, word[q]);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if (threadIdx.x + NT * q < nt) {
            d[q] = range_of(e[q].h1, nc);
            r[q] = atomicAdd(&hist[d[q]], 1u);
        }
    }
    lds_barrier();
    const uint32_t h = threadIdx.x < NB ? hist[threadIdx.x] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<NT>(h, sh, tot);
    if (threadIdx.x < NB) {
        loc[threadIdx.x] = ex;
        base[threadIdx.x] = h ? atomicAdd(&a.ccnt[J.cb + threadIdx.x], h) : 0u;
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < PER; ++q)
        if (threadIdx.x + NT * q < nt) buf[loc[d[q]] + r[q]] = e[q];
    lds_barrier();
    Entry *out = a.coarse + J.coff[side];
    for (uint32_t p = threadIdx.x; p < nt; p += NT) {
        const Entry v = buf[p];
        const uint32_t dd = range_of(v.h1, nc);
        const uint32_t slot = base[dd] + (p - loc[dd]);
        if (slot < ccap) out[(size_t)dd * ccap + slot] = v;
        else atomicOr(a.abort, 1u);
    }
}

__global__ void __launch_bounds__(PART_THREADS) k_bpart2(BPartArgs a) {
    constexpr int NT = PART_THREADS, T = PART_TILE, PER = T / NT, NB = BSPLIT;
    __shared__ Entry buf[T];
    __shared__ uint32_t hist[NB], loc[NB], base[NB], sh[NT / 64];
    if (blockIdx.x >= a.tpre[a.njobs]) return;  // a grid sized for a larger step (a replayed graph)
    const uint32_t j = job_of(a.tpre, a.njobs, blockIdx.x);
    const BJob &J = a.jobs[j];
    const int side = a.side;
    const uint32_t ccap = J.ccap[side], tpc = (ccap + T - 1) / T;
    const uint32_t local = blockIdx.x - a.tpre[j], c = local / tpc, t0 = (local % tpc) * (uint32_t)T;
    const uint32_t n = min(a.ccnt[J.cb + c], ccap);
    if (t0 >= n) return;
    const uint32_t nt = min((uint32_t)T, n - t0), nb = J.nb, f0 = c * J.m;
    const Entry *in = a.coarse + J.coff[side] + (size_t)c * ccap + t0;
    if (threadIdx.x < NB) hist[threadIdx.x] = 0;
    lds_barrier();
    Entry e[PER];
    uint32_t r[PER], d[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) e[q] = in[min(threadIdx.x + NT * q, nt - 1)];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if (threadIdx.x + NT * q < nt) {
            d[q] = range_of(e[q].h1, nb) - f0;
            r[q] = atomicAdd(&hist[d[q]], 1u);
        }
    }
    lds_barrier();
    const uint32_t h = threadIdx.x < NB ? hist[threadIdx.x] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<NT>(h, sh, tot);
    const uint32_t fb = J.fb + f0;
    if (threadIdx.x < NB) {
        loc[threadIdx.x] = ex;
        base[threadIdx.x] = h ? atomicAdd(&a.fcnt[fb + threadIdx.x], h) : 0u;
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < PER; ++q)
        if (threadIdx.x + NT * q < nt) buf[loc[d[q]] + r[q]] = e[q];
    lds_barrier();
    for (uint32_t p = threadIdx.x; p < nt; p += NT) {
        const Entry v = buf[p];
        const uint32_t dd = range_of(v.h1, nb) - f0;
        const uint32_t slot = base[dd] + (p - loc[dd]);
        const uint32_t bucket = fb + dd;
        if (slot < a.fcap) {
            a.fine[(size_t)bucket * a.fcap + slot] = v;
        } else {
            const uint32_t s = atomicAdd(a.spill_n, 1u);
            if (s < a.spill_cap) a.spill[s] = SpillRec{v, bucket};
            else atomicOr(a.abort, 1u);
        }
    }
}

__device__ __forceinline__ void to_big(const BJoinArgs &a, uint32_t b) {
    const uint32_t i = atomicAdd(a.n_big, 1u);
    if (i < a.big_cap) a.big[i] = b;
    if (a.edge_cnt) a.edge_cnt[b] = 0;
}

// ---------------------------------------------------------------------------------------------
// Dedup of one fine bucket: keep the smallest read id of every key, append the kept entries to
// the job's segment of lists[0].
template <int CAP, int NT>
__global__ void __launch_bounds__(NT) k_bdedup(BJoinArgs a) {
    constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    constexpr int PER = CAP / NT;
    __shared__ uint32_t h1s[CAP], h2s[CAP], rids[CAP], head[CAP];
    __shared__ uint16_t nxt[CAP];
    __shared__ uint32_t kcnt;
    __shared__ unsigned long long kbase;
    if (*a.abort) return;
    const uint32_t b = blockIdx.x;
    if (b >= a.fpre[a.njobs]) return;  // a grid sized for a larger step (a replayed graph)
    const uint32_t n = a.s_cnt[b];
    if (n > (uint32_t)CAP) {
        if (threadIdx.x == 0) to_big(a, b);
        return;
    }
    if (n == 0) return;
    const uint32_t j = job_of(a.fpre, a.njobs, b);
    const Entry *in = a.s_in + (size_t)b * CAP;
    for (uint32_t i = threadIdx.x; i < CAP; i += NT) head[i] = EMPTY;
    if (threadIdx.x == 0) kcnt = 0;
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
    uint32_t keepmask = 0, slot[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        slot[q] = 0;
        if (i < n) {
            const uint32_t h1 = h1s[i], h2 = h2s[i], rid = rids[i];
            bool keep = true;
            for (uint32_t k = head[h2 & (CAP - 1)]; k != EMPTY && keep; k = (nxt[k] == 0xFFFFu ? EMPTY : nxt[k]))
                if (((h2s[k] ^ h2) & KEY_MASK) == 0 && ((h1s[k] ^ h1) & KEY_MASK) == 0 && rids[k] < rid) keep = false;
            if (keep) {
                keepmask |= 1u << q;
                slot[q] = atomicAdd(&kcnt, 1u);
            }
        }
    }
    lds_barrier();
    if (threadIdx.x == 0) kbase = a.jobs[j].seg + atomicAdd(a.lcnt + j, kcnt);
    lds_barrier();
    Entry *out = a.lists[0] + kbase;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        if (keepmask & (1u << q)) {
            const uint32_t i = threadIdx.x + NT * q;
            out[slot[q]] = Entry{h1s[i], h2s[i], rids[i]};
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Join of one fine bucket of round t_j of its job j (the LDS join of emp_kernels.hip with the
// survivors appended to the job's segments and the edges counted per job).
template <int CAP, int MCAP, int NT>
__global__ void __launch_bounds__(NT) k_bjoin(BJoinArgs a) {
    constexpr uint32_t EMPTY = 0xFFFFFFFFu;
    constexpr int PER = CAP / NT;
    static_assert(MCAP <= NT, "one matched record per thread in the rank sort");
    __shared__ JoinShared<CAP, MCAP, NT> S;
    __shared__ unsigned long long sbase[2];
    if (*a.abort) return;
    const uint32_t b = blockIdx.x;
    if (b >= a.fpre[a.njobs]) return;  // a grid sized for a larger step (a replayed graph)
    const uint32_t ns = a.s_cnt[b], np = a.p_cnt[b];
    if (ns > (uint32_t)CAP || np > (uint32_t)CAP) {
        if (threadIdx.x == 0) to_big(a, b);
        return;
    }
    const uint32_t j = job_of(a.fpre, a.njobs, b);
    const Entry *sin = a.s_in + (size_t)b * CAP, *pin = a.p_in + (size_t)b * CAP;
    Entry se[PER], pe[PER];  // all loads in flight together (empty sides read slot 0, unused)
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        se[q] = sin[ns ? min(threadIdx.x + NT * q, ns - 1) : 0u];
        pe[q] = pin[np ? min(threadIdx.x + NT * q, np - 1) : 0u];
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (i < np) { S.pk1[i] = pe[q].h1; S.pk2[i] = pe[q].h2; S.prid[i] = pe[q].rid; S.used[i] = 0; }
    }
    if (threadIdx.x == 0) { S.ms = 0; S.mp = 0; S.ovf = 0; S.etot = 0; S.rcnt[0] = 0; S.rcnt[1] = 0; }
    for (uint32_t i = threadIdx.x; i < CAP; i += NT) S.u.h.head[i] = EMPTY;
    lds_barrier();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        if (i < np) S.u.h.nxt[i] = (uint16_t)atomicExch(S.u.h.head + (S.pk2[i] & (CAP - 1)), i);
    }
    lds_barrier();

    // Probe: valid partners (equal key, different read) of every suffix; flag the prefixes used.
    uint32_t sc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        sc[q] = 0;
        if (i < ns && np) {
            const uint32_t h1 = se[q].h1 & KEY_MASK, h2 = se[q].h2 & KEY_MASK, rd = se[q].rid >> 1;
            uint32_t c = 0;
            for (uint32_t k = S.u.h.head[h2 & (CAP - 1)]; k != EMPTY; k = (S.u.h.nxt[k] == 0xFFFFu ? EMPTY : S.u.h.nxt[k])) {
                if ((S.pk2[k] & KEY_MASK) == h2 && (S.pk1[k] & KEY_MASK) == h1 && (S.prid[k] >> 1) != rd) {
                    ++c;
                    S.used[k] = 1;
                }
            }
            sc[q] = c;
        }
    }
    lds_barrier();  // the hash table is dead; its space holds the matched records

    uint32_t smask = 0, pmask = 0, slot[PER], ecount = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const uint32_t i = threadIdx.x + NT * q;
        slot[q] = 0;
        if (i < ns) {
            if (sc[q]) {
                ecount += sc[q];
                const uint32_t m = atomicAdd(&S.ms, 1u);
                if (m < (uint32_t)MCAP) {
                    S.u.m.msk[m] = key_of(se[q]);
                    S.u.m.msr[m] = rid_rank(se[q].rid);
                    S.u.m.msc[m] = sc[q];
                } else {
                    S.ovf = 1;
                }
            } else {
                smask |= 1u << q;
                slot[q] = atomicAdd(&S.rcnt[0], 1u);
            }
        }
        if (i < np) {
            if (S.used[i]) {
                const uint32_t m = atomicAdd(&S.mp, 1u);
                if (m < (uint32_t)MCAP) {
                    S.u.m.mpk[m] = ((uint64_t)(S.pk1[i] & KEY_MASK) << 32) | (S.pk2[i] & KEY_MASK);
                    S.u.m.mpr[m] = rid_rank(S.prid[i]);
                } else {
                    S.ovf = 1;
                }
            } else {
                pmask |= 1u << q;
                slot[q] |= atomicAdd(&S.rcnt[1], 1u) << 16;
            }
        }
    }
    if (ecount) atomicAdd(&S.etot, ecount);
    lds_barrier();
    if (S.ovf) {
        if (threadIdx.x == 0) to_big(a, b);
        return;
    }
    const uint32_t nms = S.ms, nmp = S.mp, tot = S.etot;
    const BJob &J = a.jobs[j];

    // Reservations (read back after the sort): threads 0 and 1 the survivor runs of the two
    // sides in the job's segments, thread 2 the edges.
    unsigned long long resv = 0;
    if (threadIdx.x < 2) {
        const uint32_t cnt = S.rcnt[threadIdx.x];
        resv = cnt ? atomicAdd(a.lcnt + threadIdx.x * a.njobs + j, cnt) : 0u;
    } else if (threadIdx.x == 2 && tot) {
        resv = atomicAdd(a.pool_cursor + (b & (SUBS - 1)), (unsigned long long)tot);
    }

    // Rank sort of the matched records by (key, rank).
    if (nms) {
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
    if (threadIdx.x < 2) {
        sbase[threadIdx.x] = J.seg + resv;
    } else if (threadIdx.x == 2) {
        const uint32_t sub = b & (SUBS - 1);
        const unsigned long long base =
            tot ? place_edges(resv, tot, sub, a.pool_cursor + sub, a.pool_cap, a.ovf_cursor, a.ovf_cap) : ~0ull;
        S.base = base;
        a.edge_off[b] = base;
        a.edge_cnt[b] = base == ~0ull ? 0u : tot;  // unplaced edges are not written; the host reruns the step
        if (tot) atomicAdd(a.ejob + j, tot);
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
                    const uint32_t mid = (lo + hi) >> 1;
                    if (S.u.m.mpk[mid] < k) lo = mid + 1; else hi = mid;
                }
                unsigned long long pos = base + S.u.m.msc[i];
                const unsigned long long end = base + (i + 1 < nms ? S.u.m.msc[i + 1] : tot);
                for (uint32_t q = lo; q < nmp && S.u.m.mpk[q] == k && pos < end; ++q) {
                    const uint32_t prd = (S.u.m.mpr[q] << 1) | (S.u.m.mpr[q] >> 31);
                    if ((prd >> 1) != (srid >> 1)) {
                        a.pool[pos] = make_uint2(srid, prd);
                        ++pos;
                    }
                }
            }
        }
    }

    // Survivors into the job's segments; the next step's first pass regenerates them.
    if (J.keep) {
        Entry *so = a.lists[0] + sbase[0], *po = a.lists[1] + sbase[1];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            if (smask & (1u << q)) so[slot[q] & 0xFFFFu] = se[q];
            if (pmask & (1u << q)) {
                const uint32_t i = threadIdx.x + NT * q;
                po[slot[q] >> 16] = Entry{S.pk1[i], S.pk2[i], S.prid[i]};
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Large buckets.
__global__ void __launch_bounds__(1024) k_bjoin_big(BBigArgs a) {
    __shared__ unsigned long long sh64[16];
    __shared__ uint32_t s_pos;
    __shared__ unsigned long long s_base;
    const uint32_t b = a.ids[blockIdx.x];
    const uint32_t j = job_of(a.fpre, a.njobs, b);
    const BJob &J = a.jobs[j];
    const uint32_t ns = a.s_cnt[b], np = a.p_cnt[b];
    const uint32_t ns2 = pow2_at_least(ns), np2 = pow2_at_least(np);
    uint64_t *SK = a.skey + a.soff[blockIdx.x], *PK = a.pkey + a.poff[blockIdx.x];
    uint32_t *SR = a.srank + a.soff[blockIdx.x], *PR = a.prank + a.poff[blockIdx.x];
    uint32_t *SA = a.saux + a.soff[blockIdx.x], *PA = a.paux + a.poff[blockIdx.x];
    gather_bucket(a.s_in, ns, a.fcap, a.s_spill, a.s_spill_n, a.spill_cap, b, SK, SR, ns2, true, &s_pos);
    gather_bucket(a.p_in, np, a.fcap, a.p_spill, a.p_spill_n, a.spill_cap, b, PK, PR, np2, true, &s_pos);
    bitonic(SK, SR, nullptr, ns2);
    bitonic(PK, PR, nullptr, np2);
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
        const uint64_t k = SK[i];
        const uint32_t lo = lower_u64(PK, 0, np, k), hi = upper_u64(PK, lo, np, k);
        const uint32_t rd = SR[i] & 0x7FFFFFFFu;
        const uint32_t same = has_u32(PR, lo, hi, rd) + has_u32(PR, lo, hi, rd | 0x80000000u);
        SA[i] = (hi - lo) - same;
    }
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        const uint64_t k = PK[i];
        const uint32_t lo = lower_u64(SK, 0, ns, k), hi = upper_u64(SK, lo, ns, k);
        const uint32_t rd = PR[i] & 0x7FFFFFFFu;
        const uint32_t same = has_u32(SR, lo, hi, rd) + has_u32(SR, lo, hi, rd | 0x80000000u);
        PA[i] = ((hi - lo) - same) ? 1u : 0u;
    }
    __syncthreads();
    // exclusive scan in 64 bits: a bucket of 2^31 edges or more (the reference's int edge
    // offsets, overlap.cu:561,567) writes none and sets the step's error flag
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
            atomicAdd(a.ejob + j, tot);
        }
        s_base = base;
        a.edge_off[b] = base;
        a.edge_cnt[b] = base == ~0ull ? 0u : tot;  // unplaced edges are not written (see k_bjoin)
    }
    __syncthreads();
    if (too_many) return;
    const unsigned long long base = s_base;
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
        const uint32_t o = SA[i], e = (i + 1 < ns) ? SA[i + 1] : tot;
        const uint32_t srid = (SR[i] << 1) | (SR[i] >> 31);
        if (e > o && base != ~0ull) {
            const uint64_t k = SK[i];
            unsigned long long pos = base + o;
            for (uint32_t q = lower_u64(PK, 0, np, k); q < np && PK[q] == k && pos < base + e; ++q) {
                const uint32_t prd = (PR[q] << 1) | (PR[q] >> 31);
                if ((prd >> 1) != (srid >> 1)) {
                    a.pool[pos] = make_uint2(srid, prd);
                    ++pos;
                }
            }
        }
        if (e == o) {
            const uint32_t pos = atomicAdd(a.lcnt + j, 1u);
            if (J.keep) {
                Entry en = {(uint32_t)(SK[i] >> 32), (uint32_t)SK[i], srid};
                set_carry(en, bcarry_after(a.cols, J, srid, true));
                a.lists[0][J.seg + pos] = en;
            }
        }
    }
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        if (!PA[i]) {
            const uint32_t prd = (PR[i] << 1) | (PR[i] >> 31);
            const uint32_t pos = atomicAdd(a.lcnt + a.njobs + j, 1u);
            if (J.keep) {
                Entry en = {(uint32_t)(PK[i] >> 32), (uint32_t)PK[i], prd};
                set_carry(en, bcarry_after(a.cols, J, prd, false));
                a.lists[1][J.seg + pos] = en;
            }
        }
    }
}

__global__ void __launch_bounds__(1024) k_bdedup_big(BBigArgs a) {
    __shared__ uint32_t s_pos;
    const uint32_t b = a.ids[blockIdx.x];
    const uint32_t j = job_of(a.fpre, a.njobs, b);
    const uint32_t n = a.s_cnt[b], n2 = pow2_at_least(n);
    uint64_t *K = a.skey + a.soff[blockIdx.x];
    uint32_t *R = a.srank + a.soff[blockIdx.x];
    gather_bucket(a.s_in, n, a.fcap, a.s_spill, a.s_spill_n, a.spill_cap, b, K, R, n2, false, &s_pos);
    bitonic(K, R, nullptr, n2);
    const uint64_t seg = a.jobs[j].seg;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        if (i == 0 || K[i] != K[i - 1]) {
            const uint32_t pos = atomicAdd(a.lcnt + j, 1u);
            a.lists[0][seg + pos] = Entry{(uint32_t)(K[i] >> 32), (uint32_t)K[i], R[i]};
        }
    }
}

// ---------------------------------------------------------------------------------------------
hipError_t launch_bpart(const BPartArgs &a, int pass, bool regen, uint32_t groups, hipStream_t s) {
    if (groups == 0) return hipSuccess;
    if (pass == 2)
        hipLaunchKernelGGL(k_bpart2, dim3(groups), dim3(PART_THREADS), 0, s, a);
    else if (regen)
        hipLaunchKernelGGL(k_bpart1<true>, dim3(groups), dim3(PART_THREADS), 0, s, a);
    else
        hipLaunchKernelGGL(k_bpart1<false>, dim3(groups), dim3(PART_THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bdedup(const BJoinArgs &a, uint32_t nb, uint32_t fcap, hipStream_t s) {
    if (nb == 0) return hipSuccess;
    if (fcap == (uint32_t)BCAP_L)
        hipLaunchKernelGGL((k_bdedup<BCAP_L, BJOIN_THREADS_L>), dim3(nb), dim3(BJOIN_THREADS_L), 0, s, a);
    else if (fcap == (uint32_t)BCAP)
        hipLaunchKernelGGL((k_bdedup<BCAP, BJOIN_THREADS>), dim3(nb), dim3(BJOIN_THREADS), 0, s, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_bjoin(const BJoinArgs &a, uint32_t nb, uint32_t fcap, hipStream_t s) {
    if (nb == 0) return hipSuccess;
    if (fcap == (uint32_t)BCAP_L)
        hipLaunchKernelGGL((k_bjoin<BCAP_L, BMCAP_L, BJOIN_THREADS_L>), dim3(nb), dim3(BJOIN_THREADS_L), 0, s, a);
    else if (fcap == (uint32_t)BCAP)
        hipLaunchKernelGGL((k_bjoin<BCAP, BMCAP, BJOIN_THREADS>), dim3(nb), dim3(BJOIN_THREADS), 0, s, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_bdedup_big(const BBigArgs &a, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bdedup_big, dim3(n), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bjoin_big(const BBigArgs &a, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bjoin_big, dim3(n), dim3(1024), 0, s, a);
    return hipGetLastError();
}

}  // namespace emp
