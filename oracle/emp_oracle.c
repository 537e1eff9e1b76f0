/*
 * TEST INFRASTRUCTURE ONLY — the oracle. CPU restatement of empanada's fingerprint -> sort ->
 * overlap path; the parity checker and host-core baseline for the HIP library, never product
 * code. Every function cites the reference file:line whose behaviour it restates.
 *
 * Pinned by: Appendix-style known answers computed from the closed forms (tests/test_oracle.py),
 * the rolling-update invariants NextSfx(key(s)) == key(s[1:]) and PrevPfx(key(s)) == key(s[:-1])
 * away from the wrap, the exact u64-wrap value of PrevPfx, a brute-force string checker on small
 * read sets (every suffix/prefix equality by string comparison, greedy removal, reference edge
 * order), and hand-built FASTQ cases for the reader's block, seek-back and N-filter rules.
 */
#define _FILE_OFFSET_BITS 64
#define _POSIX_C_SOURCE 200809L
#include "emp_oracle.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/types.h>

/* assembly.cuh:29-34 */
#define PRIME1 536870909u
#define PRIME2 536870879u
#define RADIX1 5u
#define RADIX2 7u
#define MAX_RD_LEN 255
/* assembly.cuh:34 */
#define ENCODING(c) ((((c) >> 1) & 3) + 1)
/* assembly.cuh:50-52 */
#define HIGH32(x) (((x) >> 32u) & 0xffffffffu)
#define LOW32(x) ((x)&0xffffffffu)
#define APPEND64(h, l) ((((uint64_t)(h)&0xffffffffu) << 32u) | ((uint64_t)(l)&0xffffffffu))

/* preprocess.cu:18 */
static const char h_revc[8] = {0, 'T', 0, 'G', 'A', 0, 0, 'C'};

static uint32_t g_mod1[MAX_RD_LEN + 1], g_mod2[MAX_RD_LEN + 1];
static uint32_t g_inv1, g_inv2;
static int g_init = 0;

#define EO_RS_NAME sort_u64_u32
#define EO_RS_KT uint64_t
#define EO_RS_VT uint32_t
#include "eo_radix.inc"

#define EO_RS_NAME sort_u32_u64
#define EO_RS_KT uint32_t
#define EO_RS_VT uint64_t
#include "eo_radix.inc"

#define EO_RS_NAME sort_u32_u8
#define EO_RS_KT uint32_t
#define EO_RS_VT uint8_t
#include "eo_radix.inc"

char eo_revc(char c) { return h_revc[c & 7]; }

/* overlap.cu:30-39 */
uint32_t eo_modulo_exp(uint64_t a, unsigned n, unsigned p) {
    uint64_t res = 1;
    while (n > 0) {
        if (n & 1) res = (res * a) % p;
        a = (a * a) % p;
        n >>= 1;
    }
    return (uint32_t)res;
}

/* fingerprint.cu:26-34 (u32 products cannot overflow: R * (P-1) < 2^32) */
static void eo_init(void) {
    if (g_init) return;
    g_mod1[0] = 1;
    g_mod2[0] = 1;
    for (int i = 1; i <= MAX_RD_LEN; ++i) {
        g_mod1[i] = (g_mod1[i - 1] * RADIX1) % PRIME1;
        g_mod2[i] = (g_mod2[i - 1] * RADIX2) % PRIME2;
    }
    /* overlap.cu:121-122 */
    g_inv1 = eo_modulo_exp(RADIX1, PRIME1 - 2, PRIME1);
    g_inv2 = eo_modulo_exp(RADIX2, PRIME2 - 2, PRIME2);
    g_init = 1;
}

void eo_modulos(uint32_t *m1, uint32_t *m2) {
    eo_init();
    memcpy(m1, g_mod1, sizeof g_mod1);
    memcpy(m2, g_mod2, sizeof g_mod2);
}

/* FingerprintMul, fingerprint.cu:36-57: (la,a) (+) (lb,b) = (la+lb, a*R^lb + b mod P) per half. */
static inline uint64_t fp_mul(uint64_t a, uint64_t b, unsigned lb) {
    uint64_t fp1a = HIGH32(a), fp1b = HIGH32(b), fp2a = LOW32(a), fp2b = LOW32(b);
    uint64_t n1 = (((fp1a * g_mod1[lb]) % PRIME1) + fp1b) % PRIME1;
    uint64_t n2 = (((fp2a * g_mod2[lb]) % PRIME2) + fp2b) % PRIME2;
    return APPEND64(n1, n2);
}

/* EncodeBases, fingerprint.cu:59-65 */
static inline uint64_t fp_base(char c) {
    uint64_t e = (uint64_t)ENCODING(c);
    return APPEND64(e, e);
}

/*
 * The inclusive_scan_by_key of fingerprint.cu:90-92 restricted to one read; its last element is
 * the whole-read key kept at fingerprint.cu:97-105. FingerprintMul is associative and exact, so
 * the left-to-right fold equals any scan order.
 */
uint64_t eo_key_of(const char *s, int len) {
    eo_init();
    uint64_t acc = fp_base(s[0]);
    for (int i = 1; i < len; ++i) acc = fp_mul(acc, fp_base(s[i]), 1);
    return acc;
}

/* Key of the reverse complement, built as preprocess.cu:254-275 does (d_revc[c&7], reversed). */
uint64_t eo_key_of_rc(const char *s, int len) {
    eo_init();
    uint64_t acc = fp_base(h_revc[s[len - 1] & 7]);
    for (int i = 1; i < len; ++i) acc = fp_mul(acc, fp_base(h_revc[s[len - 1 - i] & 7]), 1);
    return acc;
}

/* NextSfx, overlap.cu:57-75 */
uint64_t eo_next_sfx(uint64_t key, char base, int len) {
    eo_init();
    uint64_t fp1 = HIGH32(key), fp2 = LOW32(key);
    uint32_t e = (uint32_t)ENCODING(base);
    fp1 = (fp1 + PRIME1 - (e * g_mod1[len - 1]) % PRIME1) % PRIME1;
    fp2 = (fp2 + PRIME2 - (e * g_mod2[len - 1]) % PRIME2) % PRIME2;
    return APPEND64(fp1, fp2);
}

/* PrevPfx, overlap.cu:77-91, with its uint64_t wrap when fp < ENCODING(base). */
uint64_t eo_prev_pfx(uint64_t key, char base) {
    eo_init();
    uint64_t fp1 = HIGH32(key), fp2 = LOW32(key);
    uint64_t e = (uint64_t)(int64_t)ENCODING(base);
    fp1 = ((uint64_t)g_inv1 * (fp1 - e)) % PRIME1;
    fp2 = ((uint64_t)g_inv2 * (fp2 - e)) % PRIME2;
    return APPEND64(fp1, fp2);
}

void eo_buf_free(eo_buf *b) {
    free(b->data);
    b->data = NULL;
    b->size = b->cap = 0;
}

static int buf_put(eo_buf *b, const void *p, size_t n) {
    if (b->size + n > b->cap) {
        size_t nc = b->cap ? b->cap : 4096;
        while (nc < b->size + n) nc *= 2;
        unsigned char *d = (unsigned char *)realloc(b->data, nc);
        if (!d) return EO_ERR_NOMEM;
        b->data = d;
        b->cap = nc;
    }
    memcpy(b->data + b->size, p, n);
    b->size += n;
    return EO_OK;
}

const char *eo_strerror(int code) {
    switch (code) {
        case EO_OK: return "ok";
        case EO_ERR_IO: return "I/O error";
        case EO_ERR_FORMAT: return "malformed FASTQ (block does not start with '@', or a record is longer than a block)";
        case EO_ERR_READLEN: return "kept reads differ in length or exceed 255 bases";
        case EO_ERR_BASE: return "kept read holds a byte other than ACGTacgt";
        case EO_ERR_OUT_OF_CORE: return "out-of-core order: unused";
        case EO_ERR_PARTITION: return "more than M/2 equal keys in a partition";
        case EO_ERR_NOMEM: return "out of memory";
        case EO_ERR_ARG: return "bad argument";
        case EO_ERR_TOO_MANY: return "too many reads for 32-bit read ids";
        default: return "unknown error";
    }
}

/* ------------------------------------------------------------------------------------------ */
/* FASTQ reader: read_block (preprocess.cu:195-237) and its helpers.                          */

typedef struct {
    FILE *f;
    const char *mem;
    uint64_t size, pos;
} src_t;

static size_t src_read(src_t *s, char *dst, size_t n) {
    if (s->f) return fread(dst, 1, n, s->f);
    uint64_t left = s->size - s->pos;
    if (n > left) n = (size_t)left;
    memcpy(dst, s->mem + s->pos, n);
    s->pos += n;
    return n;
}

static int src_back(src_t *s, size_t nb) {
    if (s->f) return fseeko(s->f, -(off_t)nb, SEEK_CUR) == 0 ? EO_OK : EO_ERR_IO;
    s->pos -= nb;
    return EO_OK;
}

typedef struct {
    char *reads;
    uint64_t n, cap_bytes;
    int L;
} kept_t;

/*
 * One sequence line (its bytes without the '\n'). removeLQReads (preprocess.cu:98-182): markNs
 * (c>>6)*(1+(c&8)) reaches 9 exactly for a byte in [64,127] with bit 3 set; such a read is
 * dropped; '\n' is removed; an empty line leaves no read. The build accepts only ACGTacgt in a
 * kept read (other bytes reach the reference's rounds through a block-position-dependent path).
 */
static int keep_line(kept_t *k, const char *p, size_t len) {
    if (len == 0) return EO_OK;
    for (size_t i = 0; i < len; ++i) {
        unsigned char c = (unsigned char)p[i];
        if (c >= 64 && c < 128 && (c & 8)) return EO_OK; /* dropped */
    }
    for (size_t i = 0; i < len; ++i) {
        switch (p[i]) {
            case 'A': case 'C': case 'G': case 'T':
            case 'a': case 'c': case 'g': case 't': break;
            default: return EO_ERR_BASE;
        }
    }
    /* fingerprint.cu:130-135, main.cu:48, assembly.cuh:17 */
    if (k->L == 0) {
        if (len > MAX_RD_LEN) return EO_ERR_READLEN;
        k->L = (int)len;
    } else if ((size_t)k->L != len) {
        return EO_ERR_READLEN;
    }
    if ((k->n + 1) * (uint64_t)k->L > k->cap_bytes) {
        uint64_t nc = k->cap_bytes ? k->cap_bytes * 2 : (uint64_t)k->L * 4096;
        char *d = (char *)realloc(k->reads, nc);
        if (!d) return EO_ERR_NOMEM;
        k->reads = d;
        k->cap_bytes = nc;
    }
    memcpy(k->reads + k->n * (uint64_t)k->L, p, len);
    k->n++;
    return EO_OK;
}

static int parse_src(src_t *src, uint64_t block_bytes, char **reads, uint64_t *n, int *L) {
    if (block_bytes == 0) return EO_ERR_ARG;
    char *buf = (char *)malloc(block_bytes);
    if (!buf) return EO_ERR_NOMEM;
    kept_t k = {NULL, 0, 0, 0};
    int rc = EO_OK;
    for (;;) {
        /* preprocess.cu:200-206 */
        size_t got = src_read(src, buf, block_bytes);
        int done = got != block_bytes;
        if (got == 0) break;
        /* preprocess.cu:207 */
        if (buf[0] != '@') { rc = EO_ERR_FORMAT; break; }
        /* parseLines, preprocess.cu:26-37: line index of byte i = newlines before i. */
        uint64_t nl_before_last = 0;
        for (size_t i = 0; i + 1 < got; ++i) nl_before_last += buf[i] == '\n';
        uint64_t n_lines = nl_before_last + 1;
        size_t use = got;
        /* preprocess.cu:219-224 with resizeBuffer (preprocess.cu:46-62) */
        if (!done && (n_lines % 4 != 0 || buf[got - 1] != '\n')) {
            uint64_t keep = ((n_lines - 1) / 4) * 4;
            if (keep == 0) { rc = EO_ERR_FORMAT; break; } /* the reference re-reads forever */
            uint64_t seen = 0;
            size_t i = 0;
            for (; i < got; ++i)
                if (buf[i] == '\n' && ++seen == keep) break;
            use = i + 1;
            rc = src_back(src, got - use);
            if (rc) break;
        }
        /* extractReads (preprocess.cu:70-87): keep lines with index % 4 == 1. */
        uint64_t line = 0;
        size_t start = 0;
        for (size_t i = 0; i <= use && rc == EO_OK; ++i) {
            if (i == use || buf[i] == '\n') {
                if (i > start || i < use) {
                    if (line % 4 == 1) rc = keep_line(&k, buf + start, i - start);
                }
                ++line;
                start = i + 1;
            }
        }
        if (rc) break;
        if (done) break;
    }
    free(buf);
    if (rc) {
        free(k.reads);
        return rc;
    }
    *reads = k.reads;
    *n = k.n;
    *L = k.L;
    return EO_OK;
}

int eo_parse_fastq_file(const char *path, uint64_t block_bytes, char **reads, uint64_t *n, int *L) {
    src_t s = {fopen(path, "rb"), NULL, 0, 0};
    if (!s.f) return EO_ERR_IO;
    int rc = parse_src(&s, block_bytes, reads, n, L);
    fclose(s.f);
    return rc;
}

int eo_parse_fastq_mem(const char *data, uint64_t size, uint64_t block_bytes, char **reads, uint64_t *n,
                       int *L) {
    src_t s = {NULL, data, size, 0};
    return parse_src(&s, block_bytes, reads, n, L);
}

/* ------------------------------------------------------------------------------------------ */
/* Overlap phase.                                                                              */

static size_t lower_bound_u64(const uint64_t *a, size_t n, uint64_t key) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

static size_t upper_bound_u64(const uint64_t *a, size_t n, uint64_t key) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (key < a[mid]) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

static int binary_search_u32(const uint32_t *a, size_t n, uint32_t key) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && a[lo] == key;
}

typedef struct {
    uint64_t *fp;
    uint32_t *rid;
    size_t n;
} list_t;

/* ------------------------------------------------------------------------------------------ */
/* The out-of-core sort (cqutils.cuh:36-84,139-247) and unique (cqutils.cuh:86-105,271-306).   */

/* upper_bound over a[lo, hi) */
static size_t upper_in(const uint64_t *a, size_t lo, size_t hi, uint64_t key) {
    return lo + upper_bound_u64(a + lo, hi - lo, key);
}

/*
 * cq_get_merge_partitions (cqutils.cuh:36-84): partitions of the key-sorted runs a and b, each
 * side at most M/2 per partition, in merge order. Case 1 (a's window is not above b's first key):
 * a's window alone; case 2 (b's window is not above a's first key): b's window alone, so b's
 * entries equal to a's first key come first; case 3: both windows, the one with the larger last
 * key cut by upper_bound at the other's last key; the rest once both windows are short.
 */
int eo_merge_partitions(const uint64_t *a, size_t nA, const uint64_t *b, size_t nB, size_t M, size_t *parts,
                        size_t cap, size_t *n_parts) {
    size_t np = 0, offA = 0, offB = 0;
    const size_t H = M / 2;
    *n_parts = 0;
    if (H == 0) return EO_ERR_ARG;
#define EO_PUSH(x, y) do { if (np >= cap) return EO_ERR_ARG; parts[2 * np] = (x); parts[2 * np + 1] = (y); ++np; } while (0)
    while (offA < nA && offB < nB) {
        size_t nPa = nA - offA < H ? nA - offA : H, nPb = nB - offB < H ? nB - offB : H;
        if (nPa < H && nPb < H) {
            EO_PUSH(nPa, nPb);
            offA += nPa; offB += nPb;
            break;
        }
        if (a[offA + nPa - 1] <= b[offB]) {                      /* case 1 */
            EO_PUSH(nPa, 0);
            offA += nPa;
        } else if (b[offB + nPb - 1] <= a[offA]) {               /* case 2 */
            EO_PUSH(0, nPb);
            offB += nPb;
        } else {                                                   /* case 3 */
            const uint64_t la = a[offA + nPa - 1], lb = b[offB + nPb - 1];
            if (la < lb) nPb = upper_in(b, offB, offB + nPb, la) - offB;
            else if (lb < la) nPa = upper_in(a, offA, offA + nPa, lb) - offA;
            EO_PUSH(nPa, nPb);
            offA += nPa; offB += nPb;
        }
    }
    if (offA == nA && offB < nB) EO_PUSH(0, nB - offB);
    else if (offB == nB && offA < nA) EO_PUSH(nA - offA, 0);
#undef EO_PUSH
    *n_parts = np;
    return EO_OK;
}

typedef struct {
    uint64_t *k;
    uint32_t *v;
    size_t n;
} run_t;

/*
 * cq_merge_by_key (cqutils.cuh:139-177): per partition, a mixed one is sorted as [a part; b part]
 * (thrust::sort_by_key, stable: equal keys keep a before b), a one-sided one copied.
 */
static int merge_runs(const run_t *A, const run_t *B, size_t M, run_t *C) {
    size_t cap = 4 + 4 * ((A->n + B->n) / (M / 2 ? M / 2 : 1));
    size_t *parts = (size_t *)malloc(2 * cap * sizeof(size_t)), np = 0;
    C->n = A->n + B->n;
    C->k = (uint64_t *)malloc((C->n ? C->n : 1) * sizeof(uint64_t));
    C->v = (uint32_t *)malloc((C->n ? C->n : 1) * sizeof(uint32_t));
    if (!parts || !C->k || !C->v) { free(parts); return EO_ERR_NOMEM; }
    int rc = eo_merge_partitions(A->k, A->n, B->k, B->n, M, parts, cap, &np);
    if (rc) { free(parts); return rc; }
    size_t oa = 0, ob = 0, oc = 0;
    for (size_t q = 0; q < np; ++q) {
        const size_t na = parts[2 * q], nb = parts[2 * q + 1];
        memcpy(C->k + oc, A->k + oa, na * sizeof(uint64_t));
        memcpy(C->v + oc, A->v + oa, na * sizeof(uint32_t));
        memcpy(C->k + oc + na, B->k + ob, nb * sizeof(uint64_t));
        memcpy(C->v + oc + na, B->v + ob, nb * sizeof(uint32_t));
        if (na && nb) sort_u64_u32(C->k + oc, C->v + oc, na + nb);
        oa += na; ob += nb; oc += na + nb;
    }
    free(parts);
    return EO_OK;
}

/*
 * cq_sort_by_key (cqutils.cuh:179-247) on (u64 key, u32 value) pairs: n <= M one stable sort;
 * otherwise chunks of M in input order sorted each, then the FIFO of sorted runs merged pairwise
 * (the front two, the result pushed to the back) until one run is left. Sorts keys/vals in place.
 */
int eo_cq_sort_by_key(uint64_t *keys, uint32_t *vals, size_t n, size_t M) {
    if (n <= M) {
        sort_u64_u32(keys, vals, n);
        return EO_OK;
    }
    const size_t F = (n + M - 1) / M;
    run_t *q = (run_t *)malloc((2 * F) * sizeof(run_t));
    if (!q) return EO_ERR_NOMEM;
    size_t head = 0, tail = 0;
    for (size_t c = 0; c < F; ++c) {
        const size_t lo = c * M, m = n - lo < M ? n - lo : M;
        run_t r;
        r.n = m;
        r.k = (uint64_t *)malloc(m * sizeof(uint64_t));
        r.v = (uint32_t *)malloc(m * sizeof(uint32_t));
        memcpy(r.k, keys + lo, m * sizeof(uint64_t));
        memcpy(r.v, vals + lo, m * sizeof(uint32_t));
        sort_u64_u32(r.k, r.v, m);
        q[tail++] = r;
    }
    int rc = EO_OK;
    while (tail - head > 1 && !rc) {
        run_t A = q[head++], B = q[head++], C;
        rc = merge_runs(&A, &B, M, &C);
        free(A.k); free(A.v); free(B.k); free(B.v);
        if (!rc) q[tail++] = C;
    }
    if (!rc) {
        memcpy(keys, q[head].k, n * sizeof(uint64_t));
        memcpy(vals, q[head].v, n * sizeof(uint32_t));
    }
    for (size_t i = head; i < tail; ++i) { free(q[i].k); free(q[i].v); }
    free(q);
    return rc;
}

/*
 * cq_get_partitions (cqutils.cuh:86-105) + cq_unique_by_key (cqutils.cuh:271-306) over a
 * key-sorted list: windows of M cut at lower_bound of the window's last key (a window of one key
 * never ends: the reference loops, EO_ERR_PARTITION here), the first entry of each key kept in
 * each partition. Partitions never split a key, so this keeps the first of every key.
 */
static int unique_partitioned(list_t *l, size_t M) {
    size_t off = 0, w = 0;
    while (off < l->n) {
        size_t np;
        if (l->n - off <= M) {
            np = l->n - off;
        } else {
            np = lower_bound_u64(l->fp + off, M, l->fp[off + M - 1]);
            if (np == 0) return EO_ERR_PARTITION;
        }
        for (size_t i = off; i < off + np; ++i)
            if (i == off || l->fp[i] != l->fp[i - 1]) { l->fp[w] = l->fp[i]; l->rid[w] = l->rid[i]; ++w; }
        off += np;
    }
    l->n = w;
    return EO_OK;
}

/*
 * regenerate_fingerprints (overlap.cu:605-667): sort by read id (overlap.cu:613; the ids are
 * distinct, so the chunked sort of cq_sort_by_key gives the fully sorted order too), then per
 * chunk of M entries (overlap.cu:623-659): partition even/odd and sort each half by id
 * (overlap.cu:636-642; an even/odd split of an id-sorted chunk keeps both halves id-sorted),
 * gather one base per entry from the column (filter_bases, overlap.cu:670-692: base of read
 * rid/2, complemented through d_revc for odd ids), then the rolling transform
 * (overlap.cu:647,651). The list ends as chunk 0's evens, chunk 0's odds, chunk 1's evens, ...
 */
static void regenerate(list_t *l, const char *fwd, const char *rev, int is_sfx, int len, size_t M) {
    size_t n = l->n;
    if (n == 0) return;
    sort_u32_u64(l->rid, l->fp, n);
    uint64_t *fp2 = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint32_t *rid2 = (uint32_t *)malloc(n * sizeof(uint32_t));
    for (size_t lo = 0; lo < n; lo += M) {
        const size_t hi = n - lo < M ? n : lo + M;
        size_t w = lo;
        for (size_t i = lo; i < hi; ++i)
            if (!(l->rid[i] & 1)) { fp2[w] = l->fp[i]; rid2[w] = l->rid[i]; ++w; }
        for (size_t i = lo; i < hi; ++i)
            if (l->rid[i] & 1) { fp2[w] = l->fp[i]; rid2[w] = l->rid[i]; ++w; }
    }
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        size_t i = (size_t)ii;
        uint32_t r = rid2[i];
        char b = !(r & 1) ? fwd[r / 2] : h_revc[rev[r / 2] & 7];
        fp2[i] = is_sfx ? eo_next_sfx(fp2[i], b, len) : eo_prev_pfx(fp2[i], b);
    }
    free(l->fp);
    free(l->rid);
    l->fp = fp2;
    l->rid = rid2;
}

/* remove_used (overlap.cu:694-710): sort the used ids, binary-search each id, stable compaction. */
static size_t remove_used(uint64_t *fp, uint32_t *rid, size_t n, uint32_t *used, size_t nu) {
    uint8_t *dummy = (uint8_t *)calloc(nu ? nu : 1, 1);
    sort_u32_u8(used, dummy, nu);
    free(dummy);
    uint8_t *is_there = (uint8_t *)malloc(n ? n : 1);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) is_there[i] = (uint8_t)binary_search_u32(used, nu, rid[i]);
    size_t w = 0;
    for (size_t i = 0; i < n; ++i)
        if (!is_there[i]) { rid[w] = rid[i]; fp[w] = fp[i]; ++w; }
    free(is_there);
    return w;
}

/*
 * cq_get_search_partitions (cqutils.cuh:107-137): key-aligned partitions of the key-sorted lists a
 * (suffixes) and b (prefixes), at most M/2 entries per side. While a side holds M/2 entries or
 * more, the cut is at m = min(a[offA + M/2' - 1], b[offB + M/2' - 1]) (M/2' = the side's window),
 * lower_bound of m on each side; when both sides hold fewer than M/2, the rest is the last
 * partition. A window where neither side has an entry below m (for example more than M/2 equal
 * keys, or a side whose remaining window is one key not above the other side's first) gives an
 * empty partition and the reference loops forever: that is EO_ERR_PARTITION here. A side that is
 * empty from the start gives one partition of the other side alone. parts receives (na, nb) pairs.
 */
int eo_search_partitions(const uint64_t *a, size_t nA, const uint64_t *b, size_t nB, size_t M, size_t *parts,
                         size_t cap, size_t *n_parts) {
    size_t np = 0, offA = 0, offB = 0;
    const size_t H = M / 2;
    *n_parts = 0;
    while (offA < nA && offB < nB) {
        size_t nPa = nA - offA < H ? nA - offA : H;
        size_t nPb = nB - offB < H ? nB - offB : H;
        if (np >= cap) return EO_ERR_ARG;
        if (nPa < H && nPb < H) {
            parts[2 * np] = nPa; parts[2 * np + 1] = nPb; ++np;
            offA += nPa; offB += nPb;
            break;
        }
        uint64_t ka = a[offA + nPa - 1], kb = b[offB + nPb - 1];
        uint64_t m = ka < kb ? ka : kb;
        nPa = lower_bound_u64(a + offA, nPa, m);
        nPb = lower_bound_u64(b + offB, nPb, m);
        if (nPa == 0 && nPb == 0) return EO_ERR_PARTITION;
        parts[2 * np] = nPa; parts[2 * np + 1] = nPb; ++np;
        offA += nPa; offB += nPb;
    }
    if (offA == nA || offB == nB) {
        if (np >= cap) return EO_ERR_ARG;
        if (offA == nA && offB < nB) { parts[2 * np] = 0; parts[2 * np + 1] = nB - offB; ++np; }
        else if (offB == nB && offA < nA) { parts[2 * np] = nA - offA; parts[2 * np + 1] = 0; ++np; }
    }
    *n_parts = np;
    return EO_OK;
}

/*
 * emit_edges (overlap.cu:361-456) with cq_get_search_partitions (eo_search_partitions) and
 * get_edges (overlap.cu:557-603). Suffix keys are the needles, prefix keys the haystack.
 */
static int emit_edges(list_t *S, list_t *P, int ovlp, size_t M, eo_buf *out, eo_round *st) {
    const size_t nA = S->n, nB = P->n;
    /* at most one partition per M/2 entries of either side, plus the closing one */
    size_t cap = 4 + 2 * ((nA + nB) / (M / 2 ? M / 2 : 1));
    size_t *parts = (size_t *)malloc(2 * cap * sizeof(size_t));
    size_t np = 0;
    if (!parts) return EO_ERR_NOMEM;
    {
        int prc = eo_search_partitions(S->fp, nA, P->fp, nB, M, parts, cap, &np);
        if (prc) { free(parts); return prc == EO_ERR_ARG ? EO_ERR_PARTITION : prc; }
    }

    size_t offA = 0, offB = 0, wA = 0, wB = 0;
    uint64_t ne_total = 0;
    int nrec = 0;
    for (size_t q = 0; q < np; ++q) {
        size_t na = parts[2 * q], nb = parts[2 * q + 1];
        uint64_t *sfp = S->fp + offA, *pfp = P->fp + offB;
        uint32_t *srid = S->rid + offA, *prid = P->rid + offB;
        if (na && nb) {
            /* lower_bound / upper_bound (overlap.cu:413-415) */
            uint32_t *lb = (uint32_t *)malloc(na * sizeof(uint32_t));
            uint64_t *cnt = (uint64_t *)malloc((na + 1) * sizeof(uint64_t));
#pragma omp parallel for schedule(static)
            for (int64_t ii = 0; ii < (int64_t)na; ++ii) {
                size_t l = lower_bound_u64(pfp, nb, sfp[ii]);
                size_t u = upper_bound_u64(pfp, nb, sfp[ii]);
                lb[ii] = (uint32_t)l;
                cnt[ii] = u - l;
            }
            /* get_edges counts and offsets (overlap.cu:561-567) */
            uint64_t ne = 0;
            for (size_t i = 0; i < na; ++i) { uint64_t c = cnt[i]; cnt[i] = ne; ne += c; }
            cnt[na] = ne;
            if (ne >= (1ull << 31)) { free(lb); free(cnt); free(parts); return EO_ERR_PARTITION; }
            uint32_t *from = (uint32_t *)malloc((ne ? ne : 1) * sizeof(uint32_t));
            uint32_t *to = (uint32_t *)malloc((ne ? ne : 1) * sizeof(uint32_t));
            /* expansion (overlap.cu:569-591): from = sfx rid, to = pfx rid[lb], rid[lb+1], ... */
#pragma omp parallel for schedule(dynamic, 4096)
            for (int64_t ii = 0; ii < (int64_t)na; ++ii) {
                uint64_t o = cnt[ii], c = cnt[ii + 1] - cnt[ii];
                for (uint64_t j = 0; j < c; ++j) {
                    from[o + j] = srid[ii];
                    to[o + j] = prid[lb[ii] + j];
                }
            }
            /* stable removal of self edges (overlap.cu:594-601) */
            uint64_t w = 0;
            for (uint64_t i = 0; i < ne; ++i)
                if (from[i] / 2 != to[i] / 2) { from[w] = from[i]; to[w] = to[i]; ++w; }
            /* write_edges_binary (overlap.cu:148-161) */
            uint64_t n64 = w;
            int32_t ov = ovlp;
            int rc = buf_put(out, &n64, sizeof n64);
            if (!rc) rc = buf_put(out, &ov, sizeof ov);
            if (!rc && w) rc = buf_put(out, from, w * sizeof(uint32_t));
            if (!rc && w) rc = buf_put(out, to, w * sizeof(uint32_t));
            if (rc) { free(lb); free(cnt); free(from); free(to); free(parts); return rc; }
            ne_total += w;
            ++nrec;
            /* remove_used on both sides (overlap.cu:424-429) */
            size_t ka = remove_used(sfp, srid, na, from, w);
            size_t kb = remove_used(pfp, prid, nb, to, w);
            memmove(S->fp + wA, sfp, ka * sizeof(uint64_t));
            memmove(S->rid + wA, srid, ka * sizeof(uint32_t));
            memmove(P->fp + wB, pfp, kb * sizeof(uint64_t));
            memmove(P->rid + wB, prid, kb * sizeof(uint32_t));
            wA += ka;
            wB += kb;
            free(lb); free(cnt); free(from); free(to);
        } else if (na) {
            memmove(S->fp + wA, sfp, na * sizeof(uint64_t));
            memmove(S->rid + wA, srid, na * sizeof(uint32_t));
            wA += na;
        } else {
            memmove(P->fp + wB, pfp, nb * sizeof(uint64_t));
            memmove(P->rid + wB, prid, nb * sizeof(uint32_t));
            wB += nb;
        }
        offA += na;
        offB += nb;
    }
    free(parts);
    S->n = wA;
    P->n = wB;
    if (st) {
        st->edges = ne_total;
        st->records = nrec;
    }
    return EO_OK;
}

/*
 * build_graph (overlap.cu:206-293) preceded by the map phase of main.cu:42-58 over kept reads:
 * transpose_reads (fingerprint.cu:125-153) gives the columns; generate_fingerprint_read_pairs
 * (fingerprint.cu:75-115) gives key(read_k) with id 2k and key(revcomp(read_k)) with id 2k+1.
 */
/* 2-bit code of base `pos` of oriented read `rid` (odd ids: reverse complement, code ^ 2). */
static inline unsigned ocode(const char *reads, int L, uint32_t rid, int pos) {
    const char *r = reads + (size_t)(rid >> 1) * (size_t)L;
    return (rid & 1) ? ((((unsigned char)r[L - 1 - pos] >> 1) & 3u) ^ 2u) : (((unsigned char)r[pos] >> 1) & 3u);
}

/*
 * print_false_positive_rate (overlap.cu:342-359) with fpcheck (overlap.cu:320-340): over the
 * key-sorted list, count the distinct keys, and in every key group of more than one entry the
 * distinct substrings [start, start+len) of the oriented reads (a group of one counts one).
 */
static void audit_list(const list_t *l, const char *reads, int L, int start, int len, int list, int ovlp,
                       eo_audit *audit, int audit_cap, int *n_audit) {
    if (!audit || !n_audit || *n_audit >= audit_cap) return;
    uint64_t keys = 0, strings = 0;
    size_t i = 0;
    while (i < l->n) {
        size_t e = i + 1;
        while (e < l->n && l->fp[e] == l->fp[i]) ++e;
        ++keys;
        for (size_t a = i; a < e; ++a) {
            int fresh = 1;
            for (size_t b = i; b < a && fresh; ++b) {
                int same = 1;
                for (int p = 0; p < len && same; ++p)
                    same = ocode(reads, L, l->rid[a], start + p) == ocode(reads, L, l->rid[b], start + p);
                if (same) fresh = 0;
            }
            strings += (uint64_t)fresh;
        }
        i = e;
    }
    eo_audit *r = &audit[(*n_audit)++];
    r->list = list;
    r->ovlp = ovlp;
    r->entries = l->n;
    r->keys = keys;
    r->strings = strings;
}

/*
 * The dedup of build_graph alone (overlap.cu:212-225): the 2n map entries sorted by id, then by key
 * (cq_sort_by_key with buffer M), the first of every key kept (cq_unique_by_key). ids receives
 * the kept ids in key order (2n of room), *nd their count.
 */
int eo_dedup(const char *reads, uint64_t n, int L, size_t M, uint32_t *ids, uint64_t *nd) {
    eo_init();
    *nd = 0;
    if (n == 0) return EO_OK;
    if (L <= 0 || L > MAX_RD_LEN || M < 2) return EO_ERR_ARG;
    list_t S;
    S.n = 2 * n;
    S.fp = (uint64_t *)malloc(S.n * sizeof(uint64_t));
    S.rid = (uint32_t *)malloc(S.n * sizeof(uint32_t));
    if (!S.fp || !S.rid) { free(S.fp); free(S.rid); return EO_ERR_NOMEM; }
    for (uint64_t k = 0; k < n; ++k) {
        S.fp[2 * k] = eo_key_of(reads + k * (uint64_t)L, L);
        S.fp[2 * k + 1] = eo_key_of_rc(reads + k * (uint64_t)L, L);
        S.rid[2 * k] = (uint32_t)(2 * k);
        S.rid[2 * k + 1] = (uint32_t)(2 * k + 1);
    }
    int rc = eo_cq_sort_by_key(S.fp, S.rid, S.n, M);
    if (!rc) rc = unique_partitioned(&S, M);
    if (!rc) {
        memcpy(ids, S.rid, S.n * sizeof(uint32_t));
        *nd = S.n;
    }
    free(S.fp);
    free(S.rid);
    return rc;
}

int eo_overlap_reads(const char *reads, uint64_t n, int L, int min_ovlp, uint64_t reduce_mem, eo_buf *out,
                     eo_summary *sum, eo_round *rounds) {
    return eo_overlap_reads_audit(reads, n, L, min_ovlp, reduce_mem, out, sum, rounds, NULL, 0, NULL);
}

int eo_overlap_reads_audit(const char *reads, uint64_t n, int L, int min_ovlp, uint64_t reduce_mem, eo_buf *out,
                           eo_summary *sum, eo_round *rounds, eo_audit *audit, int audit_cap, int *n_audit) {
    eo_init();
    if (n_audit) *n_audit = 0;
    if (sum) memset(sum, 0, sizeof *sum);
    if (sum) { sum->n_reads = n; sum->read_len = L; }
    if (n == 0) return EO_OK;
    if (L <= 0 || L > MAX_RD_LEN) return EO_ERR_READLEN;
    const size_t M = (size_t)(reduce_mem / 16);
    const uint64_t N0 = 2 * n;
    if (N0 >= (1ull << 32)) return EO_ERR_TOO_MANY;
    if (M < 2) return EO_ERR_ARG;

    /* transpose_reads: col[j][k] = read_k[j] */
    char *cols = (char *)malloc((size_t)L * n);
    if (!cols) return EO_ERR_NOMEM;
#pragma omp parallel for schedule(static)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk)
        for (int j = 0; j < L; ++j) cols[(size_t)j * n + (size_t)kk] = reads[(size_t)kk * L + j];

    list_t S;
    S.n = N0;
    S.fp = (uint64_t *)malloc(N0 * sizeof(uint64_t));
    S.rid = (uint32_t *)malloc(N0 * sizeof(uint32_t));
#pragma omp parallel for schedule(static)
    for (int64_t kk = 0; kk < (int64_t)n; ++kk) {
        const char *r = reads + (size_t)kk * L;
        S.fp[2 * kk] = eo_key_of(r, L);
        S.fp[2 * kk + 1] = eo_key_of_rc(r, L);
        S.rid[2 * kk] = (uint32_t)(2 * kk);
        S.rid[2 * kk + 1] = (uint32_t)(2 * kk + 1);
    }

    /* overlap.cu:212-213: sort by id (already id-ordered), then sort by key (cq_sort_by_key) */
    sort_u32_u64(S.rid, S.fp, S.n);
    int rc = eo_cq_sort_by_key(S.fp, S.rid, S.n, M);
    if (rc) { free(cols); free(S.fp); free(S.rid); return rc; }
    audit_list(&S, reads, L, 0, L, 0, L, audit, audit_cap, n_audit); /* overlap.cu:215-219 */
    /* cq_unique_by_key (cqutils.cuh:271-306): keep the first of each key */
    if ((rc = unique_partitioned(&S, M))) { free(cols); free(S.fp); free(S.rid); return rc; }
    if (sum) sum->n_dedup = S.n;

    /* overlap.cu:235-242 */
    list_t P;
    P.n = S.n;
    P.fp = (uint64_t *)malloc((S.n ? S.n : 1) * sizeof(uint64_t));
    P.rid = (uint32_t *)malloc((S.n ? S.n : 1) * sizeof(uint32_t));
    memcpy(P.fp, S.fp, S.n * sizeof(uint64_t));
    memcpy(P.rid, S.rid, S.n * sizeof(uint32_t));

    int round = 0;
    /* overlap.cu:258 */
    for (round = 0; (round < L - min_ovlp) && (S.n > 0); ++round) {
        eo_round st;
        memset(&st, 0, sizeof st);
        st.s = S.n;
        st.p = P.n;
        st.ovlp = L - 1 - round;
        const char *c_fwd = cols + (size_t)round * n;
        const char *c_bwd = cols + (size_t)(L - 1 - round) * n;
        /* overlap.cu:266-267 */
        regenerate(&S, c_fwd, c_bwd, 1, L - round, M);
        regenerate(&P, c_bwd, c_fwd, 0, 0, M);
        /* overlap.cu:271-272 */
        if ((rc = eo_cq_sort_by_key(S.fp, S.rid, S.n, M)) || (rc = eo_cq_sort_by_key(P.fp, P.rid, P.n, M))) break;
        /* overlap.cu:274-284 */
        audit_list(&S, reads, L, round + 1, L - 1 - round, 1, L - 1 - round, audit, audit_cap, n_audit);
        audit_list(&P, reads, L, 0, L - 1 - round, 2, L - 1 - round, audit, audit_cap, n_audit);
        /* overlap.cu:288-289 */
        rc = emit_edges(&S, &P, L - 1 - round, M, out, &st);
        if (rc) break;
        st.s_after = S.n;
        st.p_after = P.n;
        if (rounds && round < 256) rounds[round] = st;
        if (sum) { sum->n_edges += st.edges; sum->n_records += (uint64_t)st.records; }
    }
    if (sum) sum->n_rounds = round;
    free(cols);
    free(S.fp); free(S.rid);
    free(P.fp); free(P.rid);
    return rc;
}

int eo_run_mem(const char *fastq, uint64_t size, const eo_opts *opts, eo_buf *out, eo_summary *sum,
               eo_round *rounds) {
    char *reads = NULL;
    uint64_t n = 0;
    int L = 0;
    int rc = eo_parse_fastq_mem(fastq, size, opts->map_mem / 16, &reads, &n, &L);
    if (rc) return rc;
    rc = eo_overlap_reads(reads, n, L, opts->min_ovlp, opts->reduce_mem, out, sum, rounds);
    free(reads);
    return rc;
}

int eo_run_file(const char *in_path, const char *out_path, const eo_opts *opts, eo_summary *sum,
                eo_round *rounds) {
    char *reads = NULL;
    uint64_t n = 0;
    int L = 0;
    int rc = eo_parse_fastq_file(in_path, opts->map_mem / 16, &reads, &n, &L);
    if (rc) return rc;
    eo_buf out = {NULL, 0, 0};
    rc = eo_overlap_reads(reads, n, L, opts->min_ovlp, opts->reduce_mem, &out, sum, rounds);
    free(reads);
    if (rc) { eo_buf_free(&out); return rc; }
    FILE *f = fopen(out_path, "wb");
    if (!f) { eo_buf_free(&out); return EO_ERR_IO; }
    if (out.size && fwrite(out.data, 1, out.size, f) != out.size) rc = EO_ERR_IO;
    if (fclose(f) != 0) rc = EO_ERR_IO;
    eo_buf_free(&out);
    return rc;
}
