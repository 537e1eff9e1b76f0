/*
 * Synthetic genomes and reads for the fingerprint -> sort -> overlap path.
 *
 * The reference ships no generator and no data; every input used by the tests and the
 * benchmark comes from here. All randomness is xoshiro256** seeded through SplitMix64.
 *
 *   genome  : i.i.d. uniform ACGT. One 64-bit draw gives 32 bases, base j of a draw being
 *             bits (2j+1:2j) of the draw, mapped through "ACGT".
 *   repeats : optional families (consensus of length len_f, i.i.d.) whose copies are written
 *             over the genome at non-overlapping positions, on a random strand, each copy with
 *             independent per-base substitutions at rate div_f.
 *   reads   : read i uses its own generator, seeded with seed ^ (i * 0x9E3779B97F4A7C15), so the
 *             set is reproducible in parallel. Draw order per read: start (Lemire multiply-shift
 *             over [0, G-L]); strand (only if p_rc > 0: uniform53 < p_rc means reverse
 *             complement); per base (only if p_sub > 0): uniform53 < p_sub substitutes the base
 *             with one of the other three, (code + 1 + mulhi32(draw>>32, 3)) & 3 over ACGT codes.
 *   FASTQ   : "@r<k>\n<seq>\n+\n<'I' x L>\n".
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t sm64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

typedef struct { uint64_t s[4]; } xo_t;

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void xo_seed(xo_t *r, uint64_t seed) {
    uint64_t s = seed;
    for (int i = 0; i < 4; ++i) r->s[i] = sm64(&s);
}

static inline uint64_t xo_next(xo_t *r) {
    uint64_t *s = r->s;
    uint64_t res = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return res;
}

static inline double xo_u53(xo_t *r) { return (double)(xo_next(r) >> 11) * 0x1.0p-53; }

static inline uint64_t xo_below(xo_t *r, uint64_t m) {
    return (uint64_t)(((unsigned __int128)xo_next(r) * m) >> 64);
}

static const char ACGT[4] = {'A', 'C', 'G', 'T'};

static inline int code_of(char c) {
    switch (c) {
        case 'A': return 0;
        case 'C': return 1;
        case 'G': return 2;
        default: return 3;
    }
}

static inline char comp_of(char c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        default: return 'A';
    }
}

/* Fill out[0..G) with i.i.d. uniform ACGT. */
void empgen_genome(uint64_t seed, uint64_t G, char *out) {
    xo_t r;
    xo_seed(&r, seed);
    uint64_t i = 0;
    while (i < G) {
        uint64_t u = xo_next(&r);
        for (int j = 0; j < 32 && i < G; ++j, ++i) out[i] = ACGT[(u >> (2 * j)) & 3];
    }
}

/*
 * Overwrite a fraction `frac` of the genome with repeat copies. Families of consensus length
 * len_f, `copies` copies each, per-copy substitution rate div_f. A copy takes a slot of a grid of
 * stride len_f, chosen uniformly by rejection among the slots whose `unit`-base cells (len_f a
 * multiple of unit) are all free in `used` (one byte per cell, n_cells of them, shared by every
 * family size), so copies of any sizes never overlap. Returns the number of families written.
 */
uint64_t empgen_add_repeats(uint64_t seed, char *genome, uint64_t G, double frac, uint64_t len_f,
                            uint64_t copies, double div_f, unsigned char *used, uint64_t n_cells, uint64_t unit) {
    if (len_f == 0 || copies == 0 || frac <= 0.0 || unit == 0 || len_f % unit) return 0;
    const uint64_t per = len_f / unit;
    uint64_t slots = G / len_f;
    if (slots * per > n_cells) slots = n_cells / per;
    uint64_t free_slots = 0;
    for (uint64_t s = 0; s < slots; ++s) {
        uint64_t k = 0;
        while (k < per && !used[s * per + k]) ++k;
        free_slots += k == per;
    }
    uint64_t fams = (uint64_t)(frac * (double)G / (double)(len_f * copies));
    if (fams * copies > free_slots / 2) fams = free_slots / (2 * copies);
    xo_t r;
    xo_seed(&r, seed);
    char *cons = (char *)malloc(len_f);
    for (uint64_t f = 0; f < fams; ++f) {
        for (uint64_t j = 0; j < len_f; ++j) cons[j] = ACGT[xo_next(&r) >> 62];
        for (uint64_t c = 0; c < copies; ++c) {
            uint64_t s, k;
            do {
                s = xo_below(&r, slots);
                k = 0;
                while (k < per && !used[s * per + k]) ++k;
            } while (k < per);
            memset(used + s * per, 1, per);
            int rc = (xo_next(&r) >> 63) != 0;
            char *dst = genome + s * len_f;
            for (uint64_t j = 0; j < len_f; ++j) {
                char b = rc ? comp_of(cons[len_f - 1 - j]) : cons[j];
                if (div_f > 0.0 && xo_u53(&r) < div_f) {
                    int code = code_of(b);
                    code = (code + 1 + (int)(((xo_next(&r) >> 32) * 3) >> 32)) & 3;
                    b = ACGT[code];
                }
                dst[j] = b;
            }
        }
    }
    free(cons);
    return fams;
}

/* Sample n reads of length L into out (row-major n x L, no separators). */
void empgen_reads(uint64_t seed, const char *genome, uint64_t G, uint64_t n, int L, double p_rc,
                  double p_sub, char *out) {
    if (n == 0 || (uint64_t)L > G) return;
#pragma omp parallel for schedule(static, 4096)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint64_t i = (uint64_t)ii;
        xo_t r;
        xo_seed(&r, seed ^ (i * 0x9E3779B97F4A7C15ull));
        uint64_t start = xo_below(&r, G - (uint64_t)L + 1);
        int rc = 0;
        if (p_rc > 0.0) rc = xo_u53(&r) < p_rc;
        char *dst = out + i * (uint64_t)L;
        const char *src = genome + start;
        if (rc)
            for (int j = 0; j < L; ++j) dst[j] = comp_of(src[L - 1 - j]);
        else
            memcpy(dst, src, (size_t)L);
        if (p_sub > 0.0) {
            for (int j = 0; j < L; ++j) {
                if (xo_u53(&r) < p_sub) {
                    int code = code_of(dst[j]);
                    code = (code + 1 + (int)(((xo_next(&r) >> 32) * 3) >> 32)) & 3;
                    dst[j] = ACGT[code];
                }
            }
        }
    }
}

/* Bytes of the FASTQ rendering of n reads of length L. */
uint64_t empgen_fastq_size(uint64_t n, int L) {
    uint64_t total = 0, k = 0, p10 = 1, digits = 1;
    while (k < n) {
        uint64_t next = p10 * 10;
        uint64_t hi = next < n ? next : n;
        total += (hi - k) * (3 + digits + (uint64_t)L + 1 + 2 + (uint64_t)L + 1);
        k = hi;
        p10 = next;
        ++digits;
    }
    return total;
}

/* Render reads as FASTQ into out (size empgen_fastq_size). Returns bytes written. */
uint64_t empgen_fastq(const char *reads, uint64_t n, int L, char *out) {
    char *p = out;
    for (uint64_t k = 0; k < n; ++k) {
        p += sprintf(p, "@r%llu\n", (unsigned long long)k);
        memcpy(p, reads + k * (uint64_t)L, (size_t)L);
        p += L;
        *p++ = '\n';
        *p++ = '+';
        *p++ = '\n';
        memset(p, 'I', (size_t)L);
        p += L;
        *p++ = '\n';
    }
    return (uint64_t)(p - out);
}

/* Write reads as a FASTQ file. Returns 0 on success. */
int empgen_write_fastq(const char *path, const char *reads, uint64_t n, int L) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    char *line = (char *)malloc((size_t)(2 * L + 64));
    for (uint64_t k = 0; k < n; ++k) {
        int h = sprintf(line, "@r%llu\n", (unsigned long long)k);
        memcpy(line + h, reads + k * (uint64_t)L, (size_t)L);
        char *q = line + h + L;
        *q++ = '\n';
        *q++ = '+';
        *q++ = '\n';
        memset(q, 'I', (size_t)L);
        q += L;
        *q++ = '\n';
        if (fwrite(line, 1, (size_t)(q - line), f) != (size_t)(q - line)) {
            free(line);
            fclose(f);
            return -1;
        }
    }
    free(line);
    return fclose(f) == 0 ? 0 : -1;
}
