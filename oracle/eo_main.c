/*
 * TEST INFRASTRUCTURE ONLY — command-line front end of the oracle, with the reference's CLI
 * (config.h:120-140): emp_oracle <input fastq> <output> [-o N] [-m S] [-r S] [-b S] [-c S].
 * -b and -c are accepted and ignored (they size host blocks and do not change the output).
 * Prints the in-process end-to-end time (FASTQ open to output closed) on stderr.
 */
#define _POSIX_C_SOURCE 199309L
#include <ctype.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "emp_oracle.h"

/* hmsize2bytes, config.h:75-100 */
static int size_arg(const char *s, uint64_t *out) {
    size_t n = strlen(s);
    if (n == 0) return -1;
    char unit = s[n - 1];
    char *end = NULL;
    unsigned long long v = strtoull(s, &end, 10);
    if (unit >= '0' && unit <= '9') {
        if (*end) return -1;
        *out = v;
        return 0;
    }
    if (end != s + n - 1) return -1;
    switch (unit) {
        case 'k': case 'K': *out = v << 10; return 0;
        case 'm': case 'M': *out = v << 20; return 0;
        case 'g': case 'G': *out = v << 30; return 0;
        default: return -1;
    }
}

static void usage(void) {
    fprintf(stderr,
            "Usage: emp_oracle [input fastq] [output] [options]\n"
            "Options:\n"
            "  -o, minimum overlap length, default = 31\n"
            "  -m, device memory during the fingerprint generation phase, default = 1g\n"
            "  -r, device memory during the overlap-detection phase, default = 1g\n"
            "  -b, block size (count) of transposed bases, default = 32m\n"
            "  -c, block size (count) of key-value pairs, default = 64m\n");
}

int main(int argc, char **argv) {
    if (argc < 3) { usage(); return 2; }
    const char *in = argv[1], *out = argv[2];
    uint64_t o = 31, m = 1ull << 30, r = 1ull << 30, b = 32ull << 20, c = 64ull << 20;
    for (int i = 3; i < argc; i += 2) {
        const char *a = argv[i];
        if (a[0] != '-' || strlen(a) != 2 || i + 1 >= argc) { fprintf(stderr, "Invalid option %s\n", a); usage(); return 2; }
        uint64_t *dst = NULL;
        switch (a[1]) {
            case 'o': dst = &o; break;
            case 'm': dst = &m; break;
            case 'r': dst = &r; break;
            case 'b': dst = &b; break;
            case 'c': dst = &c; break;
            default: fprintf(stderr, "Invalid option %s\n", a); usage(); return 2;
        }
        if (size_arg(argv[i + 1], dst)) { fprintf(stderr, "Unknown unit in %s\n", argv[i + 1]); return 2; }
    }
    eo_opts opts = {(int)o, m, r};
    eo_summary sum;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int rc = eo_run_file(in, out, &opts, &sum, NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (rc) { fprintf(stderr, "emp_oracle: %s\n", eo_strerror(rc)); return 1; }
    double s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    fprintf(stderr,
            "emp_oracle: reads=%llu L=%d dedup=%llu rounds=%d records=%llu edges=%llu threads=%d t_e2e=%.3f s "
            "Gbases/s=%.4f\n",
            (unsigned long long)sum.n_reads, sum.read_len, (unsigned long long)sum.n_dedup, sum.n_rounds,
            (unsigned long long)sum.n_records, (unsigned long long)sum.n_edges, omp_get_max_threads(), s,
            s > 0 ? (double)sum.n_reads * sum.read_len / s / 1e9 : 0.0);
    return 0;
}
