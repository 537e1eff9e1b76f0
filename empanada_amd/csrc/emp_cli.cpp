// empanada: the reference's command line (config.h:120-140, README.md:15-20) over libempanada_hip.
//
//   empanada <input fastq> <output> [-o N] [-m S] [-r S] [-b S] [-c S]
//
// Sizes take k/m/g suffixes (config.h:75-100). The map phase reads the input in blocks of
// map_mem/16 bytes and resumes after the last complete record of each block, as read_block does
// (preprocess.cu:195-237), through emp_add_fastq_file; the overlap phase writes output.bin
// (overlap.cu:148-161).
// Differences from the reference: errors return a non-zero status instead of exit(0) or a
// release-mode assert, and timings go to stderr instead of timings.csv.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "../../include/empanada.h"

namespace {

double now_s() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + t.tv_nsec * 1e-9;
}

// hmsize2bytes (config.h:75-100)
bool size_arg(const char *s, uint64_t *out) {
    std::string v(s);
    if (v.empty()) return false;
    char unit = v.back();
    char *end = nullptr;
    if (unit >= '0' && unit <= '9') {
        unsigned long long x = strtoull(v.c_str(), &end, 10);
        if (*end) return false;
        *out = x;
        return true;
    }
    v.pop_back();
    unsigned long long x = strtoull(v.c_str(), &end, 10);
    if (*end || v.empty()) return false;
    switch (unit) {
        case 'k': case 'K': *out = x << 10; return true;
        case 'm': case 'M': *out = x << 20; return true;
        case 'g': case 'G': *out = x << 30; return true;
        default: return false;
    }
}

void usage() {
    printf("Usage: empanada [input fastq] [output gfa] [options]\n");
    printf("Options:\n");
    printf("  -o or --min-overlap, minimum overlap length, default = 31\n");
    printf("  -m or --map-mem, device memory during the fingerprint generation phase, default = 1g\n");
    printf("  -r or --reduce-mem, device memory during the overlap-detection phase, default = 1g\n");
    printf("  -b or --base-blocksz, block size (count) of transposed bases, default = 32m\n");
    printf("  -c or --kv-chunksz, block size (count) of key-value pairs, default = 64m\n");
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        usage();
        return 2;
    }
    emp_opts o;
    emp_opts_default(&o);
    uint64_t ovl = (uint64_t)o.min_ovlp;
    for (int i = 3; i < argc; i += 2) {
        const char *a = argv[i];
        if (a[0] != '-' || strlen(a) != 2 || i + 1 >= argc) {
            printf("Invalid option %s\n", a);
            usage();
            return 2;
        }
        uint64_t *dst = nullptr;
        switch (a[1]) {
            case 'o': dst = &ovl; break;
            case 'm': dst = &o.map_mem; break;
            case 'r': dst = &o.reduce_mem; break;
            case 'b': dst = &o.base_blocksz; break;
            case 'c': dst = &o.kv_chunksz; break;
            default:
                printf("Invalid option %s\n", a);
                usage();
                return 2;
        }
        if (!size_arg(argv[i + 1], dst)) {
            fprintf(stderr, "Unknown unit in %s\n", argv[i + 1]);
            return 2;
        }
    }
    o.min_ovlp = (int32_t)ovl;
    const size_t block = (size_t)(o.map_mem / 16);  // main.cu:16
    if (block == 0) {
        fprintf(stderr, "map memory too small\n");
        return 2;
    }
    const char *env = getenv("EMPANADA_PROFILE");
    o.profile = env && *env && *env != '0';
    // EMPANADA_AUDIT=1: the collision audit the reference builds with FPCHECK (overlap.cu:295-359)
    const char *aenv = getenv("EMPANADA_AUDIT");
    const bool audit = aenv && *aenv && *aenv != '0';

    double t_init = now_s();
    emp_ctx *ctx = nullptr;
    int rc = emp_ctx_create(&o, &ctx);
    if (rc) {
        fprintf(stderr, "empanada: cannot create a context (%d)\n", rc);
        return 1;
    }
    emp_set_audit(ctx, audit ? 1 : 0);
    double t0 = now_s();
    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        fprintf(stderr, "empanada: cannot open %s\n", argv[1]);
        emp_ctx_destroy(ctx);
        return 1;
    }
    fclose(f);
    // the block loop of main.cu:42-58 (read_block in blocks of map_mem/16 bytes, each resumed after
    // its last complete record), with the file streamed into HBM while earlier blocks are parsed
    rc = emp_add_fastq_file(ctx, argv[1]);
    double t1 = now_s();
    if (!rc) rc = emp_run_to_file(ctx, argv[2]);
    double t2 = now_s();
    if (rc) {
        fprintf(stderr, "empanada: error %d: %s\n", rc, emp_last_error(ctx));
        emp_ctx_destroy(ctx);
        return 1;
    }
    emp_summary s;
    emp_get_summary(ctx, &s);
    fprintf(stderr,
            "empanada: reads=%llu L=%d dedup=%llu rounds=%d records=%llu edges=%llu init=%.3f s map=%.3f s "
            "overlap=%.3f s t_e2e=%.3f s Gbases/s=%.4f\n",
            (unsigned long long)s.n_reads, s.read_len, (unsigned long long)s.n_dedup, s.n_rounds,
            (unsigned long long)s.n_records, (unsigned long long)s.n_edges, t0 - t_init, t1 - t0, t2 - t1, t2 - t0,
            (t2 - t0) > 0 ? (double)s.n_reads * s.read_len / (t2 - t0) / 1e9 : 0.0);
    for (int i = 0; audit && i < emp_audit_count(ctx); ++i) {
        emp_audit a;
        emp_get_audit(ctx, i, &a);
        static const char *names[3] = {"reads", "suffixes", "prefixes"};
        fprintf(stderr, "  audit %-8s len=%3d entries=%llu keys=%llu strings=%llu collision rate = %f%%\n",
                names[a.list < 0 || a.list > 2 ? 0 : a.list], a.ovlp, (unsigned long long)a.entries,
                (unsigned long long)a.keys, (unsigned long long)a.strings,
                a.strings ? (1.0 - (double)a.keys / (double)a.strings) * 100.0 : 0.0);
    }
    if (o.profile) {
        for (int i = 0; i < emp_kernel_count(ctx); ++i) {
            const char *name;
            uint64_t launches;
            double ms, bytes;
            emp_kernel_info(ctx, i, &name, &launches, &ms, &bytes);
            fprintf(stderr, "  %-16s launches=%llu ms=%.3f GB/s=%.1f\n", name, (unsigned long long)launches, ms,
                    ms > 0 ? bytes / ms / 1e6 : 0.0);
        }
    }
    emp_ctx_destroy(ctx);
    return 0;
}
