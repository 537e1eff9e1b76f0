// emp-convert: output.bin -> GFA 1.0 or CSV.
//
// output.bin is a sequence of records [u64 n][i32 ovlp][u32 from[n]][u32 to[n]]
// (write_edges_binary, overlap.cu:148-161). An edge (from, to) at overlap l says the last l bases
// of oriented read `from` equal the first l bases of oriented read `to`; oriented id 2k is read k
// and 2k+1 its reverse complement (fingerprint.cu:75-115).
//
//   GFA: one S line per read (write_vertices, overlap.cu:129-146: name = read index, no
//        sequence, LN = read length), then one L line per edge: read from/2 with orientation
//        '+' for an even id and '-' for an odd one, the same for `to`, CIGAR <l>M. The reference's
//        dormant writer (write_edges_gfa, overlap.cu:163-188) divides by 4 and marks odd ids '+';
//        both disagree with its own id assignment, so this follows the id assignment.
//   CSV: "from,to,ovlp" per edge with the oriented ids, as write_edges (overlap.cu:190-204).
//
// Usage: emp-convert output.bin out.gfa --gfa [--reads N --len L]
//        emp-convert output.bin out.csv --csv
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

namespace {

int usage() {
    fprintf(stderr,
            "usage: emp-convert output.bin OUT (--gfa [--reads N --len L] | --csv)\n"
            "  --gfa    GFA 1.0: S line per read (with --reads/--len), L line per edge\n"
            "  --csv    from,to,ovlp per edge (oriented read ids)\n");
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 4) return usage();
    const char *in = argv[1], *outp = argv[2];
    bool gfa = false, csv = false;
    unsigned long long reads = 0;
    long len = 0;
    for (int i = 3; i < argc; ++i) {
        if (!strcmp(argv[i], "--gfa")) gfa = true;
        else if (!strcmp(argv[i], "--csv")) csv = true;
        else if (!strcmp(argv[i], "--reads") && i + 1 < argc) reads = strtoull(argv[++i], nullptr, 10);
        else if (!strcmp(argv[i], "--len") && i + 1 < argc) len = strtol(argv[++i], nullptr, 10);
        else return usage();
    }
    if (gfa == csv) return usage();
    FILE *f = fopen(in, "rb");
    if (!f) {
        fprintf(stderr, "emp-convert: cannot open %s\n", in);
        return 1;
    }
    FILE *o = fopen(outp, "w");
    if (!o) {
        fprintf(stderr, "emp-convert: cannot create %s\n", outp);
        fclose(f);
        return 1;
    }
    if (gfa) {
        fprintf(o, "H\tVN:Z:1.0\n");
        for (unsigned long long k = 0; k < reads; ++k) fprintf(o, "S\t%llu\t*\tLN:i:%ld\n", k, len);
    }
    std::vector<uint32_t> from, to;
    uint64_t records = 0, edges = 0;
    int rc = 0;
    for (;;) {
        uint64_t n;
        int32_t ovlp;
        size_t got = fread(&n, sizeof n, 1, f);
        if (got == 0) break;
        if (fread(&ovlp, sizeof ovlp, 1, f) != 1 || n > (1ull << 36)) {
            fprintf(stderr, "emp-convert: truncated or malformed record %llu\n", (unsigned long long)records);
            rc = 1;
            break;
        }
        from.resize(n);
        to.resize(n);
        if (n && (fread(from.data(), 4, n, f) != n || fread(to.data(), 4, n, f) != n)) {
            fprintf(stderr, "emp-convert: truncated record %llu\n", (unsigned long long)records);
            rc = 1;
            break;
        }
        for (uint64_t i = 0; i < n; ++i) {
            if (gfa)
                fprintf(o, "L\t%u\t%c\t%u\t%c\t%dM\n", from[i] >> 1, (from[i] & 1) ? '-' : '+', to[i] >> 1,
                        (to[i] & 1) ? '-' : '+', ovlp);
            else
                fprintf(o, "%u,%u,%d\n", from[i], to[i], ovlp);
        }
        ++records;
        edges += n;
    }
    fclose(f);
    if (fclose(o) != 0) {
        fprintf(stderr, "emp-convert: writing %s failed\n", outp);
        rc = 1;
    }
    if (!rc) fprintf(stderr, "emp-convert: %llu records, %llu edges\n", (unsigned long long)records,
                     (unsigned long long)edges);
    return rc;
}
