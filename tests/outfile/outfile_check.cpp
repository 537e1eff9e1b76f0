// CPU check of the output.bin writer (empanada_amd/csrc/emp_outfile.cpp): random record sequences
// appended, flushed and settled in random order, the file compared with the records laid out as
// write_edges_binary writes them (overlap.cu:148-161: n, overlap, from[n], to[n]).
// usage: outfile_check <path> [trials]
#include "../../empanada_amd/csrc/emp_outfile.h"

#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include <random>
#include <string>
#include <vector>

#ifdef OUTFILE_FAULTS
#include "fault_shim.h"
#endif

using namespace emp;

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const char *path = argv[1];
    const int trials = argc > 2 ? atoi(argv[2]) : 12;
    int fails = 0;
    DirectIO dio;  // kept across files, as a context keeps it
    for (int trial = 0; trial < trials; ++trial) {
        std::mt19937_64 rng(1000 + trial);
        const int nrec = 1 + (int)(rng() % 40);
        std::vector<uint64_t> ns;
        uint64_t tot = 0;
        for (int i = 0; i < nrec; ++i) {
            uint64_t n;
            switch (rng() % 5) {
                case 0: n = 0; break;
                case 1: n = rng() % 10; break;
                case 2: n = rng() % 100000; break;
                case 3: n = rng() % 1500000; break;
                default: n = 262144 * (1 + rng() % 3) - 1 - rng() % 3;  // records ending near chunk edges
            }
            ns.push_back(n);
            tot += n;
        }
        std::vector<uint32_t> from(tot + 1), to(tot + 1);
        for (uint64_t i = 0; i < tot; ++i) {
            from[i] = (uint32_t)rng();
            to[i] = (uint32_t)rng();
        }
        std::string want;
        OutFile of;
        unlink(path);
        if (trial % 4 == 3) {  // a file already there, longer or shorter than the new one
            FILE *old = fopen(path, "wb");
            const size_t len = (size_t)(rng() % (32 * tot + 4)) + 1;
            std::vector<char> junk(len, (char)0xA5);
            if (!old || fwrite(junk.data(), 1, len, old) != len) return 1;
            fclose(old);
        }
        if (of.open(path, trial % 3 == 2 ? nullptr : &dio)) {
            printf("open failed\n");
            return 1;
        }
        uint64_t off = 0;
        for (int i = 0; i < nrec; ++i) {
            const uint64_t n = ns[i];
            const int32_t ov = 99 - i;
            want.append((const char *)&n, 8);
            want.append((const char *)&ov, 4);
            want.append((const char *)(from.data() + off), 4 * n);
            want.append((const char *)(to.data() + off), 4 * n);
            if (of.append(n, ov, from.data() + off, to.data() + off)) {
                printf("append failed\n");
                return 1;
            }
            if (rng() % 3 == 0 && of.flush()) {
                printf("flush failed\n");
                return 1;
            }
            if (trial % 4 == 1 && rng() % 7 == 0 && of.settle()) {
                printf("settle failed\n");
                return 1;
            }
            off += n;
        }
        const uint64_t direct = of.direct_bytes();
        if (of.finish()) {
            printf("finish failed\n");
            return 1;
        }
        FILE *f = fopen(path, "rb");
        std::string got;
        char b[1 << 16];
        size_t r;
        while (f && (r = fread(b, 1, sizeof b, f)) > 0) got.append(b, r);
        if (f) fclose(f);
        const bool ok = got == want;
        fails += ok ? 0 : 1;
        printf("trial %2d: %2d records, %10zu bytes, %10llu direct, %s\n", trial, nrec, want.size(),
               (unsigned long long)direct, ok ? "equal" : "DIFFERENT");
    }
    unlink(path);
#ifdef OUTFILE_FAULTS
    printf("injected: %ld failed submissions of %ld, %ld failed completions\n", fi_failed_submits.load(),
           fi_submits.load(), fi_failed_events.load());
#endif
    printf("%s\n", fails ? "FAILED" : "all equal");
    return fails != 0;
}
