// parallel_pack (empanada_amd/csrc/emp_pack.cpp) on the CPU: every base's 2-bit code (c >> 1) & 3
// (A 0, C 1, T 2, G 3, lower case alike) at bits 2 (i % 4) of byte i / 4, nothing written past
// byte (n + 3) / 4, and false exactly when a byte other than ACGTacgt is present, wherever it sits
// (the first or last byte of a thread's part, the 1-3 base tail, the scalar steps after the last
// 32-base step); the same from four threads calling at once (the calls share one thread pool).
// Prints "all equal" when every case holds.
#include <stdio.h>

#include <atomic>
#include <random>
#include <thread>
#include <vector>

#include "emp_hostcopy.h"

int main() {
    std::mt19937_64 rng(0x5eed);
    const char bases[] = "ACGTacgt";
    const char others[] = {'N', 'n', 'X', 'U', '@', '-', '\r', '\n', '\0', (char)0xC1, (char)0xE1, (char)0x81};
    int fails = 0, cases = 0;
    const uint64_t sizes[] = {0, 1, 2, 3, 4, 5, 31, 32, 33, 35, 1000, 262143, 262144, 262147, 4194305, 33554467};
    for (uint64_t n : sizes) {
        std::vector<char> src(n);
        for (uint64_t i = 0; i < n; ++i) src[i] = bases[rng() % 8];
        // positions of a stray byte: none, then the ends and the middle, then random ones
        std::vector<int64_t> where = {-1};
        if (n) {
            const uint64_t ends[] = {0, n - 1, n / 2, n & ~3ull};
            for (uint64_t p : ends)
                if (p < n) where.push_back((int64_t)p);
            // about where eight threads' parts meet (parts of a multiple of 32 bases)
            const uint64_t part = ((n / 4 * 4 + 7) / 8 + 31) & ~31ull;
            for (uint64_t k = 1; k < 8; k += 3)
                for (uint64_t p : {k * part - 1, k * part})
                    if (p < n) where.push_back((int64_t)p);
            for (int k = 0; k < 4; ++k) where.push_back((int64_t)(rng() % n));
        }
        for (int64_t w : where) {
            std::vector<char> s = src;
            if (w >= 0) s[w] = others[rng() % sizeof(others)];
            const uint64_t nb = (n + 3) / 4;
            std::vector<uint8_t> dst(nb + 8, 0xEE);
            const bool ok = emp::parallel_pack(dst.data(), s.data(), n);
            bool same = true;
            if (w < 0) {
                for (uint64_t i = 0; i < n && same; ++i)
                    same = ((dst[i / 4] >> (2 * (i % 4))) & 3u) == ((uint8_t(s[i]) >> 1) & 3u);
                // the last byte's unused bit pairs are zero
                if (n % 4 && same) same = (dst[nb - 1] >> (2 * (n % 4))) == 0;
            }
            for (uint64_t i = nb; i < dst.size() && same; ++i) same = dst[i] == 0xEE;
            ++cases;
            if (ok != (w < 0) || !same) {
                ++fails;
                printf("n %llu stray at %lld: returned %d, bytes %s\n", (unsigned long long)n, (long long)w, ok,
                       same ? "equal" : "differ");
            }
        }
    }
    std::atomic<int> cfails{0};
    auto caller = [&](int t) {
        std::mt19937_64 r(1000 + t);
        for (int k = 0; k < 12; ++k) {
            const uint64_t n = (3ull << 20) + r() % (6ull << 20);
            std::vector<char> s(n);
            for (auto &c : s) c = bases[r() % 8];
            const int64_t w = (k % 3 == 2) ? (int64_t)(r() % n) : -1;
            if (w >= 0) s[w] = 'N';
            std::vector<uint8_t> d((n + 3) / 4 + 8, 0xEE);
            bool same = emp::parallel_pack(d.data(), s.data(), n) == (w < 0);
            for (uint64_t i = 0; i < n && same && w < 0; ++i)
                same = ((d[i / 4] >> (2 * (i % 4))) & 3u) == ((uint8_t(s[i]) >> 1) & 3u);
            for (uint64_t i = (n + 3) / 4; i < d.size() && same; ++i) same = d[i] == 0xEE;
            if (!same) ++cfails;
        }
    };
    std::thread th[4];
    for (int t = 0; t < 4; ++t) th[t] = std::thread(caller, t);
    for (auto &t : th) t.join();
    cases += 48;
    if (cfails) {
        fails += cfails;
        printf("%d concurrent calls differ\n", cfails.load());
    }
    printf("%d cases, %s\n", cases, fails ? "FAILED" : "all equal");
    return fails != 0;
}
