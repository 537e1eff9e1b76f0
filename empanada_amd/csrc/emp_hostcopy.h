// Host memory copies split over a few threads: a pageable source into pinned staging is bound by
// one core's memcpy rate (about 10 GB/s here), below what PCIe then moves.
#pragma once

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <thread>

namespace emp {

constexpr int COPY_THREADS = 4;
constexpr uint64_t COPY_SPLIT_MIN = 4ull << 20;  // smaller copies stay on the calling thread

inline void parallel_copy(char *dst, const char *src, uint64_t bytes) {
    if (bytes < COPY_SPLIT_MIN) {
        memcpy(dst, src, (size_t)bytes);
        return;
    }
    const uint64_t per = (bytes + COPY_THREADS - 1) / COPY_THREADS;
    auto part = [&](int p) {
        const uint64_t a = std::min(bytes, per * (uint64_t)p), e = std::min(bytes, a + per);
        if (e > a) memcpy(dst + a, src + a, (size_t)(e - a));
    };
    std::thread th[COPY_THREADS - 1];
    for (int p = 1; p < COPY_THREADS; ++p) th[p - 1] = std::thread(part, p);
    part(0);
    for (auto &t : th) t.join();
}

}  // namespace emp
