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

// Kept reads packed for the link: the n bases of src (reads back to back, ASCII) as 2-bit codes
// (c >> 1) & 3, four to a byte (base i at bits 2 (i % 4) of byte i / 4; ceil(n / 4) bytes), split over
// up to PACK_THREADS threads of a pool kept for the process (a part per MB at least); callers on
// several threads take turns. Returns false when a byte is not one of ACGTacgt: the caller then sends the
// bytes as they are and the device's check reports the error as it does for any batch.
constexpr int PACK_THREADS = 12;
bool parallel_pack(uint8_t *dst, const char *src, uint64_t n);

}  // namespace emp
