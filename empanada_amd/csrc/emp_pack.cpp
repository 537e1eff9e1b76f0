// Kept reads packed for the link (emp_hostcopy.h): ASCII bases to 2-bit codes on the host cores,
// four to a byte, so that a quarter of the bytes cross PCIe; the device expands them again
// (k_unpack) before the map kernel, which reads the same codes from the expanded rows.
#include "emp_hostcopy.h"

#include <immintrin.h>

#include <unistd.h>

#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <vector>

namespace emp {

namespace {

inline bool valid_base(uint8_t c) {
    const uint8_t l = c | 0x20;
    return l == 'a' || l == 'c' || l == 'g' || l == 't';
}

// Bytes [0, n) of src into dst (n / 4 bytes, n a multiple of 4).
bool pack_scalar(uint8_t *dst, const uint8_t *src, uint64_t n) {
    bool ok = true;
    for (uint64_t i = 0; i < n; i += 4) {
        uint8_t b = 0;
        for (int k = 0; k < 4; ++k) {
            const uint8_t c = src[i + k];
            ok &= valid_base(c);
            b |= (uint8_t)(((c >> 1) & 3u) << (2 * k));
        }
        dst[i / 4] = b;
    }
    return ok;
}

// 32 bases per step: codes (c >> 1) & 3 per byte, then code0 + 4 code1 + 16 code2 + 64 code3 per
// four bytes as two multiply-adds (byte pairs, then 16-bit pairs), the eight sums narrowed to bytes.
// Validity: (c | 0x20) is one of a, c, g, t. Returns the step's 8 packed bytes; `good` gathers the
// validity of every byte.
__attribute__((target("avx2"))) inline uint64_t pack32(const uint8_t *src, __m256i &good) {
    const __m256i m3 = _mm256_set1_epi8(3), lower = _mm256_set1_epi8(0x20);
    const __m256i w_pairs = _mm256_set1_epi16(0x0401), w_quads = _mm256_set1_epi32(0x00100001);
    const __m256i v = _mm256_loadu_si256(reinterpret_cast<const __m256i *>(src));
    const __m256i l = _mm256_or_si256(v, lower);
    const __m256i ok = _mm256_or_si256(
        _mm256_or_si256(_mm256_cmpeq_epi8(l, _mm256_set1_epi8('a')), _mm256_cmpeq_epi8(l, _mm256_set1_epi8('c'))),
        _mm256_or_si256(_mm256_cmpeq_epi8(l, _mm256_set1_epi8('g')), _mm256_cmpeq_epi8(l, _mm256_set1_epi8('t'))));
    good = _mm256_and_si256(good, ok);
    const __m256i code = _mm256_and_si256(_mm256_srli_epi16(v, 1), m3);
    const __m256i p16 = _mm256_maddubs_epi16(code, w_pairs);  // c0 + 4 c1 per byte pair
    const __m256i p32 = _mm256_madd_epi16(p16, w_quads);      // (c0 + 4 c1) + 16 (c2 + 4 c3)
    __m256i q = _mm256_packus_epi32(p32, p32);                // per 128-bit half: 4 words, twice
    q = _mm256_packus_epi16(q, q);                            // per half: 4 bytes, four times
    return (uint64_t)(uint32_t)_mm256_extract_epi32(q, 0) | ((uint64_t)(uint32_t)_mm256_extract_epi32(q, 4) << 32);
}

// Whole 32-base steps, then the scalar rest. Where the output is 32-byte aligned, four steps go
// out as one non-temporal 32-byte store: the packed bytes are read by the DMA engine, not by this
// core, and a normal store would first read each output line from memory. The fence at the end
// orders the streamed stores before the caller's next step (the copy to the device).
__attribute__((target("avx2"))) bool pack_avx2(uint8_t *dst, const uint8_t *src, uint64_t n) {
    __m256i good = _mm256_set1_epi8(-1);
    uint64_t i = 0;
    for (; i + 32 <= n && ((uintptr_t)(dst + i / 4) & 31u); i += 32) {
        const uint64_t w = pack32(src + i, good);
        memcpy(dst + i / 4, &w, 8);
    }
    bool streamed = false;
    for (; i + 128 <= n; i += 128) {
        const uint64_t w0 = pack32(src + i, good), w1 = pack32(src + i + 32, good);
        const uint64_t w2 = pack32(src + i + 64, good), w3 = pack32(src + i + 96, good);
        _mm256_stream_si256(reinterpret_cast<__m256i *>(dst + i / 4),
                            _mm256_set_epi64x((long long)w3, (long long)w2, (long long)w1, (long long)w0));
        streamed = true;
    }
    for (; i + 32 <= n; i += 32) {
        const uint64_t w = pack32(src + i, good);
        memcpy(dst + i / 4, &w, 8);
    }
    if (streamed) _mm_sfence();
    bool ok = _mm256_movemask_epi8(good) == -1;
    if (i < n) ok &= pack_scalar(dst + i / 4, src + i, n - i);
    return ok;
}

// PACK_THREADS - 1 threads kept for the life of the process (spawned per call they cost about 0.3 ms
// a call on an 8-core host, as much as packing 4 MB takes). run() gives parts 1 .. parts - 1 to them,
// part 0 to the caller, and returns when all are done; calls from several threads take turns. The
// pool is never destroyed: at exit its threads are waiting on the condition variable. A child
// process after fork() has none of its threads, so there the caller packs every part itself.
class PackPool {
public:
    PackPool() : pid_(getpid()) {
        for (int i = 1; i < PACK_THREADS; ++i) std::thread([this, i] { work(i); }).detach();
    }
    void run(int parts, const std::function<void(int)> &fn) {
        if (parts <= 1 || getpid() != pid_) {
            for (int p = 0; p < parts; ++p) fn(p);
            return;
        }
        std::lock_guard<std::mutex> turn(call_mu_);
        {
            std::lock_guard<std::mutex> l(mu_);
            fn_ = &fn;
            parts_ = parts;
            left_ = parts - 1;
            ++gen_;
        }
        cv_.notify_all();
        fn(0);
        std::unique_lock<std::mutex> l(mu_);
        done_.wait(l, [this] { return left_ == 0; });
        fn_ = nullptr;
    }

private:
    void work(int id) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> l(mu_);
        for (;;) {
            cv_.wait(l, [&] { return gen_ != seen; });
            seen = gen_;
            if (id >= parts_) continue;  // not needed for this call
            const std::function<void(int)> *fn = fn_;
            l.unlock();
            (*fn)(id);
            l.lock();
            if (--left_ == 0) done_.notify_one();
        }
    }
    const pid_t pid_;
    std::mutex call_mu_, mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *fn_ = nullptr;
    int parts_ = 0, left_ = 0;
    uint64_t gen_ = 0;
};

constexpr uint64_t PACK_PART_MIN = 1ull << 20;  // bases per part, at least (but for the last)

}  // namespace

bool parallel_pack(uint8_t *dst, const char *src, uint64_t n) {
    const uint8_t *s = reinterpret_cast<const uint8_t *>(src);
    const uint64_t body = n & ~3ull;
    static const bool avx2 = __builtin_cpu_supports("avx2");
    auto run = [&](uint8_t *d, const uint8_t *x, uint64_t m) { return avx2 ? pack_avx2(d, x, m) : pack_scalar(d, x, m); };
    bool ok = true;
    const int parts = (int)std::min<uint64_t>(PACK_THREADS, body / PACK_PART_MIN);
    if (parts <= 1) {
        ok = run(dst, s, body);
    } else {
        static PackPool *pool = new PackPool();
        // parts of a multiple of 32 bases, so that every part starts on a whole packed byte
        const uint64_t per = ((body + parts - 1) / parts + 31) & ~31ull;
        std::atomic<bool> all{true};
        pool->run(parts, [&](int p) {
            const uint64_t a = std::min(body, per * (uint64_t)p), e = std::min(body, a + per);
            if (e > a && !run(dst + a / 4, s + a, e - a)) all = false;
        });
        ok = all;
    }
    if (body < n) {  // the last one to three bases
        uint8_t b = 0;
        for (uint64_t i = body; i < n; ++i) {
            ok &= valid_base(s[i]);
            b |= (uint8_t)(((s[i] >> 1) & 3u) << (2 * (i - body)));
        }
        dst[body / 4] = b;
    }
    return ok;
}

}  // namespace emp
