// Device-side constants and integer arithmetic shared by the kernels.
//
// Keys are (h1, h2): h1 = H(R=5, P1 = 2^29-3), h2 = H(R=7, P2 = 2^29-33), the two 32-bit halves of
// the reference's 64-bit key (assembly.cuh:29-34,50-52). Every reduction below is exact: with
// P = 2^29 - c, 2^29 = c (mod P), so folding the bits above 29 as hi*c + lo preserves the residue.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emp {

constexpr uint32_t P1 = 536870909u;  // 2^29 - 3
constexpr uint32_t P2 = 536870879u;  // 2^29 - 33
constexpr uint32_t C1 = 3u;
constexpr uint32_t C2 = 33u;
constexpr uint32_t R1 = 5u;
constexpr uint32_t R2 = 7u;
constexpr uint32_t INV1 = 107374182u;  // 5^(P1-2) mod P1 (overlap.cu:121)
constexpr uint32_t INV2 = 76695840u;   // 7^(P2-2) mod P2 (overlap.cu:122)
constexpr uint32_t MASK29 = (1u << 29) - 1;
constexpr int MAX_RD_LEN = 255;

// Column stores: 2-bit base codes ((c >> 1) & 3) of every read position, reads grouped in chunks of
// 2^CHUNK_SHIFT. The batched engine's job regions use columns in groups of 4: a 32-bit word holds
// 4 reads x 4 columns, read k, column 4g + c at bits 2 * (4 * (k % 4) + c) (G = ceil(L / 4) groups).
constexpr int CHUNK_SHIFT = 20;
constexpr uint32_t CHUNK_READS = 1u << CHUNK_SHIFT;
constexpr uint32_t CHUNK_GWORDS = CHUNK_READS / 4;
constexpr int GROUP_COLS = 4;
__host__ __device__ __forceinline__ uint32_t col_groups(uint32_t L) { return (L + GROUP_COLS - 1) / GROUP_COLS; }

// Wide column store (the single-job engine): columns in groups of 8, a 32-bit word holds 2 reads x
// 8 columns: read k, column 8g + c at bits 2 * (8 * (k % 2) + c). Chunk ch, group g holds
// CHUNK_WWORDS words starting at word (ch * G8 + g) * CHUNK_WWORDS, G8 = ceil(L / 8).
constexpr uint32_t CHUNK_WWORDS = CHUNK_READS / 2;
constexpr int WIDE_COLS = 8;
__host__ __device__ __forceinline__ uint32_t col_groups8(uint32_t L) { return (L + WIDE_COLS - 1) / WIDE_COLS; }

// Keys are below 2^29: bits 29..31 of h1 and h2 carry base codes an entry will need in its next
// rounds (the batched engine: up to three, the 6-bit field h1 >> 29 | (h2 >> 29) << 3 with the first
// in bits 0-1; the single-job engine's round entries: seven, below). Key comparisons mask them off
// with KEY_MASK.
constexpr uint32_t KEY_MASK = MASK29;

// Entries of the single-job engine's round buckets carry 8 more codes' bits in bits 21..28 of h1:
// a bucket set's count is a multiple of 2^WIDE_MIN_BITS (key h in bucket floor(h nb / 2^29)), so
// those key bits are the same for every key of a bucket and are restored from it (wide_key1, the
// top bits of the bucket's smallest key). The 14-bit carry is
// h1 >> 29 | (h2 >> 29) << 3 | ((h1 >> 21) & 0xFF) << 6, the next codes two bits each, first at
// bits 0-1. Keys of one bucket compare on h1 & LOW_MASK.
constexpr int WIDE_MIN_BITS = 8;
constexpr int WIDE_SHIFT = 21;
constexpr uint32_t LOW_MASK = (1u << WIDE_SHIFT) - 1;

// One (fingerprint, oriented read id) entry, 12 bytes.
struct Entry {
    uint32_t h1, h2, rid;
};

// x mod P1 for x < 2^32.
__host__ __device__ __forceinline__ uint32_t red1(uint32_t x) {
    uint32_t y = (x >> 29) * C1 + (x & MASK29);  // < 2^29 + 21 < 2*P1
    return y >= P1 ? y - P1 : y;
}

// x mod P2 for x < 2^32.
__host__ __device__ __forceinline__ uint32_t red2(uint32_t x) {
    uint32_t y = (x >> 29) * C2 + (x & MASK29);  // < 2^29 + 231 < 2*P2
    return y >= P2 ? y - P2 : y;
}

// x mod (2^29 - c) for any 64-bit x, c <= 33.
__host__ __device__ __forceinline__ uint32_t red64(uint64_t x, uint32_t c, uint32_t p) {
    uint64_t y = (x >> 29) * c + (x & MASK29);  // < 2^41 + 2^29
    y = (y >> 29) * c + (y & MASK29);           // < 2^29 + 2^18
    uint32_t z = (uint32_t)y;
    return z >= p ? z - p : z;
}

// ENCODING (assembly.cuh:34) of a 2-bit code (c>>1)&3: A0 C1 T2 G3 -> 1..4. The complement
// (preprocess.cu:18: A<->T, C<->G) of a code is code ^ 2.
__host__ __device__ __forceinline__ uint32_t enc_of_code(uint32_t code) { return code + 1; }


// Fingerprints as byte dot products (the map kernel): with W_i = R^(L-1-i) mod P split into bytes
// W_i = sum_j b_ij 256^j, H = sum_i e_i W_i = sum_j S_j 256^j (mod P) for S_j = sum_i e_i b_ij.
// For L <= 255, e <= 4 and b <= 255, S_j < 2^18 and x = sum_j S_j 256^j < 2^43; fold1 / fold2
// reduce x with 2^29 = C (mod 2^29 - C): (x >> 29) C + (x mod 2^29) < 2^29 + 2^20 < 2P, so one
// conditional subtraction ends in [0, P).
__host__ __device__ __forceinline__ uint32_t fold_bytes(uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3, uint32_t c,
                                                        uint32_t p) {
    const uint64_t x = (uint64_t)s0 + ((uint64_t)s1 << 8) + ((uint64_t)s2 << 16) + ((uint64_t)s3 << 24);
    const uint32_t y = (uint32_t)(x >> 29) * c + (uint32_t)(x & MASK29);
    return y >= p ? y - p : y;
}
__host__ __device__ __forceinline__ uint32_t fold1(uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3) {
    return fold_bytes(s0, s1, s2, s3, C1, P1);
}
__host__ __device__ __forceinline__ uint32_t fold2(uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3) {
    return fold_bytes(s0, s1, s2, s3, C2, P2);
}

// NextSfx (overlap.cu:57-75): drop the first base of a length-l suffix. m1 = R1^(l-1) mod P1,
// m2 = R2^(l-1) mod P2. e*m < 4*2^29, so the 32-bit product is exact as in the reference.
__host__ __device__ __forceinline__ void next_sfx(uint32_t &h1, uint32_t &h2, uint32_t e, uint32_t m1, uint32_t m2) {
    uint32_t t1 = red1(e * m1), t2 = red2(e * m2);
    uint32_t v1 = h1 + P1 - t1, v2 = h2 + P2 - t2;  // in (0, 2P)
    h1 = v1 >= P1 ? v1 - P1 : v1;
    h2 = v2 >= P2 ? v2 - P2 : v2;
}

// u * R^-1 mod P for u < P as an exact division: P = -1 (mod R) for (R1, P1) and (R2, P2), so
// u + (u mod R) P is a multiple of R below R 2^29 < 2^32, and its quotient by R (a product with R's
// inverse mod 2^32) is u R^-1 mod P, below P.
__host__ __device__ __forceinline__ uint32_t div_r1(uint32_t u) { return (u + (u % R1) * P1) * 0xCCCCCCCDu; }
__host__ __device__ __forceinline__ uint32_t div_r2(uint32_t u) { return (u + (u % R2) * P2) * 0xB6DB6DB7u; }

// PrevPfx (overlap.cu:77-91): ((INV * ((h - e) mod 2^64)) mod 2^64) mod P, keeping the
// reference's uint64_t wrap when h < e (the 64-bit product then; else the exact division above).
__host__ __device__ __forceinline__ void prev_pfx(uint32_t &h1, uint32_t &h2, uint32_t e) {
    if (h1 >= e) {
        h1 = div_r1(h1 - e);
    } else {
        const uint64_t x1 = (uint64_t)INV1 * ((uint64_t)h1 - (uint64_t)e);
        h1 = red64(x1, C1, P1);
    }
    if (h2 >= e) {
        h2 = div_r2(h2 - e);
    } else {
        const uint64_t x2 = (uint64_t)INV2 * ((uint64_t)h2 - (uint64_t)e);
        h2 = red64(x2, C2, P2);
    }
}

// Sort rank of a read id inside a group of equal keys: (rid & 1, rid) as one 32-bit value.
__host__ __device__ __forceinline__ uint32_t rid_rank(uint32_t rid) { return (rid >> 1) | (rid << 31); }

}  // namespace emp
