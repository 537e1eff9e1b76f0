// Allocation-free fingerprint stage (emp_stage_fingerprints_noalloc): the kernels of
// emp_stage_fingerprints over caller-owned device input and output, with every intermediate
// buffer in one statically allocated device array per GPU.
//
// The array (`scratch`) holds, from its start: the bad-base flag (256 bytes), the map kernel's
// dot-product weights for the call's read length (4 KiB), then the 12-byte (h1, h2, rid) entries
// the map kernel writes, 2 per read. Each GPU the library runs on holds its own copy for the life
// of the process; its device address comes from hipGetSymbolAddress, once per device, with that
// device current (the host-side name `scratch` is not a device address). Every call on a GPU
// uses that GPU's copy: calls that overlap on one GPU (two host threads, or two streams in
// flight) race, and each call clears what must start at zero before its kernels read it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/empanada.h"
#include "../../include/empanada_stages.h"
#include "emp_kernels.h"

namespace emp {

constexpr uint64_t NA_MAX_READS = EMP_NOALLOC_MAX_READS;  // 2^20 reads: 24 MiB of entries
constexpr size_t NA_FLAG_BYTES = 256;
constexpr size_t NA_WT_BYTES = 4096;  // ceil(255 / 4) * 16 dwords = 4,096 bytes
constexpr size_t NA_ENTRY_OFF = NA_FLAG_BYTES + NA_WT_BYTES;
constexpr size_t NA_BYTES = NA_ENTRY_OFF + NA_MAX_READS * 2 * sizeof(Entry);

static __device__ __align__(256) uint8_t scratch[NA_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex na_mu;
uint8_t *na_ptr[MAX_DEVICES];  // device address of `scratch` on each device (null: not looked up yet)

// R^e mod p.
__device__ uint32_t pow_mod(uint32_t r, uint32_t e, uint32_t p) {
    uint64_t x = 1, b = r % p;
    while (e) {
        if (e & 1u) x = x * b % p;
        b = b * b % p;
        e >>= 1;
    }
    return (uint32_t)x;
}

// The map kernel's weight table for read length L (the layout of launch_map_wt), one dword per
// thread, and the bad-base flag cleared.
__global__ void __launch_bounds__(256) k_na_setup(uint32_t *wt, uint32_t *flag, int L) {
    const int ng = (L + 3) / 4;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0) *flag = 0u;
    if (i >= ng * 16) return;
    const int d = i / 16, h = (i % 16) / 4, j = i % 4;
    uint32_t word = 0;
    for (int u = 0; u < 4; ++u) {
        const int k = 4 * d + u;
        if (k >= L) break;
        uint32_t w;
        switch (h) {
            case 0: w = pow_mod(R1, (uint32_t)(L - 1 - k), P1); break;
            case 1: w = pow_mod(R2, (uint32_t)(L - 1 - k), P2); break;
            case 2: w = pow_mod(R1, (uint32_t)k, P1); break;
            default: w = pow_mod(R2, (uint32_t)k, P2); break;
        }
        word |= ((w >> (8 * j)) & 0xFFu) << (8 * u);
    }
    wt[i] = word;
}

// Entries -> 64-bit keys (HIGH32 = h1, LOW32 = h2: assembly.cuh:50-52), and the flag -> status.
__global__ void __launch_bounds__(256) k_na_keys(const Entry *__restrict__ e, uint64_t m, uint64_t *__restrict__ keys,
                                                 const uint32_t *flag, uint32_t *status) {
    const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    if (i == 0) *status = *flag;
    if (i < m) keys[i] = ((uint64_t)e[i].h1 << 32) | e[i].h2;
}

// This device's copy of `scratch`, looked up on its first use (with the device current).
hipError_t scratch_of(uint8_t **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(na_mu);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        na_ptr[dev] = static_cast<uint8_t *>(p);
    }
    *out = na_ptr[dev];
    return hipSuccess;
}

}  // namespace
}  // namespace emp

using namespace emp;

extern "C" {

int emp_stage_fingerprints_noalloc(const void *d_reads, uint64_t n, int32_t len, uint64_t *d_keys, uint32_t *d_status,
                                   void *stream) {
    if (n > NA_MAX_READS) return EMP_ERR_TOO_MANY;
    if (len < 1 || len > MAX_RD_LEN) return EMP_ERR_READLEN;
    if (!d_status || (n && (!d_reads || !d_keys))) return EMP_ERR_ARG;
    if ((reinterpret_cast<uintptr_t>(d_reads) & 3u) || (reinterpret_cast<uintptr_t>(d_keys) & 7u) ||
        (reinterpret_cast<uintptr_t>(d_status) & 3u))
        return EMP_ERR_ARG;
    uint8_t *base = nullptr;
    if (scratch_of(&base) != hipSuccess) return EMP_ERR_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t *flag = reinterpret_cast<uint32_t *>(base);
    uint32_t *wt = reinterpret_cast<uint32_t *>(base + NA_FLAG_BYTES);
    Entry *entries = reinterpret_cast<Entry *>(base + NA_ENTRY_OFF);
    const int ng = (len + 3) / 4;
    hipLaunchKernelGGL(k_na_setup, dim3((ng * 16 + 255) / 256), dim3(256), 0, s, wt, flag, (int)len);
    if (hipGetLastError() != hipSuccess) return EMP_ERR_HIP;
    MapArgs a{};
    a.reads = static_cast<const uint8_t *>(d_reads);
    a.n = (uint32_t)n;
    a.L = len;
    map_shape(len, &a.Lpad, &a.rows);
    a.k0 = 0;
    a.cols = nullptr;  // fingerprints only
    a.entries = entries;
    a.ends = nullptr;
    a.err = flag;
    a.wt = wt;
    if (launch_map_wt(a, s) != hipSuccess) return EMP_ERR_HIP;
    const uint64_t m = 2 * n;
    hipLaunchKernelGGL(k_na_keys, dim3((uint32_t)((m + 255) / 256 > 0 ? (m + 255) / 256 : 1)), dim3(256), 0, s,
                       entries, m, d_keys, flag, d_status);
    return hipGetLastError() == hipSuccess ? EMP_OK : EMP_ERR_HIP;
}

}  // extern "C"
