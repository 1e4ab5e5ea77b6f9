// hq_internal.h -- shared device/host helpers of the hierarquicos_amd HIP library.
// gfx950 (CDNA4) only: 64-lane wavefronts, LDS, scalar loads for wave-uniform data.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/hq.h"

namespace hq {

constexpr int MAX_DEPTH = 25;     // octree_default.f90:95 (max_depth = 25)
constexpr int HI_LEVELS = 21;     // levels 1..21 live in the 63-bit high key word
constexpr int WAVE = 64;
constexpr int END = -1;           // "no node" for next pointers

// Per-thread last-error text (the C-ABI hands it out through hq_last_error).
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define HQ_CHECK_HIP(call)                                                              \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::hq::fail(HQ_ERR_HIP, "%s failed: %s (%s:%d)", #call,               \
                              hipGetErrorString(e_), __FILE__, __LINE__);               \
    } while (0)

// Simple owning device buffer that only grows (no free on shrink).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int ensure(size_t need) {
        if (need <= bytes) return HQ_OK;
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t b = need < 256 ? 256 : need;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) return fail(HQ_ERR_NOMEM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
        bytes = b;
        return HQ_OK;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

// XCD-aware block order: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs; remapping gives each XCD a contiguous range of tiles so that
// neighbouring (Morton-adjacent) groups share that XCD's L2.
__device__ __forceinline__ int xcd_swizzle(int bid, int nblocks) {
    const int nx = 8;
    int per = (nblocks + nx - 1) / nx;
    int x = bid % nx, k = bid / nx;
    int full = nblocks % nx == 0 ? nx : nblocks % nx;  // XCDs that get "per" blocks
    int r = x < full ? x * per + k : full * per + (x - full) * (per - 1) + k;
    return r < nblocks ? r : bid;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    int l = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T u = __shfl_up(v, o, 64);
        if (l >= o) v += u;
    }
    return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Exclusive scan of a uint32 array of length len in one workgroup; writes the
// grand total to a[len].
static __global__ __launch_bounds__(1024) void k_scan_excl(uint32_t* __restrict__ a, int64_t len) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int64_t base = 0; base < len; base += 1024) {
        int64_t i = base + threadIdx.x;
        uint32_t v = i < len ? a[i] : 0u;
        uint32_t incl = wave_incl_scan(v);
        int w = threadIdx.x >> 6;
        if (__lane_id() == 63) wsum[w] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t e = 0;
            for (int k = 0; k < 16; ++k) {  // exclusive wave offsets
                uint32_t c = wsum[k];
                wsum[k] = e;
                e += c;
            }
        }
        __syncthreads();
        uint32_t carry = carry_s;
        uint32_t out = carry + wsum[w] + incl - v;
        if (i < len) a[i] = out;
        __syncthreads();
        if (threadIdx.x == 1023) carry_s = out + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) a[len] = carry_s;
}

}  // namespace hq
