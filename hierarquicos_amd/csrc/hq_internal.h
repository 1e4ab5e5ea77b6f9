// hq_internal.h -- shared device/host helpers of the hierarquicos_amd HIP library.
// gfx950 (CDNA4) only: 64-lane wavefronts, LDS, scalar loads for wave-uniform data.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <atomic>
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

// True while this thread records a stream capture (set by the integrator).
bool capturing();

// Bumped whenever any DevBuf gets or frees device memory: a captured HIP graph
// holds buffer addresses, so one captured at another epoch may not be replayed.
inline std::atomic<uint64_t> g_devbuf_epoch{0};

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
        if (capturing()) return fail(HQ_ERR_ARGS, "buffer of %zu bytes would grow to %zu inside a "
                                     "captured step", bytes, need);
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        g_devbuf_epoch.fetch_add(1, std::memory_order_relaxed);
        size_t b = need < 256 ? 256 : need;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) return fail(HQ_ERR_NOMEM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
        bytes = b;
        return HQ_OK;
    }
    void release() {
        if (p) {
            hipFree(p);
            g_devbuf_epoch.fetch_add(1, std::memory_order_relaxed);
        }
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

// Device-wide exclusive scan of uint32 a[0..len) with the grand total written
// to a[len]: per-tile sums (SCAN_TILE contiguous elements per workgroup), a
// scan of the tile sums in one workgroup, then a per-tile scan with the tile
// offset.  `part` needs ceil(len / SCAN_TILE) + 1 entries.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

static __device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* ws, uint32_t& total) {
    const uint32_t incl = wave_incl_scan(v);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 63) ws[w] = incl;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
        const uint32_t c = ws[k];
        off += k < w ? c : 0u;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return off + incl - v;
}

static __global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const uint32_t* __restrict__ a,
                                                                    int64_t len,
                                                                    uint32_t* __restrict__ part) {
    __shared__ uint32_t ws[SCAN_THREADS / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int64_t i = base + it * SCAN_THREADS + threadIdx.x;
        s += i < len ? a[i] : 0u;
    }
    uint32_t tot;
    block_excl_scan(s, ws, tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// single-workgroup scan of a short array (the tile sums); total to a[len]
static __global__ __launch_bounds__(1024) void k_scan_small(uint32_t* __restrict__ a, int64_t len) {
    __shared__ uint32_t ws[16];
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < len; b0 += 1024) {
        const int64_t i = b0 + threadIdx.x;
        const uint32_t v = i < len ? a[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, ws, tot);
        if (i < len) a[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) a[len] = carry;
}

static __global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(uint32_t* __restrict__ a,
                                                                    int64_t len,
                                                                    const uint32_t* __restrict__ part) {
    __shared__ uint32_t ws[SCAN_THREADS / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int64_t i = base + it;
        v[it] = i < len ? a[i] : 0u;
        s += v[it];
    }
    uint32_t tot;
    uint32_t run = part[blockIdx.x] + block_excl_scan(s, ws, tot);
#pragma unroll
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        const int64_t i = base + it;
        if (i < len) a[i] = run;
        run += v[it];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1) a[len] = part[gridDim.x];
}

// One-workgroup scan of an array of at most 1024 * SCAN_ONE_ITEMS entries:
// thread t owns a contiguous run of ceil(len / 1024) entries.
constexpr int SCAN_ONE_ITEMS = 32;
static __global__ __launch_bounds__(1024) void k_scan_one(uint32_t* __restrict__ a, int64_t len) {
    __shared__ uint32_t ws[16];
    const int per = (int)((len + 1023) / 1024);
    const int64_t b = (int64_t)threadIdx.x * per;
    uint32_t v[SCAN_ONE_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int it = 0; it < SCAN_ONE_ITEMS; ++it) {
        const int64_t i = b + it;
        v[it] = (it < per && i < len) ? a[i] : 0u;
        sum += v[it];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(sum, ws, tot);
#pragma unroll
    for (int it = 0; it < SCAN_ONE_ITEMS; ++it) {
        const int64_t i = b + it;
        if (it < per && i < len) a[i] = run;
        run += v[it];
    }
    if (threadIdx.x == 0) a[len] = tot;
}

// Exclusive scan in place; grand total to a[len]; part: scratch (>= tiles + 1)
static inline void scan_excl(uint32_t* a, int64_t len, uint32_t* part, hipStream_t s) {
    const int tiles = (int)((len + SCAN_TILE - 1) / SCAN_TILE);
    if (len <= 1024 * SCAN_ONE_ITEMS) {
        k_scan_one<<<1, 1024, 0, s>>>(a, len);
        return;
    }
    k_scan_tiles<<<tiles, SCAN_THREADS, 0, s>>>(a, len, part);
    k_scan_small<<<1, 1024, 0, s>>>(part, tiles);
    k_scan_apply<<<tiles, SCAN_THREADS, 0, s>>>(a, len, part);
}

static inline int64_t scan_part_size(int64_t len) { return (len + SCAN_TILE - 1) / SCAN_TILE + 1; }

}  // namespace hq
