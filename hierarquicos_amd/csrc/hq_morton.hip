// hq_morton.hip -- Morton keys and sort permutation of the reference's Morton
// octree (fortran/morton/morton_octree.f90 generate_nodes :357-417), bit for
// bit:
//   * ssa = amplificator * max extent of the raw positions (:357-358)
//   * centre of mass by a left-to-right sum (:361-367): order-dependent
//     rounding, so it runs as one sequential lane (k_morton_com)
//   * q_n = (q - qcm) / ssa, then + 0.5 - centre of the normalised bbox (:370-384)
//   * ix = min(floor(q_n * 2^L), 2^L - 1), interleaved MSB first, x the high bit
//     of each octal digit (morton_key :184-213)
//   * stable argsort (mergesort :501-560) == stable LSD radix sort
// Depth: floor(q * 2^L) = floor(q * 2^21) >> (21 - L) exactly (power-of-two
// scaling), so every key is a shift of the 21-level key; with L = 0 the
// depth is choose_depth's (:217-244) smallest L with distinct keys, i.e. one
// more than the longest common prefix of neighbours in the 21-level order.
#include <math.h>

#include <algorithm>
#include <mutex>

#include "hq_tree.h"

namespace hq {

constexpr int ML = 21;

__global__ void k_morton_com(int64_t n, const double* __restrict__ m, const double* __restrict__ q,
                             double* __restrict__ out) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double cx = 0.0, cy = 0.0, cz = 0.0, mass = 0.0;
#pragma unroll 8
    for (int64_t p = 0; p < n; ++p) {
        const double mp = m[p];
        cx = cx + mp * q[3 * p];
        cy = cy + mp * q[3 * p + 1];
        cz = cz + mp * q[3 * p + 2];
        mass = mass + mp;
    }
    out[0] = cx / mass;
    out[1] = cy / mass;
    out[2] = cz / mass;
    out[3] = mass;
}

// min/max of raw positions (slots 0..5) or of normalised ones (slots 6..11)
__global__ __launch_bounds__(256) void k_morton_minmax(int64_t n, const double* __restrict__ q,
                                                       const double* __restrict__ par, int norm,
                                                       unsigned long long* __restrict__ mm) {
#pragma clang fp contract(off)
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) {
            double v = q[3 * p + c];
            if (norm) v = (v - par[c]) / par[4];
            mn[c] = fmin(mn[c], v);
            mx[c] = fmax(mx[c], v);
        }
    for (int c = 0; c < 3; ++c) {
        for (int o = 32; o > 0; o >>= 1) {
            mn[c] = fmin(mn[c], __shfl_xor(mn[c], o, 64));
            mx[c] = fmax(mx[c], __shfl_xor(mx[c], o, 64));
        }
    }
    // order-preserving integer image of a double for atomic min / max
    auto ord = [](double d) -> unsigned long long {
        unsigned long long u = __double_as_longlong(d);
        return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
    };
    if (__lane_id() == 0) {
        const int o = norm ? 6 : 0;
        for (int c = 0; c < 3; ++c) {
            atomicMin(&mm[o + 2 * c], ord(mn[c]));
            atomicMax(&mm[o + 2 * c + 1], ord(mx[c]));
        }
    }
}

__device__ __forceinline__ double unord(unsigned long long u) {
    return __longlong_as_double((u >> 63) ? (u & 0x7FFFFFFFFFFFFFFFull) : ~u);
}

__global__ void k_morton_ssa(const unsigned long long* __restrict__ mm, double amp,
                             double* __restrict__ par) {
#pragma clang fp contract(off)
    if (threadIdx.x) return;
    double ext = unord(mm[1]) - unord(mm[0]);
    ext = fmax(ext, unord(mm[3]) - unord(mm[2]));
    ext = fmax(ext, unord(mm[5]) - unord(mm[4]));
    par[4] = amp * ext;
}

__global__ void k_morton_centre(const unsigned long long* __restrict__ mm, double* __restrict__ par) {
#pragma clang fp contract(off)
    if (threadIdx.x) return;
    for (int c = 0; c < 3; ++c) par[5 + c] = 0.5 * (unord(mm[6 + 2 * c]) + unord(mm[6 + 2 * c + 1]));
}

__global__ __launch_bounds__(256) void k_morton_key21(int64_t n, const double* __restrict__ q,
                                                      const double* __restrict__ par,
                                                      uint64_t* __restrict__ key,
                                                      int* __restrict__ idx) {
#pragma clang fp contract(off)
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t scale = (int64_t)1 << ML;
    int64_t iq[3];
    for (int c = 0; c < 3; ++c) {
        double v = (q[3 * p + c] - par[c]) / par[4];
        v = v + 0.5 - par[5 + c];
        int64_t f = (int64_t)floor(v * (double)scale);
        iq[c] = f < scale - 1 ? f : scale - 1;
    }
    uint64_t k = 0;
    for (int b = ML - 1; b >= 0; --b)
        k = (k << 3) | (uint64_t)((((iq[0] >> b) & 1) << 2) | (((iq[1] >> b) & 1) << 1) | ((iq[2] >> b) & 1));
    key[p] = k;
    idx[p] = (int)p;
}

__global__ __launch_bounds__(256) void k_max_lcp(int64_t n, const uint64_t* __restrict__ key,
                                                 int* __restrict__ out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int d = 0;
    if (i + 1 < n) {
        uint64_t a = key[i] ^ key[i + 1];
        d = a ? (__clzll((long long)a) - 1) / 3 : ML;
    }
    int m = d;
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if (__lane_id() == 0) atomicMax(out, m);
}

__global__ __launch_bounds__(256) void k_shift_keys(int64_t n, uint64_t* __restrict__ key, int sh) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] >>= sh;
}

__global__ __launch_bounds__(256) void k_keys_to_orig(int64_t n, const uint64_t* __restrict__ skey,
                                                      const int* __restrict__ perm,
                                                      int64_t* __restrict__ keys_out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys_out[perm[i]] = (int64_t)skey[i];
}

// Work arrays of one key computation: par (16 doubles: centre of mass, mass,
// ssa, normalised-box centre), mm (12 order-mapped min/max words), lmax (the
// longest common prefix), the sorted keys and indices with their ping-pong
// copies, and the sort's count table (rs_hist_words(n) words).
struct MortonWork {
    double* par;
    unsigned long long* mm;
    int* lmax;
    uint64_t* key;
    int* idx;
    uint64_t* ktmp;
    int* vtmp;
    uint32_t* hist;
};

// starting values of the min/max words and the prefix maximum
__global__ void k_morton_init(unsigned long long* __restrict__ mm, int* __restrict__ lmax) {
    const int c = threadIdx.x;
    if (c < 12) mm[c] = (c & 1) ? 0ull : ~0ull;
    if (c == 0) *lmax = 0;
}

// Keys into w.key / w.idx, then out to keys (original order) and perm.  The
// ping-pong copies w.ktmp / w.vtmp may be the output arrays themselves: the
// sorted keys and indices end in w.key / w.idx before the outputs are written.
static int morton_keys_ws(int64_t n, const double* m, const double* q, int L, double ssa,
                          int64_t* keys, int* perm, int* L_out, const MortonWork& w,
                          hipStream_t s, double* par_out) {
    int rc;
    if (L < 0 || L > ML) return fail(HQ_ERR_DEPTH, "Morton depth %d outside 1..21", L);
    k_morton_init<<<1, 64, 0, s>>>(w.mm, w.lmax);
    const int nb = (int)std::min<int64_t>(1024, grid_for(n, 256));
    k_morton_minmax<<<nb, 256, 0, s>>>(n, q, w.par, 0, w.mm);
    k_morton_ssa<<<1, 64, 0, s>>>(w.mm, ssa, w.par);
    k_morton_com<<<1, 64, 0, s>>>(n, m, q, w.par);
    k_morton_minmax<<<nb, 256, 0, s>>>(n, q, w.par, 1, w.mm);
    k_morton_centre<<<1, 64, 0, s>>>(w.mm, w.par);
    k_morton_key21<<<grid_for(n, 256), 256, 0, s>>>(n, q, w.par, w.key, w.idx);
    int Lu = L;
    if (L == 0) {
        if ((rc = radix_sort_pairs_ws(w.key, w.idx, w.ktmp, w.vtmp, n, 3 * ML, w.hist, s))) return rc;
        k_max_lcp<<<grid_for(n, 256), 256, 0, s>>>(n, w.key, w.lmax);
        int mx = 0;
        HQ_CHECK_HIP(hipMemcpyAsync(&mx, w.lmax, sizeof(int), hipMemcpyDeviceToHost, s));
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        if (mx >= ML) return fail(HQ_ERR_DEPTH, "keys are not distinct at 21 levels (64-bit keys)");
        Lu = mx + 1;
        k_shift_keys<<<grid_for(n, 256), 256, 0, s>>>(n, w.key, 3 * (ML - Lu));
    } else {
        k_shift_keys<<<grid_for(n, 256), 256, 0, s>>>(n, w.key, 3 * (ML - L));
        if ((rc = radix_sort_pairs_ws(w.key, w.idx, w.ktmp, w.vtmp, n, 3 * L, w.hist, s))) return rc;
    }
    k_keys_to_orig<<<grid_for(n, 256), 256, 0, s>>>(n, w.key, w.idx, keys);
    HQ_CHECK_HIP(hipMemcpyAsync(perm, w.idx, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    if (par_out)  // qcm (3), mass, ssa, normalised-box centre (3)
        HQ_CHECK_HIP(hipMemcpyAsync(par_out, w.par, 8 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    HQ_CHECK_HIP(hipGetLastError());
    *L_out = Lu;
    return HQ_OK;
}

int morton_keys(int64_t n, const double* m, const double* q, int L, double ssa, int64_t* keys,
                int* perm, int* L_out, hipStream_t s, double* par_out) {
    int rc;
    DevBuf par, mm, key, idx, ktmp, vtmp, hist, lmax;
    if ((rc = par.ensure(16 * sizeof(double))) || (rc = mm.ensure(12 * sizeof(uint64_t))) ||
        (rc = key.ensure(n * 8)) || (rc = idx.ensure(n * 4)) || (rc = ktmp.ensure(n * 8)) ||
        (rc = vtmp.ensure(n * 4)) || (rc = lmax.ensure(sizeof(int))) ||
        (rc = hist.ensure(rs_hist_words(n) * sizeof(uint32_t))))
        return rc;
    const MortonWork w{par.as<double>(), mm.as<unsigned long long>(), lmax.as<int>(),
                       key.as<uint64_t>(), idx.as<int>(), ktmp.as<uint64_t>(), vtmp.as<int>(),
                       hist.as<uint32_t>()};
    return morton_keys_ws(n, m, q, L, ssa, keys, perm, L_out, w, s, par_out);
}

}  // namespace hq

// ---------------------------------------------------------------------------
// hq_morton_keys_noalloc: the device form of hq_morton_keys for callers that
// cannot allocate (stream capture, serving loops).  Its work arrays live in
// one fixed file-scope device array, sized for MORTON_NOALLOC_MAX keys; the
// caller's keys and perm outputs double as the sort's ping-pong copies.  Every
// call on a GPU shares that GPU's array, so calls that overlap on one GPU race.

namespace hq {

constexpr int64_t MORTON_NOALLOC_MAX = (int64_t)1 << 20;
constexpr size_t MW_PAR = 0;                    // 16 doubles
constexpr size_t MW_MM = 128;                   // 12 words
constexpr size_t MW_LMAX = 224;                 // 1 int
constexpr size_t MW_KEY = 256;                  // MORTON_NOALLOC_MAX u64
constexpr size_t MW_IDX = MW_KEY + 8 * (size_t)MORTON_NOALLOC_MAX;  // MORTON_NOALLOC_MAX int
constexpr size_t MW_HIST = MW_IDX + 4 * (size_t)MORTON_NOALLOC_MAX;
// count table at the largest size: 256 counts per 2048-key tile plus totals
// (tiles of 512 keys below 2^18 keys give fewer tiles)
constexpr size_t MW_HIST_WORDS = 256 * (size_t)(MORTON_NOALLOC_MAX / 2048) + 256;
constexpr size_t MW_BYTES = MW_HIST + 4 * MW_HIST_WORDS;  // about 12.5 MiB

static __device__ __align__(256) unsigned char scratch[MW_BYTES];

constexpr int MW_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[MW_MAX_DEVICES] = {};

// Device address of the current device's copy of `scratch`, looked up once
// per device with that device current.
static int scratch_of_current_device(unsigned char** out) {
    int dev = 0;
    HQ_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MW_MAX_DEVICES) return fail(HQ_ERR_ARGS, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HQ_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(HQ_ERR_HIP, "no device address for the key workspace");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return HQ_OK;
}

}  // namespace hq

extern "C" int hq_morton_keys_noalloc(int64_t n, const double* m, const double* q, int L,
                                      double ssa, int64_t* keys, int* perm, int* L_out) {
    using namespace hq;
    if (n < 1 || !m || !q || !keys || !perm || !L_out) return fail(HQ_ERR_ARGS, "bad argument");
    if (n > MORTON_NOALLOC_MAX)
        return fail(HQ_ERR_ARGS, "hq_morton_keys_noalloc: %lld keys exceed its fixed workspace "
                    "of %lld keys; use hq_morton_keys", (long long)n, (long long)MORTON_NOALLOC_MAX);
    if (rs_hist_words(n) > MW_HIST_WORDS) return fail(HQ_ERR_ARGS, "sort table exceeds the workspace");
    unsigned char* base = nullptr;
    int rc;
    if ((rc = scratch_of_current_device(&base))) return rc;
    const MortonWork w{reinterpret_cast<double*>(base + MW_PAR),
                       reinterpret_cast<unsigned long long*>(base + MW_MM),
                       reinterpret_cast<int*>(base + MW_LMAX),
                       reinterpret_cast<uint64_t*>(base + MW_KEY),
                       reinterpret_cast<int*>(base + MW_IDX),
                       reinterpret_cast<uint64_t*>(keys), perm,
                       reinterpret_cast<uint32_t*>(base + MW_HIST)};
    return morton_keys_ws(n, m, q, L, ssa, keys, perm, L_out, w, nullptr, nullptr);
}
