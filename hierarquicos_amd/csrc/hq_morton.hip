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

int morton_keys(int64_t n, const double* m, const double* q, int L, double ssa, int64_t* keys,
                int* perm, int* L_out, hipStream_t s, double* par_out) {
    int rc;
    DevBuf par, mm, key, idx, ktmp, vtmp, hist, part, lmax;
    if ((rc = par.ensure(16 * sizeof(double))) || (rc = mm.ensure(12 * sizeof(uint64_t))) ||
        (rc = key.ensure(n * 8)) || (rc = idx.ensure(n * 4)) || (rc = ktmp.ensure(n * 8)) ||
        (rc = vtmp.ensure(n * 4)) || (rc = lmax.ensure(sizeof(int))))
        return rc;
    unsigned long long init[12];
    for (int c = 0; c < 12; ++c) init[c] = (c & 1) ? 0ull : ~0ull;
    HQ_CHECK_HIP(hipMemcpyAsync(mm.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    HQ_CHECK_HIP(hipMemsetAsync(lmax.p, 0, sizeof(int), s));
    const int nb = (int)std::min<int64_t>(1024, grid_for(n, 256));
    k_morton_minmax<<<nb, 256, 0, s>>>(n, q, par.as<double>(), 0, mm.as<unsigned long long>());
    k_morton_ssa<<<1, 64, 0, s>>>(mm.as<unsigned long long>(), ssa, par.as<double>());
    k_morton_com<<<1, 64, 0, s>>>(n, m, q, par.as<double>());
    k_morton_minmax<<<nb, 256, 0, s>>>(n, q, par.as<double>(), 1, mm.as<unsigned long long>());
    k_morton_centre<<<1, 64, 0, s>>>(mm.as<unsigned long long>(), par.as<double>());
    k_morton_key21<<<grid_for(n, 256), 256, 0, s>>>(n, q, par.as<double>(), key.as<uint64_t>(),
                                                    idx.as<int>());
    int Lu = L;
    if (L == 0) {
        if ((rc = radix_sort_pairs(key.as<uint64_t>(), idx.as<int>(), ktmp.as<uint64_t>(),
                                   vtmp.as<int>(), n, 3 * ML, hist, part, s)))
            return rc;
        k_max_lcp<<<grid_for(n, 256), 256, 0, s>>>(n, key.as<uint64_t>(), lmax.as<int>());
        int mx = 0;
        HQ_CHECK_HIP(hipMemcpyAsync(&mx, lmax.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        if (mx >= ML) return fail(HQ_ERR_DEPTH, "keys are not distinct at 21 levels (64-bit keys)");
        Lu = mx + 1;
        k_shift_keys<<<grid_for(n, 256), 256, 0, s>>>(n, key.as<uint64_t>(), 3 * (ML - Lu));
    } else {
        if (L < 0 || L > ML) return fail(HQ_ERR_DEPTH, "Morton depth %d outside 1..21", L);
        k_shift_keys<<<grid_for(n, 256), 256, 0, s>>>(n, key.as<uint64_t>(), 3 * (ML - L));
        if ((rc = radix_sort_pairs(key.as<uint64_t>(), idx.as<int>(), ktmp.as<uint64_t>(),
                                   vtmp.as<int>(), n, 3 * L, hist, part, s)))
            return rc;
    }
    k_keys_to_orig<<<grid_for(n, 256), 256, 0, s>>>(n, key.as<uint64_t>(), idx.as<int>(), keys);
    HQ_CHECK_HIP(hipMemcpyAsync(perm, idx.p, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    if (par_out)  // qcm (3), mass, ssa, normalised-box centre (3)
        HQ_CHECK_HIP(hipMemcpyAsync(par_out, par.p, 8 * sizeof(double), hipMemcpyDeviceToDevice, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    HQ_CHECK_HIP(hipGetLastError());
    *L_out = Lu;
    return HQ_OK;
}

}  // namespace hq
