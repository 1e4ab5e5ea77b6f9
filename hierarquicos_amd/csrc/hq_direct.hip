// hq_direct.hip -- O(N^2) direct sum (fortran/direct_forces.f90:8-64).
//
// Gather form: one target per lane, source tiles staged in LDS, no atomics
// (the reference's per-thread REDUCTION(+:forces) copies disappear).  The
// pair force is the reference's f = G m_a m_b / (sqrt(d2) d2), d2 = |d|^2 +
// eps^2 (:29-36).  HQ_FP64 keeps that arithmetic; HQ_MIXED forms d in FP64
// (exact recentring on the target), rounds it to FP32, evaluates the pair in
// FP32 with rsqrt and accumulates each LDS tile into a Kahan-compensated sum.
#include "hq_internal.h"

namespace hq {

constexpr int DT = 256;  // targets per workgroup = sources per LDS tile

__global__ __launch_bounds__(DT) void k_direct_fp64(int64_t n, const double* __restrict__ m,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z, int64_t nt,
                                                    const int64_t* __restrict__ targets,
                                                    double G, double eps2,
                                                    double* __restrict__ F) {
    __shared__ double4 tile[DT];
    const int64_t t = blockIdx.x * (int64_t)DT + threadIdx.x;
    const bool ok = t < nt;
    const int64_t a = ok ? (targets ? targets[t] : t) : 0;
    const double xa = x[a], ya = y[a], za = z[a], ma = m[a];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int64_t b0 = 0; b0 < n; b0 += DT) {
        const int64_t b = b0 + threadIdx.x;
        __syncthreads();
        tile[threadIdx.x] = b < n ? make_double4(x[b], y[b], z[b], m[b]) : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int cnt = (int)((n - b0) < DT ? (n - b0) : DT);
        for (int j = 0; j < cnt; ++j) {
            const double4 s = tile[j];
            if (b0 + j == a) continue;
            const double dx = s.x - xa, dy = s.y - ya, dz = s.z - za;
            const double dist2 = dx * dx + dy * dy + dz * dz + eps2;
            const double dist = sqrt(dist2);
            const double f = G * ma * s.w / (dist * dist2);
            fx += f * dx;
            fy += f * dy;
            fz += f * dz;
        }
    }
    if (ok) {
        F[3 * t] = fx;
        F[3 * t + 1] = fy;
        F[3 * t + 2] = fz;
    }
}

__global__ __launch_bounds__(DT) void k_direct_mixed(int64_t n, const double* __restrict__ m,
                                                     const double* __restrict__ x,
                                                     const double* __restrict__ y,
                                                     const double* __restrict__ z, int64_t nt,
                                                     const int64_t* __restrict__ targets,
                                                     double G, double eps2,
                                                     double* __restrict__ F) {
    __shared__ double4 tile[DT];
    const int64_t t = blockIdx.x * (int64_t)DT + threadIdx.x;
    const bool ok = t < nt;
    const int64_t a = ok ? (targets ? targets[t] : t) : 0;
    const double xa = x[a], ya = y[a], za = z[a];
    const float e2 = (float)eps2;
    float s[3] = {0.f, 0.f, 0.f}, c[3] = {0.f, 0.f, 0.f};
    for (int64_t b0 = 0; b0 < n; b0 += DT) {
        const int64_t b = b0 + threadIdx.x;
        __syncthreads();
        tile[threadIdx.x] = b < n ? make_double4(x[b], y[b], z[b], m[b]) : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int cnt = (int)((n - b0) < DT ? (n - b0) : DT);
        float p[3] = {0.f, 0.f, 0.f};
        for (int j = 0; j < cnt; ++j) {
            const double4 q = tile[j];
            const float dx = (float)(q.x - xa), dy = (float)(q.y - ya), dz = (float)(q.z - za);
            const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, e2)));
            const float ri = __frsqrt_rn(d2);
            float f = (float)q.w * ri * ri * ri;
            f = (b0 + j == a) ? 0.0f : f;
            p[0] = fmaf(f, dx, p[0]);
            p[1] = fmaf(f, dy, p[1]);
            p[2] = fmaf(f, dz, p[2]);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float yv = p[k] - c[k];
            float tv = s[k] + yv;
            c[k] = (tv - s[k]) - yv;
            s[k] = tv;
        }
    }
    if (ok) {
        const double gm = G * m[a];
        F[3 * t] = gm * ((double)s[0] - (double)c[0]);
        F[3 * t + 1] = gm * ((double)s[1] - (double)c[1]);
        F[3 * t + 2] = gm * ((double)s[2] - (double)c[2]);
    }
}

int direct_forces(int64_t n, const double* m, const double* x, const double* y, const double* z,
                  int64_t nt, const int64_t* targets, double G, double eps2, int precision,
                  double* F, hipStream_t s) {
    if (nt <= 0) return HQ_OK;
    if (precision == HQ_MIXED)
        k_direct_mixed<<<grid_for(nt, DT), DT, 0, s>>>(n, m, x, y, z, nt, targets, G, eps2, F);
    else
        k_direct_fp64<<<grid_for(nt, DT), DT, 0, s>>>(n, m, x, y, z, nt, targets, G, eps2, F);
    HQ_CHECK_HIP(hipGetLastError());
    return HQ_OK;
}

}  // namespace hq
