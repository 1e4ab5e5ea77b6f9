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

// One workgroup per (tile of DT targets, slice of the sources): the sources
// [b_lo, b_hi) in LDS tiles.  With one slice F is the force; with several,
// part holds each slice's sum for k_direct_reduce.
__global__ __launch_bounds__(DT) void k_direct_fp64(int64_t n, const double* __restrict__ m,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z, int64_t nt,
                                                    const int64_t* __restrict__ targets,
                                                    double G, double eps2, int64_t slice,
                                                    double* __restrict__ F) {
    __shared__ double4 tile[DT];
    const int64_t t = blockIdx.x * (int64_t)DT + threadIdx.x;
    const bool ok = t < nt;
    const int64_t a = ok ? (targets ? targets[t] : t) : 0;
    const int64_t b_lo = (int64_t)blockIdx.y * slice;
    const int64_t b_hi = b_lo + slice < n ? b_lo + slice : n;
    const double xa = x[a], ya = y[a], za = z[a], ma = m[a];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int64_t b0 = b_lo; b0 < b_hi; b0 += DT) {
        const int64_t b = b0 + threadIdx.x;
        __syncthreads();
        tile[threadIdx.x] = b < b_hi ? make_double4(x[b], y[b], z[b], m[b]) : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int cnt = (int)((b_hi - b0) < DT ? (b_hi - b0) : DT);
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
        double* o = F + 3 * ((int64_t)blockIdx.y * nt + t);
        o[0] = fx;
        o[1] = fy;
        o[2] = fz;
    }
}

// F = the slices' sums in slice order
__global__ __launch_bounds__(256) void k_direct_reduce(const double* __restrict__ part,
                                                       int64_t nt, int slices,
                                                       double* __restrict__ F) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= 3 * nt) return;
    double acc = 0.0;
    for (int k = 0; k < slices; ++k) acc += part[(int64_t)k * 3 * nt + i];
    F[i] = acc;
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

// Source slices for nt targets: enough workgroups to fill the GPU when the
// targets are few (a sample), each slice at least 2048 sources.
static int direct_slices(int64_t n, int64_t nt) {
    const int64_t tiles = (nt + DT - 1) / DT;
    int64_t sl = (2048 + tiles - 1) / tiles;
    const int64_t most = (n + 2047) / 2048;
    if (sl > most) sl = most;
    if (sl > 1024) sl = 1024;
    return sl < 1 ? 1 : (int)sl;
}

int direct_forces(int64_t n, const double* m, const double* x, const double* y, const double* z,
                  int64_t nt, const int64_t* targets, double G, double eps2, int precision,
                  double* F, hipStream_t s, DevBuf* scratch) {
    if (nt <= 0) return HQ_OK;
    if (precision == HQ_MIXED) {
        k_direct_mixed<<<grid_for(nt, DT), DT, 0, s>>>(n, m, x, y, z, nt, targets, G, eps2, F);
        HQ_CHECK_HIP(hipGetLastError());
        return HQ_OK;
    }
    const int S = direct_slices(n, nt);
    if (S == 1) {
        k_direct_fp64<<<dim3(grid_for(nt, DT), 1), DT, 0, s>>>(n, m, x, y, z, nt, targets, G, eps2,
                                                              n, F);
        HQ_CHECK_HIP(hipGetLastError());
        return HQ_OK;
    }
    DevBuf local;
    DevBuf& part = scratch ? *scratch : local;
    int rc;
    if ((rc = part.ensure((size_t)S * 3 * nt * sizeof(double)))) return rc;
    const int64_t slice = ((n + S - 1) / S + DT - 1) / DT * DT;
    k_direct_fp64<<<dim3(grid_for(nt, DT), S), DT, 0, s>>>(n, m, x, y, z, nt, targets, G, eps2,
                                                          slice, part.as<double>());
    k_direct_reduce<<<grid_for(3 * nt, 256), 256, 0, s>>>(part.as<double>(), nt, S, F);
    HQ_CHECK_HIP(hipGetLastError());
    if (!scratch) HQ_CHECK_HIP(hipStreamSynchronize(s));  // before the local buffer goes
    return HQ_OK;
}

}  // namespace hq
