// hq_metrics.hip -- the benchmark harness's force error metric on the GPU.
//
// evaluate_error_on_accelerations (benchmarks/test_forces.f90:66-127):
//   errors(1)   = SQRT((NORM2(accel - accel_dir) / a_scale)**2 / N),
//                 accel = F / m, a_scale = SUM m / |q|^2  (McMillan-Aarseth)
//   errors(2:4) = SUM F
//   errors(5:7) = SUM q x F
// Each workgroup reduces a fixed slice in a fixed order and a final workgroup
// adds the partials in block order, so the result does not depend on timing.
#include <math.h>

#include <algorithm>

#include "hq_internal.h"

namespace hq {

constexpr int EB = 256;    // threads per workgroup
constexpr int ENB = 1024;  // workgroups of the partial pass (at most)

__device__ __forceinline__ double fcomp(const double* F, int layout, int64_t n, int64_t p, int c) {
    return layout == HQ_LAYOUT_3N ? F[3 * p + c] : F[(int64_t)c * n + p];
}

__global__ __launch_bounds__(EB) void k_err_partial(int64_t n, const double* __restrict__ m,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ z,
                                                    const double* __restrict__ F,
                                                    const double* __restrict__ Fd, int layout,
                                                    double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double red[8][EB];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t p = blockIdx.x * (int64_t)EB + threadIdx.x; p < n; p += (int64_t)gridDim.x * EB) {
        const double mp = m[p], q0 = x[p], q1 = y[p], q2 = z[p];
        const double f0 = fcomp(F, layout, n, p, 0), f1 = fcomp(F, layout, n, p, 1),
                     f2 = fcomp(F, layout, n, p, 2);
        const double d0 = f0 / mp - fcomp(Fd, layout, n, p, 0) / mp;
        const double d1 = f1 / mp - fcomp(Fd, layout, n, p, 1) / mp;
        const double d2 = f2 / mp - fcomp(Fd, layout, n, p, 2) / mp;
        acc[0] = acc[0] + (d0 * d0 + d1 * d1 + d2 * d2);
        acc[1] = acc[1] + mp / (q0 * q0 + q1 * q1 + q2 * q2);
        acc[2] = acc[2] + f0;
        acc[3] = acc[3] + f1;
        acc[4] = acc[4] + f2;
        acc[5] = acc[5] + (q1 * f2 - f1 * q2);    // cross_product (:58-63)
        acc[6] = acc[6] + (-q0 * f2 + f0 * q2);
        acc[7] = acc[7] + (q0 * f1 - f0 * q1);
    }
    for (int c = 0; c < 8; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int o = EB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int c = 0; c < 8; ++c) red[c][threadIdx.x] = red[c][threadIdx.x] + red[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 8) part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void k_err_final(const double* __restrict__ part, int nb,
                                                  int64_t n, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int c = threadIdx.x;
    if (c >= 8) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s = s + part[(int64_t)c * nb + b];
    out[c] = s;
}

int errors_ma(int64_t n, const double* m, const double* x, const double* y, const double* z,
              const double* F, const double* Fd, int layout, double* out8, DevBuf& part,
              hipStream_t s) {
    const int nb = (int)std::min<int64_t>(ENB, grid_for(n, EB));
    int rc;
    if ((rc = part.ensure(((size_t)8 * nb + 8) * sizeof(double)))) return rc;
    k_err_partial<<<nb, EB, 0, s>>>(n, m, x, y, z, F, Fd, layout, part.as<double>());
    k_err_final<<<1, 64, 0, s>>>(part.as<double>(), nb, n, part.as<double>() + 8 * nb);
    HQ_CHECK_HIP(hipGetLastError());
    HQ_CHECK_HIP(hipMemcpyAsync(out8, part.as<double>() + 8 * nb, 8 * sizeof(double),
                                hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    return HQ_OK;
}

}  // namespace hq

using namespace hq;

extern "C" int hq_errors_ma(int64_t n, const double* m, const double* x, const double* y,
                            const double* z, const double* F, const double* Fd, int layout,
                            int where, double* errors) {
    if (n < 1 || !m || !x || !y || !z || !F || !Fd || !errors)
        return fail(HQ_ERR_ARGS, "bad argument");
    if (layout != HQ_LAYOUT_3N && layout != HQ_LAYOUT_N3) return fail(HQ_ERR_ARGS, "bad layout");
    if (where != HQ_HOST && where != HQ_DEVICE) return fail(HQ_ERR_ARGS, "bad location");
    int rc;
    DevBuf dm, dx, dy, dz, dF, dFd, part;
    const double* p[6] = {m, x, y, z, F, Fd};
    if (where == HQ_HOST) {
        DevBuf* d[6] = {&dm, &dx, &dy, &dz, &dF, &dFd};
        const int64_t len[6] = {n, n, n, n, 3 * n, 3 * n};
        for (int k = 0; k < 6; ++k) {
            if ((rc = d[k]->ensure(len[k] * sizeof(double)))) return rc;
            HQ_CHECK_HIP(hipMemcpy(d[k]->p, p[k], len[k] * sizeof(double), hipMemcpyHostToDevice));
            p[k] = d[k]->as<double>();
        }
    }
    double out[8];
    if ((rc = errors_ma(n, p[0], p[1], p[2], p[3], p[4], p[5], layout, out, part, nullptr)))
        return rc;
    const double r = sqrt(out[0]) / out[1];
    errors[0] = sqrt(r * r / (double)n);
    for (int c = 0; c < 6; ++c) errors[1 + c] = out[2 + c];
    return HQ_OK;
}
