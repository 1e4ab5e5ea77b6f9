// hq_dehnen.hip -- Dehnen's dual-tree method, methods 20/21/22
// (fortran/octree_default.f90 evaluate_rmax :1095-1130, mutual walk
// :1132-1248, evaluate_mutual_interaction_monopole :1425-1445,
// evaluate_mutual_interaction_quadrupole :1447-1563 (Marcello's torque
// correction for method 22), dehnen_forces :1565-1678).
//
//   k_rmax        bottom-up bounding radius about the COM, one level per launch
//   k_dual_*      the dual walk as a breadth-first frontier of node pairs:
//                 each pair is dropped, accepted (mutual interaction) or split,
//                 with the reference's rules -- (a,a) expands to the child pairs
//                 i <= j in octant order, otherwise the node with the larger
//                 rmax is opened (a on ties).  Outputs are compacted by a
//                 scan, so the pair lists are deterministic.
//   k_m2l         mutual M2L of every accepted pair into node force / Hessian /
//                 third-derivative accumulators (FP64 atomics)
//   k_down        top-down L2L/L2P, one parent level per launch
#include <math.h>

#include <algorithm>

#include "hq_tree.h"

namespace hq {

struct DView {
    const double *mass, *cx, *cy, *cz, *quad, *oct;
    const int *first, *cstart;
    const int8_t* type;
    const uint8_t* digit;
    int64_t stride;
};

static DView dview(const hq_tree* t) {
    DView v;
    v.mass = t->nmass.as<double>();
    v.cx = t->ncomx.as<double>();
    v.cy = t->ncomy.as<double>();
    v.cz = t->ncomz.as<double>();
    v.quad = t->multipole >= 4 ? t->nquad.as<double>() : nullptr;
    v.oct = t->multipole > 4 ? t->noct.as<double>() : nullptr;
    v.first = t->nfirst.as<int>();
    v.cstart = t->ncstart.as<int>();
    v.type = t->ntype.as<int8_t>();
    v.digit = t->ndigit.as<uint8_t>();
    v.stride = t->cap;
    return v;
}

__device__ __forceinline__ int dslot(uint32_t dig) {
    return (int)((dig >> 2) & 1u) + 2 * (int)(1u - ((dig >> 1) & 1u)) + 4 * (int)(1u - (dig & 1u));
}

// ------------------------------------------------------------------ rmax
__global__ __launch_bounds__(128) void k_rmax(int64_t a, int64_t b, DView v,
                                              double* __restrict__ rmax) {
#pragma clang fp contract(off)
    int64_t k = a + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= b) return;
    double r = 0.0;
    if (v.type[k] != 1) {
        for (int c = v.cstart[k]; c < v.cstart[k + 1]; ++c) {
            if (v.mass[c] == 0.0) continue;
            const double dx = v.cx[c] - v.cx[k];
            const double dy = v.cy[c] - v.cy[k];
            const double dz = v.cz[c] - v.cz[k];
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            r = fmax(r, dist + rmax[c]);
        }
    }
    rmax[k] = r;
}

// ------------------------------------------------------------------ dual walk
// action of a pair: 0 drop, 1 accept, 2 self-expand, 3 open a, 4 open b
__device__ __forceinline__ int classify(const DView& v, const double* __restrict__ rmax,
                                        double theta2, int a, int b, int& nnext) {
#pragma clang fp contract(off)
    nnext = 0;
    if (v.mass[a] == 0.0 || v.mass[b] == 0.0) return 0;
    if (a == b) {
        if (v.type[a] == 1) return 0;
        const int nc = v.cstart[a + 1] - v.cstart[a];
        nnext = nc * (nc + 1) / 2;
        return 2;
    }
    const double dx = v.cx[a] - v.cx[b];
    const double dy = v.cy[a] - v.cy[b];
    const double dz = v.cz[a] - v.cz[b];
    const double dist2 = dx * dx + dy * dy + dz * dz;
    const double rs = rmax[a] + rmax[b];
    if (rs * rs <= theta2 * dist2) return 1;
    if (rmax[a] >= rmax[b]) {
        nnext = v.cstart[a + 1] - v.cstart[a];
        return 3;
    }
    nnext = v.cstart[b + 1] - v.cstart[b];
    return 4;
}

constexpr int DB = 256;

__global__ __launch_bounds__(DB) void k_dual_count(DView v, const double* __restrict__ rmax,
                                                   double theta2, const int2* __restrict__ fr,
                                                   int64_t nf, uint32_t* __restrict__ bacc,
                                                   uint32_t* __restrict__ bnext, int nblk) {
    __shared__ uint32_t sa[DB / 64], sn[DB / 64];
    int64_t i = blockIdx.x * (int64_t)DB + threadIdx.x;
    uint32_t acc = 0, nx = 0;
    if (i < nf) {
        int nn;
        int act = classify(v, rmax, theta2, fr[i].x, fr[i].y, nn);
        acc = act == 1;
        nx = (uint32_t)nn;
    }
    acc = wave_sum(acc);
    nx = wave_sum(nx);
    if (__lane_id() == 0) {
        sa[threadIdx.x >> 6] = acc;
        sn[threadIdx.x >> 6] = nx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s1 = 0, s2 = 0;
        for (int w = 0; w < DB / 64; ++w) { s1 += sa[w]; s2 += sn[w]; }
        bacc[blockIdx.x] = s1;
        bnext[blockIdx.x] = s2;
    }
    (void)nblk;
}

__global__ __launch_bounds__(DB) void k_dual_emit(DView v, const double* __restrict__ rmax,
                                                  double theta2, const int2* __restrict__ fr,
                                                  int64_t nf, const uint32_t* __restrict__ bacc,
                                                  const uint32_t* __restrict__ bnext,
                                                  int2* __restrict__ out_next,
                                                  int2* __restrict__ out_acc) {
    __shared__ uint32_t wa[DB / 64], wn[DB / 64];
    int64_t i = blockIdx.x * (int64_t)DB + threadIdx.x;
    int act = 0, nn = 0, a = 0, b = 0;
    if (i < nf) {
        a = fr[i].x;
        b = fr[i].y;
        act = classify(v, rmax, theta2, a, b, nn);
    }
    const uint32_t ca = act == 1, cn = (uint32_t)nn;
    const uint32_t ia = wave_incl_scan(ca), in = wave_incl_scan(cn);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 63) { wa[w] = ia; wn[w] = in; }
    __syncthreads();
    uint32_t oa = bacc[blockIdx.x], on = bnext[blockIdx.x];
    for (int k = 0; k < w; ++k) { oa += wa[k]; on += wn[k]; }
    oa += ia - ca;
    on += in - cn;
    if (act == 1) out_acc[oa] = make_int2(a, b);
    if (act == 2) {
        int slot[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
        for (int c = v.cstart[a]; c < v.cstart[a + 1]; ++c) slot[dslot(v.digit[c])] = c;
        uint32_t o = on;
        for (int r1 = 0; r1 < 8; ++r1) {
            if (slot[r1] < 0) continue;
            for (int r2 = r1; r2 < 8; ++r2) {
                if (slot[r2] < 0) continue;
                out_next[o++] = make_int2(slot[r1], slot[r2]);
            }
        }
    } else if (act == 3) {
        uint32_t o = on;
        for (int c = v.cstart[a]; c < v.cstart[a + 1]; ++c) out_next[o++] = make_int2(c, b);
    } else if (act == 4) {
        uint32_t o = on;
        for (int c = v.cstart[b]; c < v.cstart[b + 1]; ++c) out_next[o++] = make_int2(a, c);
    }
}

// ------------------------------------------------------------------ M2L
// ind_oct (:234-242), 0-based octupole slot of (j,k,l)
__constant__ int c_ind_oct[27] = {0, 3, 6, 3, 1, 9, 6, 9, 2, 3, 1, 9, 1, 4, 7, 9, 7, 5,
                                  6, 9, 2, 9, 7, 5, 2, 5, 8};

__device__ __forceinline__ void aadd(double* p, double v) { unsafeAtomicAdd(p, v); }

template <int MP>
__global__ __launch_bounds__(128) void k_m2l(DView v, const int2* __restrict__ pairs, int64_t np,
                                             double G, double eps2, double* __restrict__ force,
                                             double* __restrict__ hess,
                                             double* __restrict__ third, int64_t nn) {
#pragma clang fp contract(off)
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int a = pairs[i].x, b = pairs[i].y;
    const double Ma = v.mass[a], Mb = v.mass[b];
    if (MP == 1) {
        // evaluate_mutual_interaction_monopole (:1425-1445), vector from a to b
        const double dx = v.cx[b] - v.cx[a];
        const double dy = v.cy[b] - v.cy[a];
        const double dz = v.cz[b] - v.cz[a];
        const double dist2 = dx * dx + dy * dy + dz * dz;
        const double rinv3 = 1.0 / pow(dist2 + eps2, 1.5);
        const double factor = G * rinv3;
        const double f[3] = {factor * dx, factor * dy, factor * dz};
        for (int c = 0; c < 3; ++c) {
            aadd(&force[c * nn + a], Mb * f[c]);
            aadd(&force[c * nn + b], -(Ma * f[c]));
        }
        return;
    }
    const int64_t S = v.stride;
    const double dx = v.cx[a] - v.cx[b];
    const double dy = v.cy[a] - v.cy[b];
    const double dz = v.cz[a] - v.cz[b];
    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
    const double rinv3 = 1.0 / pow(r2, 1.5);
    const double rinv5 = 1.0 / pow(r2, 2.5);
    const double rinv7 = 1.0 / pow(r2, 3.5);
    double D1[3], D2[6], D3[10];
    D1[0] = -(G * dx * rinv3);
    D1[1] = -(G * dy * rinv3);
    D1[2] = -(G * dz * rinv3);
    D2[0] = G * (3.0 * dx * dx * rinv5 - rinv3);
    D2[1] = G * (3.0 * dy * dy * rinv5 - rinv3);
    D2[2] = G * (3.0 * dz * dz * rinv5 - rinv3);
    D2[3] = G * 3.0 * dx * dy * rinv5;
    D2[4] = G * 3.0 * dx * dz * rinv5;
    D2[5] = G * 3.0 * dy * dz * rinv5;
    D3[0] = G * (-(15.0 * dx * dx * dx * rinv7) + 9.0 * dx * rinv5);
    D3[1] = G * (-(15.0 * dx * dx * dy * rinv7) + 3.0 * dy * rinv5);
    D3[2] = G * (-(15.0 * dx * dx * dz * rinv7) + 3.0 * dz * rinv5);
    D3[3] = G * (-(15.0 * dx * dy * dy * rinv7) + 3.0 * dx * rinv5);
    D3[4] = G * (-(15.0 * dx * dy * dz * rinv7));
    D3[5] = G * (-(15.0 * dx * dz * dz * rinv7) + 3.0 * dx * rinv5);
    D3[6] = G * (-(15.0 * dy * dy * dy * rinv7) + 9.0 * dy * rinv5);
    D3[7] = G * (-(15.0 * dy * dy * dz * rinv7) + 3.0 * dz * rinv5);
    D3[8] = G * (-(15.0 * dy * dz * dz * rinv7) + 3.0 * dy * rinv5);
    D3[9] = G * (-(15.0 * dz * dz * dz * rinv7) + 9.0 * dz * rinv5);
    double Qa[6], Qb[6];
    for (int j = 0; j < 6; ++j) {
        Qa[j] = v.quad[j * S + a] / Ma;
        Qb[j] = v.quad[j * S + b] / Mb;
    }
    double CQa[3], CQb[3];
    CQa[0] = Qa[0] * D3[0] + 2.0 * Qa[3] * D3[1] + 2.0 * Qa[4] * D3[2] + Qa[1] * D3[3] +
             2.0 * Qa[5] * D3[4] + Qa[2] * D3[5];
    CQb[0] = Qb[0] * D3[0] + 2.0 * Qb[3] * D3[1] + 2.0 * Qb[4] * D3[2] + Qb[1] * D3[3] +
             2.0 * Qb[5] * D3[4] + Qb[2] * D3[5];
    CQa[1] = Qa[0] * D3[1] + 2.0 * Qa[3] * D3[3] + 2.0 * Qa[4] * D3[4] + Qa[1] * D3[6] +
             2.0 * Qa[5] * D3[7] + Qa[2] * D3[8];
    CQb[1] = Qb[0] * D3[1] + 2.0 * Qb[3] * D3[3] + 2.0 * Qb[4] * D3[4] + Qb[1] * D3[6] +
             2.0 * Qb[5] * D3[7] + Qb[2] * D3[8];
    CQa[2] = Qa[0] * D3[2] + 2.0 * Qa[3] * D3[4] + 2.0 * Qa[4] * D3[5] + Qa[1] * D3[7] +
             2.0 * Qa[5] * D3[8] + Qa[2] * D3[9];
    CQb[2] = Qb[0] * D3[2] + 2.0 * Qb[3] * D3[4] + 2.0 * Qb[4] * D3[5] + Qb[1] * D3[7] +
             2.0 * Qb[5] * D3[8] + Qb[2] * D3[9];
    double fa[3], fb[3], fc[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 3; ++c) {
        fa[c] = Mb * (D1[c] + 0.5 * CQb[c]);
        fb[c] = -(Ma * (D1[c] + 0.5 * CQa[c]));
    }
    if (MP == 8) {
        const double R[3] = {dx, dy, dz};
        double ci[3] = {0.0, 0.0, 0.0}, CRR[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    const int io = c_ind_oct[9 * j + 3 * k + l];
                    const double Cjkl = Mb * v.oct[io * S + a] - Ma * v.oct[io * S + b];
                    if (k == l) ci[j] = ci[j] + Cjkl;
                    CRR[j] = CRR[j] + Cjkl * R[k] * R[l];
                }
        for (int j = 0; j < 3; ++j)
            fc[j] = 0.5 * (3.0 * G * rinv5 * ci[j] - 15.0 * G * rinv7 * CRR[j]);
    }
    for (int c = 0; c < 3; ++c) {
        aadd(&force[c * nn + a], fa[c] + fc[c] / Ma);
        aadd(&force[c * nn + b], fb[c] - fc[c] / Mb);
    }
    for (int j = 0; j < 6; ++j) {
        aadd(&hess[j * nn + a], Mb * D2[j]);
        aadd(&hess[j * nn + b], Ma * D2[j]);
    }
    for (int j = 0; j < 10; ++j) {
        aadd(&third[j * nn + a], Mb * D3[j]);
        aadd(&third[j * nn + b], -(Ma * D3[j]));
    }
}

// ------------------------------------------------------------------ downward
__global__ __launch_bounds__(128) void k_down(int64_t a, int64_t b, int multipole, DView v,
                                              const double* __restrict__ sm,
                                              double* __restrict__ force,
                                              double* __restrict__ hess,
                                              double* __restrict__ third, int64_t nn,
                                              double* __restrict__ fx, double* __restrict__ fy,
                                              double* __restrict__ fz) {
#pragma clang fp contract(off)
    int64_t k = a + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= b) return;
    if (v.type[k] == 1) return;
    double H[6] = {0, 0, 0, 0, 0, 0}, T[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (multipole >= 4) {
        for (int j = 0; j < 6; ++j) H[j] = hess[j * nn + k];
        for (int j = 0; j < 10; ++j) T[j] = third[j * nn + k];
    }
    const double hxx = H[0], hyy = H[1], hzz = H[2], hxy = H[3], hxz = H[4], hyz = H[5];
    const double txxx = T[0], txxy = T[1], txxz = T[2], txyy = T[3], txyz = T[4], txzz = T[5],
                 tyyy = T[6], tyyz = T[7], tyzz = T[8], tzzz = T[9];
    const double fk[3] = {force[k], force[nn + k], force[2 * nn + k]};
    for (int c = v.cstart[k]; c < v.cstart[k + 1]; ++c) {
        if (v.mass[c] == 0.0) continue;
        double f[3];
        for (int j = 0; j < 3; ++j) f[j] = force[j * nn + c];
        for (int j = 0; j < 3; ++j) f[j] = f[j] + fk[j];
        if (multipole >= 4) {
            const double dx = v.cx[c] - v.cx[k];
            const double dy = v.cy[c] - v.cy[k];
            const double dz = v.cz[c] - v.cz[k];
            f[0] = f[0] + hxx * dx + hxy * dy + hxz * dz + txxy * dx * dy + txxz * dx * dz +
                   txyz * dy * dz + 0.5 * (txxx * dx * dx + txyy * dy * dy + txzz * dz * dz);
            f[1] = f[1] + hxy * dx + hyy * dy + hyz * dz + txyy * dx * dy + txyz * dx * dz +
                   tyyz * dy * dz + 0.5 * (txxy * dx * dx + tyyy * dy * dy + tyzz * dz * dz);
            f[2] = f[2] + hxz * dx + hyz * dy + hzz * dz + txyz * dx * dy + txzz * dx * dz +
                   tyzz * dy * dz + 0.5 * (txxz * dx * dx + tyyz * dy * dy + tzzz * dz * dz);
            if (v.type[c] != 1) {
                double* Hc = hess;
                Hc[0 * nn + c] = Hc[0 * nn + c] + hxx + txxx * dx + txxy * dy + txxz * dz;
                Hc[1 * nn + c] = Hc[1 * nn + c] + hyy + txyy * dx + tyyy * dy + tyyz * dz;
                Hc[2 * nn + c] = Hc[2 * nn + c] + hzz + txzz * dx + tyzz * dy + tzzz * dz;
                Hc[3 * nn + c] = Hc[3 * nn + c] + hxy + txxy * dx + txyy * dy + txyz * dz;
                Hc[4 * nn + c] = Hc[4 * nn + c] + hxz + txxz * dx + txyz * dy + txzz * dz;
                Hc[5 * nn + c] = Hc[5 * nn + c] + hyz + txyz * dx + tyyz * dy + tyzz * dz;
            }
            for (int j = 0; j < 10; ++j) third[j * nn + c] = third[j * nn + c] + T[j];
        }
        for (int j = 0; j < 3; ++j) force[j * nn + c] = f[j];
        if (v.type[c] == 1) {
            const int p = v.first[c];
            const double mp = sm[p];
            fx[p] = mp * f[0];
            fy[p] = mp * f[1];
            fz[p] = mp * f[2];
        }
    }
}

__global__ __launch_bounds__(256) void k_root_pairs(int64_t nsys, int2* __restrict__ pairs) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < nsys) pairs[k] = make_int2((int)k, (int)k);
}

// ------------------------------------------------------------------ driver
int tree_forces_dehnen(hq_tree* t, double theta2, double G, double eps2) {
    int rc;
    const int64_t nn = t->nn, n = t->n;
    hipStream_t s = t->stream;
    const size_t b8 = n * sizeof(double);
    if ((rc = t->fx.ensure(b8)) || (rc = t->fy.ensure(b8)) || (rc = t->fz.ensure(b8))) return rc;
    DevBuf& acc = t->counts;  // reused: rmax | force(3) | hess(6) | third(10) per node
    if ((rc = acc.ensure((size_t)nn * 20 * sizeof(double)))) return rc;
    double* rmax = acc.as<double>();
    double* force = rmax + nn;
    double* hess = force + 3 * nn;
    double* third = hess + 6 * nn;
    HQ_CHECK_HIP(hipMemsetAsync(acc.p, 0, (size_t)nn * 20 * sizeof(double), s));
    HQ_CHECK_HIP(hipMemsetAsync(t->fx.p, 0, b8, s));
    HQ_CHECK_HIP(hipMemsetAsync(t->fy.p, 0, b8, s));
    HQ_CHECK_HIP(hipMemsetAsync(t->fz.p, 0, b8, s));
    HQ_CHECK_HIP(hipEventRecord(t->ev[5], s));
    const DView v = dview(t);
    for (int d = t->most_depth; d >= 0; --d) {
        const int64_t a = t->level_start[d], b = t->level_start[d + 1];
        if (b > a) k_rmax<<<grid_for(b - a, 128), 128, 0, s>>>(a, b, v, rmax);
    }
    // breadth-first dual walk: frontier in t->clist / t->plist (ping-pong),
    // accepted pairs in t->lstat (grown on demand, contents kept)
    int64_t nf = t->nsys, nacc = 0;  // one (root, root) pair per system
    DevBuf sums;
    DevBuf* pc = &t->clist;
    DevBuf* pn = &t->plist;
    DevBuf& accp = t->lstat;
    if ((rc = pc->ensure(std::max<int64_t>(std::max<int64_t>(1024, nf), 4 * nn) * sizeof(int2))) ||
        (rc = pn->ensure(std::max<int64_t>(1024, 4 * nn) * sizeof(int2))) ||
        (rc = accp.ensure(std::max<int64_t>(1024, 8 * nn) * sizeof(int2))))
        return rc;
    k_root_pairs<<<grid_for(nf, 256), 256, 0, s>>>(nf, pc->as<int2>());
    uint32_t tot[2];
    while (nf > 0) {
        const int nblk = grid_for(nf, DB);
        if ((rc = sums.ensure(((size_t)2 * (nblk + 1)) * sizeof(uint32_t)))) return rc;
        uint32_t* bacc = sums.as<uint32_t>();
        uint32_t* bnext = bacc + nblk + 1;
        k_dual_count<<<nblk, DB, 0, s>>>(v, rmax, theta2, pc->as<int2>(), nf, bacc, bnext, nblk);
        if ((rc = scan_excl_buf(bacc, nblk, t->scanpart, s)) ||
            (rc = scan_excl_buf(bnext, nblk, t->scanpart, s)))
            return rc;
        HQ_CHECK_HIP(hipMemcpyAsync(&tot[0], bacc + nblk, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HQ_CHECK_HIP(hipMemcpyAsync(&tot[1], bnext + nblk, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        const int64_t nnew = tot[0], nnext = tot[1];
        if ((int64_t)((nacc + nnew) * sizeof(int2)) > (int64_t)accp.bytes) {
            DevBuf tmp;
            if ((rc = tmp.ensure(std::max<size_t>(2 * accp.bytes, (nacc + nnew) * sizeof(int2)))))
                return rc;
            HQ_CHECK_HIP(hipMemcpyAsync(tmp.p, accp.p, nacc * sizeof(int2), hipMemcpyDeviceToDevice, s));
            HQ_CHECK_HIP(hipStreamSynchronize(s));
            std::swap(tmp.p, accp.p);
            std::swap(tmp.bytes, accp.bytes);
        }
        if ((rc = pn->ensure(std::max<int64_t>(nnext, 1) * sizeof(int2)))) return rc;
        k_dual_emit<<<nblk, DB, 0, s>>>(v, rmax, theta2, pc->as<int2>(), nf, bacc, bnext,
                                        pn->as<int2>(), accp.as<int2>() + nacc);
        nacc += nnew;
        std::swap(pc, pn);
        nf = nnext;
    }
    if (nacc > 0) {
        const int2* pr = accp.as<int2>();
        if (t->multipole == 1)
            k_m2l<1><<<grid_for(nacc, 128), 128, 0, s>>>(v, pr, nacc, G, eps2, force, hess, third, nn);
        else if (t->multipole == 4)
            k_m2l<4><<<grid_for(nacc, 128), 128, 0, s>>>(v, pr, nacc, G, eps2, force, hess, third, nn);
        else
            k_m2l<8><<<grid_for(nacc, 128), 128, 0, s>>>(v, pr, nacc, G, eps2, force, hess, third, nn);
    }
    for (int d = 0; d < t->most_depth; ++d) {
        const int64_t a = t->level_start[d], b = t->level_start[d + 1];
        if (b > a)
            k_down<<<grid_for(b - a, 128), 128, 0, s>>>(a, b, t->multipole, v, t->sm.as<double>(),
                                                       force, hess, third, nn, t->fx.as<double>(),
                                                       t->fy.as<double>(), t->fz.as<double>());
    }
    HQ_CHECK_HIP(hipEventRecord(t->ev[6], s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    HQ_CHECK_HIP(hipGetLastError());
    hipEventElapsedTime(&t->ms_walk, t->ev[5], t->ev[6]);
    t->list_cells = nacc;
    return HQ_OK;
}

}  // namespace hq
