// hq_morton_tree.hip -- the reference's Morton octree on the GPU
// (fortran/morton/morton_octree.f90: generate_nodes :328-499, build_next
// :304-326, evaluate_quadrupoles :562-625, barnes_hut :627-766, forces_par
// :150-180).
//
// Keys, depth and the stable permutation come from hq_morton.hip (bit-exact).
// Nodes: every run of equal level-l prefix in the sorted key array is a
// level-l node, for l = 0..L (chains of one-particle nodes continue to level
// L, as in the reference).  With D_i the common levels of sorted particles i
// and i+1, particle i starts the nodes of levels D_{i-1}+1..L; two ballot
// passes number them in level order, which is the reference's numbering
// (level by level, parents in order, children in key order).  Mass, COM and
// quadrupoles are summed upward from the children (the reference sums each
// level over particles; the results agree to rounding); the root keeps the
// normalisation's sequential sums exactly as the reference does.
#include <math.h>

#include <algorithm>
#include <vector>

#include "hq_tree.h"

struct hq_morton {
    int64_t n = 0;
    int L = 0, Lreq = 0, use_quad = 1;
    double ssa_amp = 1.2;
    hipStream_t stream = nullptr;
    hq::DevBuf m, q, key, perm, sm, sq, lcp, hist, part, par;
    hq::DevBuf level, first, last, cstart, next, parent;
    hq::DevBuf prefix, mass, qcm, quad, fx;
    int64_t nn = 0;
    int level_start[23] = {};
    bool built = false;
};

namespace hq {

int morton_keys(int64_t n, const double* m, const double* q, int L, double ssa, int64_t* keys,
                int* perm, int* L_out, hipStream_t s, double* par_out);

constexpr int MT = 256;

__global__ __launch_bounds__(MT) void k_mgather(int64_t n, const int* __restrict__ perm,
                                                const int64_t* __restrict__ keys_orig,
                                                const double* __restrict__ m,
                                                const double* __restrict__ q,
                                                uint64_t* __restrict__ skey,
                                                double* __restrict__ sm, double* __restrict__ sq) {
    int64_t i = blockIdx.x * (int64_t)MT + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    skey[i] = (uint64_t)keys_orig[p];
    sm[i] = m[p];
    sq[3 * i] = q[3 * (int64_t)p];
    sq[3 * i + 1] = q[3 * (int64_t)p + 1];
    sq[3 * i + 2] = q[3 * (int64_t)p + 2];
}

__global__ __launch_bounds__(MT) void k_mlcp(int64_t n, const uint64_t* __restrict__ key, int L,
                                             int8_t* __restrict__ D) {
    int64_t i = blockIdx.x * (int64_t)MT + threadIdx.x;
    if (i >= n) return;
    int d = -1;
    if (i + 1 < n) {
        const uint64_t a = key[i] ^ key[i + 1];
        d = a ? (__clzll((long long)a) - (64 - 3 * L)) / 3 : L;
    }
    D[i] = (int8_t)d;
}

__device__ __forceinline__ bool mhas(const int8_t* __restrict__ D, int64_t i, int l) {
    const int dm = i > 0 ? (int)D[i - 1] : -1;
    return l > dm;  // particle i starts a level-l node
}

__global__ __launch_bounds__(MT) void k_mcount(const int8_t* __restrict__ D, int64_t n, int L,
                                               uint32_t* __restrict__ bcnt, int nblk) {
    __shared__ uint32_t wc[MT / 64][22];
    const int64_t i = blockIdx.x * (int64_t)MT + threadIdx.x;
    const bool ok = i < n;
    const int w = threadIdx.x >> 6;
    for (int l = 0; l <= L; ++l) {
        const uint64_t m = __ballot(ok && mhas(D, i, l));
        if (__lane_id() == 0) wc[w][l] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if ((int)threadIdx.x <= L) {
        uint32_t s = 0;
        for (int k = 0; k < MT / 64; ++k) s += wc[k][threadIdx.x];
        bcnt[(int64_t)threadIdx.x * nblk + blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(MT) void k_memit(const int8_t* __restrict__ D,
                                              const uint64_t* __restrict__ key, int64_t n, int L,
                                              const uint32_t* __restrict__ boff, int nblk,
                                              int* __restrict__ level, int* __restrict__ first,
                                              int* __restrict__ cstart,
                                              int64_t* __restrict__ prefix) {
    __shared__ uint32_t wpre[MT / 64][23];
    __shared__ uint32_t base[23];
    const int64_t i = blockIdx.x * (int64_t)MT + threadIdx.x;
    const bool ok = i < n;
    const int w = threadIdx.x >> 6;
    for (int l = 0; l <= L; ++l) {
        const uint64_t m = __ballot(ok && mhas(D, i, l));
        if (__lane_id() == 0) wpre[w][l] = (uint32_t)__popcll(m);
    }
    if (__lane_id() == 0) wpre[w][L + 1] = 0;
    __syncthreads();
    const uint32_t total = boff[(int64_t)(L + 1) * nblk];
    if ((int)threadIdx.x <= L + 1) {
        const int l = threadIdx.x;
        uint32_t run = 0;
        for (int k = 0; k < MT / 64; ++k) {
            const uint32_t c = wpre[k][l];
            wpre[k][l] = run;
            run += c;
        }
        base[l] = l <= L ? boff[(int64_t)l * nblk + blockIdx.x] : total;
    }
    __syncthreads();
    const uint64_t lt = lanemask_lt();
    uint64_t mcur = __ballot(ok && mhas(D, i, 0));
    for (int l = 0; l <= L; ++l) {
        const uint64_t mnext = l < L ? __ballot(ok && mhas(D, i, l + 1)) : 0ull;
        if (ok && mhas(D, i, l)) {
            const int64_t k = (int64_t)base[l] + wpre[w][l] + __popcll(mcur & lt);
            const int64_t cs = l < L ? (int64_t)base[l + 1] + wpre[w][l + 1] + __popcll(mnext & lt)
                                     : (int64_t)total;
            level[k] = l;
            first[k] = (int)i;
            cstart[k] = (int)cs;
            prefix[k] = (int64_t)(key[i] >> (3 * (L - l)));
        }
        mcur = mnext;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cstart[total] = (int)total;
}

__global__ __launch_bounds__(MT) void k_mlast(int64_t a, int64_t b, int64_t n,
                                              const int* __restrict__ first,
                                              int* __restrict__ last) {
    const int64_t k = a + blockIdx.x * (int64_t)MT + threadIdx.x;
    if (k >= b) return;
    last[k] = k + 1 < b ? first[k + 1] - 1 : (int)(n - 1);
}

// upward: level L nodes from their particles (P2M), others from children (M2M)
__global__ __launch_bounds__(128) void k_mup(int64_t a, int64_t b, int lvl, int L, int use_quad,
                                             const int* __restrict__ first,
                                             const int* __restrict__ last,
                                             const int* __restrict__ cstart,
                                             const double* __restrict__ sm,
                                             const double* __restrict__ sq,
                                             const double* __restrict__ par,
                                             double* __restrict__ mass, double* __restrict__ qcm,
                                             double* __restrict__ quad) {
#pragma clang fp contract(off)
    const int64_t k = a + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= b) return;
    double M = 0.0, c[3] = {0.0, 0.0, 0.0};
    const int cs = cstart[k], ce = cstart[k + 1];
    if (lvl == L) {
        for (int p = first[k]; p <= last[k]; ++p) {
            const double mp = sm[p];
            if (p == first[k]) {
                M = mp;
                for (int j = 0; j < 3; ++j) c[j] = mp * sq[3 * (int64_t)p + j];
            } else {
                M = M + mp;
                for (int j = 0; j < 3; ++j) c[j] = c[j] + mp * sq[3 * (int64_t)p + j];
            }
        }
        for (int j = 0; j < 3; ++j) c[j] = c[j] / M;
    } else {
        for (int ch = cs; ch < ce; ++ch) {
            const double mc = mass[ch];
            M = M + mc;
            for (int j = 0; j < 3; ++j) c[j] = c[j] + mc * qcm[3 * (int64_t)ch + j];
        }
        for (int j = 0; j < 3; ++j) c[j] = c[j] / M;
    }
    if (lvl == 0) {  // root: the sequential sums of the normalisation (generate_nodes :361-367, :424)
        M = par[3];
        for (int j = 0; j < 3; ++j) c[j] = par[j];
    }
    mass[k] = M;
    for (int j = 0; j < 3; ++j) qcm[3 * k + j] = c[j];
    if (!use_quad) return;
    double Q[6] = {0, 0, 0, 0, 0, 0};
    if (ce - cs <= 1 && lvl == L) {
        for (int p = first[k]; p <= last[k]; ++p) {
            const double pm = sm[p];
            const double dxi = sq[3 * (int64_t)p] - c[0];
            const double dyi = sq[3 * (int64_t)p + 1] - c[1];
            const double dzi = sq[3 * (int64_t)p + 2] - c[2];
            Q[0] = Q[0] + pm * (dxi * dxi);
            Q[1] = Q[1] + pm * (dyi * dyi);
            Q[2] = Q[2] + pm * (dzi * dzi);
            Q[3] = Q[3] + pm * dxi * dyi;
            Q[4] = Q[4] + pm * dxi * dzi;
            Q[5] = Q[5] + pm * dyi * dzi;
        }
    } else {
        for (int ch = cs; ch < ce; ++ch) {
            const double pm = mass[ch];
            const double dxi = qcm[3 * (int64_t)ch] - c[0];
            const double dyi = qcm[3 * (int64_t)ch + 1] - c[1];
            const double dzi = qcm[3 * (int64_t)ch + 2] - c[2];
            const double* Qc = quad + 6 * (int64_t)ch;
            Q[0] = Q[0] + Qc[0] + pm * dxi * dxi;
            Q[1] = Q[1] + Qc[1] + pm * dyi * dyi;
            Q[2] = Q[2] + Qc[2] + pm * dzi * dzi;
            Q[3] = Q[3] + Qc[3] + pm * dxi * dyi;
            Q[4] = Q[4] + Qc[4] + pm * dxi * dzi;
            Q[5] = Q[5] + Qc[5] + pm * dyi * dzi;
        }
    }
    for (int j = 0; j < 6; ++j) quad[6 * k + j] = Q[j];
}

__global__ __launch_bounds__(128) void k_mnext(int64_t a, int64_t b, const int* __restrict__ cstart,
                                               int* __restrict__ next, int* __restrict__ parent) {
    const int64_t k = a + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= b) return;
    if (k == 0) {
        next[0] = -1;
        parent[0] = -1;
    }
    const int cs = cstart[k], ce = cstart[k + 1];
    const int nx = next[k];
    for (int c = cs; c < ce; ++c) {
        parent[c] = (int)k;
        next[c] = c + 1 < ce ? c + 1 : nx;
    }
}

// barnes_hut (:627-766): stackless walk over next / first child
__global__ __launch_bounds__(128) void k_mwalk(int64_t n, int L, const int* __restrict__ level,
                                               const int* __restrict__ first,
                                               const int* __restrict__ last,
                                               const int* __restrict__ cstart,
                                               const int* __restrict__ next,
                                               const double* __restrict__ mass,
                                               const double* __restrict__ qcm,
                                               const double* __restrict__ quad,
                                               const double* __restrict__ sm,
                                               const double* __restrict__ sq,
                                               const int* __restrict__ perm,
                                               const double* __restrict__ par, double theta2,
                                               double G, double eps2, int mult,
                                               double* __restrict__ F, int layout) {
#pragma clang fp contract(off)
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double mp = sm[p];
    const double qp[3] = {sq[3 * p], sq[3 * p + 1], sq[3 * p + 2]};
    const double ssa = par[4];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    int idx = 0;
    while (idx != -1) {
        const int fi = first[idx], la = last[idx];
        const bool only_one = fi == la;
        const int nch = cstart[idx + 1] - cstart[idx];
        if (nch == 0) {
            for (int b = fi; b <= la; ++b) {
                if (b == (int)p) continue;
                const double dx = sq[3 * (int64_t)b] - qp[0];
                const double dy = sq[3 * (int64_t)b + 1] - qp[1];
                const double dz = sq[3 * (int64_t)b + 2] - qp[2];
                const double dist2 = dx * dx + dy * dy + dz * dz + eps2;
                double inv = 1.0 / sqrt(dist2);
                inv = inv * inv * inv;
                const double f = G * mp * sm[b] * inv;
                f0 = f0 + f * dx;
                f1 = f1 + f * dy;
                f2 = f2 + f * dz;
            }
            idx = next[idx];
            continue;
        }
        if (fi <= (int)p && (int)p <= la) {
            idx = cstart[idx];
            continue;
        }
        const double dx = qcm[3 * (int64_t)idx] - qp[0];
        const double dy = qcm[3 * (int64_t)idx + 1] - qp[1];
        const double dz = qcm[3 * (int64_t)idx + 2] - qp[2];
        double dist2 = dx * dx + dy * dy + dz * dz;
        const double sz = ssa * ldexp(1.0, -level[idx]);
        if (only_one || sz * sz <= theta2 * dist2) {
            const double M = mass[idx];
            if (M > 0) {
                dist2 = dist2 + eps2;
                if (!mult || only_one) {
                    double inv = 1.0 / sqrt(dist2);
                    inv = inv * inv * inv;
                    const double f = G * mp * M * inv;
                    f0 = f0 + f * dx;
                    f1 = f1 + f * dy;
                    f2 = f2 + f * dz;
                } else {
                    const double invR = 1.0 / sqrt(dist2);
                    const double invR2 = 1.0 / dist2;
                    const double invR3 = invR * invR2;
                    const double invR5 = invR3 * invR * invR;
                    const double mx = dx * invR3 * M;
                    const double my = dy * invR3 * M;
                    const double mz = dz * invR3 * M;
                    const double* Q = quad + 6 * (int64_t)idx;
                    const double dx2 = 15.0 * dx * dx * invR2 - 3.0;
                    const double dy2 = 15.0 * dy * dy * invR2 - 3.0;
                    const double dz2 = 15.0 * dz * dz * invR2 - 3.0;
                    const double dxyz = 15.0 * dx * dy * dz * invR2;
                    double Qx = (Q[0] * (dx2 - 6.0) + Q[1] * dy2 + Q[2] * dz2) * 0.5 * dx;
                    double Qy = (Q[1] * (dy2 - 6.0) + Q[2] * dz2 + Q[0] * dx2) * 0.5 * dy;
                    double Qz = (Q[2] * (dz2 - 6.0) + Q[0] * dx2 + Q[1] * dy2) * 0.5 * dz;
                    Qx = Qx + Q[3] * dy * dx2 + Q[4] * dz * dx2 + Q[5] * dxyz;
                    Qy = Qy + Q[5] * dz * dy2 + Q[3] * dx * dy2 + Q[4] * dxyz;
                    Qz = Qz + Q[4] * dx * dz2 + Q[5] * dy * dz2 + Q[3] * dxyz;
                    f0 = f0 + G * mp * (mx + invR5 * Qx);
                    f1 = f1 + G * mp * (my + invR5 * Qy);
                    f2 = f2 + G * mp * (mz + invR5 * Qz);
                }
            }
            idx = next[idx];
            continue;
        }
        idx = cstart[idx];
    }
    const int64_t o = perm[p];
    if (layout == HQ_LAYOUT_3N) {
        F[3 * o] = f0;
        F[3 * o + 1] = f1;
        F[3 * o + 2] = f2;
    } else {
        F[o] = f0;
        F[n + o] = f1;
        F[2 * n + o] = f2;
    }
}

int morton_build(hq_morton* t, const double* m, const double* q, int where) {
    const int64_t n = t->n;
    hipStream_t s = t->stream;
    int rc;
    if ((rc = t->m.ensure(n * 8)) || (rc = t->q.ensure(3 * n * 8)) || (rc = t->key.ensure(n * 8)) ||
        (rc = t->perm.ensure(n * 4)) || (rc = t->sm.ensure(n * 8)) || (rc = t->sq.ensure(3 * n * 8)) ||
        (rc = t->lcp.ensure(n)) || (rc = t->par.ensure(16 * 8)))
        return rc;
    const hipMemcpyKind kind = where == HQ_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    HQ_CHECK_HIP(hipMemcpyAsync(t->m.p, m, n * 8, kind, s));
    HQ_CHECK_HIP(hipMemcpyAsync(t->q.p, q, 3 * n * 8, kind, s));
    int L = 0;
    const int Lreq = t->Lreq;
    DevBuf korig;
    if ((rc = korig.ensure(n * 8))) return rc;
    if ((rc = morton_keys(n, t->m.as<double>(), t->q.as<double>(), Lreq, t->ssa_amp,
                          korig.as<int64_t>(), t->perm.as<int>(), &L, s, t->par.as<double>())))
        return rc;
    t->L = L;
    k_mgather<<<grid_for(n, MT), MT, 0, s>>>(n, t->perm.as<int>(), korig.as<int64_t>(),
                                             t->m.as<double>(), t->q.as<double>(),
                                             t->key.as<uint64_t>(), t->sm.as<double>(),
                                             t->sq.as<double>());
    k_mlcp<<<grid_for(n, MT), MT, 0, s>>>(n, t->key.as<uint64_t>(), L, t->lcp.as<int8_t>());
    const int nblk = grid_for(n, MT);
    if ((rc = t->hist.ensure(((size_t)(L + 1) * nblk + 1) * 4))) return rc;
    k_mcount<<<nblk, MT, 0, s>>>(t->lcp.as<int8_t>(), n, L, t->hist.as<uint32_t>(), nblk);
    if ((rc = scan_excl_buf(t->hist.as<uint32_t>(), (int64_t)(L + 1) * nblk, t->part, s))) return rc;
    std::vector<uint32_t> lv(L + 2);
    for (int l = 0; l <= L + 1; ++l)
        HQ_CHECK_HIP(hipMemcpyAsync(&lv[l], t->hist.as<uint32_t>() + (int64_t)l * nblk, 4,
                                    hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    const int64_t nn = lv[L + 1];
    t->nn = nn;
    for (int l = 0; l <= L + 1; ++l) t->level_start[l] = (int)lv[l];
    if ((rc = t->level.ensure(nn * 4)) || (rc = t->first.ensure(nn * 4)) ||
        (rc = t->last.ensure(nn * 4)) || (rc = t->cstart.ensure((nn + 1) * 4)) ||
        (rc = t->next.ensure(nn * 4)) || (rc = t->parent.ensure(nn * 4)) ||
        (rc = t->prefix.ensure(nn * 8)) || (rc = t->mass.ensure(nn * 8)) ||
        (rc = t->qcm.ensure(nn * 24)) || (rc = t->quad.ensure(nn * 48)))
        return rc;
    k_memit<<<nblk, MT, 0, s>>>(t->lcp.as<int8_t>(), t->key.as<uint64_t>(), n, L,
                                t->hist.as<uint32_t>(), nblk, t->level.as<int>(),
                                t->first.as<int>(), t->cstart.as<int>(), t->prefix.as<int64_t>());
    for (int l = 0; l <= L; ++l) {
        const int64_t a = lv[l], b = lv[l + 1];
        if (b > a) k_mlast<<<grid_for(b - a, MT), MT, 0, s>>>(a, b, n, t->first.as<int>(), t->last.as<int>());
    }
    for (int l = L; l >= 0; --l) {
        const int64_t a = lv[l], b = lv[l + 1];
        if (b > a)
            k_mup<<<grid_for(b - a, 128), 128, 0, s>>>(
                a, b, l, L, t->use_quad, t->first.as<int>(), t->last.as<int>(), t->cstart.as<int>(),
                t->sm.as<double>(), t->sq.as<double>(), t->par.as<double>(), t->mass.as<double>(),
                t->qcm.as<double>(), t->quad.as<double>());
    }
    if (L == 0) {
        const int endv = -1;
        HQ_CHECK_HIP(hipMemcpyAsync(t->next.p, &endv, 4, hipMemcpyHostToDevice, s));
        HQ_CHECK_HIP(hipMemcpyAsync(t->parent.p, &endv, 4, hipMemcpyHostToDevice, s));
    }
    for (int l = 0; l < L; ++l) {
        const int64_t a = lv[l], b = lv[l + 1];
        if (b > a)
            k_mnext<<<grid_for(b - a, 128), 128, 0, s>>>(a, b, t->cstart.as<int>(), t->next.as<int>(),
                                                         t->parent.as<int>());
    }
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    HQ_CHECK_HIP(hipGetLastError());
    t->built = true;
    return HQ_OK;
}

int morton_forces(hq_morton* t, double G, double eps2, double theta2, int mult, double* F,
                  int layout, int where) {
    const int64_t n = t->n;
    hipStream_t s = t->stream;
    int rc;
    double* dF = F;
    if (where == HQ_HOST) {
        if ((rc = t->fx.ensure(3 * n * 8))) return rc;
        dF = t->fx.as<double>();
    }
    k_mwalk<<<grid_for(n, 128), 128, 0, s>>>(
        n, t->L, t->level.as<int>(), t->first.as<int>(), t->last.as<int>(), t->cstart.as<int>(),
        t->next.as<int>(), t->mass.as<double>(), t->qcm.as<double>(), t->quad.as<double>(),
        t->sm.as<double>(), t->sq.as<double>(), t->perm.as<int>(), t->par.as<double>(), theta2, G,
        eps2, mult, dF, layout);
    if (where == HQ_HOST)
        HQ_CHECK_HIP(hipMemcpyAsync(F, dF, 3 * n * 8, hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    HQ_CHECK_HIP(hipGetLastError());
    return HQ_OK;
}

}  // namespace hq

using namespace hq;

extern "C" {

int hq_morton_create(hq_morton** out, int64_t n, int L, double ssa, int use_quad) {
    if (!out) return fail(HQ_ERR_ARGS, "null handle pointer");
    *out = nullptr;
    if (n < 1 || n > (int64_t)1 << 30) return fail(HQ_ERR_ARGS, "particle count out of range");
    if (L < 0 || L > 21) return fail(HQ_ERR_DEPTH, "Morton depth %d outside 0..21", L);
    hq_morton* t = new hq_morton();
    t->n = n;
    t->L = L;
    t->Lreq = L;
    t->ssa_amp = ssa;
    t->use_quad = use_quad != 0;
    hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete t;
        return fail(HQ_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    *out = t;
    return HQ_OK;
}

int hq_morton_build(hq_morton* t, const double* m, const double* q, int where) {
    if (!t || !m || !q) return fail(HQ_ERR_ARGS, "null argument");
    return morton_build(t, m, q, where);
}

int hq_morton_forces(hq_morton* t, double G, double eps2, double theta2, int mult, double* F,
                     int layout, int where) {
    if (!t || !F) return fail(HQ_ERR_ARGS, "null argument");
    if (!t->built) return fail(HQ_ERR_ARGS, "Morton tree not built");
    if (mult && !t->use_quad) return fail(HQ_ERR_ARGS, "quadrupoles were not evaluated");
    return morton_forces(t, G, eps2, theta2, mult, F, layout, where);
}

int hq_morton_info(const hq_morton* t, int* L, int64_t* nnodes) {
    if (!t) return fail(HQ_ERR_ARGS, "null handle");
    if (L) *L = t->L;
    if (nnodes) *nnodes = t->nn;
    return HQ_OK;
}

int hq_morton_export(const hq_morton* t, int* perm, int64_t* prefix, int* level, int* first,
                     int* last, int* parent, int* nchild, int* first_child, int* next,
                     double* mass, double* qcm, double* quad) {
    if (!t || !t->built) return fail(HQ_ERR_ARGS, "Morton tree not built");
    const int64_t nn = t->nn, n = t->n;
    hipStream_t s = t->stream;
    std::vector<int> cs(nn + 1);
    HQ_CHECK_HIP(hipMemcpyAsync(cs.data(), t->cstart.p, (nn + 1) * 4, hipMemcpyDeviceToHost, s));
    if (perm) HQ_CHECK_HIP(hipMemcpyAsync(perm, t->perm.p, n * 4, hipMemcpyDeviceToHost, s));
    if (prefix) HQ_CHECK_HIP(hipMemcpyAsync(prefix, t->prefix.p, nn * 8, hipMemcpyDeviceToHost, s));
    if (level) HQ_CHECK_HIP(hipMemcpyAsync(level, t->level.p, nn * 4, hipMemcpyDeviceToHost, s));
    if (first) HQ_CHECK_HIP(hipMemcpyAsync(first, t->first.p, nn * 4, hipMemcpyDeviceToHost, s));
    if (last) HQ_CHECK_HIP(hipMemcpyAsync(last, t->last.p, nn * 4, hipMemcpyDeviceToHost, s));
    if (parent) HQ_CHECK_HIP(hipMemcpyAsync(parent, t->parent.p, nn * 4, hipMemcpyDeviceToHost, s));
    if (next) HQ_CHECK_HIP(hipMemcpyAsync(next, t->next.p, nn * 4, hipMemcpyDeviceToHost, s));
    if (mass) HQ_CHECK_HIP(hipMemcpyAsync(mass, t->mass.p, nn * 8, hipMemcpyDeviceToHost, s));
    if (qcm) HQ_CHECK_HIP(hipMemcpyAsync(qcm, t->qcm.p, nn * 24, hipMemcpyDeviceToHost, s));
    if (quad) HQ_CHECK_HIP(hipMemcpyAsync(quad, t->quad.p, nn * 48, hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    for (int64_t k = 0; k < nn; ++k) {
        if (nchild) nchild[k] = cs[k + 1] - cs[k];
        if (first_child) first_child[k] = cs[k + 1] > cs[k] ? cs[k] : -1;
    }
    return HQ_OK;
}

int hq_morton_destroy(hq_morton* t) {
    if (!t) return HQ_OK;
    hipStream_t s = t->stream;
    if (s) (void)hipStreamSynchronize(s);
    delete t;
    if (s) (void)hipStreamDestroy(s);
    return HQ_OK;
}

}  // extern "C"
