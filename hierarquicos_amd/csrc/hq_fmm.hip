// hq_fmm.hip -- Cartesian fast multipole method of order p = 4 or 6 on the
// octree of hq_build.hip (beyond the reference: its Dehnen method,
// fortran/octree_default.f90:1095-1712, stops at quadrupole sources).
//
// Expansions are about the geometric cube centres, scaled by the cube side s:
//   raw moments   Q~_m = sum_j m_j ((x_j - c) / s)^m                 |m| <= p
//   local coeffs  L~_n = s^(|n|+1) L_n,  phi(c + u) = sum_n L_n u^n  |n| <= p
// so that M2M and L2L between a cube and its children use constant
// coefficients, and an M2L between two cubes of one level depends only on the
// integer offset o = (c_B - c_A) / s:
//   L~_A = T_o Q~_B,   T_o[n][m] = -(-1)^|m| C(n+m, n) a_{n+m}(o),
// a_k = D^k(1/r) / k! (Taylor coefficients, computed by their three-term
// recurrence).  A dual-tree walk (Dehnen 2002, mutual form of
// octree_default.f90:1132-1248, here directed) pairs cells with the MAC
// rmax_A + rmax_B <= theta |c_A - c_B|.  Same-level accepted pairs are binned
// by (target group, offset) and evaluated as dense GEMMs on the matrix cores
// (v_mfma_f32_16x16x4f32, FP32 in and out); pairs of different levels, or all
// pairs in the FP64 mode, use a scalar FP64 M2L.  Near pairs of leaf cells
// (<= S bodies) are summed directly; local expansions are passed down (L2L)
// and evaluated at the bodies (L2P) in FP64.  Pair lists are binned with
// atomic cursors and the scalar M2L adds with atomics, so sums are not
// bit-reproducible run to run (round-off level).
#include <math.h>

#include <algorithm>
#include <memory>
#include <type_traits>
#include <vector>

#include "hq_tree.h"

namespace hq {

constexpr int PMAX = 6;
constexpr int KMAX = 84;     // (p+1)(p+2)(p+3)/6 at p = 6
constexpr int K2MAX = 455;   // same at 2p = 12
__host__ __device__ constexpr int ncoef(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
__host__ __device__ constexpr int kpad(int p) { return (ncoef(p) + 15) / 16 * 16; }
// the M2L contraction length: the moments padded to the MFMA's k-step of 4
// (rows of the operator and the locals are padded to 16 instead)
__host__ __device__ constexpr int kq(int p) { return (ncoef(p) + 3) / 4 * 4; }
// Target cubes per group of the matrix-core M2L, whose FP32 locals live in
// LDS: at order 4 each wavefront keeps its own copy of them (k_fmm_m2l4), at
// order 6 each wavefront owns 16 rows of one shared copy (k_fmm_m2l6).  128
// at both; at order 6, 96 took 74 ms and 192 took 79 ms against 67 (4M cube).
constexpr int M2L6_GT = 128;
__host__ __device__ constexpr int group_of(int p) { return p <= 4 ? 128 : M2L6_GT; }

typedef float floatx4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ host tables
namespace {

struct MI { int x, y, z; };

std::vector<MI> graded(int p) {  // all multi-indices with |n| <= p, graded order
    std::vector<MI> v;
    for (int q = 0; q <= p; ++q)
        for (int a = q; a >= 0; --a)
            for (int b = q - a; b >= 0; --b) v.push_back({a, b, q - a - b});
    return v;
}

double binom(int n, int k) {
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

// Taylor coefficients a_k(R) = D^k(1/|R|) / k! for the multi-indices `ks`
// (graded), from |k| r^2 a_k = -(2|k|-1) sum_d R_d a_{k-e_d} - (|k|-1) sum_d a_{k-2e_d}
void taylor(const double R[3], const std::vector<MI>& ks, const std::vector<int>& rec,
            std::vector<double>& a) {
    const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
    a.assign(ks.size(), 0.0);
    a[0] = 1.0 / sqrt(r2);
    for (size_t i = 1; i < ks.size(); ++i) {
        const int n = ks[i].x + ks[i].y + ks[i].z;
        double s = 0.0;
        for (int d = 0; d < 3; ++d) {
            if (rec[6 * i + d] >= 0) s += (2 * n - 1) * R[d] * a[rec[6 * i + d]];
            if (rec[6 * i + 3 + d] >= 0) s += (n - 1) * a[rec[6 * i + 3 + d]];
        }
        a[i] = -s / (n * r2);
    }
}

}  // namespace

struct FmmTables {
    // offsets: class 0 (same level) ids [0, noff0); class 1 (target one level
    // coarser) [noff0, noff0 + noff1); class 2 (target one level finer) after
    int p = 0, K = 0, Kp = 0, K2 = 0, R = 0, nd = 0, noff0 = 0, nd1 = 0, noff1 = 0, noff = 0;
    double theta = -1.0;
    DevBuf mi, mi2, rec2, m2lidx, m2lcoef, m2m, l2l, T;
    // order 4 builds its operators on the device: the FP32 Taylor
    // coefficients a_k (|k| <= 2p) of each offset, and for each operator
    // element (n, m) of the padded Kp x Kq block its coefficient
    // -(-1)^|m| C(n+m, n) (0 in the padding) and the index of n+m
    DevBuf A, cf, ix, deg;
};

struct FmmState {
    FmmTables tab;
    DevBuf flag, scan, fid, fnode, Qd, Qf, Ld, Lf, rmax, lvl;
    DevBuf rec, code, fa, fb, cnt, lat, gen, p2p, m2p, blist, gnb, ghist, gpos, gpair, lpk;
    DevBuf phist, ppos, plist, edges;
    int64_t nf = 0;
    std::vector<int> fls;  // FMM node range per depth
};

void fmm_state_free(FmmState* s) { delete s; }

static int build_tables(FmmTables& tb, int p, double theta) {
    if (tb.p == p && tb.theta == theta) return HQ_OK;
    const std::vector<MI> k1 = graded(PMAX), k2 = graded(2 * PMAX);
    const int K = ncoef(p), Kp = kpad(p), Kq = kq(p), K2 = ncoef(2 * p);
    std::vector<int> pos(13 * 13 * 13, -1);
    for (size_t i = 0; i < k2.size(); ++i) pos[(k2[i].x * 13 + k2[i].y) * 13 + k2[i].z] = (int)i;
    auto at = [&](int x, int y, int z) { return (x < 0 || y < 0 || z < 0) ? -1 : pos[(x * 13 + y) * 13 + z]; };
    std::vector<int> rec(6 * k2.size());
    for (size_t i = 0; i < k2.size(); ++i) {
        const MI k = k2[i];
        rec[6 * i + 0] = at(k.x - 1, k.y, k.z);
        rec[6 * i + 1] = at(k.x, k.y - 1, k.z);
        rec[6 * i + 2] = at(k.x, k.y, k.z - 1);
        rec[6 * i + 3] = at(k.x - 2, k.y, k.z);
        rec[6 * i + 4] = at(k.x, k.y - 2, k.z);
        rec[6 * i + 5] = at(k.x, k.y, k.z - 2);
    }
    // M2L index and coefficient -(-1)^|m| C(n+m, n) on the K x K block
    std::vector<int> midx(K * K);
    std::vector<double> mcoef(K * K);
    for (int n = 0; n < K; ++n)
        for (int m = 0; m < K; ++m) {
            const MI a = k1[n], b = k1[m];
            midx[n * K + m] = at(a.x + b.x, a.y + b.y, a.z + b.z);
            const double sgn = ((b.x + b.y + b.z) & 1) ? 1.0 : -1.0;
            mcoef[n * K + m] = sgn * binom(a.x + b.x, a.x) * binom(a.y + b.y, a.y) * binom(a.z + b.z, a.z);
        }
    // M2M / L2L for the child with digit dg (bits 4x + 2y + z: 1 = upper half),
    // child centre = parent centre + delta s_parent / 4.  Both are stored by
    // input coefficient, then output coefficient, so that the lanes (one per
    // output coefficient) of a step over the inputs read consecutive words.
    std::vector<double> m2m(8 * K * K, 0.0), l2l(8 * K * K, 0.0);
    for (int dg = 0; dg < 8; ++dg) {
        const double del[3] = {(dg & 4) ? 1.0 : -1.0, (dg & 2) ? 1.0 : -1.0, (dg & 1) ? 1.0 : -1.0};
        for (int n = 0; n < K; ++n)
            for (int m = 0; m < K; ++m) {
                const MI a = k1[n], b = k1[m];
                if (b.x <= a.x && b.y <= a.y && b.z <= a.z) {
                    // Q~_p,n += C(n,m) (delta/4)^(n-m) 2^-|m| Q~_c,m
                    double c = binom(a.x, b.x) * binom(a.y, b.y) * binom(a.z, b.z) *
                               pow(del[0] / 4, a.x - b.x) * pow(del[1] / 4, a.y - b.y) *
                               pow(del[2] / 4, a.z - b.z);
                    m2m[(dg * K + m) * K + n] = c * ldexp(1.0, -(b.x + b.y + b.z));
                    // L~_c,m += C(n,m) (delta/4)^(n-m) 2^-(|m|+1) L~_p,n
                    l2l[(dg * K + n) * K + m] = c * ldexp(1.0, -(b.x + b.y + b.z) - 1);
                }
            }
    }
    // lattice operators.  Class 0: integer offsets |o_d| <= R in units of the
    // common side, T[n][m] = coef a_{n+m}(-o).  Classes 1 and 2: half-integer
    // offsets u (|u_d| <= R + 1/2) in units of the finer side, stored at index
    // (2u + 2R + 1) / 2 per axis; class 1 (target coarser) scales rows by
    // 2^(|n|+1), class 2 (target finer) columns by 2^|m|.
    const int R = std::max(2, (int)floor(2.0 * sqrt(3.0) / theta + sqrt(3.0)));
    const int nd = 2 * R + 1, noff0 = nd * nd * nd;
    const int nd1 = 2 * R + 2, noff1 = nd1 * nd1 * nd1;
    const int noff = noff0 + 2 * noff1;
    std::vector<MI> ks2(k2.begin(), k2.begin() + K2);
    const int K2A = (K2 + 3) / 4 * 4;  // a row of A, padded to whole float4s
    std::vector<float> T((size_t)noff * Kp * Kq, 0.0f), A((size_t)noff * K2A, 0.0f);
    std::vector<double> a;
    auto fill = [&](int o, const double u[3], int cls) {
        // u is the source's offset from the target; the expansion needs r = -u
        const double Ro[3] = {-u[0], -u[1], -u[2]};
        taylor(Ro, ks2, rec, a);
        for (int k = 0; k < K2; ++k) A[(size_t)o * K2A + k] = (float)a[k];
        float* t = T.data() + (size_t)o * Kp * Kq;
        for (int n = 0; n < K; ++n)
            for (int m = 0; m < K; ++m) {
                double c = mcoef[n * K + m] * a[midx[n * K + m]];
                if (cls == 1) c = ldexp(c, k1[n].x + k1[n].y + k1[n].z + 1);
                if (cls == 2) c = ldexp(c, k1[m].x + k1[m].y + k1[m].z);
                t[n * Kq + m] = (float)c;
            }
    };
    for (int o = 0; o < noff0; ++o) {
        const int ox = o / (nd * nd) - R, oy = (o / nd) % nd - R, oz = o % nd - R;
        if (!ox && !oy && !oz) continue;
        const double u[3] = {(double)ox, (double)oy, (double)oz};
        fill(o, u, 0);
    }
    for (int cls = 1; cls <= 2; ++cls)
        for (int o = 0; o < noff1; ++o) {
            const int ix = o / (nd1 * nd1), iy = (o / nd1) % nd1, iz = o % nd1;
            const double u[3] = {ix - R - 0.5, iy - R - 0.5, iz - R - 0.5};
            fill(noff0 + (cls - 1) * noff1 + o, u, cls);
        }
    std::vector<int> mi(4 * KMAX), mi2(4 * K2MAX);
    for (int i = 0; i < KMAX; ++i) {
        mi[4 * i] = k1[i].x, mi[4 * i + 1] = k1[i].y, mi[4 * i + 2] = k1[i].z;
        mi[4 * i + 3] = k1[i].x + k1[i].y + k1[i].z;
    }
    for (int i = 0; i < K2MAX; ++i) {
        mi2[4 * i] = k2[i].x, mi2[4 * i + 1] = k2[i].y, mi2[4 * i + 2] = k2[i].z;
        mi2[4 * i + 3] = k2[i].x + k2[i].y + k2[i].z;
    }
    std::vector<float> cf((size_t)Kp * Kq, 0.0f);
    std::vector<uint16_t> ix((size_t)Kp * Kq, 0);
    std::vector<int8_t> deg(128, 0);
    for (int n = 0; n < K; ++n) {
        deg[n] = (int8_t)(k1[n].x + k1[n].y + k1[n].z);
        for (int m = 0; m < K; ++m) {
            cf[(size_t)n * Kq + m] = (float)mcoef[n * K + m];
            ix[(size_t)n * Kq + m] = (uint16_t)midx[n * K + m];
        }
    }
    int rc;
    auto up = [&](DevBuf& d, const void* h, size_t b) -> int {
        int r;
        if ((r = d.ensure(b))) return r;
        HQ_CHECK_HIP(hipMemcpy(d.p, h, b, hipMemcpyHostToDevice));
        return HQ_OK;
    };
    if ((rc = up(tb.mi, mi.data(), mi.size() * 4)) || (rc = up(tb.mi2, mi2.data(), mi2.size() * 4)) ||
        (rc = up(tb.rec2, rec.data(), rec.size() * 4)) ||
        (rc = up(tb.m2lidx, midx.data(), midx.size() * 4)) ||
        (rc = up(tb.m2lcoef, mcoef.data(), mcoef.size() * 8)) ||
        (rc = up(tb.m2m, m2m.data(), m2m.size() * 8)) || (rc = up(tb.l2l, l2l.data(), l2l.size() * 8)) ||
        (rc = up(tb.T, T.data(), T.size() * 4)) || (rc = up(tb.A, A.data(), A.size() * 4)) ||
        (rc = up(tb.cf, cf.data(), cf.size() * 4)) || (rc = up(tb.ix, ix.data(), ix.size() * 2)) ||
        (rc = up(tb.deg, deg.data(), deg.size())))
        return rc;
    tb.p = p, tb.K = K, tb.Kp = Kp, tb.K2 = K2, tb.R = R, tb.nd = nd, tb.noff0 = noff0;
    tb.nd1 = nd1, tb.noff1 = noff1, tb.noff = noff, tb.theta = theta;
    return HQ_OK;
}

// ------------------------------------------------------------------ node sets
// flag: 0 outside the FMM tree, 1 leaf (<= S bodies, or a tree leaf), 2 internal
__global__ __launch_bounds__(256) void k_fmm_flag(int64_t nn, int S, const int* __restrict__ count,
                                                  const int8_t* __restrict__ type,
                                                  const uint8_t* __restrict__ depth,
                                                  const int* __restrict__ parent,
                                                  int8_t* __restrict__ flag,
                                                  uint32_t* __restrict__ one) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const int c = count[k], p = parent[k];
    int8_t f = 0;
    if (c > 0 && (p < 0 || count[p] > S)) f = (c <= S || type[k] == 1 || depth[k] >= MAX_DEPTH) ? 1 : 2;
    flag[k] = f;
    one[k] = f ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_fmm_compact(int64_t nn, const int8_t* __restrict__ flag,
                                                     const uint32_t* __restrict__ scan,
                                                     int* __restrict__ fid, int* __restrict__ fnode) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const int f = flag[k] ? (int)scan[k] : -1;
    fid[k] = f;
    if (f >= 0) fnode[f] = (int)k;
}

__global__ void k_fmm_levels(const uint32_t* __restrict__ scan, const int* __restrict__ ls,
                             int* __restrict__ out) {
    const int d = threadIdx.x;
    if (d <= MAX_DEPTH + 1) out[d] = (int)scan[ls[d]];
}

struct FView {
    const int *first, *count, *cstart, *fid, *fnode, *ssys;
    const int8_t* flag;
    const uint8_t *depth, *digit;
    const double *cx, *cy, *cz, *sg;  // cube centres; per-system geometry table
};

__device__ __forceinline__ double side_of(const FView& v, int k) {
    const int sys = v.ssys[v.first[k]];
    return 2.0 * v.sg[SG * sys + S_HALFD + v.depth[k]];
}

// ------------------------------------------------------------------ upward
// P2M of a leaf: one wavefront, lanes over bodies for the powers, then over
// coefficients for the sums.
template <int P>
__global__ __launch_bounds__(64) void k_fmm_p2m(int nf, FView v, const double* __restrict__ sx,
                                                const double* __restrict__ sy,
                                                const double* __restrict__ sz,
                                                const double* __restrict__ sm,
                                                const int* __restrict__ mi, double* __restrict__ Qd,
                                                double* __restrict__ rmax) {
    constexpr int K = ncoef(P);
    __shared__ double pw[3][P + 1][64];
    __shared__ double mass[64];
    const int f = blockIdx.x;
    if (f >= nf) return;
    const int k = v.fnode[f];
    if (v.flag[k] != 1) return;
    const int lane = threadIdx.x;
    const int first = v.first[k], cnt = v.count[k];
    const double s = side_of(v, k), c[3] = {v.cx[k], v.cy[k], v.cz[k]};
    double r = 0.0;
    if (lane < cnt) {
        const int j = first + lane;
        const double y[3] = {sx[j] - c[0], sy[j] - c[1], sz[j] - c[2]};
        r = sqrt(y[0] * y[0] + y[1] * y[1] + y[2] * y[2]);
        for (int d = 0; d < 3; ++d) {
            const double t = s > 0.0 ? y[d] / s : 0.0;
            double w = 1.0;
            for (int e = 0; e <= P; ++e) { pw[d][e][lane] = w; w *= t; }
        }
        mass[lane] = sm[j];
    }
    for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o, 64));
    __syncthreads();
    for (int n = lane; n < K; n += 64) {
        const int a = mi[4 * n], b = mi[4 * n + 1], cz = mi[4 * n + 2];
        double acc = 0.0;
        for (int j = 0; j < cnt; ++j) acc += mass[j] * pw[0][a][j] * pw[1][b][j] * pw[2][cz][j];
        Qd[(int64_t)f * K + n] = acc;
    }
    if (lane == 0) rmax[f] = r;
}

// M2M of the internal FMM nodes of one level, one wavefront per node
template <int P>
__global__ __launch_bounds__(64) void k_fmm_m2m(int f0, int f1, FView v, const double* __restrict__ m2m,
                                                double* __restrict__ Qd, double* __restrict__ rmax) {
    constexpr int K = ncoef(P);
    const int f = f0 + blockIdx.x;
    if (f >= f1) return;
    const int k = v.fnode[f];
    if (v.flag[k] != 2) return;
    const int lane = threadIdx.x;
    const double s = side_of(v, k);
    double acc[2] = {0.0, 0.0}, r = 0.0;
    for (int c = v.cstart[k]; c < v.cstart[k + 1]; ++c) {
        if (v.count[c] == 0) continue;
        const int fc = v.fid[c];
        const double* q = Qd + (int64_t)fc * K;
        const double* t = m2m + (int64_t)v.digit[c] * K * K;
        // input m to output n (zero unless m <= n), in the order of m
        for (int m = 0; m < K; ++m) {
            const double qm = q[m];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int n = lane + 64 * h;
                if (n < K) acc[h] += t[m * K + n] * qm;
            }
        }
        r = fmax(r, rmax[fc] + 0.4330127018922193 * s);  // |c_child - c| = sqrt(3) s / 4
    }
    for (int h = 0; h < 2; ++h)
        if (lane + 64 * h < K) Qd[(int64_t)f * K + lane + 64 * h] = acc[h];
    if (lane == 0) rmax[f] = fmin(r, 0.8660254037844386 * s);
}

template <int P>
__global__ __launch_bounds__(256) void k_fmm_qf(int64_t nf, const double* __restrict__ Qd,
                                                float* __restrict__ Qf) {
    constexpr int K = ncoef(P), Kq = kq(P);
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= (nf + 1) * Kq) return;
    const int64_t f = i / Kq;
    const int n = (int)(i % Kq);
    // record nf is all zero: the M2L reads it for the empty columns of a tile
    Qf[i] = n < K && f < nf ? (float)Qd[f * K + n] : 0.0f;
}

// ------------------------------------------------------------------ dual walk
enum { ACT_EXPAND = 0, ACT_LATTICE = 1, ACT_GENERIC = 2, ACT_P2P = 3, ACT_M2P = 4, ACT_NONE = 5 };
constexpr int NACT = 5;

// The dual walk's per-category totals (the category starts of the scanned
// [category][block] counts, then the grand total), gathered into one small
// array so that they return to the host in one copy.
__global__ void k_fmm_edges(const uint32_t* __restrict__ cnt, int nblk, uint32_t* __restrict__ out) {
    const int q = threadIdx.x;
    if (q <= NACT) out[q] = cnt[(int64_t)q * (nblk + 1)];
}


struct WalkPar {
    double theta2;
    int R, nd, nd1, noff0, noff1;
};

// One FMM cell as the walk reads it (64 bytes, four 16-byte loads): centre,
// bounding radius, side and its reciprocal, children's FMM ids [kid0, kid1)
// (the non-empty children of an internal cell are FMM cells with
// consecutive ids), body count, flag | depth << 2.
struct __align__(16) FRec {
    double cx, cy, cz, rmax, side, rside;
    int kid0, kid1, count;
    uint32_t info;
};

__global__ __launch_bounds__(256) void k_fmm_records(int nf, FView v, const uint32_t* __restrict__ scan,
                                                     const double* __restrict__ rmax,
                                                     FRec* __restrict__ rec) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int k = v.fnode[f];
    const int flag = v.flag[k];
    FRec r;
    r.cx = v.cx[k];
    r.cy = v.cy[k];
    r.cz = v.cz[k];
    r.rmax = rmax[f];
    r.side = side_of(v, k);
    r.rside = 1.0 / r.side;
    r.kid0 = flag == 2 ? (int)scan[v.cstart[k]] : 0;
    r.kid1 = flag == 2 ? (int)scan[v.cstart[k + 1]] : 0;
    r.count = v.count[k];
    r.info = (uint32_t)flag | ((uint32_t)v.depth[k] << 2);
    rec[f] = r;
}

// Classifies the directed pair (a <- b) of FMM cells: accept with the MAC,
// sum directly when both are leaves, else split (both when of one level,
// else the larger or the non-leaf).  Returns the action code: the action in
// bits 0-2; for a split, the split (1 = a, 2 = b, 3 = both) in bits 3-4 and
// the number of child pairs in bits 5-11; for a lattice accept, the lattice
// offset id from bit 3.
__device__ __forceinline__ uint32_t fmm_classify(const FRec& ra, const FRec& rb, const WalkPar& w,
                                                 int a, int b) {
    if (a < 0) return ACT_NONE;
    const bool la = (ra.info & 3) == 1, lb = (rb.info & 3) == 1;
    const int da = (int)(ra.info >> 2), db = (int)(rb.info >> 2);
    if (a != b) {
        const double dx = rb.cx - ra.cx, dy = rb.cy - ra.cy, dz = rb.cz - ra.cz;
        const double rs = ra.rmax + rb.rmax;
        // rs <= theta d, squared
        if (rs * rs <= w.theta2 * (dx * dx + dy * dy + dz * dz)) {
            // a leaf target much larger than its source: the local expansion
            // would carry the larger error, so the source acts on the bodies
            // directly (its multipole, or its bodies when it is a leaf)
            if (la && (ra.rmax > 2.0 * rb.rmax || (ra.rmax > rb.rmax && ra.count <= 8)))
                return lb ? ACT_P2P : ACT_M2P;
            if (da == db) {
                // the cubes share a lattice: the offset is an integer of sides
                const double rsd = ra.rside;
                const int ox = (int)rint(dx * rsd), oy = (int)rint(dy * rsd), oz = (int)rint(dz * rsd);
                if (abs(ox) <= w.R && abs(oy) <= w.R && abs(oz) <= w.R)
                    return ACT_LATTICE |
                           ((uint32_t)(((ox + w.R) * w.nd + (oy + w.R)) * w.nd + (oz + w.R)) << 3);
            } else if (da == db - 1 || da == db + 1) {
                // half-integer offset in units of the finer cube
                const double rsd = 2.0 * (da > db ? ra.rside : rb.rside);
                const int wx = (int)rint(dx * rsd), wy = (int)rint(dy * rsd), wz = (int)rint(dz * rsd);
                const int lim = 2 * w.R + 1;
                if ((wx & 1) && (wy & 1) && (wz & 1) && abs(wx) <= lim && abs(wy) <= lim &&
                    abs(wz) <= lim) {
                    const int ix = (wx + lim) / 2, iy = (wy + lim) / 2, iz = (wz + lim) / 2;
                    const int off = w.noff0 + (da < db ? 0 : w.noff1) + (ix * w.nd1 + iy) * w.nd1 + iz;
                    return ACT_LATTICE | ((uint32_t)off << 3);
                }
            }
            return ACT_GENERIC;
        }
    }
    if (la && lb) return ACT_P2P;
    int sp;
    if (a == b) sp = 3;
    else if (la) sp = 2;
    else if (lb) sp = 1;
    else if (da == db) sp = 3;
    else sp = da < db ? 1 : 2;
    const int na = (sp & 1) ? ra.kid1 - ra.kid0 : 1, nb = (sp & 2) ? rb.kid1 - rb.kid0 : 1;
    return ACT_EXPAND | ((uint32_t)sp << 3) | ((uint32_t)(na * nb) << 5);
}

__device__ __forceinline__ int code_act(uint32_t c) { return (int)(c & 7u); }
__device__ __forceinline__ uint32_t code_count(uint32_t c) {
    const int act = code_act(c);
    return act == ACT_EXPAND ? (c >> 5) & 127u : (act == ACT_NONE ? 0u : 1u);
}

constexpr int WB = 256;

// Classifies every frontier pair once, keeps its code, and stores per chunk
// of WB pairs (the emit kernel's workgroups) the counts of next-frontier
// pairs, lattice, generic, near and M2P pairs.  A workgroup takes WALK_U
// consecutive chunks, a thread one pair of each, with the loads of all its
// pairs in flight together (the pass is bound by the latency of the pair and
// cell-record loads).
constexpr int WALK_U = 2;

__global__ __launch_bounds__(WB) void k_fmm_walk_count(const FRec* __restrict__ rec, WalkPar w,
                                                       const int2* __restrict__ fr, int64_t nfr,
                                                       uint32_t* __restrict__ code,
                                                       uint32_t* __restrict__ cnt, int nblk) {
    __shared__ uint32_t ws[WALK_U][NACT][WB / 64];
    const int64_t i0 = blockIdx.x * (int64_t)(WALK_U * WB) + threadIdx.x;
    int2 pr[WALK_U];
#pragma unroll
    for (int u = 0; u < WALK_U; ++u) {
        const int64_t i = i0 + u * WB;
        pr[u] = i < nfr ? fr[i] : make_int2(-1, 0);
    }
    FRec ra[WALK_U], rb[WALK_U];
#pragma unroll
    for (int u = 0; u < WALK_U; ++u)
        if (pr[u].x >= 0) {
            ra[u] = rec[pr[u].x];
            rb[u] = rec[pr[u].y];
        }
#pragma unroll
    for (int u = 0; u < WALK_U; ++u) {
        const int64_t i = i0 + u * WB;
        uint32_t c = ACT_NONE;
        if (pr[u].x >= 0) c = fmm_classify(ra[u], rb[u], w, pr[u].x, pr[u].y);
        if (i < nfr) code[i] = c;
        const int act = code_act(c);
        const uint32_t cv = code_count(c);
        // a split counts its child pairs, any other pair one entry of its list
        const uint32_t te = wave_sum(act == ACT_EXPAND ? cv : 0u);
        if (__lane_id() == 0) ws[u][ACT_EXPAND][threadIdx.x >> 6] = te;
#pragma unroll
        for (int q = 1; q < NACT; ++q) {
            const uint32_t t = (uint32_t)__popcll(__ballot(act == q));
            if (__lane_id() == 0) ws[u][q][threadIdx.x >> 6] = t;
        }
    }
    __syncthreads();
    if (threadIdx.x < WALK_U * NACT) {
        const int u = threadIdx.x / NACT, q = threadIdx.x % NACT;
        const int64_t blk = (int64_t)blockIdx.x * WALK_U + u;
        uint32_t t = 0;
        for (int k = 0; k < WB / 64; ++k) t += ws[u][q][k];
        if (blk < nblk) cnt[(int64_t)q * (nblk + 1) + blk] = t;
        if (blk == 0) cnt[(int64_t)q * (nblk + 1) + nblk] = 0;
    }
}

// Writes each pair's output at its place (chunk prefix from the counts,
// scanned as one array [category][chunk], then the wavefront's prefix): a
// split's child pairs to the next frontier, an accept or near pair to its
// list.  A workgroup takes WALK_U consecutive chunks of WB pairs, as the
// classify pass does, with the loads of all of them issued first.
__global__ __launch_bounds__(WB) void k_fmm_walk_emit(const FRec* __restrict__ rec,
                                                      const int2* __restrict__ fr, int64_t nfr,
                                                      const uint32_t* __restrict__ code,
                                                      const uint32_t* __restrict__ cnt,
                                                      int nblk, int2* __restrict__ next,
                                                      int3* __restrict__ lat, int64_t lat0,
                                                      int2* __restrict__ gen, int64_t gen0,
                                                      int2* __restrict__ p2p, int64_t p2p0,
                                                      int2* __restrict__ m2p, int64_t m2p0) {
    __shared__ uint32_t ws[WALK_U][NACT][WB / 64];
    __shared__ uint32_t xstart[WB / 64][64];
    __shared__ int4 xpar[WB / 64][64];
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    const int64_t i0 = blockIdx.x * (int64_t)(WALK_U * WB) + threadIdx.x;
    uint32_t c[WALK_U];
    int a[WALK_U], b[WALK_U];
#pragma unroll
    for (int u = 0; u < WALK_U; ++u) {
        const int64_t i = i0 + u * WB;
        c[u] = ACT_NONE;
        a[u] = b[u] = 0;
        if (i < nfr) {
            c[u] = code[i];
            const int2 pr = fr[i];
            a[u] = pr.x;
            b[u] = pr.y;
        }
    }
    // a split's first child ids and the source's child count
    int a0[WALK_U], b0[WALK_U], nbk[WALK_U];
#pragma unroll
    for (int u = 0; u < WALK_U; ++u) {
        a0[u] = a[u];
        b0[u] = b[u];
        nbk[u] = 1;
        if (code_act(c[u]) == ACT_EXPAND) {
            const int sp = (int)((c[u] >> 3) & 3u);
            if (sp & 1) a0[u] = rec[a[u]].kid0;
            if (sp & 2) {
                b0[u] = rec[b[u]].kid0;
                nbk[u] = rec[b[u]].kid1 - b0[u];
            }
        }
    }
    uint32_t pre[WALK_U][NACT];
#pragma unroll
    for (int u = 0; u < WALK_U; ++u) {
        const int act = code_act(c[u]);
        const uint32_t v = act == ACT_EXPAND ? code_count(c[u]) : 0u;
        const uint32_t incl = wave_incl_scan(v);
        pre[u][ACT_EXPAND] = incl - v;
        if (lane == 63) ws[u][ACT_EXPAND][wv] = incl;
#pragma unroll
        for (int q = 1; q < NACT; ++q) {
            const uint64_t m = __ballot(act == q);
            pre[u][q] = (uint32_t)__popcll(m & lt);
            if (lane == 0) ws[u][q][wv] = (uint32_t)__popcll(m);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < WALK_U; ++u) {
        const int64_t blk = (int64_t)blockIdx.x * WALK_U + u;
        if (blk >= nblk) break;  // the same in the whole workgroup
        for (int q = 0; q < NACT; ++q) {
            uint32_t base = cnt[(int64_t)q * (nblk + 1) + blk] - cnt[(int64_t)q * (nblk + 1)];
            for (int k = 0; k < wv; ++k) base += ws[u][q][k];
            pre[u][q] += base;
        }
        const int act = code_act(c[u]);
        const uint32_t cv = code_count(c[u]);
        // a split's child pairs: the wavefront writes them together,
        // consecutive lanes at consecutive places (a pair may have up to 64
        // children): lane j's pairs are [ex[j], ex[j] + cv) of the
        // wavefront's run, which starts at pre[0] of its first lane
        const uint32_t ne = act == ACT_EXPAND ? cv : 0u;
        const uint32_t incl = wave_incl_scan(ne);
        const uint32_t tot = __shfl(incl, 63, 64);
        if (tot > 0) {
            uint32_t* xs = &xstart[wv][0];
            int4* xp = &xpar[wv][0];
            __builtin_amdgcn_wave_barrier();  // the previous chunk's reads are done
            xs[lane] = incl - ne;
            xp[lane] = make_int4(a0[u], b0[u], nbk[u], (int)ne);
            const int64_t o0 = (int64_t)__shfl((int)(pre[u][0] - (incl - ne)), 0, 64);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t k = lane; k < tot; k += 64) {
                // the last lane whose run starts at or before k (lanes with no
                // children have empty runs and are skipped)
                int lo = 0, hi = 63;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (xs[mid] <= k) lo = mid;
                    else hi = mid - 1;
                }
                while (xp[lo].w == 0 || xs[lo] + (uint32_t)xp[lo].w <= k) --lo;
                const int4 q = xp[lo];
                const int r = (int)(k - xs[lo]);
                next[o0 + k] = make_int2(q.x + r / q.z, q.y + r % q.z);
            }
        }
        if (act == ACT_LATTICE) {
            lat[lat0 + pre[u][1]] = make_int3(a[u], b[u], (int)(c[u] >> 3));
        } else if (act == ACT_GENERIC) {
            gen[gen0 + pre[u][2]] = make_int2(a[u], b[u]);
        } else if (act == ACT_P2P) {
            p2p[p2p0 + pre[u][3]] = make_int2(a[u], b[u]);
        } else if (act == ACT_M2P) {
            m2p[m2p0 + pre[u][4]] = make_int2(a[u], b[u]);
        }
    }
}

// ------------------------------------------------------------------ binning
// Lattice pairs -> bins (target group, offset), without contended global
// atomics: first by target group (counted and placed a tile at a time in
// LDS), then one workgroup per group sorts the group's pairs by offset with
// an LDS histogram and writes the group's row of bin starts.  Near pairs ->
// bins (target) with one atomic per run of equal targets in a wavefront.
// A run of equal keys within a wavefront: one atomic, made by its first lane.
struct RunInfo {
    bool valid, head;
    int count, hlane;  // run length (at the head), the run's head lane
};

__device__ __forceinline__ RunInfo run_of(int key, int64_t base, int64_t n) {
    const int l = __lane_id();
    const bool valid = base + l < n;
    const int prev = __shfl_up(key, 1, 64);
    const bool head = valid && (l == 0 || key != prev);
    const uint64_t heads = __ballot(head);
    const int nvalid = (int)min<int64_t>(64, n - base);
    const uint64_t above = l == 63 ? 0ull : heads & (~0ull << (l + 1));
    const int next = above ? __ffsll((unsigned long long)above) - 1 : nvalid;
    const uint64_t le = l == 63 ? heads : heads & ((2ull << l) - 1);
    const int h = le ? 63 - __clzll((long long)le) : 0;
    return RunInfo{valid, head, next - l, h};
}

// Lattice pairs by target group, a tile of BT_T * BT_U consecutive pairs at a
// time.  The walk emits a split pair's children together and its pairs in
// the order of their ancestors, so a tile's targets fall in a few groups: a
// tile's runs of equal groups are counted in LDS (a window of win <= BT_W
// groups from the tile's smallest, HQ_TUNE_FMM_BIN_WINDOW) and the window's
// non-zero counts reach global memory with one atomic each.  A group outside
// the window takes a global atomic per run.
constexpr int BT_T = 256, BT_U = 8, BT_TILE = BT_T * BT_U, BT_W = 4096;

// the smallest and largest valid key of a workgroup's tile, in every thread
__device__ __forceinline__ int2 tile_key_range(const int (&key)[BT_U], int* __restrict__ red) {
    int lo = INT_MAX, hi = -1;
#pragma unroll
    for (int u = 0; u < BT_U; ++u)
        if (key[u] >= 0) lo = min(lo, key[u]), hi = max(hi, key[u]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
    }
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0) red[2 * w] = lo, red[2 * w + 1] = hi;
    __syncthreads();
#pragma unroll
    for (int v = 0; v < BT_T / 64; ++v) lo = min(lo, red[2 * v]), hi = max(hi, red[2 * v + 1]);
    __syncthreads();
    return make_int2(lo, hi);
}

__global__ __launch_bounds__(BT_T) void k_fmm_lat_ghist(const int3* __restrict__ lat, int64_t n, int gt,
                                                        int win, uint32_t* __restrict__ ghist) {
    __shared__ uint32_t h[BT_W];
    __shared__ int red[2 * BT_T / 64];
    for (int j = threadIdx.x; j < BT_W; j += BT_T) h[j] = 0;
    const int64_t ntile = (n + BT_TILE - 1) / BT_TILE;
    const int wbase = threadIdx.x & ~63;
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int64_t t0 = tile * BT_TILE;
        int key[BT_U];
#pragma unroll
        for (int u = 0; u < BT_U; ++u) {
            const int64_t i = t0 + u * BT_T + threadIdx.x;
            key[u] = i < n ? lat[i].x / gt : -1;
        }
        const int2 kr = tile_key_range(key, red);  // also orders the clears before the counts
#pragma unroll
        for (int u = 0; u < BT_U; ++u) {
            const int64_t base = t0 + u * BT_T + wbase;  // the same in the whole wavefront
            if (base >= n) break;
            const RunInfo r = run_of(key[u], base, n);
            if (r.head) {
                const int w = key[u] - kr.x;
                if (w < win) atomicAdd(&h[w], (uint32_t)r.count);
                else atomicAdd(&ghist[key[u]], (uint32_t)r.count);
            }
        }
        __syncthreads();
        const int span = min(win, kr.y - kr.x + 1);
        for (int j = threadIdx.x; j < span; j += BT_T) {
            const uint32_t c = h[j];
            if (c) {
                atomicAdd(&ghist[kr.x + j], c);
                h[j] = 0;
            }
        }
    }
}

// pair -> (source, column << 24 | offset) at its group's next free place:
// a tile's runs take their places within (tile, group) from LDS counters,
// then each of the tile's groups reserves its places with one global atomic
__global__ __launch_bounds__(BT_T) void k_fmm_lat_gscatter(const int3* __restrict__ lat, int64_t n,
                                                           int gt, int win, uint32_t* __restrict__ gpos,
                                                           int2* __restrict__ out) {
    __shared__ uint32_t h[BT_W];
    __shared__ int red[2 * BT_T / 64];
    for (int j = threadIdx.x; j < BT_W; j += BT_T) h[j] = 0;
    const int64_t ntile = (n + BT_TILE - 1) / BT_TILE;
    const int wbase = threadIdx.x & ~63;
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int64_t t0 = tile * BT_TILE;
        int3 e[BT_U];
        int key[BT_U];
#pragma unroll
        for (int u = 0; u < BT_U; ++u) {
            const int64_t i = t0 + u * BT_T + threadIdx.x;
            e[u] = i < n ? lat[i] : make_int3(0, 0, 0);
            key[u] = i < n ? e[u].x / gt : -1;
        }
        const int2 kr = tile_key_range(key, red);
        // place within (tile, group) for keys in the window, else the final place
        uint32_t at[BT_U];
#pragma unroll
        for (int u = 0; u < BT_U; ++u) at[u] = 0;
#pragma unroll
        for (int u = 0; u < BT_U; ++u) {
            const int64_t base = t0 + u * BT_T + wbase;
            if (base >= n) break;
            const RunInfo r = run_of(key[u], base, n);
            uint32_t o = 0;
            if (r.head) {
                const int w = key[u] - kr.x;
                o = w < win ? atomicAdd(&h[w], (uint32_t)r.count)
                             : atomicAdd(&gpos[key[u]], (uint32_t)r.count);
            }
            at[u] = __shfl(o, r.hlane, 64) + (uint32_t)(__lane_id() - r.hlane);
        }
        __syncthreads();
        const int span = min(win, kr.y - kr.x + 1);
        for (int j = threadIdx.x; j < span; j += BT_T) {
            const uint32_t c = h[j];
            if (c) h[j] = atomicAdd(&gpos[kr.x + j], c);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < BT_U; ++u) {
            if (key[u] < 0) continue;
            const int w = key[u] - kr.x;
            const uint32_t o = w < win ? h[w] + at[u] : at[u];
            out[o] = make_int2(e[u].y, (int)(((uint32_t)(e[u].x - key[u] * gt) << 24) | (uint32_t)e[u].z));
        }
        __syncthreads();
        // clear the window for the next tile (an entry that stayed zero is zero)
        for (int j = threadIdx.x; j < span; j += BT_T) h[j] = 0;
    }
}

// One workgroup per target group: its pairs [gstart[g], gstart[g+1]) counted
// by offset in LDS (LSORT_CH offsets per pass), the counts scanned, the
// group's non-empty bins listed in offset order as (offset, first pair) at
// blist[g * noff + r], r < gnb[g] (a bin ends where the next one starts, the
// last at gstart[g+1]), and the pairs placed by offset as one word each:
// source | column << 24 (sources below 2^24, columns below 128).
// The offset counts live in dynamic LDS, min(noff, LSORT_CH) entries: at
// theta = 0.5 the 16,577 offsets take 66 KB, so two workgroups share a CU.
constexpr int LSORT_T = 1024;
constexpr int LSORT_CH = 32768;
constexpr int LSORT_U = 8;
__host__ __device__ inline int lsort_chunk(int noff) { return noff < LSORT_CH ? noff : LSORT_CH; }

__global__ __launch_bounds__(LSORT_T) void k_fmm_lat_lsort(const int2* __restrict__ gp,
                                                           const uint32_t* __restrict__ gstart,
                                                           int noff, int2* __restrict__ blist,
                                                           uint32_t* __restrict__ gnb,
                                                           int* __restrict__ pk) {
    extern __shared__ uint32_t h[];  // lsort_chunk(noff) counts
    __shared__ uint32_t ws[LSORT_T / 64];
    const int ch = lsort_chunk(noff);
    const int g = blockIdx.x, tid = threadIdx.x;
    const uint32_t s0 = gstart[g], s1 = gstart[g + 1];
    int2* row = blist + (int64_t)g * noff;
    uint32_t done = 0, listed = 0;  // pairs and non-empty bins of the earlier passes
    for (int c0 = 0; c0 < noff; c0 += ch) {
        const int nc = min(ch, noff - c0);
        for (int o = tid; o < nc; o += LSORT_T) h[o] = 0;
        __syncthreads();
        // LSORT_U pairs per thread in flight: the loads of a round are
        // issued before its atomics
        for (uint32_t e0 = s0 + tid; e0 < s1; e0 += LSORT_T * LSORT_U) {
            int o[LSORT_U];
#pragma unroll
            for (int u = 0; u < LSORT_U; ++u) {
                const uint32_t e = e0 + u * LSORT_T;
                o[u] = e < s1 ? (int)((uint32_t)gp[e].y & 0xFFFFFFu) - c0 : -1;
            }
#pragma unroll
            for (int u = 0; u < LSORT_U; ++u)
                if (o[u] >= 0 && o[u] < nc) atomicAdd(&h[o[u]], 1u);
        }
        __syncthreads();
        // exclusive scans of the counts and of the non-empty flags of
        // h[0, nc), each thread over a contiguous run
        const int per = (nc + LSORT_T - 1) / LSORT_T;
        const int oa = min(nc, tid * per), ob = min(nc, oa + per);
        uint32_t sum = 0, ne = 0;
        for (int o = oa; o < ob; ++o) {
            sum += h[o];
            ne += h[o] ? 1u : 0u;
        }
        uint32_t tot, ntot;
        uint32_t run = block_excl_scan(sum, ws, tot) + done;
        uint32_t rank = block_excl_scan(ne, ws, ntot) + listed;
        for (int o = oa; o < ob; ++o) {
            const uint32_t c = h[o];
            h[o] = run;
            if (c) row[rank++] = make_int2(c0 + o, (int)(s0 + run));
            run += c;
        }
        __syncthreads();
        for (uint32_t e0 = s0 + tid; e0 < s1; e0 += LSORT_T * LSORT_U) {
            int2 v[LSORT_U];
#pragma unroll
            for (int u = 0; u < LSORT_U; ++u) {
                const uint32_t e = e0 + u * LSORT_T;
                v[u] = e < s1 ? gp[e] : make_int2(0, 0);
            }
#pragma unroll
            for (int u = 0; u < LSORT_U; ++u) {
                // (column << 24 | offset) may have its sign bit set: the
                // pair's place decides whether it is one
                const bool ok = e0 + u * LSORT_T < s1;
                const int o = ok ? (int)((uint32_t)v[u].y & 0xFFFFFFu) - c0 : -1;
                if (o >= 0 && o < nc) {
                    const uint32_t p = s0 + atomicAdd(&h[o], 1u);
                    pk[p] = v[u].x | (int)((uint32_t)v[u].y & 0xFF000000u);
                }
            }
        }
        done += tot;
        listed += ntot;
        __syncthreads();
    }
    if (tid == 0) gnb[g] = listed;
}

// The near-pair binning kernels take RUN_U wavefront-chunks of 64 entries
// per loop round: the round's loads are issued before its atomics, so that
// each thread has RUN_U loads in flight.
constexpr int RUN_U = 4;

__global__ __launch_bounds__(256) void k_fmm_p2p_hist(const int2* __restrict__ pr, int64_t n,
                                                      uint32_t* __restrict__ hist) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); base0 < n;
         base0 += stride * RUN_U) {
        int key[RUN_U];
#pragma unroll
        for (int u = 0; u < RUN_U; ++u) {
            const int64_t i = base0 + u * stride + __lane_id();
            key[u] = i < n ? pr[i].x : -1;
        }
#pragma unroll
        for (int u = 0; u < RUN_U; ++u) {
            const int64_t base = base0 + u * stride;
            if (base >= n) break;
            const RunInfo r = run_of(key[u], base, n);
            if (r.head) atomicAdd(&hist[key[u]], (uint32_t)r.count);
        }
    }
}

__global__ __launch_bounds__(256) void k_fmm_p2p_scatter(const int2* __restrict__ pr, int64_t n,
                                                         uint32_t* __restrict__ pos,
                                                         int* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); base0 < n;
         base0 += stride * RUN_U) {
        int2 e[RUN_U];
#pragma unroll
        for (int u = 0; u < RUN_U; ++u) {
            const int64_t i = base0 + u * stride + __lane_id();
            e[u] = i < n ? pr[i] : make_int2(-1, 0);
        }
#pragma unroll
        for (int u = 0; u < RUN_U; ++u) {
            const int64_t base = base0 + u * stride;
            if (base >= n) break;
            const RunInfo r = run_of(e[u].x, base, n);
            uint32_t o = 0;
            if (r.head) o = atomicAdd(&pos[e[u].x], (uint32_t)r.count);
            o = __shfl(o, r.hlane, 64) + (uint32_t)(__lane_id() - r.hlane);
            if (r.valid) out[o] = e[u].y;
        }
    }
}

// ------------------------------------------------------------------ M2L
// Lattice M2L on the matrix cores.  One workgroup per group of target cubes
// keeps their FP32 locals in LDS.  The group's pairs are binned by operator
// (offset); a bin's product D = T_o Q runs 16 columns (pairs) at a time on
// v_mfma_f32_16x16x4f32 and each column is added into its target's locals.
// A target appears once per bin, so the adds need no atomics as long as no
// two wavefronts write the same locals at once: at order 4 each wavefront
// takes its own bins and keeps a private copy of the locals (summed at the
// end); at order 6 (too large for private copies) every wavefront takes
// every tile but only its own 16 rows of the product.  (LDS float atomics
// into one shared copy cost about 135 cycles per instruction.)

// Adds a tile's column into a private copy in two halves: the reads are
// issued first, the adds and writes later, so that the read latency
// overlaps other work.
template <int RT, int Kq, int LS>
__device__ __forceinline__ void m2l_acc_read(const float* mine, int c, int lk, float4* v) {
    if (c < 0) return;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int row = 16 * rt + 4 * lk;
        if (row < Kq) v[rt] = *reinterpret_cast<const float4*>(mine + c * LS + row);
    }
}

template <int RT, int Kq, int LS>
__device__ __forceinline__ void m2l_acc_write(float* mine, int c, int lk, float4* v,
                                              const floatx4* d) {
    if (c < 0) return;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int row = 16 * rt + 4 * lk;
        if (row < Kq) {
            v[rt].x += d[rt][0];
            v[rt].y += d[rt][1];
            v[rt].z += d[rt][2];
            v[rt].w += d[rt][3];
            *reinterpret_cast<float4*>(mine + c * LS + row) = v[rt];
        }
    }
}

// Order-4 lattice M2L on the matrix cores.  One workgroup of M2L4_WAVES
// wavefronts per group of 128 target cubes (with nsplit > 1, per share of
// the group's bins); wavefront w takes a contiguous share of the bins (in
// offset order) and cuts each into tiles of 16 pairs, the columns.  A tile's
// product D = T_o Q (48 x 16: the operator padded to 16-row tiles, the
// contraction to 36) runs on v_mfma_f32_16x16x4f32 and is added into the
// wavefront's private FP32 copy of the group's locals in LDS (a plain 16-byte
// read-add-write: a target appears once per bin); the copies are summed in
// order at the end.  A bin's operator is built in registers from the FP32
// Taylor coefficients a_k of its offset, staged in LDS:
// T[n][m] = c[n][m] a_{n+m} 2^e, with e the class scaling (rows by 2^(|n|+1)
// for a coarser target, columns by 2^|m| for a finer one).  The loads are
// software pipelined over the wavefront's stream of tiles: the step of tile
// j fetches the moments, columns and Taylor coefficients of tile j+2 and the
// source ids of tile j+4, so every load has two steps to arrive.  Every step
// issues the same loads (past the end of the stream they read valid dummy
// records), which keeps the wait counts the same from step to step.
constexpr int M2L4_WAVES = 4;
constexpr int M2L4_GT = 128;

struct M2lTile {   // one tile of a wavefront's stream (wave-uniform)
    int o;         // offset id; -1 past the end of the stream
    uint32_t e0, b1;  // its pairs: [e0, min(e0 + 16, b1))
    int first;     // the first tile of its bin
};

struct M2lStream {  // the wavefront's bins [k, k1) of a group's list, tile by tile
    const int2* list;
    int k, k1, nbg;
    uint32_t s1, b0, b1;
    int o, t, nt;
    __device__ __forceinline__ M2lTile next() {
        if (t >= nt) {
            if (k >= k1) return M2lTile{-1, 0u, 0u, 0};
            // the same for every lane: kept in scalar registers
            o = __builtin_amdgcn_readfirstlane(list[k].x);
            b0 = (uint32_t)__builtin_amdgcn_readfirstlane(list[k].y);
            b1 = k + 1 < nbg ? (uint32_t)__builtin_amdgcn_readfirstlane(list[k + 1].y) : s1;
            nt = (int)((b1 - b0 + 15) / 16);
            t = 0;
            ++k;
        }
        const M2lTile r{o, b0 + 16u * (uint32_t)t, b1, t == 0};
        ++t;
        return r;
    }
};

template <int P>
__global__ __launch_bounds__(64 * M2L4_WAVES) void k_fmm_m2l4(
    int64_t nf, int noff, int noff0, int noff1, int nsplit, const int2* __restrict__ blist,
    const uint32_t* __restrict__ gnb, const uint32_t* __restrict__ gstart,
    const int* __restrict__ pk, const float* __restrict__ A,
    const float* __restrict__ cf, const uint16_t* __restrict__ ix, const int8_t* __restrict__ deg,
    const float* __restrict__ Qf, float* __restrict__ Lf) {
    constexpr int K = ncoef(P), K2 = ncoef(2 * P), Kp = kpad(P), Kq = kq(P);
    constexpr int RT = Kp / 16, KT = Kq / 4, NW = M2L4_WAVES, NT = 64 * NW, GT = M2L4_GT, LS = Kq;
    constexpr int K2V = (K2 + 3) / 4;  // float4 slots of one offset's Taylor coefficients
    static_assert(K2V <= 64, "one float4 of Taylor coefficients per lane");
    __shared__ float4 acc4[NW * GT * LS / 4];
    __shared__ float4 stage4[NW][K2V];
    float* acc = reinterpret_cast<float*>(acc4);
    const int g = blockIdx.x / nsplit, sp = blockIdx.x % nsplit;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lk = lane >> 4;
    for (int i = tid; i < NW * GT * LS / 4; i += NT) acc4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    // this lane's operator elements (n, m) = (16 rt + li, lk KT + kt):
    // coefficients, indices of n+m (four 8-bit indices per word), degrees
    float cfl[RT][KT];
    uint32_t ixw[(RT * KT + 3) / 4];
    int rdeg[RT], cdeg[KT];
#pragma unroll
    for (int w = 0; w < (RT * KT + 3) / 4; ++w) ixw[w] = 0u;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        rdeg[rt] = deg[16 * rt + li] + 1;
#pragma unroll
        for (int kt = 0; kt < KT; ++kt) {
            const int n = 16 * rt + li, m = lk * KT + kt, q = rt * KT + kt;
            cfl[rt][kt] = cf[n * Kq + m];
            ixw[q / 4] |= (uint32_t)ix[n * Kq + m] << (8 * (q % 4));
        }
    }
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) cdeg[kt] = deg[lk * KT + kt];
    __syncthreads();
    float* __restrict__ mine = acc + wv * GT * LS;
    float4* stg4 = stage4[wv];
    const float* stg = reinterpret_cast<const float*>(stage4[wv]);
    // this wavefront's bins: its share of the workgroup's share of the group
    const int nbg = __builtin_amdgcn_readfirstlane((int)gnb[g]);
    const int64_t kb = (int64_t)nbg * sp / nsplit, ke = (int64_t)nbg * (sp + 1) / nsplit;
    M2lStream st;
    st.list = blist + (int64_t)g * noff;
    st.k = (int)(kb + (ke - kb) * wv / NW);
    st.k1 = (int)(kb + (ke - kb) * (wv + 1) / NW);
    st.nbg = nbg;
    st.s1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)gstart[g + 1]);
    st.b0 = st.b1 = 0u;
    st.o = 0;
    st.t = st.nt = 0;
    M2lTile t0 = st.next(), t1 = st.next(), t2 = st.next(), t3 = st.next(), t4 = st.next();
    // a lane's pair in tile t: index, or -1 (past the bin or the stream)
    auto pair_of = [&](const M2lTile& t) -> int64_t {
        const uint32_t e = t.e0 + (uint32_t)li;
        return (t.o >= 0 && e < t.b1) ? (int64_t)e : -1;
    };
    // loads of a step: moments, column and Taylor coefficients of a tile (its
    // source id fetched two steps before), the source id of a later tile.
    // They go to rings of three slots; the loop is unrolled three times so
    // that a slot is never copied while its load is in flight.
    float bq0[KT], bq1[KT], bq2[KT];
    float4 ak0, ak1, ak2;
    int cl0, cl1, cl2, sr0, sr1, sr2;
    // ring slot R of a ring of three named registers (R known at compile time)
    auto slot = [](auto RS, auto& x0, auto& x1, auto& x2) -> auto& {
        constexpr int R = decltype(RS)::value % 3;
        if constexpr (R == 0) return x0;
        else if constexpr (R == 1) return x1;
        else return x2;
    };
    // sid: the tile's packed entry (source | column << 24)
    auto load_q = [&](const M2lTile& t, int sid, float (&q3)[KT], int& c) {
        const int64_t e = pair_of(t);
        const float* q = Qf + (e >= 0 ? (int64_t)(sid & 0xFFFFFF) : nf) * Kq + lk * KT;
#pragma unroll
        for (int kt = 0; kt < KT; ++kt) q3[kt] = q[kt];
        c = sid >> 24;
    };
    auto load_a = [&](const M2lTile& t) -> float4 {
        return reinterpret_cast<const float4*>(A + (int64_t)(t.o >= 0 ? t.o : 0) * (4 * K2V))[lane < K2V ? lane : 0];
    };
    auto load_s = [&](const M2lTile& t) -> int {
        const int64_t e = pair_of(t);
        return pk[e >= 0 ? e : 0];
    };
    // the operator of tile t from its Taylor coefficients, staged in LDS: the
    // elements are read first (build_read), then scaled in place (build_scale)
    const auto build_read = [&](float (&avx)[RT][KT]) {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int kt = 0; kt < KT; ++kt) {
                const int q = rt * KT + kt;
                avx[rt][kt] = stg[(ixw[q / 4] >> (8 * (q % 4))) & 0xFFu];
            }
    };
    const auto build_scale = [&](const M2lTile& t, float (&avx)[RT][KT]) {
        if (t.o < noff0) {
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int kt = 0; kt < KT; ++kt) avx[rt][kt] *= cfl[rt][kt];
        } else {
            const int f1 = t.o < noff0 + noff1, f2 = 1 - f1;
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                for (int kt = 0; kt < KT; ++kt)
                    avx[rt][kt] = ldexpf(cfl[rt][kt] * avx[rt][kt], f1 * rdeg[rt] + f2 * cdeg[kt]);
        }
    };
    const auto stage = [&](const float4& a) {
        __builtin_amdgcn_wave_barrier();
        if (lane < K2V) stg4[lane] = a;
        __builtin_amdgcn_wave_barrier();
    };
    float av[RT][KT];
    {
        // in the order of two steps, so that the loop's first wait is the
        // same as its later ones
        const int s0 = load_s(t0), s1 = load_s(t1);
        __builtin_amdgcn_sched_barrier(0);
        load_q(t0, s0, bq0, cl0);
        ak0 = load_a(t0);
        sr2 = load_s(t2);
        __builtin_amdgcn_sched_barrier(0);
        load_q(t1, s1, bq1, cl1);
        ak1 = load_a(t1);
        sr0 = load_s(t3);
        __builtin_amdgcn_sched_barrier(0);
    }
    floatx4 dp[RT];
    int cp = -1;
    bool have_prev = false;
    // step for tile t0 in ring slot R (its loads made two steps before):
    // fetch tile t2's moments, column and Taylor coefficients into slot R + 2
    // (its source id in slot R + 2 of the source ring) and the source id of
    // tile t4 into slot R + 1 (mod 3)
    auto step = [&](auto RS) {
        using C0 = std::integral_constant<int, decltype(RS)::value>;
        using C1 = std::integral_constant<int, decltype(RS)::value + 1>;
        using C2 = std::integral_constant<int, decltype(RS)::value + 2>;
        float(&bqR)[KT] = slot(C0{}, bq0, bq1, bq2);
        load_q(t2, slot(C2{}, sr0, sr1, sr2), slot(C2{}, bq0, bq1, bq2), slot(C2{}, cl0, cl1, cl2));
        slot(C2{}, ak0, ak1, ak2) = load_a(t2);
        slot(C1{}, sr0, sr1, sr2) = load_s(t4);
        __builtin_amdgcn_sched_barrier(0);
        if (t0.first) {
            stage(slot(C0{}, ak0, ak1, ak2));
            build_read(av);
            build_scale(t0, av);
        }
        float4 pv[RT];
        if (have_prev) m2l_acc_read<RT, Kq, LS>(mine, cp, lk, pv);
        floatx4 d[RT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) d[rt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < KT / 2; ++kt)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
                d[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[rt][kt], bqR[kt], d[rt], 0, 0, 0);
        // the previous tile's sums go into the locals while the second half
        // of this tile's products runs
        __builtin_amdgcn_sched_barrier(0);
        if (have_prev) m2l_acc_write<RT, Kq, LS>(mine, cp, lk, pv, dp);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int kt = KT / 2; kt < KT; ++kt)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
                d[rt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[rt][kt], bqR[kt], d[rt], 0, 0, 0);
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) dp[rt] = d[rt];
        cp = pair_of(t0) >= 0 ? slot(C0{}, cl0, cl1, cl2) : -1;
        have_prev = true;
        t0 = t1;
        t1 = t2;
        t2 = t3;
        t3 = t4;
        t4 = st.next();
    };
    // three steps per pass, checked at the end: up to two steps past the end
    // of the stream run on the zero record and add nothing
    if (t0.o >= 0) do {
        step(std::integral_constant<int, 0>{});
        step(std::integral_constant<int, 1>{});
        step(std::integral_constant<int, 2>{});
    } while (t0.o >= 0);
    if (have_prev) {
        float4 pv[RT];
        m2l_acc_read<RT, Kq, LS>(mine, cp, lk, pv);
        m2l_acc_write<RT, Kq, LS>(mine, cp, lk, pv, dp);
    }
    __syncthreads();
    const int64_t f0 = (int64_t)g * GT;
    for (int i = tid; i < GT * K; i += NT) {
        const int tc = i / K, k = i % K;
        const int64_t f = f0 + tc;
        if (f >= nf) break;
        float v = 0.0f;
#pragma unroll
        for (int w = 0; w < NW; ++w) v += acc[(w * GT + tc) * LS + k];
        float* dst = &Lf[f * Kp + k];
        if (nsplit == 1) *dst = v;
        else if (v != 0.0f) atomicAdd(dst, v);
    }
}

// Operand layout (wave64): A[i][k] from lane i + 16k, B[k][j] from lane
// j + 16k, D[4q + r][j] in lane j + 16q, element r.  The contraction index of
// step kt in lane group lk is lk * KT + kt, so that each lane reads KT
// consecutive moments of its source (wide loads).

// Order-6 lattice M2L on the matrix cores.  One workgroup of RT = 6
// wavefronts per group of M2L6_GT target cubes (with nsplit > 1, per share of
// the group's bins); all six follow the same stream of tiles of 16 pairs
// (bins in offset order), and wavefront r computes rows 16 r .. 16 r + 15 of
// each tile's product D = T_o Q (96 x 16, contraction 84) with
// v_mfma_f32_16x16x4f32.  The wavefronts thus own disjoint rows of the
// group's FP32 locals in LDS and add their rows with plain read-add-writes,
// no atomics (the LDS float atomics of a shared copy kept the LDS busy 87% of
// the time).  The tile's 16 source records are loaded once for all six into
// LDS, double-buffered: the next tile's records are fetched while the
// current one runs.  A wavefront's operator rows (21 values per lane) stay
// in registers over a bin; the next bin's are read from the operator table
// one tile ahead.  (Building them from the offset's 455 Taylor coefficients
// staged in LDS, as at order 4, took 85 ms instead of 67 at 4M: the build
// then sits between the barrier and the products.)
template <int P>
__global__ __launch_bounds__(64 * (kpad(P) / 16)) void k_fmm_m2l6(
    int64_t nf, int noff, int nsplit, const int2* __restrict__ blist,
    const uint32_t* __restrict__ gnb, const uint32_t* __restrict__ gstart,
    const int* __restrict__ pk, const float* __restrict__ T,
    const float* __restrict__ Qf, float* __restrict__ Lf) {
    constexpr int K = ncoef(P), Kp = kpad(P), Kq = kq(P), RT = Kp / 16, KT = Kq / 4;
    constexpr int NT = 64 * RT, GT = M2L6_GT, KQ4 = Kq / 4, NS = 16 * KQ4;
    static_assert(K == Kq && K % 4 == 0 && NS <= NT, "whole float4 rows; one record slice per thread");
    __shared__ float4 loc4[GT * K / 4];
    __shared__ float4 bst4[2][NS];
    __shared__ int bcol[2][16];
    float* loc = reinterpret_cast<float*>(loc4);
    const float* bstf = reinterpret_cast<const float*>(&bst4[0][0]);
    const int g = blockIdx.x / nsplit, sp = blockIdx.x % nsplit;
    const int tid = threadIdx.x, lane = tid & 63;
    const int rt = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lk = lane >> 4;
    const int sj = tid / KQ4, sq = tid % KQ4;  // staging: this thread's column and record slice
    for (int i = tid; i < GT * K / 4; i += NT) loc4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    const int nbg = __builtin_amdgcn_readfirstlane((int)gnb[g]);
    M2lStream st;
    st.list = blist + (int64_t)g * noff;
    st.k = (int)((int64_t)nbg * sp / nsplit);
    st.k1 = (int)((int64_t)nbg * (sp + 1) / nsplit);
    st.nbg = nbg;
    st.s1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)gstart[g + 1]);
    st.b0 = st.b1 = 0u;
    st.o = 0;
    st.t = st.nt = 0;
    // the staging thread's pair of tile t, or -1 (not a staging thread, past
    // the bin or past the stream)
    auto pair_of = [&](const M2lTile& t) -> int64_t {
        const uint32_t e = t.e0 + (uint32_t)sj;
        return (tid < NS && t.o >= 0 && e < t.b1) ? (int64_t)e : -1;
    };
    // the packed entry (source | column << 24), -1 for an empty column,
    // which reads the zero record nf and adds to no column
    auto load_src = [&](const M2lTile& t) -> int {
        const int64_t e = pair_of(t);
        return e >= 0 ? pk[e] : -1;
    };
    auto src_of = [&](int p) -> int { return p >= 0 ? (p & 0xFFFFFF) : (int)nf; };
    auto col_of = [](int p) -> int { return p >= 0 ? (p >> 24) : -1; };
    auto load_rec = [&](int sid) -> float4 {
        return reinterpret_cast<const float4*>(Qf + (int64_t)sid * Kq)[tid < NS ? sq : 0];
    };
    // this wavefront's operator rows of offset o: lane (li, lk) holds
    // T[16 rt + li][lk KT + kt], kt < KT
    auto load_op = [&](int o, float (&a)[KT]) {
        const float* to = T + ((int64_t)o * Kp + 16 * rt + li) * Kq + lk * KT;
#pragma unroll
        for (int kt = 0; kt < KT; ++kt) a[kt] = to[kt];
    };
    M2lTile t0 = st.next(), t1 = st.next(), t2 = st.next();
    float av[KT], an[KT];
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) av[kt] = an[kt] = 0.0f;
    if (t0.o >= 0) load_op(t0.o, av);
    {
        const int s0 = load_src(t0);
        const int c0 = col_of(s0);
        const float4 r0 = load_rec(src_of(s0));
        if (tid < NS) bst4[0][tid] = r0;
        if (tid < NS && sq == 0) bcol[0][sj] = c0;
    }
    int s1 = load_src(t1);
    int b = 0;
    while (t0.o >= 0) {
        // the next tile's records and column (and operator rows at a bin
        // start), the source ids of the tile after it
        const float4 r1 = load_rec(src_of(s1));
        const int c1 = col_of(s1);
        const int s2 = load_src(t2);
        if (t1.o >= 0 && t1.first) load_op(t1.o, an);
        __syncthreads();  // buffer b holds tile t0; buffer 1-b is no longer read
        floatx4 d = {0.f, 0.f, 0.f, 0.f};
        const float* bq = bstf + b * NS * 4 + li * Kq + lk * KT;
#pragma unroll
        for (int kt = 0; kt < KT; ++kt)
            d = __builtin_amdgcn_mfma_f32_16x16x4f32(av[kt], bq[kt], d, 0, 0, 0);
        const int c = bcol[b][li];
        const int row = 16 * rt + 4 * lk;
        if (c >= 0 && row < K) {
            float4* p = reinterpret_cast<float4*>(loc + c * K + row);
            float4 v = *p;
            v.x += d[0];
            v.y += d[1];
            v.z += d[2];
            v.w += d[3];
            *p = v;
        }
        if (tid < NS) bst4[1 - b][tid] = r1;
        if (tid < NS && sq == 0) bcol[1 - b][sj] = c1;
        if (t1.first) {
#pragma unroll
            for (int kt = 0; kt < KT; ++kt) av[kt] = an[kt];
        }
        t0 = t1;
        t1 = t2;
        t2 = st.next();
        s1 = s2;
        b = 1 - b;
    }
    __syncthreads();
    const int64_t f0 = (int64_t)g * GT;
    for (int i = tid; i < GT * K; i += NT) {
        const int tc = i / K, k = i % K;
        const int64_t f = f0 + tc;
        if (f >= nf) break;
        const float v = loc[tc * K + k];
        float* dst = &Lf[f * Kp + k];
        if (nsplit == 1) *dst = v;
        else if (v != 0.0f) atomicAdd(dst, v);
    }
}

template <int P>
__global__ __launch_bounds__(64) void k_fmm_m2l_fp64(const int2* __restrict__ pr, int64_t npr,
                                                     const int3* __restrict__ lat, int64_t nlat,
                                                     FView v, const int* __restrict__ mi,
                                                     const int* __restrict__ mi2,
                                                     const int* __restrict__ rec2,
                                                     const int* __restrict__ midx,
                                                     const double* __restrict__ mcoef,
                                                     const double* __restrict__ Qd,
                                                     double* __restrict__ Ld) {
    constexpr int K = ncoef(P), K2 = ncoef(2 * P);
    __shared__ double a[K2];
    __shared__ double q[K];
    for (int64_t i = blockIdx.x; i < npr + nlat; i += gridDim.x) {
        int fa, fb;
        if (i < npr) fa = pr[i].x, fb = pr[i].y;
        else fa = lat[i - npr].x, fb = lat[i - npr].y;
        const int lane = threadIdx.x;
        __syncthreads();  // the previous pair is done with a[] and q[]
        const int ka = v.fnode[fa], kb = v.fnode[fb];
        const double sa = side_of(v, ka), sb = side_of(v, kb);
        const double R[3] = {(v.cx[ka] - v.cx[kb]) / sa, (v.cy[ka] - v.cy[kb]) / sa,
                             (v.cz[ka] - v.cz[kb]) / sa};
        const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
        const double ratio = sb / sa;
        for (int m = lane; m < K; m += 64) q[m] = Qd[(int64_t)fb * K + m] * pow(ratio, mi[4 * m + 3]);
        if (lane == 0) a[0] = 1.0 / sqrt(r2);
        __syncthreads();
        int lo = 1;
        for (int ord = 1; ord <= 2 * P; ++ord) {
            const int hi = lo + (ord + 1) * (ord + 2) / 2;
            for (int j = lo + lane; j < hi; j += 64) {
                double s = 0.0;
                for (int d = 0; d < 3; ++d) {
                    const int e1 = rec2[6 * j + d], e2 = rec2[6 * j + 3 + d];
                    if (e1 >= 0) s += (2 * ord - 1) * R[d] * a[e1];
                    if (e2 >= 0) s += (ord - 1) * a[e2];
                }
                a[j] = -s / (ord * r2);
            }
            __syncthreads();
            lo = hi;
        }
        for (int n = lane; n < K; n += 64) {
            double acc = 0.0;
            for (int m = 0; m < K; ++m) acc += mcoef[n * K + m] * q[m] * a[midx[n * K + m]];
            unsafeAtomicAdd(&Ld[(int64_t)fa * K + n], acc);
        }
    }
}

template <int P>
__global__ __launch_bounds__(256) void k_fmm_merge(int64_t nf, const float* __restrict__ Lf,
                                                   double* __restrict__ Ld) {
    constexpr int K = ncoef(P), Kp = kpad(P);
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nf * K) return;
    const int64_t f = i / K;
    const int n = (int)(i % K);
    Ld[i] += (double)Lf[f * Kp + n];
}

// ------------------------------------------------------------------ downward
template <int P>
__global__ __launch_bounds__(64) void k_fmm_l2l(int f0, int f1, FView v, const int* __restrict__ par,
                                                const double* __restrict__ l2l,
                                                double* __restrict__ Ld) {
    constexpr int K = ncoef(P);
    const int f = f0 + blockIdx.x;
    if (f >= f1) return;
    const int k = v.fnode[f];
    const int p = par[k];
    if (p < 0) return;
    const int fp = v.fid[p];
    if (fp < 0) return;
    const double* t = l2l + (int64_t)v.digit[k] * K * K;
    const double* lp = Ld + (int64_t)fp * K;
    // parent coefficient n to child coefficient m, in the order of n
    double acc[2] = {0.0, 0.0};
    for (int n = 0; n < K; ++n) {
        const double ln = lp[n];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int m = threadIdx.x + 64 * h;
            if (m < K) acc[h] += t[n * K + m] * ln;
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int m = threadIdx.x + 64 * h;
        if (m < K) Ld[(int64_t)f * K + m] += acc[h];
    }
}

// Leaves: L2P (FP64) plus the near field from the leaf's near list, one
// wavefront per leaf.  The wavefront splits into 64 / W lane groups of W
// lanes, W the smallest power of two >= the leaf's count: lane t of group g
// holds body t, takes the L2P terms n = 1 + g, 1 + g + 64/W, ... and every
// (64/W)-th staged source body from the g-th, and the groups' sums meet at
// the end by cross-lane adds (so a leaf of 16 bodies keeps all 64 lanes
// busy).  Source bodies are staged in LDS, up to NB at a time: the near-list
// entries are read 64 at a time, one per lane, and their bodies are copied by
// 16-lane quarters, one entry per quarter.  FP32 mode sums the near field in
// FP32 on coordinates recentred on the leaf centre, Kahan-compensated per
// staged batch; FP64 mode in FP64.  A leaf holds at most 64 bodies (cells of
// more are split, and a tree leaf holds one).
template <int P, bool F32>
__global__ __launch_bounds__(64) void k_fmm_leaf(int nf, FView v, const double* __restrict__ sx,
                                                 const double* __restrict__ sy,
                                                 const double* __restrict__ sz,
                                                 const double* __restrict__ sm,
                                                 const double4* __restrict__ sb,
                                                 const int* __restrict__ mi,
                                                 const double* __restrict__ Ld,
                                                 const uint32_t* __restrict__ pstart,
                                                 const int* __restrict__ plist, double G,
                                                 double eps2, double* __restrict__ fx,
                                                 double* __restrict__ fy, double* __restrict__ fz) {
    constexpr int K = ncoef(P);
    constexpr int NB = 256;  // staged source bodies (LDS)
    using V4 = typename std::conditional<F32, float4, double4>::type;
    __shared__ V4 sbuf[NB];
    __shared__ int sidx[NB];
    const int f = blockIdx.x;
    if (f >= nf) return;
    const int k = v.fnode[f];
    if (v.flag[k] != 1) return;
    const int lane = threadIdx.x;
    const int first = v.first[k], cnt = min(v.count[k], 64);
    int W = 1;
    while (W < cnt) W <<= 1;
    const int NG = 64 / W, tl = lane & (W - 1), grp = lane / W;
    const bool ok = tl < cnt;
    const int i = first + (ok ? tl : 0);
    const double s = side_of(v, k), c[3] = {v.cx[k], v.cy[k], v.cz[k]};
    const double xi = sx[i], yi = sy[i], zi = sz[i];
    // L2P: a = -(1/s^2) sum_n L~_n n_d u~^(n - e_d); a zero-size cube (one
    // body, or bodies at one point) has no far field
    const double si = s > 0.0 ? 1.0 / s : 0.0;
    const double u[3] = {(xi - c[0]) * si, (yi - c[1]) * si, (zi - c[2]) * si};
    double pw[3][P + 1];
    for (int d = 0; d < 3; ++d) {
        pw[d][0] = 1.0;
        for (int e = 1; e <= P; ++e) pw[d][e] = pw[d][e - 1] * u[d];
    }
    double g[3] = {0.0, 0.0, 0.0};
    const double* L = Ld + (int64_t)f * K;
    for (int n = 1 + grp; n < K; n += NG) {
        const int nx = mi[4 * n], ny = mi[4 * n + 1], nz = mi[4 * n + 2];
        const double l = L[n];
        if (nx) g[0] += l * nx * pw[0][nx - 1] * pw[1][ny] * pw[2][nz];
        if (ny) g[1] += l * ny * pw[0][nx] * pw[1][ny - 1] * pw[2][nz];
        if (nz) g[2] += l * nz * pw[0][nx] * pw[1][ny] * pw[2][nz - 1];
    }
    const double sc = -si * si;
    double ax = sc * g[0], ay = sc * g[1], az = sc * g[2];
    float ks[3] = {0.f, 0.f, 0.f}, kc[3] = {0.f, 0.f, 0.f};
    const float tx = (float)(xi - c[0]), ty = (float)(yi - c[1]), tz = (float)(zi - c[2]);
    const uint32_t e1 = pstart[f + 1];
    const int q = lane >> 4, ql = lane & 15;
    for (uint32_t e = pstart[f]; e < e1;) {
        // this lane's entry e + lane: its body range
        const uint32_t ee = e + (uint32_t)lane;
        int bf = 0, bc = 0;
        if (ee < e1) {
            const int kb = v.fnode[plist[ee]];
            bf = v.first[kb];
            bc = min(v.count[kb], 64);
        }
        // the longest run of entries whose bodies fit in NB (at least one)
        const int incl = wave_incl_scan(bc);
        const int take = max(1, __popcll(__ballot(ee < e1 && incl <= NB)));
        const int fill = min(NB, __shfl(incl, take - 1, 64));
        __syncthreads();  // the previous batch is no longer read
        for (int j0 = 0; j0 < take; j0 += 4) {
            const int j = j0 + q;
            const int jf = __shfl(bf, j & 63, 64), jc = __shfl(bc, j & 63, 64);
            const int jp = __shfl(incl - bc, j & 63, 64);
            if (j < take)
                for (int b = ql; b < jc && jp + b < NB; b += 16) {
                    const int jj = jf + b;
                    const double4 r = sb[jj];  // the sorted body as one (x, y, z, m) record
                    if constexpr (F32) {
                        sbuf[jp + b] = make_float4((float)(r.x - c[0]), (float)(r.y - c[1]),
                                                   (float)(r.z - c[2]), (float)r.w);
                    } else {
                        sbuf[jp + b] = r;
                        sidx[jp + b] = jj;
                    }
                }
        }
        __syncthreads();
        if constexpr (F32) {
            float p[3] = {0.f, 0.f, 0.f};
            const float e2 = (float)eps2;
#pragma unroll 4
            for (int j = grp; j < fill; j += NG) {
                const float4 sq = sbuf[j];
                const float dx = sq.x - tx, dy = sq.y - ty, dz = sq.z - tz;
                const float d2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, e2)));
                // the target itself (staged by the same recentring, so dx = dy =
                // dz = 0 exactly) adds nothing: with eps2 > 0 through the zero
                // differences, with eps2 = 0 through ri = 0
                const float ri = d2 > 0.0f ? __frsqrt_rn(d2) : 0.0f;
                const float w = sq.w * ri * ri * ri;
                p[0] = fmaf(w, dx, p[0]);
                p[1] = fmaf(w, dy, p[1]);
                p[2] = fmaf(w, dz, p[2]);
            }
            for (int d = 0; d < 3; ++d) {
                const float yv = p[d] - kc[d];
                const float tv = ks[d] + yv;
                kc[d] = (tv - ks[d]) - yv;
                ks[d] = tv;
            }
        } else {
            for (int j = grp; j < fill; j += NG) {
                if (sidx[j] == i) continue;
                const V4 sq = sbuf[j];
                const double dx = sq.x - xi, dy = sq.y - yi, dz = sq.z - zi;
                const double d2 = dx * dx + dy * dy + dz * dz + eps2;
                const double w = sq.w / (sqrt(d2) * d2);
                ax += w * dx;
                ay += w * dy;
                az += w * dz;
            }
        }
        e += (uint32_t)take;
    }
    if (F32) {
        ax += (double)ks[0] - (double)kc[0];
        ay += (double)ks[1] - (double)kc[1];
        az += (double)ks[2] - (double)kc[2];
    }
    for (int m = W; m < 64; m <<= 1) {
        ax += __shfl_xor(ax, m, 64);
        ay += __shfl_xor(ay, m, 64);
        az += __shfl_xor(az, m, 64);
    }
    if (ok && grp == 0) {
        const double gm = G * sm[i];
        fx[i] = gm * ax;
        fy[i] = gm * ay;
        fz[i] = gm * az;
    }
}

// M2P of the lopsided pairs (a leaf target much larger than its source cell):
// one wavefront per pair, one lane per target body, added to its force with
// atomic adds (a target's pairs run in parallel).
template <int P>
__global__ __launch_bounds__(64) void k_fmm_m2p(const int2* __restrict__ pr, int64_t npr, FView v,
                                                const double* __restrict__ sx,
                                                const double* __restrict__ sy,
                                                const double* __restrict__ sz,
                                                const double* __restrict__ sm,
                                                const int* __restrict__ mi2,
                                                const int* __restrict__ rec2,
                                                const double* __restrict__ Qd, double G,
                                                double* __restrict__ fx, double* __restrict__ fy,
                                                double* __restrict__ fz) {
    constexpr int K = ncoef(P), K1 = ncoef(P + 1);
    __shared__ double at[K1][64];  // per-lane Taylor coefficients
    const int lane = threadIdx.x;
    for (int64_t e = blockIdx.x; e < npr; e += gridDim.x) {
        const int fa = pr[e].x, fb = pr[e].y;
        const int k = v.fnode[fa], kb = v.fnode[fb];
        const int first = v.first[k], cnt = v.count[k];
        const bool ok = lane < cnt;
        const int i = first + (ok ? lane : 0);
        // M2P: a_d = s^-2 sum_m (-1)^|m| Q~_m (m_d + 1) a_{m+e_d}((x - c_B) / s),
        // accumulated as each a_k (|k| <= p+1) is produced
        const double sb = side_of(v, kb);
        const double R[3] = {(sx[i] - v.cx[kb]) / sb, (sy[i] - v.cy[kb]) / sb, (sz[i] - v.cz[kb]) / sb};
        const double r2 = R[0] * R[0] + R[1] * R[1] + R[2] * R[2];
        const double* q = Qd + (int64_t)fb * K;
        double g2[3] = {0.0, 0.0, 0.0};
        at[0][lane] = 1.0 / sqrt(r2);
        for (int j = 1; j < K1; ++j) {
            const int ord = mi2[4 * j + 3];
            double sacc = 0.0;
            for (int d = 0; d < 3; ++d) {
                const int e1 = rec2[6 * j + d], e2 = rec2[6 * j + 3 + d];
                if (e1 >= 0) sacc += (2 * ord - 1) * R[d] * at[e1][lane];
                if (e2 >= 0) sacc += (ord - 1) * at[e2][lane];
            }
            const double aj = -sacc / (ord * r2);
            at[j][lane] = aj;
            for (int d = 0; d < 3; ++d) {
                const int mm = rec2[6 * j + d];  // index of k - e_d
                if (mm >= 0 && ord - 1 <= P) {
                    const double sg = ((ord - 1) & 1) ? -1.0 : 1.0;
                    g2[d] += sg * q[mm] * (double)mi2[4 * j + d] * aj;
                }
            }
        }
        if (ok) {
            const double gm = G * sm[i] / (sb * sb);
            atomicAdd(&fx[i], gm * g2[0]);
            atomicAdd(&fy[i], gm * g2[1]);
            atomicAdd(&fz[i], gm * g2[2]);
        }
    }
}

// the roots (tree nodes 0 .. nsys-1) as FMM cell pairs; an empty system's
// root is no FMM cell (id -1) and produces nothing
__global__ __launch_bounds__(256) void k_fmm_roots(int nsys, const int* __restrict__ fid,
                                                   int2* __restrict__ fr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nsys) fr[k] = make_int2(fid[k], fid[k]);
}

// ------------------------------------------------------------------ driver
// workgroups for a grid-stride loop over n items, 256 per group
static inline int capped(int64_t n) { return (int)std::min<int64_t>(grid_for(std::max<int64_t>(n, 1), 256), 1 << 16); }
// tiles of lattice pairs: eight workgroups per CU, each over every 2048th tile
static inline int bin_grid(int64_t n) {
    return (int)std::min<int64_t>((std::max<int64_t>(n, 1) + BT_TILE - 1) / BT_TILE, 2048);
}

template <int P>
static int fmm_run(hq_tree* t, FmmState& st, double theta, double G, double eps2, bool mfma,
                   int S) {
    constexpr int K = ncoef(P), Kp = kpad(P);
    hipStream_t s = t->stream;
    int rc;
    FmmTables& tb = st.tab;
    if ((rc = build_tables(tb, P, theta))) return rc;
    const int64_t nn = t->nn, n = t->n;
    HQ_CHECK_HIP(hipEventRecord(t->ev[5], s));
    // 1. FMM node set
    if ((rc = st.flag.ensure(nn)) || (rc = st.scan.ensure((nn + 1) * 4)) ||
        (rc = st.fid.ensure(nn * 4)) || (rc = st.lvl.ensure(64 * 4)))
        return rc;
    k_fmm_flag<<<grid_for(nn, 256), 256, 0, s>>>(nn, S, t->ncount.as<int>(), t->ntype.as<int8_t>(),
                                                 t->ndepth.as<uint8_t>(), t->nparent.as<int>(),
                                                 st.flag.as<int8_t>(), st.scan.as<uint32_t>());
    if ((rc = scan_excl_buf(st.scan.as<uint32_t>(), nn, t->scanpart, s))) return rc;
    uint32_t nf32 = 0;
    HQ_CHECK_HIP(hipMemcpyAsync(&nf32, st.scan.as<uint32_t>() + nn, 4, hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    const int64_t nf = nf32;
    st.nf = nf;
    if ((rc = st.fnode.ensure(std::max<int64_t>(nf, 1) * 4))) return rc;
    k_fmm_compact<<<grid_for(nn, 256), 256, 0, s>>>(nn, st.flag.as<int8_t>(), st.scan.as<uint32_t>(),
                                                    st.fid.as<int>(), st.fnode.as<int>());
    {
        std::vector<int> ls(MAX_DEPTH + 2);
        for (int d = 0; d <= MAX_DEPTH + 1; ++d) ls[d] = std::min<int64_t>(t->level_start[d], nn);
        DevBuf dls;
        if ((rc = dls.ensure(ls.size() * 4))) return rc;
        HQ_CHECK_HIP(hipMemcpyAsync(dls.p, ls.data(), ls.size() * 4, hipMemcpyHostToDevice, s));
        k_fmm_levels<<<1, 64, 0, s>>>(st.scan.as<uint32_t>(), dls.as<int>(), st.lvl.as<int>());
        st.fls.assign(MAX_DEPTH + 2, 0);
        HQ_CHECK_HIP(hipMemcpyAsync(st.fls.data(), st.lvl.p, (MAX_DEPTH + 2) * 4, hipMemcpyDeviceToHost, s));
        HQ_CHECK_HIP(hipStreamSynchronize(s));
    }
    FView v;
    v.first = t->nfirst.as<int>();
    v.count = t->ncount.as<int>();
    v.cstart = t->ncstart.as<int>();
    v.fid = st.fid.as<int>();
    v.fnode = st.fnode.as<int>();
    v.ssys = t->ssys.as<int>();
    v.flag = st.flag.as<int8_t>();
    v.depth = t->ndepth.as<uint8_t>();
    v.digit = t->ndigit.as<uint8_t>();
    v.cx = t->ncx.as<double>();
    v.cy = t->ncy.as<double>();
    v.cz = t->ncz.as<double>();
    v.sg = t->geom.as<double>() + G_SYS;
    // 2. upward
    if ((rc = st.Qd.ensure(nf * K * 8)) || (rc = st.Qf.ensure((nf + 1) * kq(P) * 4)) ||
        (rc = st.Ld.ensure(nf * K * 8)) || (rc = st.Lf.ensure(nf * Kp * 4)) ||
        (rc = st.rmax.ensure(nf * 8)))
        return rc;
    k_fmm_p2m<P><<<(int)nf, 64, 0, s>>>((int)nf, v, t->sx.as<double>(), t->sy.as<double>(),
                                        t->sz.as<double>(), t->sm.as<double>(), tb.mi.as<int>(),
                                        st.Qd.as<double>(), st.rmax.as<double>());
    for (int d = MAX_DEPTH; d >= 0; --d) {
        const int f0 = st.fls[d], f1 = st.fls[d + 1];
        if (f1 > f0)
            k_fmm_m2m<P><<<f1 - f0, 64, 0, s>>>(f0, f1, v, tb.m2m.as<double>(), st.Qd.as<double>(),
                                                st.rmax.as<double>());
    }
    k_fmm_qf<P><<<grid_for((nf + 1) * kq(P), 256), 256, 0, s>>>(nf, st.Qd.as<double>(), st.Qf.as<float>());
    if ((rc = st.rec.ensure(std::max<int64_t>(nf, 1) * sizeof(FRec)))) return rc;
    k_fmm_records<<<grid_for(std::max<int64_t>(nf, 1), 256), 256, 0, s>>>(
        (int)nf, v, st.scan.as<uint32_t>(), st.rmax.as<double>(), st.rec.as<FRec>());
    HQ_CHECK_HIP(hipMemsetAsync(st.Ld.p, 0, nf * K * 8, s));
    // 3. dual walk: frontier ping-pong in fa / fb; lists appended
    WalkPar w{theta * theta, tb.R, tb.nd, tb.nd1, tb.noff0, tb.noff1};
    int64_t nfr = t->nsys, nlat = 0, ngen = 0, np2p = 0, nm2p = 0;
    // the lists keep their capacity from one evaluation to the next
    size_t lat_cap = st.lat.bytes / sizeof(int3), gen_cap = st.gen.bytes / sizeof(int2),
           p2p_cap = st.p2p.bytes / sizeof(int2), m2p_cap = st.m2p.bytes / sizeof(int2);
    if ((rc = st.fa.ensure(std::max<int64_t>(nfr, 1024) * sizeof(int2))) ||
        (rc = st.edges.ensure(8 * sizeof(uint32_t))))
        return rc;
    k_fmm_roots<<<grid_for(nfr, 256), 256, 0, s>>>((int)nfr, st.fid.as<int>(), st.fa.as<int2>());
    DevBuf* cur = &st.fa;
    DevBuf* nxt = &st.fb;
    auto grow = [&](DevBuf& b, size_t& cap, int64_t used, int64_t need, size_t esz) -> int {
        if ((size_t)need <= cap) return HQ_OK;
        size_t nc = std::max<size_t>(need, 2 * cap);
        DevBuf tmp;
        int r;
        if ((r = tmp.ensure(nc * esz))) return r;
        if (used) HQ_CHECK_HIP(hipMemcpyAsync(tmp.p, b.p, used * esz, hipMemcpyDeviceToDevice, s));
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        std::swap(tmp.p, b.p);
        std::swap(tmp.bytes, b.bytes);
        cap = nc;
        return HQ_OK;
    };
    while (nfr > 0) {
        if (nfr > ((int64_t)1 << 31))
            return fail(HQ_ERR_OVERFLOW, "FMM pair frontier of %lld pairs", (long long)nfr);
        const int nblk = grid_for(nfr, WB);
        // [category][block], one end slot per category, and the grand total
        if ((rc = st.cnt.ensure(((size_t)NACT * (nblk + 1) + 1) * 4)) || (rc = st.code.ensure(nfr * 4)))
            return rc;
        uint32_t* cnt = st.cnt.as<uint32_t>();
        k_fmm_walk_count<<<(nblk + WALK_U - 1) / WALK_U, WB, 0, s>>>(st.rec.as<FRec>(), w, cur->as<int2>(), nfr,
                                             st.code.as<uint32_t>(), cnt, nblk);
        // one scan over [category][block]: category q's slice starts at its base
        if ((rc = scan_excl_buf(cnt, (int64_t)NACT * (nblk + 1), t->scanpart, s))) return rc;
        // the totals come back in one copy, into host-only slots of the
        // handle's pinned status block
        uint32_t* edge = reinterpret_cast<uint32_t*>(t->hstatus + ST_HOST_FMM);
        k_fmm_edges<<<1, 64, 0, s>>>(cnt, nblk, st.edges.as<uint32_t>());
        HQ_CHECK_HIP(hipMemcpyAsync(edge, st.edges.p, (NACT + 1) * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, s));
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        uint32_t tot[NACT];
        for (int q = 0; q < NACT; ++q) tot[q] = edge[q + 1] - edge[q];
        if ((rc = nxt->ensure(std::max<int64_t>(tot[0], 1) * sizeof(int2))) ||
            (rc = grow(st.lat, lat_cap, nlat, nlat + tot[1] + 1, sizeof(int3))) ||
            (rc = grow(st.gen, gen_cap, ngen, ngen + tot[2] + 1, sizeof(int2))) ||
            (rc = grow(st.p2p, p2p_cap, np2p, np2p + tot[3] + 1, sizeof(int2))) ||
            (rc = grow(st.m2p, m2p_cap, nm2p, nm2p + tot[4] + 1, sizeof(int2))))
            return rc;
        k_fmm_walk_emit<<<(nblk + WALK_U - 1) / WALK_U, WB, 0, s>>>(st.rec.as<FRec>(), cur->as<int2>(), nfr,
                                            st.code.as<uint32_t>(), cnt, nblk, nxt->as<int2>(),
                                            st.lat.as<int3>(), nlat, st.gen.as<int2>(), ngen,
                                            st.p2p.as<int2>(), np2p, st.m2p.as<int2>(), nm2p);
        nlat += tot[1];
        ngen += tot[2];
        np2p += tot[3];
        nm2p += tot[4];
        std::swap(cur, nxt);
        nfr = tot[0];
    }
    st.nf = nf;
    t->list_cells = nlat + ngen;
    t->list_parts = np2p;
    // 4. M2L
    if (mfma) {
        const int GT = group_of(P);
        static_assert(P != 4 || group_of(4) == M2L4_GT, "order-4 groups");
        // a binned pair is one word: source | column << 24
        static_assert(group_of(P) <= 128, "columns of 7 bits");
        if (nf >= (1 << 24))
            return fail(HQ_ERR_OVERFLOW, "FMM matrix-core M2L of %lld cells (at most 2^24)", (long long)nf);
        const int ngroups = grid_for(nf, GT);
        const int64_t nbin = (int64_t)ngroups * tb.noff;
        if ((rc = st.blist.ensure(std::max<int64_t>(nbin, 1) * sizeof(int2))) ||
            (rc = st.gnb.ensure((ngroups + 1) * 4)) || (rc = st.ghist.ensure((ngroups + 1) * 4)) ||
            (rc = st.gpos.ensure((ngroups + 1) * 4)) ||
            (rc = st.gpair.ensure(std::max<int64_t>(nlat, 1) * sizeof(int2))) ||
            (rc = st.lpk.ensure(std::max<int64_t>(nlat, 1) * 4)))
            return rc;
        HQ_CHECK_HIP(hipMemsetAsync(st.ghist.p, 0, (ngroups + 1) * 4, s));
        const int win = std::min(BT_W, std::max(1, t->fmm_binw));  // groups counted in LDS
        if (nlat)
            k_fmm_lat_ghist<<<bin_grid(nlat), BT_T, 0, s>>>(st.lat.as<int3>(), nlat, GT, win,
                                                         st.ghist.as<uint32_t>());
        if ((rc = scan_excl_buf(st.ghist.as<uint32_t>(), ngroups, t->scanpart, s))) return rc;
        HQ_CHECK_HIP(hipMemcpyAsync(st.gpos.p, st.ghist.p, (ngroups + 1) * 4, hipMemcpyDeviceToDevice, s));
        if (nlat)
            k_fmm_lat_gscatter<<<bin_grid(nlat), BT_T, 0, s>>>(st.lat.as<int3>(), nlat, GT, win,
                                                            st.gpos.as<uint32_t>(), st.gpair.as<int2>());
        const int lsh = lsort_chunk(tb.noff) * (int)sizeof(uint32_t);
        HQ_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fmm_lat_lsort),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, lsh));
        k_fmm_lat_lsort<<<ngroups, LSORT_T, lsh, s>>>(st.gpair.as<int2>(), st.ghist.as<uint32_t>(),
                                                    tb.noff, st.blist.as<int2>(), st.gnb.as<uint32_t>(),
                                                    st.lpk.as<int>());
        HQ_CHECK_HIP(hipGetLastError());
        // enough workgroups to fill the device: split the bins of a group
        const int nsplit = std::max(1, std::min(64, 2048 / ngroups));
        if (nsplit > 1) HQ_CHECK_HIP(hipMemsetAsync(st.Lf.p, 0, nf * Kp * 4, s));
        HQ_CHECK_HIP(hipEventRecord(t->ev[6], s));
        if constexpr (P == 4)
            k_fmm_m2l4<P><<<ngroups * nsplit, 64 * M2L4_WAVES, 0, s>>>(
                nf, tb.noff, tb.noff0, tb.noff1, nsplit, st.blist.as<int2>(), st.gnb.as<uint32_t>(),
                st.ghist.as<uint32_t>(), st.lpk.as<int>(), tb.A.as<float>(),
                tb.cf.as<float>(), tb.ix.as<uint16_t>(), tb.deg.as<int8_t>(), st.Qf.as<float>(),
                st.Lf.as<float>());
        else
            k_fmm_m2l6<P><<<ngroups * nsplit, 64 * (Kp / 16), 0, s>>>(
                nf, tb.noff, nsplit, st.blist.as<int2>(), st.gnb.as<uint32_t>(), st.ghist.as<uint32_t>(),
                st.lpk.as<int>(), tb.T.as<float>(), st.Qf.as<float>(),
                st.Lf.as<float>());
        HQ_CHECK_HIP(hipEventRecord(t->ev[7], s));
        if (ngen)
            k_fmm_m2l_fp64<P><<<(int)std::min<int64_t>(ngen, 1 << 18), 64, 0, s>>>(
                st.gen.as<int2>(), ngen, nullptr, 0, v, tb.mi.as<int>(), tb.mi2.as<int>(),
                tb.rec2.as<int>(), tb.m2lidx.as<int>(), tb.m2lcoef.as<double>(), st.Qd.as<double>(),
                st.Ld.as<double>());
        k_fmm_merge<P><<<grid_for(nf * K, 256), 256, 0, s>>>(nf, st.Lf.as<float>(), st.Ld.as<double>());
    } else {
        HQ_CHECK_HIP(hipEventRecord(t->ev[6], s));
        if (ngen + nlat)
            k_fmm_m2l_fp64<P><<<(int)std::min<int64_t>(ngen + nlat, 1 << 18), 64, 0, s>>>(
                st.gen.as<int2>(), ngen, st.lat.as<int3>(), nlat, v, tb.mi.as<int>(),
                tb.mi2.as<int>(), tb.rec2.as<int>(), tb.m2lidx.as<int>(), tb.m2lcoef.as<double>(),
                st.Qd.as<double>(), st.Ld.as<double>());
        HQ_CHECK_HIP(hipEventRecord(t->ev[7], s));
    }
    // 5. L2L, top-down
    for (int d = 1; d <= MAX_DEPTH; ++d) {
        const int f0 = st.fls[d], f1 = st.fls[d + 1];
        if (f1 > f0)
            k_fmm_l2l<P><<<f1 - f0, 64, 0, s>>>(f0, f1, v, t->nparent.as<int>(), tb.l2l.as<double>(),
                                                st.Ld.as<double>());
    }
    // 6. near lists by target, then leaves
    if ((rc = st.phist.ensure((nf + 1) * 4)) || (rc = st.ppos.ensure((nf + 1) * 4)) ||
        (rc = st.plist.ensure(std::max<int64_t>(np2p, 1) * 4)))
        return rc;
    HQ_CHECK_HIP(hipMemsetAsync(st.phist.p, 0, (nf + 1) * 4, s));
    if (np2p)
        k_fmm_p2p_hist<<<capped(np2p), 256, 0, s>>>(st.p2p.as<int2>(), np2p,
                                                          st.phist.as<uint32_t>());
    if ((rc = scan_excl_buf(st.phist.as<uint32_t>(), nf, t->scanpart, s))) return rc;
    HQ_CHECK_HIP(hipMemcpyAsync(st.ppos.p, st.phist.p, (nf + 1) * 4, hipMemcpyDeviceToDevice, s));
    if (np2p) {
        k_fmm_p2p_scatter<<<capped(np2p), 256, 0, s>>>(st.p2p.as<int2>(), np2p,
                                                             st.ppos.as<uint32_t>(), st.plist.as<int>());
    }
    const size_t b8 = n * sizeof(double);
    if ((rc = t->fx.ensure(b8)) || (rc = t->fy.ensure(b8)) || (rc = t->fz.ensure(b8))) return rc;
    auto leaf = [&](auto kern) {
        kern<<<(int)nf, 64, 0, s>>>((int)nf, v, t->sx.as<double>(), t->sy.as<double>(),
                                    t->sz.as<double>(), t->sm.as<double>(),
                                    t->sbody.as<double4>(), tb.mi.as<int>(),
                                    st.Ld.as<double>(), st.phist.as<uint32_t>(), st.plist.as<int>(),
                                    G, eps2, t->fx.as<double>(), t->fy.as<double>(),
                                    t->fz.as<double>());
    };
    if (mfma) leaf(k_fmm_leaf<P, true>);
    else leaf(k_fmm_leaf<P, false>);
    if (nm2p)
        k_fmm_m2p<P><<<(int)std::min<int64_t>(nm2p, 1 << 16), 64, 0, s>>>(
            st.m2p.as<int2>(), nm2p, v, t->sx.as<double>(), t->sy.as<double>(), t->sz.as<double>(),
            t->sm.as<double>(), tb.mi2.as<int>(), tb.rec2.as<int>(), st.Qd.as<double>(), G,
            t->fx.as<double>(), t->fy.as<double>(), t->fz.as<double>());
    HQ_CHECK_HIP(hipEventRecord(t->ev[8], s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    HQ_CHECK_HIP(hipGetLastError());
    hipEventElapsedTime(&t->ms_walk, t->ev[5], t->ev[8]);   // the whole evaluation
    hipEventElapsedTime(&t->ms_lists, t->ev[6], t->ev[7]);  // the same-level M2L kernel
    hipEventElapsedTime(&t->ms_eval, t->ev[7], t->ev[8]);   // L2L, near field, L2P
    t->n_m2p = nlat + ngen;
    t->n_p2p = np2p;
    t->exec_cells = nlat;  // same-level (lattice) M2L pairs
    t->exec_parts = ngen;  // M2L pairs of two levels
    t->max_cells = nm2p;   // M2P pairs
    return HQ_OK;
}

int tree_forces_fmm(hq_tree* t, int order, double theta, double G, double eps2, bool mfma, int S) {
    if (!t->fmm) t->fmm.reset(new FmmState());
    if (order == 4) return fmm_run<4>(t, *t->fmm, theta, G, eps2, mfma, S);
    return fmm_run<6>(t, *t->fmm, theta, G, eps2, mfma, S);
}

}  // namespace hq
