// hq_walk.hip -- force walks over the level-order octree.
//
//   k_bh_walk        methods 10/11/12: bh_forces_over_p (octree_default.f90
//                    :895-1078) for one target per lane, FP64, stackless
//                    (first child = child_start[k], skip = next[k]); same
//                    acceptance test and expressions as the reference
//   k_lists<MODE>    one wavefront per group: the group's interaction lists
//                    (group_worker :1874-1952), 64 candidate nodes per step,
//                    lists compacted by ballot, opened children pushed on an
//                    LDS stack by a wavefront prefix sum.
//                      MODE_GROUP: the reference's groups (maximal nodes with
//                        <= group_size bodies), node-cube box, ancestor test
//                      MODE_BUNCH / MODE_BUNCH32: bunches of <= 64 particles made
//                        of adjacent sibling subtrees (so a bunch stays inside
//                        one parent cube), with their tight bounding box; the
//                        MODE_BUNCH32 lists hold FP32 records recentred on the
//                        bunch centre (FP64 subtraction, then rounding)
//   k_eval_fp64      group_worker's per-target sums (:1955-2090), FP64
//   k_eval_mixed     the same sums in FP32 on recentred coordinates with a
//                    lean (algebraically identical) multipole evaluation and
//                    block-compensated (Kahan) accumulation
//   k_walk_mixed     the mixed walk fused: one wavefront per bunch walks and
//                    evaluates 64-entry batches from LDS rings as they fill
//                    (record mode: writes the batches to a log instead);
//                    far accepted cells go to the bunch's far list.  A
//                    1152-entry LDS stack first, the 1536-entry one when a
//                    bunch overflows it
//   k_far<MP, P>     the far list's tier of order set P (5 outer, 6 inner):
//                    one local expansion per bunch, evaluated at its bodies
//   k_eval_batches,  split evaluation for few bunches: one wavefront per
//   k_combine        logged batch, then the partials summed in batch order
//                    (the fused kernel's arithmetic, bit for bit)
//   k_scatter        sorted forces -> caller order and layout
#include <math.h>

#include <algorithm>

#include "hq_tree.h"
#include "hq_local.h"

namespace hq {

struct NodeView {
    const double *mass, *cx, *cy, *cz, *quad, *oct, *ncx, *ncy, *ncz;
    const int *first, *count, *cstart, *next;
    const int8_t* type;
    const uint8_t* depth;
    const int* sys;  // system of each sorted particle; its tree's root is node sys
    int64_t stride;  // SoA stride of quad/oct
};

static NodeView node_view(const hq_tree* t) {
    NodeView v;
    v.mass = t->nmass.as<double>();
    v.cx = t->ncomx.as<double>();
    v.cy = t->ncomy.as<double>();
    v.cz = t->ncomz.as<double>();
    v.quad = t->multipole >= 4 ? t->nquad.as<double>() : nullptr;
    v.oct = t->multipole > 4 ? t->noct.as<double>() : nullptr;
    v.ncx = t->ncx.as<double>();
    v.ncy = t->ncy.as<double>();
    v.ncz = t->ncz.as<double>();
    v.first = t->nfirst.as<int>();
    v.count = t->ncount.as<int>();
    v.cstart = t->ncstart.as<int>();
    v.next = t->nnext.as<int>();
    v.type = t->ntype.as<int8_t>();
    v.depth = t->ndepth.as<uint8_t>();
    v.sys = t->ssys.as<int>();
    v.stride = t->cap;
    return v;
}

// ------------------------------------------------------------------ BH walk
// bh_forces_over_p (octree_default.f90:895-1078) for every target, one target
// per lane, the 64 targets of a wavefront (neighbours in tree order) walking
// together: the wavefront follows the union of its lanes' depth-first walks.
// At each node every lane still on its own walk decides as the reference does
// (skip an empty node or its own leaf; accept a leaf, or L^2 <= theta^2 d^2,
// or depth 25; else open).  If any lane opens, the wavefront descends; a lane
// that skipped or accepted the node waits until the wavefront reaches the
// node's skip pointer, where its own walk resumes.  Each lane therefore meets
// exactly the nodes of its own walk, in its order, and adds the same terms in
// the same order as a walk of its own; the node records are read once per
// wavefront (scalar loads) instead of once per lane.
__global__ __launch_bounds__(128) void k_bh_walk(NodeView nv, const double* __restrict__ g,
                                                 const double* __restrict__ sx,
                                                 const double* __restrict__ sy,
                                                 const double* __restrict__ sz,
                                                 const double* __restrict__ sm, int64_t n,
                                                 int multipole, double theta2, double G,
                                                 double eps2, double* __restrict__ fx,
                                                 double* __restrict__ fy, double* __restrict__ fz,
                                                 int64_t* __restrict__ status) {
#pragma clang fp contract(off)
    constexpr int ACTIVE = -2, NEVER = -3;
    const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool valid = i0 < n;
    const int64_t i = valid ? i0 : n - 1;
    const double pm = sm[i], px = sx[i], py = sy[i], pz = sz[i];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    long long np2p = 0, nm2p = 0;
    const int64_t S = nv.stride;
    const int sys = nv.sys[i];
    // the systems of the wavefront's targets, one tree walk each (a batch's
    // systems are contiguous in tree order, so a wavefront meets few)
    uint64_t todo = __ballot(valid);
    while (todo) {
        const int s = __builtin_amdgcn_readlane(sys, __ffsll((long long)todo) - 1);
        const bool mine = valid && sys == s;
        todo &= ~__ballot(mine);
        const double* __restrict__ gl2 = g + SG * s + S_L2;
        int wait = mine ? ACTIVE : NEVER;
        int k = s;
        while (k != END) {
            k = __builtin_amdgcn_readfirstlane(k);
            if (wait == k) wait = ACTIVE;
            const bool act = wait == ACTIVE;
            bool open = false;
            const double mass = nv.mass[k];
            const bool leaf = nv.type[k] == 1;
            if (act && !(mass == 0.0 || (leaf && nv.first[k] == (int)i))) {
                const double dx = nv.cx[k] - px;
                const double dy = nv.cy[k] - py;
                const double dz = nv.cz[k] - pz;
                const double dist2 = dx * dx + dy * dy + dz * dz;
                const int depth = nv.depth[k];
                if (leaf || gl2[depth] <= theta2 * dist2 || depth == MAX_DEPTH) {
                    if (leaf || multipole == 1) {
                        double rinv = 1.0 / sqrt(dist2 + eps2);
                        rinv = rinv * rinv * rinv;
                        const double f = G * pm * mass * rinv;
                        f0 = f0 + f * dx;
                        f1 = f1 + f * dy;
                        f2 = f2 + f * dz;
                        ++np2p;
                    } else {
                        const double invR = 1.0 / sqrt(dist2 + eps2);
                        const double invR2 = 1.0 / (dist2 + eps2);
                        const double invR3 = invR * invR2;
                        const double invR5 = invR3 * invR * invR;
                        const double invR7 = invR5 * invR * invR;
                        const double invR9 = invR7 * invR * invR;
                        const double Mx = dx * invR3 * mass;
                        const double My = dy * invR3 * mass;
                        const double Mz = dz * invR3 * mass;
                        const double mxi2 = nv.quad[0 * S + k], myi2 = nv.quad[1 * S + k],
                                     mzi2 = nv.quad[2 * S + k], mxyi = nv.quad[3 * S + k],
                                     mxzi = nv.quad[4 * S + k], myzi = nv.quad[5 * S + k];
                        const double dx2 = 15.0 * dx * dx * invR2 - 3.0;
                        const double dy2 = 15.0 * dy * dy * invR2 - 3.0;
                        const double dz2 = 15.0 * dz * dz * invR2 - 3.0;
                        const double dxyz = 15.0 * dx * dy * dz * invR2;
                        double Qx = (mxi2 * (dx2 - 6.0) + myi2 * dy2 + mzi2 * dz2) * 0.5 * dx;
                        double Qy = (myi2 * (dy2 - 6.0) + mzi2 * dz2 + mxi2 * dx2) * 0.5 * dy;
                        double Qz = (mzi2 * (dz2 - 6.0) + mxi2 * dx2 + myi2 * dy2) * 0.5 * dz;
                        Qx = Qx + mxyi * dy * dx2 + mxzi * dz * dx2 + myzi * dxyz;
                        Qy = Qy + myzi * dz * dy2 + mxyi * dx * dy2 + mxzi * dxyz;
                        Qz = Qz + mxzi * dx * dz2 + myzi * dy * dz2 + mxyi * dxyz;
                        if (multipole == 4) {
                            f0 = f0 + G * pm * (Mx + invR5 * Qx);
                            f1 = f1 + G * pm * (My + invR5 * Qy);
                            f2 = f2 + G * pm * (Mz + invR5 * Qz);
                        } else {
                            double O[14];
        #pragma unroll
                            for (int j = 0; j < 14; ++j) O[j] = nv.oct[j * S + k];
                            double Ox = -(1.5 * invR5 * O[11]);
                            double Oy = -(1.5 * invR5 * O[12]);
                            double Oz = -(1.5 * invR5 * O[13]);
                            double Oc = 7.5 * invR7 * (dx * O[11] + dy * O[12] + dz * O[13]);
                            Ox = Ox + dx * Oc;
                            Oy = Oy + dy * Oc;
                            Oz = Oz + dz * Oc;
                            const double ex2 = dx * dx, ey2 = dy * dy, ez2 = dz * dz;
                            Ox = Ox + 7.5 * invR7 *
                                          (ex2 * O[0] + ey2 * O[1] + ez2 * O[2] + 2.0 * dx * dy * O[3] +
                                           2.0 * dx * dz * O[6] + 2.0 * dy * dz * O[9]);
                            Oy = Oy + 7.5 * invR7 *
                                          (ex2 * O[3] + ey2 * O[4] + ez2 * O[5] + 2.0 * dx * dy * O[1] +
                                           2.0 * dx * dz * O[9] + 2.0 * dy * dz * O[7]);
                            Oz = Oz + 7.5 * invR7 *
                                          (ex2 * O[6] + ey2 * O[7] + ez2 * O[8] + 2.0 * dx * dy * O[9] +
                                           2.0 * dx * dz * O[2] + 2.0 * dy * dz * O[5]);
                            Oc = -(17.5 * invR9 *
                                   (ex2 * dx * O[0] + ey2 * dy * O[4] + ez2 * dz * O[8] +
                                    3.0 * ex2 * dy * O[3] + 3.0 * ex2 * dz * O[6] + 3.0 * ey2 * dx * O[1] +
                                    3.0 * ey2 * dz * O[7] + 3.0 * ez2 * dx * O[2] + 3.0 * ez2 * dy * O[5] +
                                    6.0 * dx * dy * dz * O[9]));
                            Ox = Ox + dx * Oc;
                            Oy = Oy + dy * Oc;
                            Oz = Oz + dz * Oc;
                            f0 = f0 + G * pm * (Mx + invR5 * Qx + Ox);
                            f1 = f1 + G * pm * (My + invR5 * Qy + Oy);
                            f2 = f2 + G * pm * (Mz + invR5 * Qz + Oz);
                        }
                        ++nm2p;
                    }
                } else {
                    open = true;
                }
            }
            const int nxt = nv.next[k];
            if (__ballot(open)) {
                if (act && !open) wait = nxt;  // its own walk goes on after this subtree
                k = nv.cstart[k];
            } else {
                k = nxt;
            }
        }
    }
    if (valid) {
        fx[i] = f0;
        fy[i] = f1;
        fz[i] = f2;
    }
    long long a = wave_sum(valid ? np2p : 0ll), b = wave_sum(valid ? nm2p : 0ll);
    if (__lane_id() == 0) {
        atomicAdd((unsigned long long*)&status[ST_P2P], (unsigned long long)a);
        atomicAdd((unsigned long long*)&status[ST_M2P], (unsigned long long)b);
    }
}

// ------------------------------------------------------------------ lists
enum { MODE_GROUP = 0, MODE_BUNCH = 1, MODE_BUNCH32 = 2 };
constexpr int SCAP = 4096;   // LDS walk stack (ints) per wavefront
constexpr int BUNCH = 64;
constexpr double LIST_GIB_MAX = 48.0;  // device memory allowed for interaction lists
constexpr int SCAP_FUSED = 1536;       // LDS walk stack of the fused mixed walk
constexpr int SCAP_FAST = 1152;        // its first try outside graphs: 8.5 KB of LDS, 18 wavefronts per CU
constexpr int SCAP_DEEP = 4096;        // its last try (small angles): 20 KB of LDS, 8 wavefronts per CU
constexpr int SCAP_LEVEL[3] = {SCAP_FAST, SCAP_FUSED, SCAP_DEEP};

struct ListOut {
    int* cidx;      // MODE_GROUP / MODE_BUNCH: node indices
    int* pidx;      // particle indices
    float4* c32;    // MODE_BUNCH32: {rc.x, rc.y, rc.z, node bits}
    float4* p32;    // {r.x, r.y, r.z, m}
    int64_t ccap, pcap;
    int* lstat;     // [2*g] = cells, [2*g+1] = particles
};

template <int MODE>
__global__ __launch_bounds__(64) void k_lists(NodeView nv, const double* __restrict__ g,
                                              const double* __restrict__ sx,
                                              const double* __restrict__ sy,
                                              const double* __restrict__ sz,
                                              const double* __restrict__ sm, int64_t n,
                                              const int* __restrict__ groups, int ngroups,
                                              double theta2, double mac_off, int lbucket,
                                              ListOut out, int64_t* __restrict__ status) {
#pragma clang fp contract(off)
    __shared__ int stk[SCAP];
    const int gi = xcd_swizzle(blockIdx.x, gridDim.x);
    if (gi >= ngroups) return;
    const int lane = __lane_id();
    const uint64_t lt = lanemask_lt();
    int gf, gl, gnode = -1;
    double bmin[3], bmax[3];
    if (MODE == MODE_GROUP) {
        gnode = groups[gi];
        gf = nv.first[gnode];
        gl = gf + nv.count[gnode] - 1;
        const double h = g[SG * nv.sys[gf] + S_HALFD + nv.depth[gnode]];
        bmin[0] = nv.ncx[gnode] - h; bmax[0] = nv.ncx[gnode] + h;
        bmin[1] = nv.ncy[gnode] - h; bmax[1] = nv.ncy[gnode] + h;
        bmin[2] = nv.ncz[gnode] - h; bmax[2] = nv.ncz[gnode] + h;
    } else {
        gf = groups[gi];
        gl = groups[gi + 1] - 1;
        int p = gf + lane;
        bool ok = p <= gl;
        double v[3] = {ok ? sx[p] : INFINITY, ok ? sy[p] : INFINITY, ok ? sz[p] : INFINITY};
        double w[3] = {ok ? sx[p] : -INFINITY, ok ? sy[p] : -INFINITY, ok ? sz[p] : -INFINITY};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                v[c] = fmin(v[c], __shfl_xor(v[c], o, 64));
                w[c] = fmax(w[c], __shfl_xor(w[c], o, 64));
            }
            bmin[c] = v[c];
            bmax[c] = w[c];
        }
    }
    const double bc[3] = {0.5 * (bmin[0] + bmax[0]), 0.5 * (bmin[1] + bmax[1]),
                          0.5 * (bmin[2] + bmax[2])};
    int top = 1, nc = 0, np = 0, maxtop = 1;
    bool overflow = false;
    const int sys = nv.sys[gf];
    const double* __restrict__ gl2 = g + SG * sys + S_L2;
    if (lane == 0) stk[0] = sys;
    __builtin_amdgcn_wave_barrier();
    while (top > 0) {
        const int take = top < 64 ? top : 64;
        const int base = top - take;
        const int k = lane < take ? stk[base + lane] : -1;
        top = base;
        int action = 0;  // 0 skip, 1 open, 2 particle, 3 cell, 4 bucket
        int bfirst = 0, bcount = 0;
        if (k >= 0) {
            const double mass = nv.mass[k];
            if (k != gnode && mass != 0.0) {
                const int f = nv.first[k];
                const int l = f + nv.count[k] - 1;
                const bool leaf = nv.type[k] == 1;
                bool anc;
                if (MODE == MODE_GROUP) anc = f <= gf && l >= gl;
                else anc = f <= gl && l >= gf;
                if (anc) {
                    action = leaf ? 0 : 1;  // a bunch's own leaves are its in-bunch sum
                } else if (leaf) {
                    action = 2;
                } else if (MODE != MODE_GROUP && l - f + 1 <= lbucket) {
                    action = 4;  // small cell: its bodies, exactly
                    bfirst = f;
                    bcount = l - f + 1;
                } else {
                    const double cc[3] = {nv.cx[k], nv.cy[k], nv.cz[k]};
                    double dist2 = 0.0;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        double dd = fmax(fmax(bmin[c] - cc[c], 0.0), cc[c] - bmax[c]);
                        dist2 = dist2 + dd * dd;
                    }
                    const int depth = nv.depth[k];
                    double lim = gl2[depth];
                    if (MODE != MODE_GROUP && mac_off > 0.0) {  // the offset test
                        const double side = ldexp(g[SG * sys + S_SIDE], -depth);
                        const double fe = 1.0 + mac_off * (double)mac_offset_q(
                            cc[0], cc[1], cc[2], nv.ncx[k], nv.ncy[k], nv.ncz[k], side);
                        lim = lim * fe * fe;
                    }
                    action = (lim <= theta2 * dist2 || depth >= MAX_DEPTH) ? 3 : 1;
                }
            }
        }
        const uint64_t mp = __ballot(action == 2);
        if (action == 2) {
            const int64_t pos = np + __popcll(mp & lt);
            if (pos < out.pcap) {
                const int p = nv.first[k];
                const int64_t o = (int64_t)gi * out.pcap + pos;
                if (MODE == MODE_BUNCH32)
                    out.p32[o] = make_float4((float)(sx[p] - bc[0]), (float)(sy[p] - bc[1]),
                                             (float)(sz[p] - bc[2]), (float)sm[p]);
                else
                    out.pidx[o] = p;
            }
        }
        np += __popcll(mp);
        if (MODE != MODE_GROUP) {
            // buckets: the wavefront copies one small cell's bodies at a time
            uint64_t mb = __ballot(action == 4);
            while (mb) {
                const int j = __ffsll((unsigned long long)mb) - 1;
                mb &= mb - 1;
                const int f = __shfl(bfirst, j, 64), c = __shfl(bcount, j, 64);
                if (lane < c) {
                    const int64_t pos = np + lane;
                    if (pos < out.pcap) {
                        const int p = f + lane;
                        const int64_t o = (int64_t)gi * out.pcap + pos;
                        if (MODE == MODE_BUNCH32)
                            out.p32[o] = make_float4((float)(sx[p] - bc[0]), (float)(sy[p] - bc[1]),
                                                     (float)(sz[p] - bc[2]), (float)sm[p]);
                        else
                            out.pidx[o] = p;
                    }
                }
                np += c;
            }
        }
        const uint64_t mc = __ballot(action == 3);
        if (action == 3) {
            const int64_t pos = nc + __popcll(mc & lt);
            if (pos < out.ccap) {
                const int64_t o = (int64_t)gi * out.ccap + pos;
                if (MODE == MODE_BUNCH32)
                    out.c32[o] = make_float4((float)(nv.cx[k] - bc[0]), (float)(nv.cy[k] - bc[1]),
                                             (float)(nv.cz[k] - bc[2]), __int_as_float(k));
                else
                    out.cidx[o] = k;
            }
        }
        nc += __popcll(mc);
        const int cs = action == 1 ? nv.cstart[k] : 0;
        const int nch = action == 1 ? nv.cstart[k + 1] - cs : 0;
        const int incl = wave_incl_scan(nch);
        const int tot = __shfl(incl, 63, 64);
        if (top + tot > SCAP) {
            overflow = true;
            break;
        }
        const int off = top + incl - nch;
        for (int c = 0; c < nch; ++c) stk[off + c] = cs + c;
        top += tot;
        maxtop = top > maxtop ? top : maxtop;
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) {
        out.lstat[2 * gi] = nc;
        out.lstat[2 * gi + 1] = np;
        if (overflow) atomicOr((unsigned long long*)&status[ST_FLAGS], 16ull);
        if (nc > out.ccap || np > out.pcap) atomicOr((unsigned long long*)&status[ST_FLAGS], 8ull);
        atomicAdd((unsigned long long*)&status[ST_LISTC], (unsigned long long)nc);
        atomicAdd((unsigned long long*)&status[ST_LISTP], (unsigned long long)np);
        const unsigned long long nb = (unsigned long long)(gl - gf + 1);
        atomicAdd((unsigned long long*)&status[ST_EXECC], (unsigned long long)nc * nb);
        atomicAdd((unsigned long long*)&status[ST_EXECP],
                  (unsigned long long)np * nb + nb * (nb - 1));
        atomicMax((long long*)&status[ST_NEEDC], (long long)nc);
        atomicMax((long long*)&status[ST_MAXSTK], (long long)maxtop);
        atomicMax((long long*)&status[ST_NEEDP], (long long)np);
    }
}

// ------------------------------------------------------------------ FP64 evaluation
// group_worker's per-target loops (:1955-2090) with the reference's
// expressions; one wavefront per group, one target per lane.
template <int MODE>
__global__ __launch_bounds__(64) void k_eval_fp64(NodeView nv, const double* __restrict__ sx,
                                                  const double* __restrict__ sy,
                                                  const double* __restrict__ sz,
                                                  const double* __restrict__ sm, int64_t n,
                                                  const int* __restrict__ groups, int ngroups,
                                                  int multipole, double G, double eps2,
                                                  ListOut lists, double* __restrict__ fx,
                                                  double* __restrict__ fy,
                                                  double* __restrict__ fz) {
#pragma clang fp contract(off)
    const int gi = xcd_swizzle(blockIdx.x, gridDim.x);
    if (gi >= ngroups) return;
    int gf, gl;
    if (MODE == MODE_GROUP) {
        const int gn = groups[gi];
        gf = nv.first[gn];
        gl = gf + nv.count[gn] - 1;
    } else {
        gf = groups[gi];
        gl = groups[gi + 1] - 1;
    }
    const int nc = lists.lstat[2 * gi], np = lists.lstat[2 * gi + 1];
    const int* __restrict__ cl = lists.cidx + (int64_t)gi * lists.ccap;
    const int* __restrict__ pl = lists.pidx + (int64_t)gi * lists.pcap;
    const int64_t S = nv.stride;
    for (int i = gf + (int)__lane_id(); i <= gl; i += 64) {
        const double xi = sx[i], yi = sy[i], zi = sz[i];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int j = gf; j <= gl; ++j) {
            if (j == i) continue;
            const double dx = sx[j] - xi, dy = sy[j] - yi, dz = sz[j] - zi;
            const double dist2 = dx * dx + dy * dy + dz * dz + eps2;
            const double invr = 1.0 / sqrt(dist2);
            const double invr3 = invr * invr * invr;
            const double f = sm[j] * invr3;
            ax = ax + f * dx;
            ay = ay + f * dy;
            az = az + f * dz;
        }
        for (int jj = 0; jj < np; ++jj) {
            const int j = pl[jj];
            const double dx = sx[j] - xi, dy = sy[j] - yi, dz = sz[j] - zi;
            const double dist2 = dx * dx + dy * dy + dz * dz + eps2;
            const double invr = 1.0 / sqrt(dist2);
            const double invr3 = invr * invr * invr;
            const double f = sm[j] * invr3;
            ax = ax + f * dx;
            ay = ay + f * dy;
            az = az + f * dz;
        }
        for (int jj = 0; jj < nc; ++jj) {
            const int k = cl[jj];
            const double dx = nv.cx[k] - xi, dy = nv.cy[k] - yi, dz = nv.cz[k] - zi;
            const double dist2 = dx * dx + dy * dy + dz * dz + eps2;
            const double invr = 1.0 / sqrt(dist2);
            const double invr3 = invr * invr * invr;
            const double M = nv.mass[k];
            if (multipole == 1) {
                const double f = M * invr3;
                ax = ax + f * dx;
                ay = ay + f * dy;
                az = az + f * dz;
                continue;
            }
            const double invr5 = invr3 * invr * invr;
            const double invr7 = invr5 * invr * invr;
            const double qxx = nv.quad[0 * S + k], qyy = nv.quad[1 * S + k],
                         qzz = nv.quad[2 * S + k], qxy = nv.quad[3 * S + k],
                         qxz = nv.quad[4 * S + k], qyz = nv.quad[5 * S + k];
            const double qx = qxx * dx + qxy * dy + qxz * dz;
            const double qy = qxy * dx + qyy * dy + qyz * dz;
            const double qz = qxz * dx + qyz * dy + qzz * dz;
            const double dqd = dx * qx + dy * qy + dz * qz;
            const double trq = qxx + qyy + qzz;
            const double s = M * invr3 - 1.5 * trq * invr5 + 7.5 * dqd * invr7;
            ax = ax + s * dx - 3.0 * qx * invr5;
            ay = ay + s * dy - 3.0 * qy * invr5;
            az = az + s * dz - 3.0 * qz * invr5;
            if (multipole == 8) {
                double c[14];
#pragma unroll
                for (int j = 0; j < 14; ++j) c[j] = nv.oct[j * S + k];
                const double invr9 = invr7 * invr * invr;
                double oc = 7.5 * invr7 * (dx * c[11] + dy * c[12] + dz * c[13]);
                double ox = -(1.5 * invr5 * c[11]) + dx * oc;
                double oy = -(1.5 * invr5 * c[12]) + dy * oc;
                double oz = -(1.5 * invr5 * c[13]) + dz * oc;
                ox = ox + 7.5 * invr7 * (dx * dx * c[0] + dy * dy * c[1] + dz * dz * c[2]);
                ox = ox + 15.0 * invr7 * (dx * dy * c[3] + dx * dz * c[6] + dy * dz * c[9]);
                oy = oy + 7.5 * invr7 * (dx * dx * c[3] + dy * dy * c[4] + dz * dz * c[5]);
                oy = oy + 15.0 * invr7 * (dx * dy * c[1] + dx * dz * c[9] + dy * dz * c[7]);
                oz = oz + 7.5 * invr7 * (dx * dx * c[6] + dy * dy * c[7] + dz * dz * c[8]);
                oz = oz + 15.0 * invr7 * (dx * dy * c[9] + dx * dz * c[2] + dy * dz * c[5]);
                oc = dx * dx * dx * c[0] + dy * dy * dy * c[4] + dz * dz * dz * c[8];
                oc = oc + 3.0 * (dx * dx * dy * c[3] + dx * dx * dz * c[6] + dx * dy * dy * c[1]);
                oc = oc + 3.0 * (dy * dy * dz * c[7] + dx * dz * dz * c[2] + dy * dz * dz * c[5]);
                oc = oc + 6.0 * dx * dy * dz * c[9];
                oc = -(17.5 * invr9 * oc);
                ox = ox + dx * oc;
                oy = oy + dy * oc;
                oz = oz + dz * oc;
                ax = ax + ox;
                ay = ay + oy;
                az = az + oz;
            }
        }
        fx[i] = G * sm[i] * ax;
        fy[i] = G * sm[i] * ay;
        fz[i] = G * sm[i] * az;
    }
}

// ------------------------------------------------------------------ mixed evaluation
// The FP32 moment records and the packed traversal records (REC, WalkNode,
// store_walk_records in hq_tree.h), from the node store, before each mixed walk.
__global__ __launch_bounds__(256) void k_mom32(int64_t nn, int multipole, NodeView nv,
                                               const double* __restrict__ g,
                                               float* __restrict__ rec,
                                               WalkNode* __restrict__ wn,
                                               const int64_t* __restrict__ dnn) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (dnn) nn = *dnn;  // graph mode: the count is on the device
    if (k >= nn) return;
    const int64_t S = nv.stride;
    const int type = (int)(uint8_t)nv.type[k];
    double Q[6] = {0, 0, 0, 0, 0, 0}, O[14] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (multipole >= 4 && type != 1)
        for (int j = 0; j < 6; ++j) Q[j] = nv.quad[j * S + k];
    if (multipole > 4 && type != 1)
        for (int j = 0; j < 14; ++j) O[j] = nv.oct[j * S + k];
    const int cs = nv.cstart[k];
    const int depth = (int)nv.depth[k];
    const double side = ldexp(g[SG * nv.sys[nv.first[k]] + S_SIDE], -depth);
    const int macq = mac_offset_q(nv.cx[k], nv.cy[k], nv.cz[k], nv.ncx[k], nv.ncy[k], nv.ncz[k], side);
    store_walk_records(k, multipole, nv.mass[k], nv.cx[k], nv.cy[k], nv.cz[k], nv.first[k],
                       nv.count[k], cs, nv.cstart[k + 1] - cs, type, depth,
                       macq, Q, O, rec, wn);
}

struct Kahan3 {
    float s[3] = {0.f, 0.f, 0.f}, c[3] = {0.f, 0.f, 0.f};
    __device__ __forceinline__ void add(const float p[3]) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float y = p[j] - c[j];
            float t = s[j] + y;
            c[j] = (t - s[j]) - y;
            s[j] = t;
        }
    }
};

// Force of one cell on one target, d = com - target, s^2 = |d|^2 + eps^2,
// with r_k = s^-k.  The reference's expansion (octree_default.f90 :2032-2082)
//   acc = d [M r3 - 1.5 trQ r5 + 7.5 r7 (d.Qd + d.T) - 17.5 r9 d.(O:dd)]
//         + r5 (-3 Qd - 1.5 T) + 7.5 r7 O:dd
// is grouped, with the record's scaled moments (w = Q~d + T~, v = O~:dd), as
//   acc = B d + r5 (w + v / s^2),
//   B   = r3 (M + (tq - 2.5 (d.(w + T~) + (14/15) d.v / s^2) / s^2) / s^2)
// which is the same polynomial in fewer operations (68 VALU for the octupole
// term, 32 for the quadrupole).
template <int MP>
__device__ __forceinline__ void m2p32(const float4 e, const float* __restrict__ r, float tx,
                                      float ty, float tz, float eps2, float acc[3]) {
    const float dx = e.x - tx, dy = e.y - ty, dz = e.z - tz;
    const float s2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
    const float ri = __frsqrt_rn(s2);
    const float ri2 = ri * ri;
    const float r3 = ri * ri2;
    if (MP == 1) {
        const float cd = r[0] * r3;  // M / s^3
        acc[0] = fmaf(cd, dx, acc[0]);
        acc[1] = fmaf(cd, dy, acc[1]);
        acc[2] = fmaf(cd, dz, acc[2]);
        return;
    }
    const float r5 = r3 * ri2;
    // w = Q~ d (+ T~)
    const float t0 = MP == 8 ? r[8] : 0.0f, t1 = MP == 8 ? r[9] : 0.0f, t2 = MP == 8 ? r[10] : 0.0f;
    float wx = fmaf(r[1], dx, fmaf(r[4], dy, fmaf(r[5], dz, t0)));
    float wy = fmaf(r[4], dx, fmaf(r[2], dy, fmaf(r[6], dz, t1)));
    float wz = fmaf(r[5], dx, fmaf(r[6], dy, fmaf(r[3], dz, t2)));
    float inner;  // the bracket that multiplies -2.5 / s^4
    if (MP == 8) {
        const float a = fmaf(dx, wx + t0, fmaf(dy, wy + t1, dz * (wz + t2)));
        const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
        const float dx2 = dx + dx;
        const float xy2 = dx2 * dy, xz2 = dx2 * dz, yz = dy * dz;
        const float vx = fmaf(xx, r[11], fmaf(yy, r[12], fmaf(zz, r[13], fmaf(xy2, r[14], fmaf(xz2, r[17], yz * r[21])))));
        const float vy = fmaf(xx, r[14], fmaf(yy, r[15], fmaf(zz, r[16], fmaf(xy2, r[12], fmaf(xz2, r[20], yz * r[22])))));
        const float vz = fmaf(xx, r[17], fmaf(yy, r[18], fmaf(zz, r[19], fmaf(xy2, r[20], fmaf(xz2, r[13], yz * r[23])))));
        const float b = fmaf(dx, vx, fmaf(dy, vy, dz * vz));
        inner = fmaf((14.0f / 15.0f) * ri2, b, a);
        wx = fmaf(ri2, vx, wx);
        wy = fmaf(ri2, vy, wy);
        wz = fmaf(ri2, vz, wz);
    } else {
        inner = fmaf(dx, wx, fmaf(dy, wy, dz * wz));
    }
    const float B = r3 * fmaf(ri2, fmaf(-2.5f * ri2, inner, r[7]), r[0]);
    acc[0] = fmaf(B, dx, fmaf(r5, wx, acc[0]));
    acc[1] = fmaf(B, dy, fmaf(r5, wy, acc[1]));
    acc[2] = fmaf(B, dz, fmaf(r5, wz, acc[2]));
}

__device__ __forceinline__ void p2p32(const float4 e, float tx, float ty, float tz, float eps2,
                                      float acc[3]) {
    const float dx = e.x - tx, dy = e.y - ty, dz = e.z - tz;
    const float s2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
    const float ri = __frsqrt_rn(s2);
    const float f = e.w * ri * ri * ri;
    acc[0] = fmaf(f, dx, acc[0]);
    acc[1] = fmaf(f, dy, acc[1]);
    acc[2] = fmaf(f, dz, acc[2]);
}

// Sum of v over the wavefront, returned to every lane.  Row shifts within the
// four 16-lane rows, then row broadcasts, in a fixed order: the rounding does
// not depend on timing, so every kernel that calls it on the same values gets
// the same bits.  All 64 lanes must be active.
__device__ __forceinline__ float wave_sum_dpp(float v) {
    int x = __float_as_int(v);
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false)));
    x = __float_as_int(__int_as_float(x) + __int_as_float(__builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false)));
    return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}

// Inclusive prefix sum of v over the wavefront (lane l gets the sum of lanes
// 0..l), by the same row shifts and broadcasts.  All 64 lanes must be active.
__device__ __forceinline__ int wave_incl_scan_dpp(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
    return x;
}

// Radius of a bunch for its far-field expansion: the largest distance of its
// bodies (lanes < nb) from the bunch centre, in the FP32 offsets the walk
// uses, at least 1e-30.
__device__ __forceinline__ float bunch_radius(float tx, float ty, float tz, int lane, int nb) {
    float r2 = lane < nb ? fmaf(tx, tx, fmaf(ty, ty, tz * tz)) : 0.0f;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r2 = fmaxf(r2, __shfl_xor(r2, o, 64));
    return fmaxf(__fsqrt_rn(r2), 1e-30f);
}

// Local expansion of this lane's far cell (node >= 0; -1 adds nothing) about
// the bunch centre bc, added to the lane's coefficients L (scaled by h^|n|).
template <int MP, int P>
__device__ __forceinline__ void far_m2l(int node, const double* __restrict__ cx,
                                        const double* __restrict__ cy,
                                        const double* __restrict__ cz,
                                        const float* __restrict__ mom, const double bc[3], float h,
                                        float eps2, float* __restrict__ L) {
#pragma clang fp contract(off)
    float r[REC];
#pragma unroll
    for (int j = 0; j < REC; ++j) r[j] = 0.0f;
    float Rx = 1.0f, Ry = 1.0f, Rz = 1.0f;  // idle lane: no mass, finite geometry
    if (node >= 0) {
        Rx = (float)(bc[0] - cx[node]);
        Ry = (float)(bc[1] - cy[node]);
        Rz = (float)(bc[2] - cz[node]);
        const float4* __restrict__ rp = reinterpret_cast<const float4*>(mom + (int64_t)node * REC);
        constexpr int NQ = MP == 1 ? 1 : (MP == 4 ? 2 : REC / 4);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const float4 v4 = rp[q];
            r[4 * q] = v4.x;
            r[4 * q + 1] = v4.y;
            r[4 * q + 2] = v4.z;
            r[4 * q + 3] = v4.w;
        }
    }
    if constexpr (P == 5) local_m2l_p5<MP>(Rx, Ry, Rz, eps2, h, r, L);
    else local_m2l_p6<MP>(Rx, Ry, Rz, eps2, h, r, L);
}

constexpr int KBLK = 32;  // interactions per compensated partial sum

template <int MP>
__global__ __launch_bounds__(64) void k_eval_mixed(const double* __restrict__ sx,
                                                   const double* __restrict__ sy,
                                                   const double* __restrict__ sz,
                                                   const double* __restrict__ sm, int64_t n,
                                                   const int* __restrict__ bstart,
                                                   int ngroups, double G, float eps2,
                                                   const float4* __restrict__ c32,
                                                   const float4* __restrict__ p32,
                                                   const int* __restrict__ lstat, int64_t ccap,
                                                   int64_t pcap, const float* __restrict__ mom,
                                                   double* __restrict__ fx,
                                                   double* __restrict__ fy,
                                                   double* __restrict__ fz) {
    __shared__ float4 own[BUNCH];
    const int gi = xcd_swizzle(blockIdx.x, gridDim.x);
    if (gi >= ngroups) return;
    const int lane = __lane_id();
    const int gf = bstart[gi];
    const int gl = bstart[gi + 1] - 1;
    const int nb = gl - gf + 1;
    // bunch centre, as the list builder computed it
    const int i = gf + (lane < nb ? lane : 0);
    double v[3] = {sx[i], sy[i], sz[i]}, w[3] = {v[0], v[1], v[2]};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            v[c] = fmin(v[c], __shfl_xor(v[c], o, 64));
            w[c] = fmax(w[c], __shfl_xor(w[c], o, 64));
        }
    double bc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) bc[c] = 0.5 * (v[c] + w[c]);
    const float tx = (float)(sx[i] - bc[0]), ty = (float)(sy[i] - bc[1]), tz = (float)(sz[i] - bc[2]);
    own[lane] = make_float4(tx, ty, tz, lane < nb ? (float)sm[i] : 0.0f);
    __syncthreads();
    Kahan3 sum;
    float acc[3] = {0.f, 0.f, 0.f};
    // in-bunch particles (self skipped)
    for (int j = 0; j < nb; ++j) {
        float4 e = own[j];
        if (j == lane) e.w = 0.0f, e.x = tx + 1.0f;  // zero mass, non-zero distance
        p2p32(e, tx, ty, tz, eps2, acc);
    }
    sum.add(acc);
    const int nc = lstat[2 * gi], np = lstat[2 * gi + 1];
    // particle list: entries prefetched one ahead
    const float4* __restrict__ pl = p32 + (int64_t)gi * pcap;
    if (np > 0) {
        float4 e0 = pl[0];
        for (int j0 = 0; j0 < np; j0 += KBLK) {
            acc[0] = acc[1] = acc[2] = 0.f;
            const int j1 = std::min(np, j0 + KBLK);
            for (int j = j0; j < j1; ++j) {
                const float4 e1 = pl[j + 1 < np ? j + 1 : j];
                p2p32(e0, tx, ty, tz, eps2, acc);
                e0 = e1;
            }
            sum.add(acc);
        }
    }
    // cell list: entries prefetched two ahead, node records one ahead, so the
    // scalar loads of the next cells overlap this cell's arithmetic
    const float4* __restrict__ cl = c32 + (int64_t)gi * ccap;
    if (nc > 0) {
        float4 e0 = cl[0];
        float4 e1 = cl[nc > 1 ? 1 : 0];
        float r0[REC];
        {
            const float4* rp = reinterpret_cast<const float4*>(mom + (int64_t)__float_as_int(e0.w) * REC);
#pragma unroll
            for (int q = 0; q < REC / 4; ++q) {
                const float4 v4 = rp[q];
                r0[4 * q] = v4.x; r0[4 * q + 1] = v4.y; r0[4 * q + 2] = v4.z; r0[4 * q + 3] = v4.w;
            }
        }
        for (int j0 = 0; j0 < nc; j0 += KBLK) {
            acc[0] = acc[1] = acc[2] = 0.f;
            const int j1 = std::min(nc, j0 + KBLK);
            for (int j = j0; j < j1; ++j) {
                const float4 e2 = cl[j + 2 < nc ? j + 2 : nc - 1];
                float r1[REC];
                const float4* rp = reinterpret_cast<const float4*>(mom + (int64_t)__float_as_int(e1.w) * REC);
#pragma unroll
                for (int q = 0; q < REC / 4; ++q) {
                    const float4 v4 = rp[q];
                    r1[4 * q] = v4.x; r1[4 * q + 1] = v4.y; r1[4 * q + 2] = v4.z; r1[4 * q + 3] = v4.w;
                }
                m2p32<MP>(e0, r0, tx, ty, tz, eps2, acc);
                e0 = e1;
                e1 = e2;
#pragma unroll
                for (int q = 0; q < REC; ++q) r0[q] = r1[q];
            }
            sum.add(acc);
        }
    }
    if (lane < nb) {
        const double gm = G * sm[i];
        fx[i] = gm * ((double)sum.s[0] - (double)sum.c[0]);
        fy[i] = gm * ((double)sum.s[1] - (double)sum.c[1]);
        fz[i] = gm * ((double)sum.s[2] - (double)sum.c[2]);
    }
}

// ------------------------------------------------------------------ fused mixed walk
// One wavefront per bunch walks the tree (as k_lists<MODE_BUNCH32>) and
// evaluates accepted cells and particles in chunks of 64 as they arrive,
// from two LDS ring buffers: no interaction lists in global memory, and the
// latency-bound walk of some wavefronts overlaps the arithmetic of others.
constexpr int RING = 128;

template <int MP, int MASK = RING - 1>
__device__ __forceinline__ void load_rec(const float4* __restrict__ ring, int pos,
                                         const float* __restrict__ mom, float4& e, float r[REC]) {
    e = ring[pos & MASK];
    const int node = __builtin_amdgcn_readfirstlane(__float_as_int(e.w));
    const float4* __restrict__ rp = reinterpret_cast<const float4*>(mom + (int64_t)node * REC);
    constexpr int NQ = MP == 1 ? 1 : (MP == 4 ? 2 : REC / 4);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const float4 v4 = rp[q];
        r[4 * q] = v4.x;
        r[4 * q + 1] = v4.y;
        r[4 * q + 2] = v4.z;
        r[4 * q + 3] = v4.w;
    }
}

// One batch of <= 64 cells (entries head .. head+cnt-1 of a ring, or of a
// batch log with MASK = -1) into acc, which starts at zero.  Ring batches
// start at multiples of 64 (heads advance by whole batches), so a batch never
// wraps: it is read from one base with immediate offsets.
template <int MP, int MASK = RING - 1>
__device__ __forceinline__ void eval_cells_acc(const float4* __restrict__ ring, int head, int cnt,
                                               const float* __restrict__ mom, float tx, float ty,
                                               float tz, float eps2, float acc[3]) {
    const float4* __restrict__ b = ring + (head & MASK);
    float4 eA, eB;
    float rA[REC], rB[REC];
    load_rec<MP, -1>(b, 0, mom, eA, rA);
    for (int j = 0; j < cnt; j += 2) {
        const bool two = j + 1 < cnt;
        if (two) load_rec<MP, -1>(b, j + 1, mom, eB, rB);
        m2p32<MP>(eA, rA, tx, ty, tz, eps2, acc);
        if (two) {
            if (j + 2 < cnt) load_rec<MP, -1>(b, j + 2, mom, eA, rA);
            m2p32<MP>(eB, rB, tx, ty, tz, eps2, acc);
        }
    }
}

template <int MASK = RING - 1>
__device__ __forceinline__ void eval_parts_acc(const float4* __restrict__ ring, int head, int cnt,
                                               float tx, float ty, float tz, float eps2,
                                               float acc[3]) {
    const float4* __restrict__ b = ring + (head & MASK);
    for (int j = 0; j < cnt; ++j) p2p32(b[j], tx, ty, tz, eps2, acc);
}

template <int MP>
__device__ __forceinline__ void eval_cells(const float4* __restrict__ ring, int head, int cnt,
                                           const float* __restrict__ mom, float tx, float ty,
                                           float tz, float eps2, Kahan3& sum) {
    float acc[3] = {0.f, 0.f, 0.f};
    eval_cells_acc<MP>(ring, head, cnt, mom, tx, ty, tz, eps2, acc);
    sum.add(acc);
}

__device__ __forceinline__ void eval_parts(const float4* __restrict__ ring, int head, int cnt,
                                           float tx, float ty, float tz, float eps2, Kahan3& sum) {
    float acc[3] = {0.f, 0.f, 0.f};
    eval_parts_acc(ring, head, cnt, tx, ty, tz, eps2, acc);
    sum.add(acc);
}

// Split evaluation of the mixed walk (for few bunches): the walk records each
// bunch's batches, in the order the fused kernel evaluates them, into a log;
// one wavefront per batch computes its partial sum; one wavefront per bunch
// adds the partials in order with the same compensated sum.  Every float
// operation is the fused kernel's, in its order, so the forces are the same
// bit for bit.
struct SplitOut {
    float4* log;    // per bunch, lcap entries: the batches' cells and bodies in order
    int2* desc;     // per bunch, bcap batches: x first entry, y count | kind << 16 (0 cells, 1 bodies)
    int* nbat;      // batches per bunch (-1: its log overflowed)
    int64_t lcap;
    int bcap;
};

__device__ __forceinline__ void record_batch(const SplitOut& so, int gi, int kind,
                                             const float4* __restrict__ ring, int head, int cnt,
                                             int& lpos, int& nbat, bool& over, int lane) {
    if (over || nbat >= so.bcap || lpos + cnt > so.lcap) {
        over = true;
        return;
    }
    float4* __restrict__ dst = so.log + (int64_t)gi * so.lcap + lpos;
    for (int j = lane; j < cnt; j += 64) dst[j] = ring[(head + j) & (RING - 1)];
    if (lane == 0) so.desc[(int64_t)gi * so.bcap + nbat] = make_int2(lpos, cnt | (kind << 16));
    lpos += cnt;
    ++nbat;
}

template <int MP, int SC, bool RECORD = false>
__global__ __launch_bounds__(64) void k_walk_mixed(NodeView nv, const double* __restrict__ g,
                                                   const double* __restrict__ sx,
                                                   const double* __restrict__ sy,
                                                   const double* __restrict__ sz,
                                                   const double* __restrict__ sm, int64_t n,
                                                   const int* __restrict__ bstart, int nbunch,
                                                   double theta2, double mac_off, double G,
                                                   float eps2, int lbucket,
                                                   float far2, float far2i, int* __restrict__ flist,
                                                   int* __restrict__ fnum, int fcap,
                                                   const WalkNode* __restrict__ wn,
                                                   const double4* __restrict__ sb,
                                                   const float* __restrict__ mom,
                                                   double* __restrict__ fx, double* __restrict__ fy,
                                                   double* __restrict__ fz,
                                                   int64_t* __restrict__ status,
                                                   const int64_t* __restrict__ dnbun,
                                                   SplitOut so) {
    __shared__ int stk[SC];
    __shared__ float4 cring[RING], pring[RING];  // pring also holds the bunch's own bodies at the end
    // graph mode: the grid is an upper bound and the count is on the device;
    // the swizzle spreads the blocks that have a bunch over all XCDs
    if (dnbun) nbunch = (int)*dnbun;
    if ((int)blockIdx.x >= nbunch) return;
    const int gi = xcd_swizzle(blockIdx.x, nbunch);
    const int lane = __lane_id();
    const uint64_t lt = lanemask_lt();
    const int gf = bstart[gi], gl = bstart[gi + 1] - 1;
    const int nb = gl - gf + 1;
    const int i = gf + (lane < nb ? lane : 0);
    double bmin[3], bmax[3];
    {
        double v[3] = {sx[i], sy[i], sz[i]}, w[3] = {v[0], v[1], v[2]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                v[c] = fmin(v[c], __shfl_xor(v[c], o, 64));
                w[c] = fmax(w[c], __shfl_xor(w[c], o, 64));
            }
            bmin[c] = v[c];
            bmax[c] = w[c];
        }
    }
    const double bc[3] = {0.5 * (bmin[0] + bmax[0]), 0.5 * (bmin[1] + bmax[1]),
                          0.5 * (bmin[2] + bmax[2])};
    const float tx = (float)(sx[i] - bc[0]), ty = (float)(sy[i] - bc[1]), tz = (float)(sz[i] - bc[2]);
    // far field: accepted cells whose centre of mass is at least sqrt(far2i)
    // bunch radii from the bunch centre go to the bunch's far list for k_far,
    // as the node index at sqrt(far2) radii or more (outer tier) and as its
    // complement below (inner tier); at most fcap of them, any more are
    // evaluated here like the others
    const float hb = bunch_radius(tx, ty, tz, lane, nb);
    const double hb2 = (double)hb * (double)hb;
    const double farR2 = far2 > 0.0f ? (double)far2 * hb2 : INFINITY;
    const double farR2i = far2i > 0.0f ? (double)far2i * hb2 : farR2;
    int* __restrict__ fl = flist + (int64_t)gi * fcap;
    Kahan3 sum;
    int top = 1, maxtop = 1, chead = 0, ccnt = 0, phead = 0, pcnt = 0, nfar = 0;
    long long ntot_c = 0, ntot_p = 0;
    bool overflow = false;
    int lpos = 0, nbat = 0;  // record mode: log position and batches
    bool rover = false;      // record mode: the log overflowed
    auto batch_cells = [&](int head, int cnt) {
        if constexpr (RECORD) record_batch(so, gi, 0, cring, head, cnt, lpos, nbat, rover, lane);
        else eval_cells<MP>(cring, head, cnt, mom, tx, ty, tz, eps2, sum);
    };
    auto batch_parts = [&](int head, int cnt) {
        if constexpr (RECORD) record_batch(so, gi, 1, pring, head, cnt, lpos, nbat, rover, lane);
        else eval_parts(pring, head, cnt, tx, ty, tz, eps2, sum);
    };
    const int sys = nv.sys[gf];
    // squared side of a cell at depth d: the root's scaled by 4^-d, the same
    // double as the build's per-depth table (powers of two scale exactly)
    const double l2root = g[SG * sys + S_L2];
    if (lane == 0) stk[0] = sys;
    __builtin_amdgcn_wave_barrier();
    while (top > 0) {
        const int take = top < 64 ? top : 64;
        const int base = top - take;
        const int k = lane < take ? stk[base + lane] : -1;
        top = base;
        int action = 0;  // 0 skip, 1 open, 2 particle, 3 cell, 4 bucket, 5 far
        int bfirst = 0, bcount = 0, fent = 0;
        double cc[3] = {0.0, 0.0, 0.0};
        float kmass = 0.0f;
        int kcs = 0, knch = 0;
        if (k >= 0) {
            const double2 w0 = reinterpret_cast<const double2*>(wn + k)[0];
            const double2 w2 = reinterpret_cast<const double2*>(wn + k)[1];
            const int4 w1 = reinterpret_cast<const int4*>(wn + k)[2];
            cc[0] = w0.x;
            cc[1] = w0.y;
            cc[2] = w2.x;
            kcs = w1.z;
            knch = w1.w & 15;
            kmass = (float)w2.y;
            if (w2.y != 0.0) {
                const int f = w1.x;
                const int l = f + w1.y - 1;
                const bool leaf = ((w1.w >> 4) & 0xff) == 1;
                if (f <= gl && l >= gf) {
                    action = leaf ? 0 : 1;
                } else if (leaf) {
                    action = 2;
                } else {
                    double dist2 = 0.0;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double dd = fmax(fmax(bmin[c] - cc[c], 0.0), cc[c] - bmax[c]);
                        dist2 = fma(dd, dd, dist2);
                    }
                    const int depth = (w1.w >> 12) & 31;
                    double lim = ldexp(l2root, -2 * depth);
                    if (mac_off > 0.0) {  // the offset test: side + theta |com - centre|
                        const double fe = 1.0 + mac_off * (double)((unsigned)w1.w >> 17);
                        lim = lim * fe * fe;
                    }
                    action = (lim <= theta2 * dist2 || depth >= MAX_DEPTH) ? 3 : 1;
                    if (action == 3) {
                        double R2 = 0.0;
#pragma unroll
                        for (int c = 0; c < 3; ++c) R2 = fma(cc[c] - bc[c], cc[c] - bc[c], R2);
                        if (R2 >= farR2i) {  // far cell, if its list has room
                            action = 5;
                            fent = R2 >= farR2 ? k : ~k;
                        }
                    }
                    if (action != 5 && l - f + 1 <= lbucket) {
                        action = 4;  // small cell, not far: its bodies, exactly
                        bfirst = f;
                        bcount = l - f + 1;
                    }
                }
            }
        }
        // a one-body leaf: its record holds the body's position and mass
        const uint64_t mp = __ballot(action == 2);
        if (action == 2)
            pring[(phead + pcnt + __popcll(mp & lt)) & (RING - 1)] =
                make_float4((float)(cc[0] - bc[0]), (float)(cc[1] - bc[1]), (float)(cc[2] - bc[2]),
                            kmass);
        pcnt += __popcll(mp);
        if (__ballot(action == 4)) {
            // buckets: the small cells' bodies into the particle ring, the
            // cells in lane order.  Each round drains whole batches, then
            // every lane whose cell fits entirely in the free space copies its
            // bodies at its offset; the first cell left always fits (at most
            // 64 bodies, at least 65 free), so every round makes progress
            __builtin_amdgcn_wave_barrier();
            const int bn = action == 4 ? bcount : 0;
            const int bincl = wave_incl_scan_dpp(bn);
            const int bexcl = bincl - bn;
            const int btot = __builtin_amdgcn_readlane(bincl, 63);
            int done = 0;
            while (done < btot) {
                while (pcnt >= 64) {
                    batch_parts(phead, 64);
                    phead += 64;
                    pcnt -= 64;
                    ntot_p += 64;
                }
                const bool mine = bn > 0 && bexcl >= done && bincl <= done + (RING - pcnt);
                const uint64_t mm = __ballot(mine);
                const int end = __builtin_amdgcn_readlane(bincl, 63 - __clzll(mm));
                if (mine) {
                    const int at = phead + pcnt + bexcl - done;
                    for (int q = 0; q < bn; ++q) {
                        const double4 bp = sb[bfirst + q];
                        pring[(at + q) & (RING - 1)] =
                            make_float4((float)(bp.x - bc[0]), (float)(bp.y - bc[1]),
                                        (float)(bp.z - bc[2]), (float)bp.w);
                    }
                }
                pcnt += end - done;
                done = end;
                __builtin_amdgcn_wave_barrier();
            }
        }
        {
            const uint64_t mfar = __ballot(action == 5);
            if (action == 5) {
                const int slot = nfar + __popcll(mfar & lt);
                if (slot < fcap) fl[slot] = fent;
                else action = 3;
            }
            nfar = min(nfar + __popcll(mfar), fcap);
        }
        const uint64_t mc = __ballot(action == 3);
        if (action == 3)
            cring[(chead + ccnt + __popcll(mc & lt)) & (RING - 1)] =
                make_float4((float)(cc[0] - bc[0]), (float)(cc[1] - bc[1]),
                            (float)(cc[2] - bc[2]), __int_as_float(k));
        ccnt += __popcll(mc);
        const int cs = action == 1 ? kcs : 0;
        const int nch = action == 1 ? knch : 0;
        const int incl = wave_incl_scan_dpp(nch);
        const int tot = __builtin_amdgcn_readlane(incl, 63);
        if (top + tot > SC) {
            overflow = true;
            break;
        }
        const int off = top + incl - nch;
        for (int c = 0; c < nch; ++c) stk[off + c] = cs + c;
        top += tot;
        maxtop = top > maxtop ? top : maxtop;
        __builtin_amdgcn_wave_barrier();
        if (ccnt >= 64) {
            batch_cells(chead, 64);
            chead += 64;
            ccnt -= 64;
            ntot_c += 64;
        }
        while (pcnt >= 64) {
            batch_parts(phead, 64);
            phead += 64;
            pcnt -= 64;
            ntot_p += 64;
        }
    }
    if (!overflow) {
        if (ccnt > 0) batch_cells(chead, ccnt);
        if (pcnt > 0) batch_parts(phead, pcnt);
        ntot_c += ccnt;
        ntot_p += pcnt;
        if constexpr (!RECORD) {
            // the rings are drained: the particle ring takes the bunch's bodies
            __builtin_amdgcn_wave_barrier();
            pring[lane] = make_float4(tx, ty, tz, lane < nb ? (float)sm[i] : 0.0f);
            __builtin_amdgcn_wave_barrier();
            float acc[3] = {0.f, 0.f, 0.f};
            for (int j = 0; j < nb; ++j) {
                float4 e = pring[j];
                if (j == lane) e.w = 0.0f, e.x = tx + 1.0f;  // zero mass, non-zero distance
                p2p32(e, tx, ty, tz, eps2, acc);
            }
            sum.add(acc);
            if (lane < nb) {
                const double gm = G * sm[i];
                fx[i] = gm * ((double)sum.s[0] - (double)sum.c[0]);
                fy[i] = gm * ((double)sum.s[1] - (double)sum.c[1]);
                fz[i] = gm * ((double)sum.s[2] - (double)sum.c[2]);
            }
        }
    }
    if constexpr (RECORD) {
        // every bunch's count is written (-1 on any overflow), so the later
        // kernels never read batches this launch did not record
        if (lane == 0) so.nbat[gi] = (rover || overflow) ? -1 : nbat;
    }
    if (lane == 0) {
        fnum[gi] = nfar;
        if (overflow) atomicOr((unsigned long long*)&status[ST_FLAGS], 16ull);
        if (rover) atomicOr((unsigned long long*)&status[ST_FLAGS], 32ull);
        const unsigned long long unb = (unsigned long long)nb;
        atomicAdd((unsigned long long*)&status[ST_LISTC], (unsigned long long)(ntot_c + nfar));
        atomicAdd((unsigned long long*)&status[ST_LISTP], (unsigned long long)ntot_p);
        atomicAdd((unsigned long long*)&status[ST_LISTF], (unsigned long long)nfar);
        atomicAdd((unsigned long long*)&status[ST_EXECC], (unsigned long long)ntot_c * unb);
        atomicAdd((unsigned long long*)&status[ST_EXECP], (unsigned long long)ntot_p * unb + unb * (unb - 1));
        atomicMax((long long*)&status[ST_NEEDC], (long long)ntot_c);
        atomicMax((long long*)&status[ST_NEEDP], (long long)ntot_p);
        atomicMax((long long*)&status[ST_MAXSTK], (long long)maxtop);
    }
}

// Target coordinates of bunch gi, relative to the centre of its bounding box,
// exactly as the walk computes them.
struct BunchTargets {
    int i, nb;
    float tx, ty, tz;
    double bc[3];
    float h, hinv;  // radius for the far-field expansion, and its inverse
};

__device__ __forceinline__ BunchTargets bunch_targets(const double* __restrict__ sx,
                                                      const double* __restrict__ sy,
                                                      const double* __restrict__ sz,
                                                      const int* __restrict__ bstart, int gi,
                                                      int lane) {
    BunchTargets b;
    const int gf = bstart[gi], gl = bstart[gi + 1] - 1;
    b.nb = gl - gf + 1;
    b.i = gf + (lane < b.nb ? lane : 0);
    double v[3] = {sx[b.i], sy[b.i], sz[b.i]}, w[3] = {v[0], v[1], v[2]};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            v[c] = fmin(v[c], __shfl_xor(v[c], o, 64));
            w[c] = fmax(w[c], __shfl_xor(w[c], o, 64));
        }
    }
    const double bc[3] = {0.5 * (v[0] + w[0]), 0.5 * (v[1] + w[1]), 0.5 * (v[2] + w[2])};
    b.tx = (float)(sx[b.i] - bc[0]);
    b.ty = (float)(sy[b.i] - bc[1]);
    b.tz = (float)(sz[b.i] - bc[2]);
    for (int c = 0; c < 3; ++c) b.bc[c] = bc[c];
    b.h = bunch_radius(b.tx, b.ty, b.tz, lane, b.nb);
    b.hinv = 1.0f / b.h;
    return b;
}

// Split evaluation, step 2: block (gi, k) evaluates batch k of bunch gi.
template <int MP>
__global__ __launch_bounds__(64) void k_eval_batches(const double* __restrict__ sx,
                                                     const double* __restrict__ sy,
                                                     const double* __restrict__ sz,
                                                     const int* __restrict__ bstart, int nbunch,
                                                     const int64_t* __restrict__ dnbun,
                                                     SplitOut so, const float* __restrict__ mom,
                                                     float eps2, float4* __restrict__ part) {
    if (dnbun) nbunch = (int)*dnbun;
    const int gi = (int)(blockIdx.x / so.bcap), k = (int)(blockIdx.x % so.bcap);
    if (gi >= nbunch || k >= so.nbat[gi]) return;
    const int lane = __lane_id();
    const BunchTargets b = bunch_targets(sx, sy, sz, bstart, gi, lane);
    const int2 d = so.desc[(int64_t)gi * so.bcap + k];
    const int first = d.x, cnt = d.y & 0xFFFF, kind = d.y >> 16;
    const float4* __restrict__ lg = so.log + (int64_t)gi * so.lcap;
    float acc[3] = {0.f, 0.f, 0.f};
    if (kind == 0) eval_cells_acc<MP, -1>(lg, first, cnt, mom, b.tx, b.ty, b.tz, eps2, acc);
    else eval_parts_acc<-1>(lg, first, cnt, b.tx, b.ty, b.tz, eps2, acc);
    part[((int64_t)gi * so.bcap + k) * 64 + lane] = make_float4(acc[0], acc[1], acc[2], 0.0f);
}

// Split evaluation, step 3: the partials of bunch gi in batch order, then the
// bunch's own bodies, in the walk kernel's compensated sum.
__global__ __launch_bounds__(64) void k_combine(const double* __restrict__ sx,
                                                const double* __restrict__ sy,
                                                const double* __restrict__ sz,
                                                const double* __restrict__ sm,
                                                const int* __restrict__ bstart, int nbunch,
                                                const int64_t* __restrict__ dnbun, SplitOut so,
                                                const float4* __restrict__ part, double G,
                                                float eps2, double* __restrict__ fx,
                                                double* __restrict__ fy, double* __restrict__ fz) {
    __shared__ float4 own[BUNCH];
    if (dnbun) nbunch = (int)*dnbun;
    const int gi = blockIdx.x;
    if (gi >= nbunch) return;
    const int lane = __lane_id();
    const BunchTargets b = bunch_targets(sx, sy, sz, bstart, gi, lane);
    const float tx = b.tx, ty = b.ty, tz = b.tz;
    own[lane] = make_float4(tx, ty, tz, lane < b.nb ? (float)sm[b.i] : 0.0f);
    __syncthreads();
    Kahan3 sum;
    const int nbat = so.nbat[gi];
    for (int k = 0; k < nbat; ++k) {
        const float4 a = part[((int64_t)gi * so.bcap + k) * 64 + lane];
        float acc[3] = {a.x, a.y, a.z};
        sum.add(acc);
    }
    float acc[3] = {0.f, 0.f, 0.f};
    for (int j = 0; j < b.nb; ++j) {
        float4 e = own[j];
        if (j == lane) e.w = 0.0f, e.x = tx + 1.0f;  // zero mass, non-zero distance
        p2p32(e, tx, ty, tz, eps2, acc);
    }
    sum.add(acc);
    if (lane < b.nb) {
        const double gm = G * sm[b.i];
        fx[b.i] = gm * ((double)sum.s[0] - (double)sum.c[0]);
        fy[b.i] = gm * ((double)sum.s[1] - (double)sum.c[1]);
        fz[b.i] = gm * ((double)sum.s[2] - (double)sum.c[2]);
    }
}

// Far field of the mixed walk: one wavefront per bunch forms the local
// expansion of one tier of its far list (which 0: the outer cells, stored as
// node indices, order set P = 5; which 1: the inner ones, stored as their
// complements, P = 6) lane-parallel over the cells, sums it over the
// wavefront, evaluates it at its bodies and adds G m times the result to the
// force the walk wrote.  Runs after the walk (fused, or split and combined).
template <int MP, int P>
__global__ __launch_bounds__(64) void k_far(const double* __restrict__ sx,
                                            const double* __restrict__ sy,
                                            const double* __restrict__ sz,
                                            const double* __restrict__ sm,
                                            const int* __restrict__ bstart, int nbunch,
                                            const int64_t* __restrict__ dnbun,
                                            const int* __restrict__ flist,
                                            const int* __restrict__ fnum, int fcap, int which,
                                            const double* __restrict__ ccx,
                                            const double* __restrict__ ccy,
                                            const double* __restrict__ ccz,
                                            const float* __restrict__ mom, double G, float eps2,
                                            double* __restrict__ fx, double* __restrict__ fy,
                                            double* __restrict__ fz) {
    if (dnbun) nbunch = (int)*dnbun;
    if ((int)blockIdx.x >= nbunch) return;
    const int gi = xcd_swizzle(blockIdx.x, nbunch);
    const int nf = min(fnum[gi], fcap);
    if (nf <= 0) return;
    const int lane = __lane_id();
    const uint64_t lt = lanemask_lt();
    const int* __restrict__ fl = flist + (int64_t)gi * fcap;
    // the tier's cells, compacted through an LDS ring of 128 (each pass adds
    // at most 64 and 64 are taken whenever that many are held)
    __shared__ int sel[128];
    int head = 0, cnt = 0, tot = 0;
    BunchTargets b;
    constexpr int NL = P == 5 ? NLOC_P5 : NLOC_P6;
    float L[NL];
#pragma unroll
    for (int j = 0; j < NL; ++j) L[j] = 0.0f;
    // lane j takes cells j, j + 64, ... of the ring into its own
    // coefficients; the wavefront sums them once and every lane evaluates the
    // sum at its body
    for (int base = 0; base < nf || cnt > 0; base += 64) {
        if (base < nf) {
            const int e = base + lane < nf ? fl[base + lane] : 0;
            const bool mine = base + lane < nf && (which ? e < 0 : e >= 0);
            const uint64_t m = __ballot(mine);
            if (mine) sel[(head + cnt + __popcll(m & lt)) & 127] = which ? ~e : e;
            cnt += __popcll(m);
            __builtin_amdgcn_wave_barrier();
        }
        if (cnt >= 64 || (base + 64 >= nf && cnt > 0)) {
            if (tot == 0) b = bunch_targets(sx, sy, sz, bstart, gi, lane);
            const int take = cnt < 64 ? cnt : 64;
            const int node = lane < take ? sel[(head + lane) & 127] : -1;
            __builtin_amdgcn_wave_barrier();
            head += take;
            cnt -= take;
            tot += take;
            far_m2l<MP, P>(node, ccx, ccy, ccz, mom, b.bc, b.h, eps2, L);
        }
    }
    if (tot == 0) return;
#pragma unroll
    for (int j = 0; j < NL; ++j) L[j] = wave_sum_dpp(L[j]);
    float a[3];
    if constexpr (P == 5) local_l2p_p5(L, b.tx * b.hinv, b.ty * b.hinv, b.tz * b.hinv, a);
    else local_l2p_p6(L, b.tx * b.hinv, b.ty * b.hinv, b.tz * b.hinv, a);
    if (lane < b.nb) {
        const double gm = G * sm[b.i];
        fx[b.i] = fx[b.i] + gm * (double)(-b.hinv * a[0]);
        fy[b.i] = fy[b.i] + gm * (double)(-b.hinv * a[1]);
        fz[b.i] = fz[b.i] + gm * (double)(-b.hinv * a[2]);
    }
}

// ------------------------------------------------------------------ scatter
__global__ __launch_bounds__(256) void k_scatter(const int* __restrict__ idx,
                                                 const double* __restrict__ fx,
                                                 const double* __restrict__ fy,
                                                 const double* __restrict__ fz, int64_t n,
                                                 int layout, double* __restrict__ F) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t p = idx[i];
    if (layout == HQ_LAYOUT_3N) {
        F[3 * p] = fx[i];
        F[3 * p + 1] = fy[i];
        F[3 * p + 2] = fz[i];
    } else {
        F[p] = fx[i];
        F[n + p] = fy[i];
        F[2 * n + p] = fz[i];
    }
}

// lost particles (no leaf) still get their walk result, as in the reference;
// nothing else to do here.

// ------------------------------------------------------------------ drivers
static int ensure_forces(hq_tree* t) {
    int rc;
    size_t b = t->n * sizeof(double);
    if ((rc = t->fx.ensure(b)) || (rc = t->fy.ensure(b)) || (rc = t->fz.ensure(b))) return rc;
    return HQ_OK;
}

int tree_forces_bh(hq_tree* t, double theta2, double G, double eps2) {
    int rc;
    if ((rc = ensure_forces(t)) || (rc = ensure_next(t))) return rc;
    hipStream_t s = t->stream;
    HQ_CHECK_HIP(hipMemsetAsync(t->status.p, 0, ST_SIZE * sizeof(int64_t), s));
    HQ_CHECK_HIP(hipEventRecord(t->ev[5], s));
    k_bh_walk<<<grid_for(t->n, 128), 128, 0, s>>>(
        node_view(t), t->geom.as<double>() + G_SYS, t->sx.as<double>(), t->sy.as<double>(),
        t->sz.as<double>(), t->sm.as<double>(), t->n, t->multipole, theta2, G, eps2,
        t->fx.as<double>(), t->fy.as<double>(), t->fz.as<double>(), t->status.as<int64_t>());
    HQ_CHECK_HIP(hipEventRecord(t->ev[6], s));
    if ((rc = read_status(t))) return rc;
    HQ_CHECK_HIP(hipGetLastError());
    hipEventElapsedTime(&t->ms_walk, t->ev[5], t->ev[6]);
    t->n_p2p = t->hstatus[ST_P2P];
    t->n_m2p = t->hstatus[ST_M2P];
    return HQ_OK;
}

// groups (methods 30-32): maximal nodes with count <= group_size, or leaves,
// found in tree order (prepare_groups :1789-1808)
__global__ __launch_bounds__(256) void k_mark_groups(NodeView nv, const int* __restrict__ nparent,
                                                     int64_t nn, int gs, int* __restrict__ gmark) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const int cnt = nv.count[k];
    if (cnt == 0) return;
    if (!(cnt <= gs || nv.type[k] == 1)) return;
    const int p = nparent[k];
    if (p >= 0 && nv.count[p] <= gs) return;  // the parent is (inside) a group already
    gmark[nv.first[k]] = (int)k;
}

// bunches: in every twig with more than `gs` bodies, runs of adjacent children
// that hold <= gs bodies each are packed greedily into super-groups of <= gs
// bodies; a tree with <= gs bodies is one super-group.  Each super-group of
// c bodies is cut into ceil(c / B) bunches of near-equal size, and the first
// particle of each bunch is marked.  (Super-groups tile the bodies: bodies
// lost at max depth, in nodes of count 0, stop a bunched walk before this.)
__global__ __launch_bounds__(256) void k_mark_bunches(NodeView nv, int64_t nn, int gs, int B,
                                                      int* __restrict__ bmark,
                                                      const int64_t* __restrict__ dnn) {
    int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (dnn) nn = *dnn;
    if (k >= nn) return;
    auto emit = [&](int s0, int c) {
        const int nb = (c + B - 1) / B;
        if (nb <= 0) return;
        const int base = c / nb, rem = c % nb;
        for (int j = 0; j < nb; ++j) {
            const int p = s0 + j * base + (j < rem ? j : rem);
            bmark[p] = p;
        }
    };
    const int cnt = nv.count[k];
    if (nv.depth[k] == 0 && cnt <= gs) {  // a root (one per system) small enough to bunch whole
        emit(nv.first[k], cnt);
        return;
    }
    if (nv.type[k] != 2 || cnt <= gs) return;
    int run = 0, rs = 0;
    bool prev = false;
    for (int c = nv.cstart[k]; c < nv.cstart[k + 1]; ++c) {
        const int cc = nv.count[c];
        if (cc == 0 || cc > gs) {
            if (prev) emit(rs, run);
            prev = false;
            continue;
        }
        if (prev && run + cc <= gs) {
            run += cc;
        } else {
            if (prev) emit(rs, run);
            rs = nv.first[c];
            run = cc;
        }
        prev = true;
    }
    if (prev) emit(rs, run);
}
__global__ __launch_bounds__(256) void k_flag_count(const int* __restrict__ gmark, int64_t n,
                                                    uint32_t* __restrict__ bcnt) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    uint64_t m = __ballot(i < n && gmark[i] >= 0);
    __shared__ uint32_t wc[4];
    if (__lane_id() == 0) wc[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) bcnt[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ __launch_bounds__(256) void k_flag_write(const int* __restrict__ gmark, int64_t n,
                                                    const uint32_t* __restrict__ boff,
                                                    int* __restrict__ out) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool f = i < n && gmark[i] >= 0;
    uint64_t m = __ballot(f);
    __shared__ uint32_t wc[4];
    int w = threadIdx.x >> 6;
    if (__lane_id() == 0) wc[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t pre = boff[blockIdx.x];
    for (int k = 0; k < w; ++k) pre += wc[k];
    if (f) out[pre + __popcll(m & lanemask_lt())] = gmark[i];
}

// a[*pos] = v: the end marker after a table whose length is on the device
__global__ void k_set_end(int* __restrict__ a, const uint32_t* __restrict__ pos, int v) {
    if (threadIdx.x == 0) a[*pos] = v;
}

// Graph mode: the bunch count for the walk's grid test, and a flag when the
// grid (sized from a launched step) is too small for it.  The count the
// kernels read is clamped to the grid, for which every per-bunch buffer is
// sized; the flag makes the integrator redo the steps launched.
__global__ void k_graph_bunches(const uint32_t* __restrict__ nb, int64_t grid,
                                int64_t* __restrict__ gs) {
    if (threadIdx.x != 0) return;
    const int64_t c = *nb;
    gs[GS_NBUN] = c < grid ? c : grid;
    if (c > grid) gs[GS_STICKY] |= GF_BUNCHES;
}

// Graph mode: the walk's stack (or split log) overflow, kept over replayed steps.
__global__ void k_graph_walk_check(const int64_t* __restrict__ st, int64_t* __restrict__ gs) {
    if (threadIdx.x == 0 && (st[ST_FLAGS] & (16 | 32))) gs[GS_STICKY] |= GF_STACK;
}

int tree_make_groups(hq_tree* t, bool bunched);
int finish_groups(hq_tree* t);

// Split evaluation (record, per-batch partials, combine) is used for few
// bunches, once a walk has measured the list lengths (split_lcap).
constexpr int SPLIT_MAX = 2048;                      // bunches (auto mode)
constexpr double SPLIT_MEM_MAX = 1024.0 * 1024 * 1024;  // bytes of log and partials

static int split_bcap(int64_t lcap) { return (int)(lcap / 64 + 3); }

// Square of the far-field factor of the mixed walk at opening angle theta
// (theta2 = theta^2; 0: no far field), and of the inner one (0: no inner
// tier; used only below the outer factor).  Below theta = 0.5 both squares
// grow as 0.5 / theta: the expansions' truncation error then falls with the
// multipole error of the cells they take.
static bool inner_on(const hq_tree* t) {
    return t->far > 0 && t->far_inner > 0 && t->far_inner < t->far;
}
// Scale of the offset acceptance test of the bunched walks: theta / 32767
// (the walk records hold the offset in multiples of 1/32767 of the side);
// 0 for the plain cell-side test.
static double mac_off(const hq_tree* t, double theta2) {
    return t->mac ? std::sqrt(theta2) / 32767.0 : 0.0;
}

static float far_scale(double theta2) {
    return theta2 > 0.0 && theta2 < 0.25 ? (float)(0.5 / std::sqrt(theta2)) : 1.0f;
}
static float far2(const hq_tree* t, double theta2) {
    return (float)(t->far * t->far) / 100.0f * far_scale(theta2);
}
static float far2i(const hq_tree* t, double theta2) {
    return inner_on(t) ? (float)(t->far_inner * t->far_inner) / 100.0f * far_scale(theta2) : 0.0f;
}

// Far cells a bunch's list holds (more are evaluated in the walk as ordinary
// cells), twice as many with the inner tier; 0 when the far field is off.
constexpr int FAR_CAP = 4096;
static int far_cap(const hq_tree* t) { return t->far > 0 ? (inner_on(t) ? 2 : 1) * FAR_CAP : 0; }

static int far_buffers(hq_tree* t, int64_t nbunch) {
    int rc;
    if ((rc = t->fnum.ensure((size_t)(nbunch + 1) * sizeof(int))) ||
        (rc = t->flist.ensure((size_t)(nbunch + 1) * (far_cap(t) + 1) * sizeof(int))))
        return rc;
    return HQ_OK;
}

// The far field of a mixed walk over grid bunches (after the walk).
static void launch_far(hq_tree* t, int grid, const int64_t* dnbun, double G, double eps2) {
    if (t->far <= 0) return;
    // the outer tier (order set p5), then the inner one (p6)
    auto launch = [&](auto kern, int which) {
        kern<<<grid, 64, 0, t->stream>>>(t->sx.as<double>(), t->sy.as<double>(), t->sz.as<double>(),
                                         t->sm.as<double>(), t->bstart.as<int>(), grid, dnbun,
                                         t->flist.as<int>(), t->fnum.as<int>(), far_cap(t), which,
                                         t->ncomx.as<double>(), t->ncomy.as<double>(),
                                         t->ncomz.as<double>(), t->mom32.as<float>(), G,
                                         (float)eps2, t->fx.as<double>(), t->fy.as<double>(),
                                         t->fz.as<double>());
    };
    if (t->multipole == 1) launch(k_far<1, 5>, 0);
    else if (t->multipole == 4) launch(k_far<4, 5>, 0);
    else launch(k_far<8, 5>, 0);
    if (!inner_on(t)) return;
    if (t->multipole == 1) launch(k_far<1, 6>, 1);
    else if (t->multipole == 4) launch(k_far<4, 6>, 1);
    else launch(k_far<8, 6>, 1);
}

static bool use_split(const hq_tree* t, int64_t nbunch) {
    if (t->split == 0 || t->split_lcap <= 0) return false;
    if (t->split == 1 && nbunch > SPLIT_MAX) return false;
    const double bcap = split_bcap(t->split_lcap);
    const double bytes = (double)nbunch * (t->split_lcap * 16.0 + bcap * (8.0 + 64 * 16.0) + 4.0);
    return bytes <= SPLIT_MEM_MAX;
}

static void note_lists(hq_tree* t) {  // log size for the next split walk, from this walk's lists
    const int64_t need = t->hstatus[ST_NEEDC] + t->hstatus[ST_NEEDP];
    t->split_lcap = (need * 5 / 4 + 128 + 63) / 64 * 64;
}

static int split_buffers(hq_tree* t, int64_t nbunch) {
    int rc;
    const int64_t lcap = t->split_lcap, bcap = split_bcap(lcap);
    if ((rc = t->slog.ensure((size_t)nbunch * lcap * sizeof(float4))) ||
        (rc = t->sdesc.ensure((size_t)nbunch * bcap * sizeof(int2))) ||
        (rc = t->snbat.ensure((size_t)nbunch * sizeof(int))) ||
        (rc = t->spart.ensure((size_t)nbunch * bcap * 64 * sizeof(float4))))
        return rc;
    return HQ_OK;
}

// The three launches of a split mixed walk over grid bunches (the count on the
// device when dnbun is given).
static int launch_split(hq_tree* t, int grid, const int64_t* dnbun, double theta2, double G,
                        double eps2) {
    int rc;
    if ((rc = split_buffers(t, grid)) || (rc = far_buffers(t, grid))) return rc;
    hipStream_t s = t->stream;
    const SplitOut so{t->slog.as<float4>(), t->sdesc.as<int2>(), t->snbat.as<int>(),
                      t->split_lcap, split_bcap(t->split_lcap)};
    auto record = [&](auto kern) {
        kern<<<grid, 64, 0, s>>>(node_view(t), t->geom.as<double>() + G_SYS, t->sx.as<double>(),
                                 t->sy.as<double>(), t->sz.as<double>(), t->sm.as<double>(), t->n,
                                 t->bstart.as<int>(), grid, theta2, mac_off(t, theta2), G, (float)eps2, t->leaf,
                                 far2(t, theta2), far2i(t, theta2), t->flist.as<int>(), t->fnum.as<int>(), far_cap(t),
                                 t->wnode.as<WalkNode>(), t->sbody.as<double4>(), t->mom32.as<float>(),
                                 t->fx.as<double>(), t->fy.as<double>(),
                                 t->fz.as<double>(), t->status.as<int64_t>(), dnbun, so);
    };
    auto eval = [&](auto kern) {
        kern<<<grid * so.bcap, 64, 0, s>>>(t->sx.as<double>(), t->sy.as<double>(),
                                           t->sz.as<double>(), t->bstart.as<int>(), grid, dnbun,
                                           so, t->mom32.as<float>(), (float)eps2,
                                           t->spart.as<float4>());
    };
    if (t->multipole == 1) record(k_walk_mixed<1, SCAP_FUSED, true>), eval(k_eval_batches<1>);
    else if (t->multipole == 4) record(k_walk_mixed<4, SCAP_FUSED, true>), eval(k_eval_batches<4>);
    else record(k_walk_mixed<8, SCAP_FUSED, true>), eval(k_eval_batches<8>);
    k_combine<<<grid, 64, 0, s>>>(t->sx.as<double>(), t->sy.as<double>(), t->sz.as<double>(),
                                  t->sm.as<double>(), t->bstart.as<int>(), grid, dnbun, so,
                                  t->spart.as<float4>(), G, (float)eps2, t->fx.as<double>(),
                                  t->fy.as<double>(), t->fz.as<double>());
    launch_far(t, grid, dnbun, G, eps2);
    HQ_CHECK_HIP(hipGetLastError());
    return HQ_OK;
}

// Buffers at the sizes graph mode uses (upper bounds), so that nothing grows
// while a step is captured.
int graph_prepare(hq_tree* t) {
    int rc;
    const int64_t n = t->n;
    if ((rc = ensure_forces(t)) || (rc = t->gstate.ensure(GS_SIZE * sizeof(int64_t))) ||
        (rc = t->gsave.ensure((size_t)6 * n * sizeof(double))) ||
        (rc = t->bstart.ensure(((size_t)n + 1) * sizeof(int))) ||
        (rc = t->gflag.ensure(((size_t)n + 1) * sizeof(int))) ||
        (rc = t->groups.ensure(((size_t)n + 1) * sizeof(int))) ||
        (rc = t->mom32.ensure((size_t)t->cap * REC * sizeof(float))) ||
        (rc = t->wnode.ensure((size_t)t->cap * sizeof(WalkNode))) ||
        (rc = t->scanpart.ensure(scan_part_size(2 * n + 65536) * sizeof(uint32_t))))
        return rc;
    t->graph_split = use_split(t, t->graph_nbun);
    if (t->graph_split && (rc = split_buffers(t, t->graph_nbun))) return rc;
    if ((rc = far_buffers(t, t->graph_nbun))) return rc;
    return HQ_OK;
}

int tree_forces_groups(hq_tree* t, double theta2, double G, double eps2, int precision,
                       bool bunched) {
    int rc;
    if ((rc = ensure_forces(t))) return rc;
    hipStream_t s = t->stream;
    int ngroups;
    if (bunched) tune_from_table(t, theta2, precision);
    if ((rc = tree_make_groups(t, bunched))) return rc;
    const bool mixed = bunched && precision == HQ_MIXED;
    if (mixed && !t->graph_mode) {
        // the walk records need no bunch count: queued before the wait for it
        if ((rc = t->mom32.ensure((size_t)t->nn * REC * sizeof(float)))) return rc;
        if ((rc = t->wnode.ensure((size_t)t->nn * sizeof(WalkNode)))) return rc;
        k_mom32<<<grid_for(t->nn, 256), 256, 0, s>>>(t->nn, t->multipole, node_view(t),
                                                     t->geom.as<double>() + G_SYS, t->mom32.as<float>(),
                                                     t->wnode.as<WalkNode>(), nullptr);
    }
    if ((rc = finish_groups(t))) return rc;
    ngroups = (int)t->ngroups;
    if (t->graph_mode) {
        // captured step: node and bunch counts stay on the device, grids are
        // upper bounds (capacity, bunches of the launched step with margin)
        if (!mixed || !t->fused) return fail(HQ_ERR_ARGS, "graph mode needs the fused mixed walk");
        int64_t* gs = t->gstate.as<int64_t>();
        if ((rc = t->mom32.ensure((size_t)t->cap * REC * sizeof(float)))) return rc;
        if ((rc = t->wnode.ensure((size_t)t->cap * sizeof(WalkNode)))) return rc;
        k_mom32<<<grid_for(t->cap, 256), 256, 0, s>>>(t->cap, t->multipole, node_view(t),
                                                      t->geom.as<double>() + G_SYS, t->mom32.as<float>(),
                                                      t->wnode.as<WalkNode>(), gs + GS_NN);
        k_fill_i64<<<1, 64, 0, s>>>(t->status.as<int64_t>(), 0, ST_SIZE);
        if (t->graph_split) {
            if ((rc = launch_split(t, (int)t->graph_nbun, gs + GS_NBUN, theta2, G, eps2))) return rc;
        } else {
            auto launch = [&](auto kern) {
                kern<<<(int)t->graph_nbun, 64, 0, s>>>(
                    node_view(t), t->geom.as<double>() + G_SYS, t->sx.as<double>(),
                    t->sy.as<double>(), t->sz.as<double>(), t->sm.as<double>(), t->n,
                    t->bstart.as<int>(), (int)t->graph_nbun, theta2, mac_off(t, theta2), G, (float)eps2, t->leaf,
                    far2(t, theta2), far2i(t, theta2), t->flist.as<int>(), t->fnum.as<int>(), far_cap(t),
                    t->wnode.as<WalkNode>(), t->sbody.as<double4>(), t->mom32.as<float>(), t->fx.as<double>(),
                    t->fy.as<double>(),
                    t->fz.as<double>(), t->status.as<int64_t>(), gs + GS_NBUN, SplitOut{});
            };
            if (t->multipole == 1) launch(k_walk_mixed<1, SCAP_FUSED>);
            else if (t->multipole == 4) launch(k_walk_mixed<4, SCAP_FUSED>);
            else launch(k_walk_mixed<8, SCAP_FUSED>);
            launch_far(t, (int)t->graph_nbun, gs + GS_NBUN, G, eps2);
        }
        k_graph_walk_check<<<1, 64, 0, s>>>(t->status.as<int64_t>(), gs);
        t->walk_stack = SCAP_FUSED;
        HQ_CHECK_HIP(hipGetLastError());
        return HQ_OK;
    }
    if (mixed) {
        if (t->fused && use_split(t, ngroups)) {
            HQ_CHECK_HIP(hipMemsetAsync(t->status.p, 0, ST_SIZE * sizeof(int64_t), s));
            HQ_CHECK_HIP(hipEventRecord(t->ev[5], s));
            if ((rc = launch_split(t, ngroups, nullptr, theta2, G, eps2))) return rc;
            HQ_CHECK_HIP(hipEventRecord(t->ev[8], s));
            if ((rc = read_status(t))) return rc;
            HQ_CHECK_HIP(hipGetLastError());
            const int64_t fl = t->hstatus[ST_FLAGS];
            note_lists(t);
            if (!(fl & (16 | 32))) {
                ++t->split_walks;
                t->walk_stack = SCAP_FUSED;
                hipEventElapsedTime(&t->ms_eval, t->ev[5], t->ev[8]);
                t->ms_lists = 0.0f;
                t->ms_walk = t->ms_eval;
                t->list_cells = t->hstatus[ST_LISTC];
                t->far_cells = t->hstatus[ST_LISTF];
                t->list_parts = t->hstatus[ST_LISTP];
                t->exec_cells = t->hstatus[ST_EXECC];
                t->exec_parts = t->hstatus[ST_EXECP];
                t->max_cells = t->hstatus[ST_NEEDC];
                t->max_parts = t->hstatus[ST_NEEDP];
                t->max_stack = t->hstatus[ST_MAXSTK];
                return HQ_OK;
            }
            // the log or the stack overflowed: the fused walk below
        }
        // the fused walk with the short stack, then (if a bunch's stack
        // overflowed) with the long one, then the deep one; it writes every
        // force and far list afresh, so a retry starts clean.  The handle
        // starts at the stack that last sufficed and steps back down once a
        // walk needs well under the next smaller one; when even the deep
        // stack overflowed it goes straight to the two-kernel walk until that
        // walk needs less.
        for (int attempt = t->stack_level; t->fused && attempt < 3; ++attempt) {
            if ((rc = far_buffers(t, ngroups))) return rc;
            HQ_CHECK_HIP(hipMemsetAsync(t->status.p, 0, ST_SIZE * sizeof(int64_t), s));
            HQ_CHECK_HIP(hipEventRecord(t->ev[5], s));
            auto launch = [&](auto kern) {
                kern<<<ngroups, 64, 0, s>>>(node_view(t), t->geom.as<double>() + G_SYS, t->sx.as<double>(),
                                            t->sy.as<double>(), t->sz.as<double>(),
                                            t->sm.as<double>(), t->n, t->bstart.as<int>(), ngroups,
                                            theta2, mac_off(t, theta2), G, (float)eps2, t->leaf, far2(t, theta2), far2i(t, theta2),
                                            t->flist.as<int>(), t->fnum.as<int>(), far_cap(t),
                                            t->wnode.as<WalkNode>(), t->sbody.as<double4>(), t->mom32.as<float>(),
                                            t->fx.as<double>(), t->fy.as<double>(),
                                            t->fz.as<double>(), t->status.as<int64_t>(), nullptr,
                                            SplitOut{});
            };
            if (attempt == 0) {
                if (t->multipole == 1) launch(k_walk_mixed<1, SCAP_FAST>);
                else if (t->multipole == 4) launch(k_walk_mixed<4, SCAP_FAST>);
                else launch(k_walk_mixed<8, SCAP_FAST>);
            } else if (attempt == 1) {
                if (t->multipole == 1) launch(k_walk_mixed<1, SCAP_FUSED>);
                else if (t->multipole == 4) launch(k_walk_mixed<4, SCAP_FUSED>);
                else launch(k_walk_mixed<8, SCAP_FUSED>);
            } else {
                if (t->multipole == 1) launch(k_walk_mixed<1, SCAP_DEEP>);
                else if (t->multipole == 4) launch(k_walk_mixed<4, SCAP_DEEP>);
                else launch(k_walk_mixed<8, SCAP_DEEP>);
            }
            launch_far(t, ngroups, nullptr, G, eps2);
            HQ_CHECK_HIP(hipEventRecord(t->ev[8], s));
            if ((rc = read_status(t))) return rc;
            HQ_CHECK_HIP(hipGetLastError());
            const bool over = t->hstatus[ST_FLAGS] & 16;
            if (over) t->stack_level = std::min(attempt + 1, 2);
            else if (attempt > 0 && t->hstatus[ST_MAXSTK] <= SCAP_LEVEL[attempt - 1] - 128)
                t->stack_level = attempt - 1;
            if (!over) {
                t->walk_stack = SCAP_LEVEL[attempt];
                note_lists(t);
                hipEventElapsedTime(&t->ms_eval, t->ev[5], t->ev[8]);
                t->ms_lists = 0.0f;
                t->ms_walk = t->ms_eval;
                t->list_cells = t->hstatus[ST_LISTC];
                t->far_cells = t->hstatus[ST_LISTF];
                t->list_parts = t->hstatus[ST_LISTP];
                t->exec_cells = t->hstatus[ST_EXECC];
                t->exec_parts = t->hstatus[ST_EXECP];
                t->max_cells = t->hstatus[ST_NEEDC];
                t->max_parts = t->hstatus[ST_NEEDP];
                t->max_stack = t->hstatus[ST_MAXSTK];
                return HQ_OK;
            }
            // a stack overflowed: the next larger one; past the deep stack the
            // two-kernel walk, whose stack is as deep, reports the overflow
        }
    }
    if ((rc = t->lstat.ensure((size_t)ngroups * 2 * sizeof(int)))) return rc;
    if (t->list_ccap == 0) t->list_ccap = 4096;
    if (t->list_pcap == 0) t->list_pcap = 1024;
    for (int attempt = 0; attempt < 4; ++attempt) {
        const size_t esz = mixed ? sizeof(float4) : sizeof(int);
        if ((rc = t->clist.ensure((size_t)ngroups * t->list_ccap * esz)) ||
            (rc = t->plist.ensure((size_t)ngroups * t->list_pcap * esz)))
            return rc;
        ListOut lo;
        lo.cidx = t->clist.as<int>();
        lo.pidx = t->plist.as<int>();
        lo.c32 = t->clist.as<float4>();
        lo.p32 = t->plist.as<float4>();
        lo.ccap = t->list_ccap;
        lo.pcap = t->list_pcap;
        lo.lstat = t->lstat.as<int>();
        HQ_CHECK_HIP(hipMemsetAsync(t->status.p, 0, ST_SIZE * sizeof(int64_t), s));
        HQ_CHECK_HIP(hipEventRecord(t->ev[5], s));
        const NodeView nv = node_view(t);
        if (!bunched)
            k_lists<MODE_GROUP><<<ngroups, 64, 0, s>>>(nv, t->geom.as<double>() + G_SYS, t->sx.as<double>(),
                                                       t->sy.as<double>(), t->sz.as<double>(),
                                                       t->sm.as<double>(), t->n,
                                                       t->groups.as<int>(), ngroups, theta2, 0.0, 0, lo,
                                                       t->status.as<int64_t>());
        else if (!mixed)
            k_lists<MODE_BUNCH><<<ngroups, 64, 0, s>>>(nv, t->geom.as<double>() + G_SYS, t->sx.as<double>(),
                                                       t->sy.as<double>(), t->sz.as<double>(),
                                                       t->sm.as<double>(), t->n, t->bstart.as<int>(),
                                                       ngroups, theta2, mac_off(t, theta2), t->leaf, lo, t->status.as<int64_t>());
        else
            k_lists<MODE_BUNCH32><<<ngroups, 64, 0, s>>>(nv, t->geom.as<double>() + G_SYS,
                                                         t->sx.as<double>(), t->sy.as<double>(),
                                                         t->sz.as<double>(), t->sm.as<double>(),
                                                         t->n, t->bstart.as<int>(), ngroups, theta2,
                                                         mac_off(t, theta2), t->leaf, lo,
                                                         t->status.as<int64_t>());
        HQ_CHECK_HIP(hipEventRecord(t->ev[7], s));
        if ((rc = read_status(t))) return rc;
        const int64_t flags = t->hstatus[ST_FLAGS];
        if (flags & 16) return fail(HQ_ERR_OVERFLOW, "walk stack overflow (more than %d pending nodes)", SCAP);
        if (flags & 8) {
            int64_t needc = t->hstatus[ST_NEEDC], needp = t->hstatus[ST_NEEDP];
            t->list_ccap = std::max<int64_t>(t->list_ccap, (needc + 1023) / 1024 * 1024);
            t->list_pcap = std::max<int64_t>(t->list_pcap, (needp + 255) / 256 * 256);
            const double gib = (double)ngroups * (t->list_ccap + t->list_pcap) * esz / (1 << 30);
            if (gib > LIST_GIB_MAX)
                return fail(HQ_ERR_OVERFLOW, "interaction lists would need %.1f GiB (largest: %lld "
                            "cells, %lld particles)", gib, (long long)needc, (long long)needp);
            continue;  // lists were truncated: rebuild with room for the largest
        }
        t->walk_stack = 0;
        t->list_cells = t->hstatus[ST_LISTC];
        t->list_parts = t->hstatus[ST_LISTP];
        t->far_cells = t->hstatus[ST_LISTF];
        t->exec_cells = t->hstatus[ST_EXECC];
        t->exec_parts = t->hstatus[ST_EXECP];
        t->max_cells = t->hstatus[ST_NEEDC];
        t->max_parts = t->hstatus[ST_NEEDP];
        t->max_stack = t->hstatus[ST_MAXSTK];
        if (!bunched)
            k_eval_fp64<MODE_GROUP><<<ngroups, 64, 0, s>>>(
                nv, t->sx.as<double>(), t->sy.as<double>(), t->sz.as<double>(), t->sm.as<double>(),
                t->n, t->groups.as<int>(), ngroups, t->multipole, G, eps2, lo, t->fx.as<double>(),
                t->fy.as<double>(), t->fz.as<double>());
        else if (!mixed)
            k_eval_fp64<MODE_BUNCH><<<ngroups, 64, 0, s>>>(
                nv, t->sx.as<double>(), t->sy.as<double>(), t->sz.as<double>(), t->sm.as<double>(),
                t->n, t->bstart.as<int>(), ngroups, t->multipole, G, eps2, lo, t->fx.as<double>(),
                t->fy.as<double>(), t->fz.as<double>());
        else {
            auto launch = [&](auto kern) {
                kern<<<ngroups, 64, 0, s>>>(t->sx.as<double>(), t->sy.as<double>(),
                                            t->sz.as<double>(), t->sm.as<double>(), t->n,
                                            t->bstart.as<int>(), ngroups,
                                            G, (float)eps2, lo.c32, lo.p32, lo.lstat, lo.ccap,
                                            lo.pcap, t->mom32.as<float>(), t->fx.as<double>(),
                                            t->fy.as<double>(), t->fz.as<double>());
            };
            if (t->multipole == 1) launch(k_eval_mixed<1>);
            else if (t->multipole == 4) launch(k_eval_mixed<4>);
            else launch(k_eval_mixed<8>);
        }
        HQ_CHECK_HIP(hipEventRecord(t->ev[8], s));
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        HQ_CHECK_HIP(hipGetLastError());
        hipEventElapsedTime(&t->ms_lists, t->ev[5], t->ev[7]);
        hipEventElapsedTime(&t->ms_eval, t->ev[7], t->ev[8]);
        t->ms_walk = t->ms_lists + t->ms_eval;
        return HQ_OK;
    }
    return fail(HQ_ERR_OVERFLOW, "interaction lists kept overflowing");
}

// groups (methods 30-32) or bunches, as a table of node indices (groups) or
// of first particles followed by n (bunches), in tree order
int tree_make_groups(hq_tree* t, bool bunched) {
    int rc;
    const int64_t n = t->n;
    hipStream_t s = t->stream;
    const bool gm = t->graph_mode;
    if (!gm && t->lost > 0)  // graph mode: the build flags it (GF_LOST)
        return fail(HQ_ERR_LOST, "[prepare_groups] body lost: %lld particles at max depth",
                    (long long)t->lost);
    if ((rc = t->gflag.ensure(n * sizeof(int))) || (rc = t->groups.ensure((n + 1) * sizeof(int))))
        return rc;
    if (gm) k_fill_i32<<<grid_for(n, 256 * 8), 256, 0, s>>>(t->gflag.as<int>(), -1, n);
    else HQ_CHECK_HIP(hipMemsetAsync(t->gflag.p, 0xFF, n * sizeof(int), s));
    const int B = std::min(t->group_size, BUNCH);
    const int S = std::max(B, t->super_size);
    const int64_t nnb = gm ? t->cap : t->nn;  // graph mode: the count is on the device
    const int64_t* dnn = gm ? t->gstate.as<int64_t>() + GS_NN : nullptr;
    if (bunched)
        k_mark_bunches<<<grid_for(nnb, 256), 256, 0, s>>>(node_view(t), nnb, S, B,
                                                         t->gflag.as<int>(), dnn);
    else
        k_mark_groups<<<grid_for(t->nn, 256), 256, 0, s>>>(node_view(t), t->nparent.as<int>(),
                                                           t->nn, t->group_size, t->gflag.as<int>());
    const int nb = grid_for(n, 256);
    if ((rc = t->hist.ensure(((size_t)nb + 1) * sizeof(uint32_t)))) return rc;
    if (bunched && (rc = t->bstart.ensure(((size_t)n + 1) * sizeof(int)))) return rc;
    // the marked first particles in order: group starts, or bunch starts
    int* out = bunched ? t->bstart.as<int>() : t->groups.as<int>();
    k_flag_count<<<nb, 256, 0, s>>>(t->gflag.as<int>(), n, t->hist.as<uint32_t>());
    if ((rc = scan_excl_buf(t->hist.as<uint32_t>(), nb, t->scanpart, s))) return rc;
    k_flag_write<<<nb, 256, 0, s>>>(t->gflag.as<int>(), n, t->hist.as<uint32_t>(), out);
    if (bunched) {
        // the bunch count stays on the device (one read-back at the end when
        // launched, none in graph mode); the table ends with n
        const uint32_t* dnb = t->hist.as<uint32_t>() + nb;
        k_set_end<<<1, 64, 0, s>>>(t->bstart.as<int>(), dnb, (int)n);
        if (gm) {
            k_graph_bunches<<<1, 64, 0, s>>>(dnb, t->graph_nbun, t->gstate.as<int64_t>());
            HQ_CHECK_HIP(hipGetLastError());
            return HQ_OK;
        }
        // the count comes back through a pinned slot; the caller may queue
        // work that does not need it (k_mom32) before finish_groups waits
        uint32_t* slot = reinterpret_cast<uint32_t*>(t->hstatus + ST_HOST_BUNCHES);
        HQ_CHECK_HIP(hipMemcpyAsync(slot, dnb, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HQ_CHECK_HIP(hipEventRecord(t->ev[12], s));
        t->groups_pending = true;
        return HQ_OK;
    }
    uint32_t total = 0;
    HQ_CHECK_HIP(hipMemcpyAsync(&total, t->hist.as<uint32_t>() + nb, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    t->ngroups = total;
    return HQ_OK;
}

// Waits for the bunch count that tree_make_groups left in flight.
int finish_groups(hq_tree* t) {
    if (!t->groups_pending) return HQ_OK;
    t->groups_pending = false;
    HQ_CHECK_HIP(hipEventSynchronize(t->ev[12]));
    HQ_CHECK_HIP(hipGetLastError());
    t->ngroups = *reinterpret_cast<const uint32_t*>(t->hstatus + ST_HOST_BUNCHES);
    return HQ_OK;
}

int tree_scatter_forces(hq_tree* t, double* F, int layout, int where) {
    hipStream_t s = t->stream;
    int rc;
    double* dF = F;
    if (where == HQ_HOST) {
        if ((rc = t->counts.ensure(3 * t->n * sizeof(double)))) return rc;
        dF = t->counts.as<double>();
    }
    k_scatter<<<grid_for(t->n, 256), 256, 0, s>>>(t->idx.as<int>(), t->fx.as<double>(),
                                                  t->fy.as<double>(), t->fz.as<double>(), t->n,
                                                  layout, dF);
    if (where == HQ_HOST)
        HQ_CHECK_HIP(hipMemcpyAsync(F, dF, 3 * t->n * sizeof(double), hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    HQ_CHECK_HIP(hipGetLastError());
    return HQ_OK;
}

}  // namespace hq
