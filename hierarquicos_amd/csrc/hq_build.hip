// hq_build.hip -- octree construction on the GPU from sorted octant-path keys.
//
// Reproduces the topology of the reference insertion octree
// (fortran/octree_default.f90 init :320-387, add :547-580, add_to_subnode
// :515-545, allocate_subnode :481-513) and its moments (incorporate_particles
// :582-628, evaluate_multipole :744-882) bit for bit, without pointer
// chasing:
//   1. k_bbox_*      root cube: bbox centre, side 1.2 * extent (:345-347, :424-443)
//   2. k_keys        per particle, the octant path of 25 levels computed with
//                    the reference's floating-point centre recurrence
//                    (c + sign*h/2, x >= cx, y < cy, z < cz), so boundary
//                    cases fall exactly as in the reference
//   3. radix sort    stable LSD radix sort (8-bit digits, wavefront ballot
//                    ranking) of the 63-bit high key word above its lowest
//                    16 bits; runs equal there are then ordered by the rest
//                    of the high word and the 12-bit low word (levels 22-25)
//   4. k_node_*      nodes from the sorted keys: with D_i the number of
//                    common levels of particles i and i+1, particle i starts
//                    the nodes of depths D_{i-1}+1 .. max(D_{i-1},D_i)+1 (the
//                    last one its leaf).  Two ballot passes place every node
//                    in level order; children of a node are contiguous.
//   5. k_parents,    parent pointers and child masks; "next" (skip) pointers
//      k_next        for the stackless per-target walk, by an ancestor walk
//                    per node (a launched build leaves them to that walk)
//   6. k_upward      level-synchronous M2M from depth 25 to 0, a wavefront per
//                    64 nodes staging their children in LDS; children summed
//                    in the reference's child order (octant index 1..8), with
//                    contraction off, so mass/COM/quad/oct match bit for bit
#include <math.h>

#include <algorithm>
#include <utility>

#include "hq_tree.h"

namespace hq {

// geom buffer layout (doubles)
enum { G_CX = 0, G_CY, G_CZ, G_SIDE, G_HALF, G_XMIN, G_XMAX, G_YMIN, G_YMAX, G_ZMIN, G_ZMAX };

// ------------------------------------------------------------------ bbox
// A depth-d node has side ldexp(side, -d): add_node gets the parent's
// halfside as its side (:508), and L2 = side^2 (:475).  Lane d of the calling
// wavefront writes depth d of a system's table g.
__device__ __forceinline__ void geom_table(double* __restrict__ g, double side, int d) {
#pragma clang fp contract(off)
    if (d <= MAX_DEPTH) {
        const double sd = ldexp(side, -d);
        g[S_L2 + d] = sd * sd;
        g[S_HALFD + d] = sd / 2.0;
    }
}

// Bounding box of one system: each workgroup reduces a grid-stride share to
// part[6 * block]; k_bbox_final reduces the shares.  (Finishing in the last
// workgroup to arrive instead took 74 us at 1M: the device-wide fence each
// workgroup needs before its arrival writes back its L2.)
__global__ __launch_bounds__(256) void k_bbox_partial(const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z, int64_t n,
                                                      double* __restrict__ part,
                                                      int64_t* __restrict__ status) {
#pragma clang fp contract(off)
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool bad = false;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            bad |= !isfinite(v[c]);
            mn[c] = fmin(mn[c], v[c]);
            mx[c] = fmax(mx[c], v[c]);
        }
    }
    __shared__ double sh[4][6];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mn[c] = fmin(mn[c], __shfl_xor(mn[c], o, 64));
            mx[c] = fmax(mx[c], __shfl_xor(mx[c], o, 64));
        }
    }
    if (__any(bad) && __lane_id() == 0) atomicOr((unsigned long long*)&status[ST_FLAGS], 1ull);
    int w = threadIdx.x >> 6;
    if (__lane_id() == 0)
        for (int c = 0; c < 3; ++c) { sh[w][2 * c] = mn[c]; sh[w][2 * c + 1] = mx[c]; }
    __syncthreads();
    if (threadIdx.x < 6) {
        double v = sh[0][threadIdx.x];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
            v = (threadIdx.x & 1) ? fmax(v, sh[k][threadIdx.x]) : fmin(v, sh[k][threadIdx.x]);
        part[blockIdx.x * 6 + threadIdx.x] = v;
    }
}

// The shares of k_bbox_partial (one wavefront): the root cube --
// node_size_center (octree_default.f90:424-443) and the root add_node
// (:345-347) -- and the per-depth tables.  min and max are exact, so the
// result does not depend on the order of the shares.
__global__ __launch_bounds__(64) void k_bbox_final(const double* __restrict__ part, int nb,
                                                   double* __restrict__ g) {
#pragma clang fp contract(off)
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += 64)
        for (int c = 0; c < 3; ++c) {
            mn[c] = fmin(mn[c], part[6 * b + 2 * c]);
            mx[c] = fmax(mx[c], part[6 * b + 2 * c + 1]);
        }
    for (int c = 0; c < 3; ++c)
        for (int o = 32; o > 0; o >>= 1) {
            mn[c] = fmin(mn[c], __shfl_xor(mn[c], o, 64));
            mx[c] = fmax(mx[c], __shfl_xor(mx[c], o, 64));
        }
    double ext = mx[0] - mn[0];
    ext = fmax(ext, mx[1] - mn[1]);
    ext = fmax(ext, mx[2] - mn[2]);
    const double side = 1.2 * ext;
    double* g0 = g + G_SYS;  // system 0 of the per-system table
    if (threadIdx.x == 0) {
        g[G_CX] = 0.5 * (mn[0] + mx[0]);
        g[G_CY] = 0.5 * (mn[1] + mx[1]);
        g[G_CZ] = 0.5 * (mn[2] + mx[2]);
        g[G_SIDE] = side;
        g[G_HALF] = side / 2.0;
        g[G_XMIN] = mn[0]; g[G_XMAX] = mx[0];
        g[G_YMIN] = mn[1]; g[G_YMAX] = mx[1];
        g[G_ZMIN] = mn[2]; g[G_ZMAX] = mx[2];
        g0[S_CX] = g[G_CX]; g0[S_CY] = g[G_CY]; g0[S_CZ] = g[G_CZ];
        g0[S_SIDE] = side; g0[S_HALF] = side / 2.0;
    }
    geom_table(g0, side, threadIdx.x);
}

// One workgroup per system of a batch: the system's root cube (same
// arithmetic as k_bbox) and per-depth tables into sg[SG * s + ...].
__global__ __launch_bounds__(256) void k_sys_bbox(const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  const double* __restrict__ z,
                                                  const int64_t* __restrict__ off,
                                                  double* __restrict__ sg,
                                                  int64_t* __restrict__ status) {
#pragma clang fp contract(off)
    const int sys = blockIdx.x;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool bad = false;
    for (int64_t i = off[sys] + threadIdx.x; i < off[sys + 1]; i += blockDim.x) {
        const double v[3] = {x[i], y[i], z[i]};
        for (int c = 0; c < 3; ++c) {
            bad |= !isfinite(v[c]);
            mn[c] = fmin(mn[c], v[c]);
            mx[c] = fmax(mx[c], v[c]);
        }
    }
    __shared__ double sh[4][6];
    for (int c = 0; c < 3; ++c)
        for (int o = 32; o > 0; o >>= 1) {
            mn[c] = fmin(mn[c], __shfl_xor(mn[c], o, 64));
            mx[c] = fmax(mx[c], __shfl_xor(mx[c], o, 64));
        }
    if (__any(bad) && __lane_id() == 0) atomicOr((unsigned long long*)&status[ST_FLAGS], 1ull);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0)
        for (int c = 0; c < 3; ++c) { sh[w][2 * c] = mn[c]; sh[w][2 * c + 1] = mx[c]; }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    for (int k = 1; k < 4; ++k)
        for (int c = 0; c < 3; ++c) {
            mn[c] = fmin(mn[c], sh[k][2 * c]);
            mx[c] = fmax(mx[c], sh[k][2 * c + 1]);
        }
    double* g = sg + SG * sys;
    double ext = mx[0] - mn[0];
    ext = fmax(ext, mx[1] - mn[1]);
    ext = fmax(ext, mx[2] - mn[2]);
    const double side = 1.2 * ext;
    if (threadIdx.x == 0) {
        g[S_CX] = 0.5 * (mn[0] + mx[0]);
        g[S_CY] = 0.5 * (mn[1] + mx[1]);
        g[S_CZ] = 0.5 * (mn[2] + mx[2]);
        g[S_SIDE] = side;
        g[S_HALF] = side / 2.0;
    }
    geom_table(g, side, threadIdx.x);
}

// system id of every particle of a batch (original order)
__global__ __launch_bounds__(256) void k_sysid(const int64_t* __restrict__ off,
                                               int* __restrict__ sysid) {
    const int sys = blockIdx.x;
    for (int64_t i = off[sys] + threadIdx.x; i < off[sys + 1]; i += blockDim.x) sysid[i] = sys;
}

// ------------------------------------------------------------------ keys
// One particle per thread: the reference decides the octant with x >= cx,
// y < cy, z < cz (:527-531) and computes the child centre as c + sign*h/2
// from the parent's halfside h (:495-505).  The digit stored here is
// 4*xb + 2*yb + zb with xb = (x >= cx), yb = !(y < cy), zb = !(z < cz).
__global__ __launch_bounds__(256) void k_keys(const double* __restrict__ x,
                                              const double* __restrict__ y,
                                              const double* __restrict__ z, int64_t n,
                                              const double* __restrict__ sg,
                                              const int* __restrict__ sysid,
                                              uint64_t* __restrict__ khi,
                                              uint32_t* __restrict__ klo,
                                              int* __restrict__ idx,
                                              uint64_t* __restrict__ khio,
                                              const double* __restrict__ m,
                                              double4* __restrict__ rec) {
#pragma clang fp contract(off)
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    double px = x[p], py = y[p], pz = z[p];
    const int sp = sysid ? sysid[p] : 0;
    const double* g = sg + SG * sp;
    double cx = g[S_CX], cy = g[S_CY], cz = g[S_CZ], h = g[S_HALF];
    uint64_t hi = 0;
    uint32_t lo = 0;
#pragma unroll
    for (int d = 1; d <= MAX_DEPTH; ++d) {
        int xb = px >= cx;
        int yb = !(py < cy);
        int zb = !(pz < cz);
        double hh = h / 2.0;
        cx = cx + (xb ? hh : -hh);
        cy = cy + (yb ? hh : -hh);
        cz = cz + (zb ? hh : -hh);
        h = hh;
        uint32_t dig = (uint32_t)(4 * xb + 2 * yb + zb);
        if (d <= HI_LEVELS) hi = (hi << 3) | dig;
        else lo = (lo << 3) | dig;
    }
    khi[p] = hi;
    klo[p] = lo;
    idx[p] = (int)p;
    if (khio) khio[p] = hi;
    // one packed record per particle, so that the gather after the sort reads
    // one line per particle instead of one per array
    rec[p] = make_double4(px, py, pz, m[p]);
}

// ------------------------------------------------------------------ radix sort
// A workgroup sorts a tile of RS_THREADS * ITEMS keys; small inputs use short
// tiles so that enough workgroups run (rs_items below).
constexpr int RS_THREADS = 256;

template <int ITEMS>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t* __restrict__ keys,
                                                        int64_t n, int shift,
                                                        uint32_t* __restrict__ hist, int nblk) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * (RS_THREADS * ITEMS);
#pragma unroll 4
    for (int it = 0; it < ITEMS; ++it) {
        int64_t i = base + it * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
}

// Per digit d (one workgroup each): exclusive scan over the tiles of the
// counts hist[d * nblk + b] in place, the digit's total to dtot[d].  The
// scatter adds the totals of the smaller digits.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scan(uint32_t* __restrict__ hist, int nblk,
                                                        uint32_t* __restrict__ dtot) {
    __shared__ uint32_t ws[RS_THREADS / 64];
    uint32_t* __restrict__ h = hist + (int64_t)blockIdx.x * nblk;
    uint32_t carry = 0;
    for (int b0 = 0; b0 < nblk; b0 += RS_THREADS) {
        const int i = b0 + threadIdx.x;
        const uint32_t v = i < nblk ? h[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, ws, tot);
        if (i < nblk) h[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

// Stable scatter: wave w of the workgroup owns the contiguous slice
// [w*64*ITEMS, (w+1)*64*ITEMS) of the tile and walks it 64 keys at a time;
// keys of equal digit inside one 64-key step are ranked by ballot (lanes in
// index order), across steps by per-wave running counters, across waves by a
// prefix over the per-wave totals.  The tile is then placed in LDS in sorted
// order and written out in that order, so that consecutive lanes write
// consecutive addresses of one digit's run (1.2x faster than each key written
// from its own lane).
template <int ITEMS>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(
    const uint64_t* __restrict__ kin, const int* __restrict__ vin, uint64_t* __restrict__ kout,
    int* __restrict__ vout, int64_t n, int shift, const uint32_t* __restrict__ hist, int nblk,
    const uint32_t* __restrict__ dtot) {
    constexpr int TILE = RS_THREADS * ITEMS;
    __shared__ uint32_t wcnt[RS_THREADS / 64][256];
    __shared__ uint32_t goff[256];
    __shared__ uint32_t lstart[256];
    __shared__ uint32_t ws[RS_THREADS / 64];
    __shared__ uint64_t skey[TILE];
    __shared__ int sval[TILE];
    const int w = threadIdx.x >> 6;
    const int lane = __lane_id();
    for (int d = threadIdx.x; d < 256 * (RS_THREADS / 64); d += RS_THREADS) (&wcnt[0][0])[d] = 0;
    uint32_t all;
    const uint32_t below = block_excl_scan(dtot[threadIdx.x], ws, all);  // keys of smaller digits
    const uint32_t gd = below + hist[(int64_t)threadIdx.x * nblk + blockIdx.x];
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * TILE;
    const int64_t base = t0 + (int64_t)w * (ITEMS * 64);
    uint64_t key[ITEMS];
    int val[ITEMS];
    uint32_t rank[ITEMS];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        int64_t i = base + it * 64 + lane;
        bool ok = i < n;
        key[it] = ok ? kin[i] : ~0ull;
        val[it] = ok ? vin[i] : 0;
        uint32_t dg = (uint32_t)((key[it] >> shift) & 255u);
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bb = __ballot((dg >> b) & 1u);
            peers &= ((dg >> b) & 1u) ? bb : ~bb;
        }
        uint32_t before = ok ? wcnt[w][dg] : 0u;
        __builtin_amdgcn_wave_barrier();
        rank[it] = before + (uint32_t)__popcll(peers & lt);
        bool leader = ok && (peers & lt) == 0;
        if (leader) wcnt[w][dg] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        const int d = threadIdx.x;  // one digit per thread
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < RS_THREADS / 64; ++k) {
            uint32_t c = wcnt[k][d];
            wcnt[k][d] = run;
            run += c;
        }
        uint32_t tt;
        const uint32_t ls = block_excl_scan(run, ws, tt);  // the digit's start in the tile
        lstart[d] = ls;
        goff[d] = gd - ls;  // output position of tile position ls is gd
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        int64_t i = base + it * 64 + lane;
        if (i < n) {
            uint32_t dg = (uint32_t)((key[it] >> shift) & 255u);
            uint32_t lp = lstart[dg] + wcnt[w][dg] + rank[it];
            skey[lp] = key[it];
            sval[lp] = val[it];
        }
    }
    __syncthreads();
    const int cnt = (int)min<int64_t>(TILE, n - t0);
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const int i = it * RS_THREADS + threadIdx.x;
        if (i < cnt) {
            const uint64_t k = skey[i];
            // goff = gd - ls wraps modulo 2^32 when gd < ls; the sum is the
            // position (< n < 2^32) again
            const int64_t pos = (int64_t)(uint32_t)(goff[(k >> shift) & 255u] + (uint32_t)i);
            kout[pos] = k;
            vout[pos] = sval[i];
        }
    }
}

// The radix sort of the tree's keys orders them by the high word without its
// lowest TIE_SKIP_BITS bits (levels 1..16 and a third of 17: two passes fewer).
// Runs of keys equal in those bits (particles sharing a cube 2^-16 of the root
// wide, rare and short) are then ordered by the rest of the high word and the
// low word (levels 22..25), stably, one run per thread: the order of a full
// sort by (high word, low word).  A run longer than TIE_RUN_MAX flags the
// build, which then sorts by the full key instead (full_key_sort).
constexpr int TIE_SKIP_BITS = 16;
constexpr int64_t TIE_RUN_MAX = 64;
__global__ void k_tie_fix(uint64_t* __restrict__ khi, const uint32_t* __restrict__ klo,
                          int* __restrict__ idx, const int* __restrict__ sysid, int64_t n,
                          int64_t* __restrict__ status) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i + 1 >= n) return;
    // the sorted bits of a run's keys are equal, so these tests read the same
    // whatever the run's own thread has moved so far
    auto same = [&](int64_t a, int64_t b) {
        return (khi[a] >> TIE_SKIP_BITS) == (khi[b] >> TIE_SKIP_BITS) &&
               (!sysid || sysid[idx[a]] == sysid[idx[b]]);
    };
    if (!same(i, i + 1)) return;
    if (i > 0 && same(i - 1, i)) return;
    int64_t j = i + 1;
    while (j + 1 < n && same(j + 1, i)) ++j;
    if (j - i + 1 > TIE_RUN_MAX) {
        atomicOr((unsigned long long*)&status[ST_FLAGS], 4ull);
        return;
    }
    for (int64_t a = i + 1; a <= j; ++a) {
        const int v = idx[a];
        const uint64_t kh = khi[a];
        const uint32_t kl = klo[v];
        int64_t b = a - 1;
        while (b >= i && (khi[b] > kh || (khi[b] == kh && klo[idx[b]] > kl))) {
            khi[b + 1] = khi[b];
            idx[b + 1] = idx[b];
            --b;
        }
        khi[b + 1] = kh;
        idx[b + 1] = v;
    }
}

// Sorted bodies from the packed records.  The system of sorted slot i is the
// one whose input range holds i (the sort leaves systems in order, each with
// as many slots as bodies), and the low key word (levels 22..25) is the
// comparisons of k_keys below the level-21 cube, whose centre follows from
// the high word's digits by the same c +- h/2 steps: both are recomputed here
// rather than gathered.
__global__ __launch_bounds__(256) void k_gather(const int* __restrict__ idx,
                                                const double4* __restrict__ rec,
                                                const uint64_t* __restrict__ khi,
                                                const double* __restrict__ sg,
                                                const int64_t* __restrict__ sysoff, int nsys,
                                                int64_t n, double* __restrict__ sx,
                                                double* __restrict__ sy, double* __restrict__ sz,
                                                double* __restrict__ sm,
                                                uint32_t* __restrict__ sklo,
                                                int* __restrict__ ssys,
                                                double4* __restrict__ sbody) {
#pragma clang fp contract(off)
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = idx[i];
    const double4 r = rec[p];
    int sys = 0;
    if (nsys > 1) {  // last system with sysoff[sys] <= i
        int lo = 0, hi = nsys - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sysoff[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        sys = lo;
    }
    const double* g = sg + SG * sys;
    double cx = g[S_CX], cy = g[S_CY], cz = g[S_CZ], h = g[S_HALF];
    const uint64_t hk = khi[i];
#pragma unroll
    for (int d = 1; d <= HI_LEVELS; ++d) {
        const uint32_t dig = (uint32_t)(hk >> (3 * (HI_LEVELS - d))) & 7u;
        const double hh = h / 2.0;
        cx = cx + ((dig & 4u) ? hh : -hh);
        cy = cy + ((dig & 2u) ? hh : -hh);
        cz = cz + ((dig & 1u) ? hh : -hh);
        h = hh;
    }
    uint32_t lo = 0;
#pragma unroll
    for (int d = HI_LEVELS + 1; d <= MAX_DEPTH; ++d) {
        const int xb = r.x >= cx;
        const int yb = !(r.y < cy);
        const int zb = !(r.z < cz);
        const double hh = h / 2.0;
        cx = cx + (xb ? hh : -hh);
        cy = cy + (yb ? hh : -hh);
        cz = cz + (zb ? hh : -hh);
        h = hh;
        lo = (lo << 3) | (uint32_t)(4 * xb + 2 * yb + zb);
    }
    sbody[i] = r;
    sx[i] = r.x;
    sy[i] = r.y;
    sz[i] = r.z;
    sm[i] = r.w;
    sklo[i] = lo;
    ssys[i] = sys;
}

// A batch is sorted by octant key, then stably by system: the second sort's
// key is the system id of each sorted slot, and afterwards the octant keys are
// read back from the copy k_keys kept in input order.
__global__ __launch_bounds__(256) void k_syskey(const int* __restrict__ idx,
                                                const int* __restrict__ sysid, int64_t n,
                                                uint64_t* __restrict__ key) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) key[i] = (uint64_t)sysid[idx[i]];
}
__global__ __launch_bounds__(256) void k_restore_keys(const int* __restrict__ idx,
                                                      const uint64_t* __restrict__ khio, int64_t n,
                                                      uint64_t* __restrict__ khi) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) khi[i] = khio[idx[i]];
}

// D_i = number of common leading levels of the keys of sorted particles i, i+1.
__global__ __launch_bounds__(256) void k_lcp(const uint64_t* __restrict__ khi,
                                             const uint32_t* __restrict__ klo,
                                             const int* __restrict__ ssys, int64_t n,
                                             int8_t* __restrict__ D) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int d = -1;
    if (i + 1 < n && ssys[i] == ssys[i + 1]) {
        uint64_t a = khi[i] ^ khi[i + 1];
        if (a) {
            d = (__clzll((long long)a) - 1) / 3;
        } else {
            uint32_t u = klo[i] ^ klo[i + 1];
            d = u ? HI_LEVELS + (__clz((int)u) - 20) / 3 : MAX_DEPTH;
        }
    }
    D[i] = (int8_t)d;
}

// ------------------------------------------------------------------ nodes
struct PartRange {
    int lo, e, hi;  // node depths lo..e start at this particle; leaf depth hi
};

__device__ __forceinline__ PartRange part_range(const int8_t* __restrict__ D, int64_t i,
                                                int64_t n) {
    PartRange r;
    int dm = i > 0 ? (int)D[i - 1] : -1;
    int dp = (int)D[i];  // D[n-1] = -1
    r.lo = dm + 1;
    r.hi = (dm > dp ? dm : dp) + 1;
    r.e = r.hi <= MAX_DEPTH ? r.hi : dp;
    return r;
}

constexpr int NB_THREADS = 256;
constexpr int NDEP = MAX_DEPTH + 1;  // depths 0..25

__global__ __launch_bounds__(NB_THREADS) void k_node_count(const int8_t* __restrict__ D,
                                                           int64_t n,
                                                           uint32_t* __restrict__ bcnt, int nblk,
                                                           int64_t* __restrict__ status) {
    __shared__ uint32_t wc[NB_THREADS / 64][NDEP];
    int64_t i = blockIdx.x * (int64_t)NB_THREADS + threadIdx.x;
    bool ok = i < n;
    PartRange r = ok ? part_range(D, i, n) : PartRange{1, 0, 0};
    int w = threadIdx.x >> 6;
    for (int d = 0; d < NDEP; ++d) {
        uint64_t msk = __ballot(ok && r.lo <= d && d <= r.e);
        if (__lane_id() == 0) wc[w][d] = (uint32_t)__popcll(msk);
    }
    uint64_t lostm = __ballot(ok && r.hi > MAX_DEPTH);
    if (__lane_id() == 0 && lostm) atomicAdd((unsigned long long*)&status[ST_LOST],
                                             (unsigned long long)__popcll(lostm));
    __syncthreads();
    if (threadIdx.x < NDEP) {
        uint32_t s = 0;
        for (int k = 0; k < NB_THREADS / 64; ++k) s += wc[k][threadIdx.x];
        bcnt[(int64_t)threadIdx.x * nblk + blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(NB_THREADS) void k_node_emit(
    const int8_t* __restrict__ D, const uint64_t* __restrict__ khi,
    const uint32_t* __restrict__ klo, const double* __restrict__ sg,
    const int* __restrict__ ssys, int64_t n,
    const uint32_t* __restrict__ boff, int nblk, int64_t cap, uint8_t* __restrict__ ndepth,
    int8_t* __restrict__ ntype, int* __restrict__ nfirst, int* __restrict__ ncstart,
    uint8_t* __restrict__ ndigit, double* __restrict__ ncx, double* __restrict__ ncy,
    double* __restrict__ ncz, uint64_t* __restrict__ nkhi, uint32_t* __restrict__ nklo,
    int64_t* __restrict__ status) {
#pragma clang fp contract(off)
    __shared__ uint32_t wpre[NB_THREADS / 64][NDEP + 1];
    __shared__ uint32_t base[NDEP + 1];
    int64_t i = blockIdx.x * (int64_t)NB_THREADS + threadIdx.x;
    bool ok = i < n;
    PartRange r = ok ? part_range(D, i, n) : PartRange{1, 0, 0};
    int w = threadIdx.x >> 6;
    for (int d = 0; d < NDEP; ++d) {
        uint64_t msk = __ballot(ok && r.lo <= d && d <= r.e);
        if (__lane_id() == 0) wpre[w][d] = (uint32_t)__popcll(msk);
    }
    if (__lane_id() == 0) wpre[w][NDEP] = 0;
    __syncthreads();
    const uint32_t total = boff[(int64_t)NDEP * nblk];
    if (threadIdx.x <= NDEP) {
        int d = threadIdx.x;
        uint32_t run = 0;
        for (int k = 0; k < NB_THREADS / 64; ++k) {
            uint32_t c = wpre[k][d];
            wpre[k][d] = run;
            run += c;
        }
        base[d] = d < NDEP ? boff[(int64_t)d * nblk + blockIdx.x] : total;
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        status[ST_NNODES] = total;
        if ((int64_t)total > cap) atomicOr((unsigned long long*)&status[ST_FLAGS], 2ull);
    }
    const uint64_t lt = lanemask_lt();
    uint64_t hi_key = ok ? khi[i] : 0;
    uint32_t lo_key = ok ? klo[i] : 0;
    const double* g = sg + SG * (ok ? ssys[i] : 0);
    double cx = g[S_CX], cy = g[S_CY], cz = g[S_CZ], h = g[S_HALF];
    uint64_t mcur = __ballot(ok && r.lo <= 0 && 0 <= r.e);
    for (int d = 0; d < NDEP; ++d) {
        uint64_t mnext = (d + 1 < NDEP) ? __ballot(ok && r.lo <= d + 1 && d + 1 <= r.e) : 0ull;
        if (ok && r.lo <= d && d <= r.e) {
            int64_t k = (int64_t)base[d] + wpre[w][d] + __popcll(mcur & lt);
            int64_t cs = (d + 1 < NDEP) ? (int64_t)base[d + 1] + wpre[w][d + 1] + __popcll(mnext & lt)
                                        : (int64_t)total;
            if (k < cap) {
                ndepth[k] = (uint8_t)d;
                ntype[k] = (int8_t)(d == r.hi ? 1 : 2);
                nfirst[k] = (int)i;
                ncstart[k] = (int)cs;
                uint32_t dig = 0;
                if (d >= 1 && d <= HI_LEVELS) dig = (uint32_t)(hi_key >> (3 * (HI_LEVELS - d))) & 7u;
                else if (d > HI_LEVELS) dig = (lo_key >> (3 * (MAX_DEPTH - d))) & 7u;
                ndigit[k] = (uint8_t)dig;
                ncx[k] = cx;
                ncy[k] = cy;
                ncz[k] = cz;
                uint64_t kh = d >= HI_LEVELS ? hi_key
                                             : (d == 0 ? 0ull : hi_key & (~0ull << (3 * (HI_LEVELS - d))));
                uint32_t kl = d <= HI_LEVELS ? 0u : (lo_key & (0xFFFu << (3 * (MAX_DEPTH - d))) & 0xFFFu);
                nkhi[k] = kh;
                nklo[k] = kl;
            }
        }
        // descend one level along this particle's own path (same recurrence as k_keys)
        if (d + 1 < NDEP) {
            int dn = d + 1;
            uint32_t dig = dn <= HI_LEVELS ? (uint32_t)(hi_key >> (3 * (HI_LEVELS - dn))) & 7u
                                           : (lo_key >> (3 * (MAX_DEPTH - dn))) & 7u;
            double hh = h / 2.0;
            cx = cx + ((dig & 4u) ? hh : -hh);
            cy = cy + ((dig & 2u) ? hh : -hh);
            cz = cz + ((dig & 1u) ? hh : -hh);
            h = hh;
        }
        mcur = mnext;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (int64_t)total <= cap) ncstart[total] = (int)total;
}

__global__ void k_levels(const uint32_t* __restrict__ boff, int nblk, int64_t* __restrict__ status) {
    int d = threadIdx.x;
    if (d <= NDEP) status[ST_LEVEL0 + d] = d < NDEP ? boff[(int64_t)d * nblk] : boff[(int64_t)NDEP * nblk];
}

// ------------------------------------------------------------------ moments
// Arguments of the upward pass (node store and sorted particles).
struct UpArgs {
    int multipole;
    const int8_t* ntype;
    const int* nfirst;
    const int* ncstart;
    const uint8_t* nmask;
    const uint8_t* ntwig;
    const double4* sbody;  // sorted bodies as (x, y, z, m)
    double *nmass, *ncomx, *ncomy, *ncomz, *nquad, *noct;
    int* ncount;
    int64_t nn;  // stride of the moment arrays (the capacity)
};

// Children of one round of parents, staged in LDS by the wavefront: every
// child's mass, centre of mass and count (slot i = the round's i-th child in
// level order), and the moments of its twig children only (compacted; tslot
// maps a child's slot to its twig slot, -1 for a leaf).
constexpr int UP_KIDS = 128, UP_TWIGS = 64;
template <int MP>
struct UpStage {
    double m[UP_KIDS], x[UP_KIDS], y[UP_KIDS], z[UP_KIDS];
    int n[UP_KIDS];
    int8_t tslot[UP_KIDS];
    int tnode[UP_TWIGS];
    double q[MP >= 4 ? 6 : 1][MP >= 4 ? UP_TWIGS : 1];
    double o[MP > 4 ? 10 : 1][MP > 4 ? UP_TWIGS : 1];
};

// Segment size (log2, 8..64 nodes) of the upward pass at a level of `nodes`
// nodes whose children number `kids`: the largest that keeps a segment's
// children near 50 on average (at the top levels, 5-8 us per level against
// 13-15 us with segments of twice as many children).
__host__ __device__ inline int up_segsh(int64_t nodes, int64_t kids) {
    int sh = 6;
    while (sh > 3 && (kids << sh) > 50 * nodes) --sh;
    return sh;
}

// Moments of the nodes of one level, one 64-thread workgroup (one wavefront)
// per segment of 2^segsh consecutive nodes.  A segment is done in rounds: a
// round takes the longest run of its remaining nodes with at most UP_KIDS
// children and UP_TWIGS twig children (at least 8 nodes).  The lanes load the
// round's children -- mass, centre of mass and count of every child, the
// moments of the twig children, consecutive lanes at consecutive addresses,
// all in flight together -- into LDS (which children are twigs comes from
// the parents' masks, so no load waits on another); lane j then forms node
// j's moments from its children there, and loads its body if the node is a
// leaf.  The sums are those of evaluate_multipole (fortran/octree_default.f90
// :744-882) and incorporate_particles (:582-628): children in the reference's
// child order r = 0..7 (octant index 1 + x + 2(1-y) + 4(1-z), :527-531, of
// digit 4x + 2y + z), contraction off, a leaf child's (zero) higher moments
// skipped, so every bit matches.  Bounds come from (a, b), or from lv[depth],
// lv[depth + 1] on the device when lv is given (graph mode, which also takes
// the segment size from the level sizes); the workgroups stride over the
// segments.
template <int MP>
__global__ __launch_bounds__(64) void k_upward(int64_t a, int64_t b,
                                               const int64_t* __restrict__ lv, int depth,
                                               int segsh, UpArgs u) {
#pragma clang fp contract(off)
    __shared__ UpStage<MP> st;
    if (lv) {
        a = lv[depth];
        b = lv[depth + 1];
        segsh = up_segsh(b - a, depth + 2 <= MAX_DEPTH + 1 ? lv[depth + 2] - b : 0);
    }
    const int lane = threadIdx.x;
    const int64_t S = u.nn;
    const int64_t nseg = (b - a + (1 << segsh) - 1) >> segsh;
    for (int64_t seg = blockIdx.x; seg < nseg; seg += gridDim.x) {
        int64_t p0 = a + (seg << segsh);
        const int64_t pend = min(b, p0 + (1 << segsh));
        while (p0 < pend) {
            const int64_t p = p0 + lane;
            const bool act = p < pend;
            int type = 0, cs = 0, ce = 0, first = 0;
            unsigned present = 0u, tw = 0u;
            if (act) {
                type = u.ntype[p];
                cs = u.ncstart[p];
                ce = u.ncstart[p + 1];
                first = u.nfirst[p];
                present = u.nmask[p];
                tw = u.ntwig[p];
            }
            const int c0 = __shfl(cs, 0);
            const int tend = wave_incl_scan((int)__popc(tw));  // twig children up to this node
            // a prefix of the lanes (both counts grow with the lane)
            const int J = __popcll(__ballot(act && ce - c0 <= UP_KIDS && tend <= UP_TWIGS));
            const int nc = __shfl(ce, J - 1) - c0;
            const int ntw = __shfl(tend, J - 1);
            const bool mine = lane < J;
            double4 body = make_double4(0.0, 0.0, 0.0, 0.0);
            if (mine && type == 1) body = u.sbody[first];
            if (mine) {
                int ts = tend - (int)__popc(tw);
                for (int i = 0; i < ce - cs; ++i) {
                    const bool t = (tw >> i) & 1u;
                    st.tslot[cs - c0 + i] = (int8_t)(t ? ts : -1);
                    if (t) st.tnode[ts++] = cs + i;
                }
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < UP_KIDS / 64; ++h) {
                const int i = lane + 64 * h;
                if (i < nc) {
                    const int64_t c = c0 + i;
                    st.m[i] = u.nmass[c];
                    st.x[i] = u.ncomx[c];
                    st.y[i] = u.ncomy[c];
                    st.z[i] = u.ncomz[c];
                    st.n[i] = u.ncount[c];
                }
            }
            if (MP >= 4 && lane < ntw) {
                const int64_t c = st.tnode[lane];
#pragma unroll
                for (int j = 0; j < 6; ++j) st.q[j][lane] = u.nquad[j * S + c];
                if (MP > 4)
#pragma unroll
                    for (int j = 0; j < 10; ++j) st.o[MP > 4 ? j : 0][lane] = u.noct[j * S + c];
            }
            __syncthreads();
            if (mine) {
                double Q[6] = {0, 0, 0, 0, 0, 0};
                double O[14] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
                double mass = 0.0, qx = 0.0, qy = 0.0, qz = 0.0;
                int count = 0;
                if (type == 1) {
                    mass = body.w;
                    qx = body.x;
                    qy = body.y;
                    qz = body.z;
                    count = 1;
                } else if (depth == MAX_DEPTH) {
                    // max-depth twig: evaluate_multipole skips it (:772-774,
                    // stays 0); incorporate_particles divides 0/0 (:614-616)
                    if (MP == 1) {
                        qx = qx / mass;
                        qy = qy / mass;
                        qz = qz / mass;
                    }
                } else {
                    const int base = cs - c0;
                    int sl[8];
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        const unsigned d = 4u * (unsigned)(r & 1) + 2u * (1u - ((r >> 1) & 1u)) +
                                           (1u - ((r >> 2) & 1u));
                        sl[r] = ((present >> d) & 1u) ? base + __popc(present & ((1u << d) - 1u))
                                                      : -1;
                    }
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
                        const int c = sl[r];
                        if (c < 0) continue;
                        const double mc = st.m[c];
                        mass = mass + mc;
                        qx = qx + mc * st.x[c];
                        qy = qy + mc * st.y[c];
                        qz = qz + mc * st.z[c];
                        count += st.n[c];
                    }
                    qx = qx / mass;
                    qy = qy / mass;
                    qz = qz / mass;
                    if (MP >= 4) {
                        constexpr bool oct = MP > 4;
#pragma unroll
                        for (int r = 0; r < 8; ++r) {
                            const int c = sl[r];
                            if (c < 0) continue;
                            const double pm = st.m[c];
                            const double dxi = st.x[c] - qx;
                            const double dyi = st.y[c] - qy;
                            const double dzi = st.z[c] - qz;
                            // O[10..13] are recomputed from O[0..9] below, so
                            // children's are not summed
                            const int ts = st.tslot[c];
                            const bool twig = ts >= 0;
                            double Qc[6] = {0, 0, 0, 0, 0, 0};
                            if (twig) {
#pragma unroll
                                for (int j = 0; j < 6; ++j) Qc[j] = st.q[j][ts];
#pragma unroll
                                for (int j = 0; j < 6; ++j) Q[j] = Q[j] + Qc[j];
                                if (oct)
#pragma unroll
                                    for (int j = 0; j < 10; ++j) O[j] = O[j] + st.o[oct ? j : 0][ts];
                            }
                            Q[0] = Q[0] + pm * (dxi * dxi);
                            Q[1] = Q[1] + pm * (dyi * dyi);
                            Q[2] = Q[2] + pm * (dzi * dzi);
                            Q[3] = Q[3] + pm * dxi * dyi;
                            Q[4] = Q[4] + pm * dxi * dzi;
                            Q[5] = Q[5] + pm * dyi * dzi;
                            if (oct) {
                                O[0] = O[0] + pm * dxi * (dxi * dxi);
                                O[1] = O[1] + pm * dxi * (dyi * dyi);
                                O[2] = O[2] + pm * dxi * (dzi * dzi);
                                O[3] = O[3] + pm * dyi * (dxi * dxi);
                                O[4] = O[4] + pm * dyi * (dyi * dyi);
                                O[5] = O[5] + pm * dyi * (dzi * dzi);
                                O[6] = O[6] + pm * dzi * (dxi * dxi);
                                O[7] = O[7] + pm * dzi * (dyi * dyi);
                                O[8] = O[8] + pm * dzi * (dzi * dzi);
                                O[9] = O[9] + pm * dxi * dyi * dzi;
                                if (twig) {
                                    O[0] = O[0] + 3.0 * dxi * Qc[0];
                                    O[1] = O[1] + 2.0 * dyi * Qc[3] + dxi * Qc[1];
                                    O[2] = O[2] + 2.0 * dzi * Qc[4] + dxi * Qc[2];
                                    O[3] = O[3] + 2.0 * dxi * Qc[3] + dyi * Qc[0];
                                    O[4] = O[4] + 3.0 * dyi * Qc[1];
                                    O[5] = O[5] + 2.0 * dzi * Qc[5] + dyi * Qc[2];
                                    O[6] = O[6] + 2.0 * dxi * Qc[4] + dzi * Qc[0];
                                    O[7] = O[7] + 2.0 * dyi * Qc[5] + dzi * Qc[1];
                                    O[8] = O[8] + 3.0 * dzi * Qc[2];
                                    O[9] = O[9] + dxi * Qc[5] + dyi * Qc[4] + dzi * Qc[3];
                                }
                            }
                        }
                        if (oct) {
                            double s = 0.0;
                            for (int j = 0; j < 10; ++j) s = s + O[j];
                            O[10] = s;
                            O[11] = (O[0] + O[1]) + O[2];
                            O[12] = (O[3] + O[4]) + O[5];
                            O[13] = (O[6] + O[7]) + O[8];
                        }
                    }
                }
                u.nmass[p] = mass;
                u.ncomx[p] = qx;
                u.ncomy[p] = qy;
                u.ncomz[p] = qz;
                u.ncount[p] = count;
                if (MP >= 4)
#pragma unroll
                    for (int j = 0; j < 6; ++j) u.nquad[j * S + p] = Q[j];
                if (MP > 4)
#pragma unroll
                    for (int j = 0; j < 14; ++j) u.noct[j * S + p] = O[j];
            }
            __syncthreads();
            p0 += J;
        }
    }
}

// Thread-per-node form of the same sums, for levels whose nodes have more
// children than nodes (a thread's loads of its children are then mostly
// distinct and the arithmetic keeps its lanes busy): node k at `depth` from
// its children (or its body).
__device__ __forceinline__ void upward_node(const UpArgs& u, int64_t k, int depth) {
#pragma clang fp contract(off)
    const int multipole = u.multipole;
    const int8_t* __restrict__ ntype = u.ntype;
    const int* __restrict__ nfirst = u.nfirst;
    const int* __restrict__ ncstart = u.ncstart;
    const uint8_t* __restrict__ nmask = u.nmask;
    double* __restrict__ nmass = u.nmass;
    double* __restrict__ ncomx = u.ncomx;
    double* __restrict__ ncomy = u.ncomy;
    double* __restrict__ ncomz = u.ncomz;
    double* __restrict__ nquad = u.nquad;
    double* __restrict__ noct = u.noct;
    int* __restrict__ ncount = u.ncount;
    const int64_t nn = u.nn;
    double Q[6] = {0, 0, 0, 0, 0, 0};
    double O[14] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double mass = 0.0, qx = 0.0, qy = 0.0, qz = 0.0;
    int count = 0;
    if (ntype[k] == 1) {
        const double4 body = u.sbody[nfirst[k]];
        mass = body.w;
        qx = body.x;
        qy = body.y;
        qz = body.z;
        count = 1;
    } else if (depth == MAX_DEPTH) {
        // max-depth twig: evaluate_multipole skips it (:772-774, stays 0);
        // incorporate_particles divides 0/0 (:614-616)
        if (multipole == 1) {
            qx = qx / mass;
            qy = qy / mass;
            qz = qz / mass;
        }
    } else {
        // children in the reference's order r = 0..7 (octant index
        // 1 + x + 2(1-y) + 4(1-z), :527-531, of digit 4x + 2y + z): the present
        // digits as a mask (written with the parent pointers); the child of
        // digit d is the (number of present digits below d)-th of the
        // contiguous range.  Every child's record is loaded up front (absent
        // slots read the first child, unused), so the loads are in flight
        // together; the sums below keep the reference's order.
        const int cs = ncstart[k];
        const unsigned present = nmask[k];
        int cidx[8];
        double mcv[8], cxv[8], cyv[8], czv[8];
        int cnv[8];
        bool twv[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const unsigned d = 4u * (unsigned)(r & 1) + 2u * (1u - ((r >> 1) & 1u)) +
                               (1u - ((r >> 2) & 1u));
            cidx[r] = ((present >> d) & 1u) ? cs + __popc(present & ((1u << d) - 1u)) : -1;
            const int c = cidx[r] >= 0 ? cidx[r] : cs;
            mcv[r] = nmass[c];
            cxv[r] = ncomx[c];
            cyv[r] = ncomy[c];
            czv[r] = ncomz[c];
            cnv[r] = ncount[c];
            twv[r] = ntype[c] == 2;
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (cidx[r] < 0) continue;
            double mc = mcv[r];
            mass = mass + mc;
            qx = qx + mc * cxv[r];
            qy = qy + mc * cyv[r];
            qz = qz + mc * czv[r];
            count += cnv[r];
        }
        qx = qx / mass;
        qy = qy / mass;
        qz = qz / mass;
        if (multipole >= 4) {
            const bool oct = multipole > 4;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const int c = cidx[r];
                if (c < 0) continue;
                double pm = mcv[r];
                double dxi = cxv[r] - qx;
                double dyi = cyv[r] - qy;
                double dzi = czv[r] - qz;
                // a leaf child's moments are zero: only twigs' are read.  O[10..13]
                // are recomputed from O[0..9] below, so children's are not summed
                const bool twig = twv[r];
                double Qc[6] = {0, 0, 0, 0, 0, 0};
                if (twig) {
                    for (int j = 0; j < 6; ++j) Qc[j] = nquad[(int64_t)j * nn + c];
                    for (int j = 0; j < 6; ++j) Q[j] = Q[j] + Qc[j];
                    if (oct)
                        for (int j = 0; j < 10; ++j) O[j] = O[j] + noct[(int64_t)j * nn + c];
                }
                Q[0] = Q[0] + pm * (dxi * dxi);
                Q[1] = Q[1] + pm * (dyi * dyi);
                Q[2] = Q[2] + pm * (dzi * dzi);
                Q[3] = Q[3] + pm * dxi * dyi;
                Q[4] = Q[4] + pm * dxi * dzi;
                Q[5] = Q[5] + pm * dyi * dzi;
                if (oct) {
                    O[0] = O[0] + pm * dxi * (dxi * dxi);
                    O[1] = O[1] + pm * dxi * (dyi * dyi);
                    O[2] = O[2] + pm * dxi * (dzi * dzi);
                    O[3] = O[3] + pm * dyi * (dxi * dxi);
                    O[4] = O[4] + pm * dyi * (dyi * dyi);
                    O[5] = O[5] + pm * dyi * (dzi * dzi);
                    O[6] = O[6] + pm * dzi * (dxi * dxi);
                    O[7] = O[7] + pm * dzi * (dyi * dyi);
                    O[8] = O[8] + pm * dzi * (dzi * dzi);
                    O[9] = O[9] + pm * dxi * dyi * dzi;
                    if (twig) {
                        O[0] = O[0] + 3.0 * dxi * Qc[0];
                        O[1] = O[1] + 2.0 * dyi * Qc[3] + dxi * Qc[1];
                        O[2] = O[2] + 2.0 * dzi * Qc[4] + dxi * Qc[2];
                        O[3] = O[3] + 2.0 * dxi * Qc[3] + dyi * Qc[0];
                        O[4] = O[4] + 3.0 * dyi * Qc[1];
                        O[5] = O[5] + 2.0 * dzi * Qc[5] + dyi * Qc[2];
                        O[6] = O[6] + 2.0 * dxi * Qc[4] + dzi * Qc[0];
                        O[7] = O[7] + 2.0 * dyi * Qc[5] + dzi * Qc[1];
                        O[8] = O[8] + 3.0 * dzi * Qc[2];
                        O[9] = O[9] + dxi * Qc[5] + dyi * Qc[4] + dzi * Qc[3];
                    }
                }
            }
            if (oct) {
                double s = 0.0;
                for (int j = 0; j < 10; ++j) s = s + O[j];
                O[10] = s;
                O[11] = (O[0] + O[1]) + O[2];
                O[12] = (O[3] + O[4]) + O[5];
                O[13] = (O[6] + O[7]) + O[8];
            }
        }
    }
    nmass[k] = mass;
    ncomx[k] = qx;
    ncomy[k] = qy;
    ncomz[k] = qz;
    ncount[k] = count;
    if (multipole >= 4)
        for (int j = 0; j < 6; ++j) nquad[(int64_t)j * nn + k] = Q[j];
    if (multipole > 4)
        for (int j = 0; j < 14; ++j) noct[(int64_t)j * nn + k] = O[j];
}


// One launch per level, a thread per node: bounds (a, b), or lv[depth],
// lv[depth + 1] from the device (graph mode), grid-stride.
__global__ __launch_bounds__(128) void k_upward_t(int64_t a, int64_t b,
                                                  const int64_t* __restrict__ lv, int depth,
                                                  UpArgs u) {
    if (lv) {
        a = lv[depth];
        b = lv[depth + 1];
    }
    for (int64_t k = a + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < b;
         k += (int64_t)gridDim.x * blockDim.x)
        upward_node(u, k, depth);
}

// One launch per level: a segment per workgroup, or (graph mode) `grid`
// workgroups striding over the level's device-held bounds.
// Which form of the upward pass a level of `nodes` nodes with `kids`
// children takes: the staged one (k_upward) where most nodes are leaves or
// the level is small, the thread-per-node one (k_upward_t) elsewhere.
// Measured per level at 1M Plummer: the staged form 44 against 52 us at the
// densest leaf-heavy level and 5-8 against 7-11 us at the top levels; the
// thread-per-node form 60 against 64-76 us where nodes have 1.4 children.
static bool up_staged(int64_t nodes, int64_t kids) { return kids < nodes || nodes < 4096; }

// grid: for graph mode (lv given), in workgroups of the chosen form
static void launch_upward(const UpArgs& u, int64_t a, int64_t b, int64_t kids,
                          const int64_t* lv, int depth, bool staged, int grid, hipStream_t s) {
    if (!staged) {
        if (!lv) grid = grid_for(b - a, 128);
        k_upward_t<<<grid, 128, 0, s>>>(a, b, lv, depth, u);
        return;
    }
    int segsh = 6;
    if (!lv) {
        segsh = up_segsh(b - a, kids);
        grid = (int)((b - a + (1 << segsh) - 1) >> segsh);
    }
    if (u.multipole == 1) k_upward<1><<<grid, 64, 0, s>>>(a, b, lv, depth, segsh, u);
    else if (u.multipole == 4) k_upward<4><<<grid, 64, 0, s>>>(a, b, lv, depth, segsh, u);
    else k_upward<8><<<grid, 64, 0, s>>>(a, b, lv, depth, segsh, u);
}

// Graph mode: the checks the host makes after node emission, as flags that
// stay set over replayed steps; node count and level bounds for the rest of
// the build (0 when the tree is unusable, so that no later kernel reads it).
__global__ void k_graph_build_check(const int64_t* __restrict__ st, int64_t cap, int nsys,
                                    int64_t* __restrict__ gs) {
    if (threadIdx.x != 0) return;
    const int64_t fl = st[ST_FLAGS], nn = st[ST_NNODES];
    int64_t f = 0;
    if (fl & 1) f |= GF_INPUT;
    if (fl & 4) f |= GF_TIES;
    if ((fl & 2) || nn > cap) f |= GF_CAPACITY;
    if (st[ST_LEVEL0 + 1] != nsys) f |= GF_ROOTS;
    if (st[ST_LOST] > 0) f |= GF_LOST;
    gs[GS_STICKY] |= f;
    const bool ok = (f & (GF_INPUT | GF_TIES | GF_CAPACITY | GF_ROOTS)) == 0;
    gs[GS_NN] = ok ? nn : 0;
    for (int d = 0; d <= NDEP; ++d) gs[GS_LEVEL0 + d] = ok ? min(st[ST_LEVEL0 + d], cap) : 0;
}

// Launched build, after node emission: the node count of the pointer kernels
// queued while the host waits for the status (0 when the tree is unusable:
// non-finite input, capacity, a long tie run, a missing root).
__global__ void k_safe_nn(int64_t* __restrict__ st, int64_t cap, int nsys) {
    if (threadIdx.x != 0) return;
    const int64_t fl = st[ST_FLAGS], nn = st[ST_NNODES];
    const bool ok = (fl & (1 | 2 | 4)) == 0 && nn <= cap && st[ST_LEVEL0 + 1] == nsys;
    st[ST_SAFENN] = ok ? nn : 0;
}

// Parent pointers of every node (roots, the first nroots nodes, get -1), the
// mask of the octant digits of each node's children and the mask of which of
// them (bit i: the i-th child in level order) are twigs.
__global__ __launch_bounds__(256) void k_parents(int64_t nn, int64_t nroots,
                                                 const int* __restrict__ ncstart,
                                                 const uint8_t* __restrict__ ndigit,
                                                 const int8_t* __restrict__ ntype,
                                                 int* __restrict__ nparent,
                                                 uint8_t* __restrict__ nmask,
                                                 uint8_t* __restrict__ ntwig,
                                                 const int64_t* __restrict__ dnn) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (dnn) nn = *dnn;  // the count is on the device (0 when the tree is unusable)
    if (k >= nn) return;
    if (k < nroots) nparent[k] = -1;
    unsigned m = 0u, tw = 0u;
    const int cs = ncstart[k], ce = ncstart[k + 1];
    for (int c = cs; c < ce; ++c) {
        nparent[c] = (int)k;
        m |= 1u << ndigit[c];
        tw |= (ntype[c] == 2 ? 1u : 0u) << (c - cs);
    }
    nmask[k] = (uint8_t)m;
    ntwig[k] = (uint8_t)tw;
}

// Skip pointer of every node: its next sibling, else that of the nearest
// ancestor that has one, else END (the threading of a depth-first walk).
__global__ __launch_bounds__(256) void k_next(int64_t nn, const int* __restrict__ ncstart,
                                              const int* __restrict__ nparent,
                                              int* __restrict__ nnext,
                                              const int64_t* __restrict__ dnn) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (dnn) nn = *dnn;
    if (k >= nn) return;
    int x = (int)k;
    int nx = END;
    for (;;) {
        const int p = nparent[x];
        if (p < 0) break;
        if (x + 1 < ncstart[p + 1]) {
            nx = x + 1;
            break;
        }
        x = p;
    }
    nnext[k] = nx;
}

// ------------------------------------------------------------------ driver
static UpArgs up_args(const hq_tree* t) {
    return UpArgs{t->multipole, t->ntype.as<int8_t>(), t->nfirst.as<int>(), t->ncstart.as<int>(),
                  t->nmask.as<uint8_t>(), t->ntwig.as<uint8_t>(), t->sbody.as<double4>(),
                  t->nmass.as<double>(),
                  t->ncomx.as<double>(), t->ncomy.as<double>(), t->ncomz.as<double>(),
                  t->multipole >= 4 ? t->nquad.as<double>() : nullptr,
                  t->multipole > 4 ? t->noct.as<double>() : nullptr, t->ncount.as<int>(), t->cap};
}

// Parent pointers and child masks of the nodes, and (with_next) their skip
// pointers; with dnn the count is read on the device (grids over the
// capacity).  A launched build leaves the skip pointers, which only the
// per-target walk reads, to ensure_next.
static void launch_pointers(hq_tree* t, int64_t nn, int nsys, const int64_t* dnn, bool with_next) {
    hipStream_t s = t->stream;
    k_parents<<<grid_for(nn, 256), 256, 0, s>>>(nn, nsys, t->ncstart.as<int>(),
                                                t->ndigit.as<uint8_t>(), t->ntype.as<int8_t>(),
                                                t->nparent.as<int>(), t->nmask.as<uint8_t>(),
                                                t->ntwig.as<uint8_t>(), dnn);
    if (with_next)
        k_next<<<grid_for(nn, 256), 256, 0, s>>>(nn, t->ncstart.as<int>(), t->nparent.as<int>(),
                                                 t->nnext.as<int>(), dnn);
    t->next_valid = with_next;
}

// The skip pointers of the built tree, once per build (on its stream).
int ensure_next(hq_tree* t) {
    if (t->next_valid || t->nn <= 0) return HQ_OK;
    k_next<<<grid_for(t->nn, 256), 256, 0, t->stream>>>(t->nn, t->ncstart.as<int>(),
                                                        t->nparent.as<int>(), t->nnext.as<int>(),
                                                        nullptr);
    HQ_CHECK_HIP(hipGetLastError());
    t->next_valid = true;
    return HQ_OK;
}

// Graph mode, after node emission: the checks on the device, then pointers and
// moments over the device-held node count and level bounds (grids sized by the
// capacity).  The same kernels' arithmetic as the launched build.
static int build_graph_tail(hq_tree* t, int64_t cap, int nsys) {
    hipStream_t s = t->stream;
    int64_t* gs = t->gstate.as<int64_t>();
    k_graph_build_check<<<1, 64, 0, s>>>(t->status.as<int64_t>(), cap, nsys, gs);
    launch_pointers(t, cap, nsys, gs + GS_NN, true);
    // grids set by graph_grids before the capture (the kernel strides over
    // whatever the level holds at replay)
    const UpArgs ua = up_args(t);
    for (int d = NDEP - 1; d >= 0; --d)
        launch_upward(ua, 0, 0, 0, gs + GS_LEVEL0, d, t->graph_lgrid[d] < 0,
                      std::max(1, std::abs(t->graph_lgrid[d])), s);
    HQ_CHECK_HIP(hipGetLastError());
    return HQ_OK;
}

// Grids of a captured build's upward pass: each level's size in the last
// launched build with a margin, never smaller than before, so that calls on
// similar systems capture the same graph.
void graph_grids(hq_tree* t) {
    for (int d = 0; d < NDEP; ++d) {
        const int64_t sz = t->level_start[d + 1] - t->level_start[d];
        const int64_t kids = d + 2 <= NDEP ? t->level_start[d + 2] - t->level_start[d + 1] : 0;
        // the form is fixed by the first capture (the sign: negative for the
        // staged form); the grid only grows
        const bool staged = t->graph_lgrid[d] != 0 ? t->graph_lgrid[d] < 0 : up_staged(sz, kids);
        const int unit = staged ? (1 << up_segsh(sz, kids)) : 128;
        const int g = (int)std::min<int64_t>(grid_for(sz + sz / 4 + 128, unit),
                                             grid_for(t->cap, staged ? 8 : 128));
        const int prev = std::abs(t->graph_lgrid[d]);
        t->graph_lgrid[d] = (staged ? -1 : 1) * std::max(prev, g);
    }
}

void collect_build_times(hq_tree* t) {
    if (!t->build_timed || capturing()) return;
    t->build_timed = false;
    if (hipEventSynchronize(t->ev[4]) != hipSuccess) return;
    hipEventElapsedTime(&t->ms_keys, t->ev[0], t->ev[1]);
    hipEventElapsedTime(&t->ms_sort, t->ev[1], t->ev[2]);
    hipEventElapsedTime(&t->ms_nodes, t->ev[2], t->ev[3]);
    hipEventElapsedTime(&t->ms_upward, t->ev[3], t->ev[4]);
    hipEventElapsedTime(&t->ms_build, t->ev[0], t->ev[4]);
}

int read_status(hq_tree* t) {
    HQ_CHECK_HIP(hipMemcpyAsync(t->hstatus, t->status.p, ST_SIZE * sizeof(int64_t),
                                hipMemcpyDeviceToHost, t->stream));
    HQ_CHECK_HIP(hipStreamSynchronize(t->stream));
    collect_build_times(t);
    return HQ_OK;
}

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the
// keys; the result is left in keys / vals (ktmp / vtmp are scratch).
int scan_excl_buf(uint32_t* a, int64_t len, DevBuf& part, hipStream_t s) {
    int rc;
    if ((rc = part.ensure(scan_part_size(len) * sizeof(uint32_t)))) return rc;
    scan_excl(a, len, part.as<uint32_t>(), s);
    return HQ_OK;
}

template <int ITEMS>
static int rs_passes(uint64_t* keys, int* vals, uint64_t* ktmp, int* vtmp, int64_t n, int bits,
                     int shift0, uint32_t* hist, hipStream_t s) {
    const int nblk = grid_for(n, RS_THREADS * ITEMS);
    uint32_t* dtot = hist + (size_t)256 * nblk;  // per-digit totals
    uint64_t *ka = keys, *kb = ktmp;
    int *va = vals, *vb = vtmp;
    const int passes = (bits - shift0 + 7) / 8;
    for (int pass = 0; pass < passes; ++pass) {
        const int shift = shift0 + 8 * pass;
        k_rs_hist<ITEMS><<<nblk, RS_THREADS, 0, s>>>(ka, n, shift, hist, nblk);
        k_rs_scan<<<256, RS_THREADS, 0, s>>>(hist, nblk, dtot);
        k_rs_scatter<ITEMS><<<nblk, RS_THREADS, 0, s>>>(ka, va, kb, vb, n, shift, hist, nblk, dtot);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        HQ_CHECK_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        HQ_CHECK_HIP(hipMemcpyAsync(vals, va, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    }
    return HQ_OK;
}

// Keys per thread of a sort workgroup for n keys.
// Measured on MI355X: 2 up to 2^18 keys, 8 above (per build at 10k keys
// 0.09 ms against 0.19 ms with 16; at 1M 0.36 against 0.41 ms).
int rs_items(int64_t n) { return n <= (1 << 18) ? 2 : 8; }

// Words of the digit-count table the sort of n keys uses: 256 counts per
// tile, then the 256 digit totals.
size_t rs_hist_words(int64_t n) {
    return (size_t)256 * grid_for(n, RS_THREADS * rs_items(n)) + 256;
}

// Bits [shift0, bits) of the keys; the lower ones are left in input order.
static int radix_sort_bits(uint64_t* keys, int* vals, uint64_t* ktmp, int* vtmp, int64_t n,
                           int bits, int shift0, uint32_t* hist, hipStream_t s) {
    if (n <= 1) return HQ_OK;
    if (rs_items(n) == 2) return rs_passes<2>(keys, vals, ktmp, vtmp, n, bits, shift0, hist, s);
    return rs_passes<8>(keys, vals, ktmp, vtmp, n, bits, shift0, hist, s);
}

int radix_sort_pairs_ws(uint64_t* keys, int* vals, uint64_t* ktmp, int* vtmp, int64_t n, int bits,
                        uint32_t* hist, hipStream_t s) {
    return radix_sort_bits(keys, vals, ktmp, vtmp, n, bits, 0, hist, s);
}

int radix_sort_pairs(uint64_t* keys, int* vals, uint64_t* ktmp, int* vtmp, int64_t n, int bits,
                     DevBuf& hist, DevBuf& part, hipStream_t s) {
    (void)part;
    if (n <= 1) return HQ_OK;
    int rc;
    if ((rc = hist.ensure(rs_hist_words(n) * sizeof(uint32_t)))) return rc;
    return radix_sort_pairs_ws(keys, vals, ktmp, vtmp, n, bits, hist.as<uint32_t>(), s);
}

// The tree's keys by their high word above the lowest TIE_SKIP_BITS bits;
// k_tie_fix orders the rest.
static int radix_sort(hq_tree* t) {
    int rc;
    if ((rc = t->khi2.ensure(t->n * sizeof(uint64_t)))) return rc;
    if ((rc = t->idx2.ensure(t->n * sizeof(int)))) return rc;
    if (t->n <= 1) return HQ_OK;
    if ((rc = t->hist.ensure(rs_hist_words(t->n) * sizeof(uint32_t)))) return rc;
    return radix_sort_bits(t->khi.as<uint64_t>(), t->idx.as<int>(), t->khi2.as<uint64_t>(),
                           t->idx2.as<int>(), t->n, 63, TIE_SKIP_BITS, t->hist.as<uint32_t>(),
                           t->stream);
}

// Gather the sorted bodies, then the nodes from the common prefixes of
// neighbouring keys (levels in order, tree order within a level).
template <class Mark>
static int gather_and_emit(hq_tree* t, int64_t cap, Mark&& mark) {
    hipStream_t s = t->stream;
    const int64_t n = t->n;
    const double* sg = t->geom.as<double>() + G_SYS;
    int rc;
    k_gather<<<grid_for(n, 256), 256, 0, s>>>(t->idx.as<int>(), t->prec.as<double4>(),
                                              t->khi.as<uint64_t>(), sg, t->sysoff.as<int64_t>(),
                                              t->nsys, n, t->sx.as<double>(), t->sy.as<double>(),
                                              t->sz.as<double>(), t->sm.as<double>(),
                                              t->sklo.as<uint32_t>(), t->ssys.as<int>(),
                                              t->sbody.as<double4>());
    HQ_CHECK_HIP(mark(2));
    // 4. nodes
    k_lcp<<<grid_for(n, 256), 256, 0, s>>>(t->khi.as<uint64_t>(), t->sklo.as<uint32_t>(),
                                           t->ssys.as<int>(), n, t->lcp.as<int8_t>());
    const int nblk = grid_for(n, NB_THREADS);
    if ((rc = t->hist.ensure(((size_t)NDEP * nblk + 1) * sizeof(uint32_t)))) return rc;
    k_node_count<<<nblk, NB_THREADS, 0, s>>>(t->lcp.as<int8_t>(), n, t->hist.as<uint32_t>(), nblk,
                                             t->status.as<int64_t>());
    if ((rc = scan_excl_buf(t->hist.as<uint32_t>(), (int64_t)NDEP * nblk, t->scanpart, s))) return rc;
    k_node_emit<<<nblk, NB_THREADS, 0, s>>>(
        t->lcp.as<int8_t>(), t->khi.as<uint64_t>(), t->sklo.as<uint32_t>(), sg,
        t->ssys.as<int>(), n, t->hist.as<uint32_t>(), nblk, cap, t->ndepth.as<uint8_t>(), t->ntype.as<int8_t>(),
        t->nfirst.as<int>(), t->ncstart.as<int>(), t->ndigit.as<uint8_t>(), t->ncx.as<double>(),
        t->ncy.as<double>(), t->ncz.as<double>(), t->nkhi.as<uint64_t>(), t->nklo.as<uint32_t>(),
        t->status.as<int64_t>());
    k_levels<<<1, 64, 0, s>>>(t->hist.as<uint32_t>(), nblk, t->status.as<int64_t>());
    HQ_CHECK_HIP(mark(3));
    HQ_CHECK_HIP(hipGetLastError());
    return HQ_OK;
}

// A batch: stably by system after the key sort (the keys ride along from khio).
static int sort_by_system(hq_tree* t) {
    hipStream_t s = t->stream;
    const int64_t n = t->n;
    int bits = 1;
    while ((1ll << bits) < t->nsys) ++bits;
    k_syskey<<<grid_for(n, 256), 256, 0, s>>>(t->idx.as<int>(), t->sysid.as<int>(), n,
                                              t->khi.as<uint64_t>());
    int rc;
    if ((rc = radix_sort_pairs(t->khi.as<uint64_t>(), t->idx.as<int>(), t->khi2.as<uint64_t>(),
                               t->idx2.as<int>(), n, bits, t->hist, t->scanpart, s)))
        return rc;
    k_restore_keys<<<grid_for(n, 256), 256, 0, s>>>(t->idx.as<int>(), t->khio.as<uint64_t>(), n,
                                                    t->khi.as<uint64_t>());
    HQ_CHECK_HIP(hipGetLastError());
    return HQ_OK;
}

// the low words (levels 22..25) as sort keys, with the identity permutation
__global__ __launch_bounds__(256) void k_lo_keys(const uint32_t* __restrict__ klo, int64_t n,
                                                 uint64_t* __restrict__ key,
                                                 int* __restrict__ idx) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = klo[i];
    idx[i] = (int)i;
}

// The full 25-level order, for inputs whose runs of equal high words are too
// long for k_tie_fix: a stable sort by the low word, then (keys carried from
// khio, the original-order high words) a stable sort by the high word, then
// by system for a batch.  The same order k_tie_fix gives to shorter runs.
static int full_key_sort(hq_tree* t) {
    hipStream_t s = t->stream;
    const int64_t n = t->n;
    const int* sysid = t->nsys > 1 ? t->sysid.as<int>() : nullptr;
    const double* sg = t->geom.as<double>() + G_SYS;
    int rc;
    if ((rc = t->khio.ensure(n * 8))) return rc;
    k_keys<<<grid_for(n, 256), 256, 0, s>>>(t->x.as<double>(), t->y.as<double>(),
                                            t->z.as<double>(), n, sg, sysid,
                                            t->khi.as<uint64_t>(), t->klo.as<uint32_t>(),
                                            t->idx.as<int>(), t->khio.as<uint64_t>(),
                                            t->m.as<double>(), t->prec.as<double4>());
    k_lo_keys<<<grid_for(n, 256), 256, 0, s>>>(t->klo.as<uint32_t>(), n, t->khi.as<uint64_t>(),
                                               t->idx.as<int>());
    if ((rc = radix_sort_pairs(t->khi.as<uint64_t>(), t->idx.as<int>(), t->khi2.as<uint64_t>(),
                               t->idx2.as<int>(), n, 12, t->hist, t->scanpart, s)))
        return rc;
    k_restore_keys<<<grid_for(n, 256), 256, 0, s>>>(t->idx.as<int>(), t->khio.as<uint64_t>(), n,
                                                    t->khi.as<uint64_t>());
    if ((rc = radix_sort_pairs(t->khi.as<uint64_t>(), t->idx.as<int>(), t->khi2.as<uint64_t>(),
                               t->idx2.as<int>(), n, 63, t->hist, t->scanpart, s)))
        return rc;
    if (t->nsys > 1 && (rc = sort_by_system(t))) return rc;
    HQ_CHECK_HIP(hipGetLastError());
    return HQ_OK;
}

int tree_build(hq_tree* t) {
    // the previous build's phase times, before this build records its events
    if (!capturing()) collect_build_times(t);
    const int64_t n = t->n;
    const int nsys = t->nsys;
    const bool batch = nsys > 1;
    int rc;
    size_t n8 = n * sizeof(double);
    if ((rc = t->geom.ensure((G_SYS + (size_t)SG * nsys) * sizeof(double)))) return rc;
    if ((rc = t->khi.ensure(n * 8)) || (rc = t->klo.ensure(n * 4)) || (rc = t->idx.ensure(n * 4)) ||
        (rc = t->prec.ensure(n * 32)) ||
        (rc = t->sbody.ensure(n * 32)))
        return rc;
    if ((rc = t->sx.ensure(n8)) || (rc = t->sy.ensure(n8)) || (rc = t->sz.ensure(n8)) ||
        (rc = t->sm.ensure(n8)) || (rc = t->sklo.ensure(n * 4)) || (rc = t->lcp.ensure(n)) ||
        (rc = t->ssys.ensure(n * 4)))
        return rc;
    if (batch && ((rc = t->sysid.ensure(n * 4)) || (rc = t->khio.ensure(n * 8)))) return rc;
    // node capacity of pre_init (:219-223): 3N for N >= 1000, else 10N, per system
    if (!batch) {
        t->cap = n >= 1000 ? 3 * n : 10 * n;
    } else {
        t->cap = 0;
        for (int k = 0; k < nsys; ++k) {
            const int64_t ns = t->hsysoff[k + 1] - t->hsysoff[k];
            t->cap += ns >= 1000 ? 3 * ns : 10 * ns;
        }
    }
    int64_t cap = t->cap;
    if ((rc = t->ndepth.ensure(cap)) || (rc = t->ntype.ensure(cap)) ||
        (rc = t->nfirst.ensure(cap * 4)) || (rc = t->ncount.ensure(cap * 4)) ||
        (rc = t->ncstart.ensure((cap + 1) * 4)) || (rc = t->nnext.ensure(cap * 4)) ||
        (rc = t->nparent.ensure(cap * 4)) || (rc = t->ndigit.ensure(cap)) ||
        (rc = t->nmask.ensure(cap)) || (rc = t->ntwig.ensure(cap)) ||
        (rc = t->nmass.ensure(cap * 8)) || (rc = t->ncomx.ensure(cap * 8)) ||
        (rc = t->ncomy.ensure(cap * 8)) || (rc = t->ncomz.ensure(cap * 8)) ||
        (rc = t->ncx.ensure(cap * 8)) || (rc = t->ncy.ensure(cap * 8)) ||
        (rc = t->ncz.ensure(cap * 8)) || (rc = t->nkhi.ensure(cap * 8)) ||
        (rc = t->nklo.ensure(cap * 4)))
        return rc;
    if (t->multipole >= 4 && (rc = t->nquad.ensure(cap * 6 * 8))) return rc;
    if (t->multipole > 4 && (rc = t->noct.ensure(cap * 14 * 8))) return rc;

    hipStream_t s = t->stream;
    double* sg = t->geom.as<double>() + G_SYS;
    const int* sysid = batch ? t->sysid.as<int>() : nullptr;
    const bool gm = t->graph_mode;
    auto mark = [&](int e) { return gm ? hipSuccess : hipEventRecord(t->ev[e], s); };
    if (gm) k_fill_i64<<<1, 64, 0, s>>>(t->status.as<int64_t>(), 0, ST_SIZE);
    else HQ_CHECK_HIP(hipMemsetAsync(t->status.p, 0, ST_SIZE * sizeof(int64_t), s));
    HQ_CHECK_HIP(mark(0));
    // 1. root cube(s)
    if (!batch) {
        const int nbb = (int)std::min<int64_t>(1024, grid_for(n, 256));
        if ((rc = t->hist.ensure((size_t)nbb * 6 * sizeof(double)))) return rc;
        k_bbox_partial<<<nbb, 256, 0, s>>>(t->x.as<double>(), t->y.as<double>(),
                                           t->z.as<double>(), n, t->hist.as<double>(),
                                           t->status.as<int64_t>());
        k_bbox_final<<<1, 64, 0, s>>>(t->hist.as<double>(), nbb, t->geom.as<double>());
    } else {
        k_sysid<<<nsys, 256, 0, s>>>(t->sysoff.as<int64_t>(), t->sysid.as<int>());
        k_sys_bbox<<<nsys, 256, 0, s>>>(t->x.as<double>(), t->y.as<double>(), t->z.as<double>(),
                                        t->sysoff.as<int64_t>(), sg, t->status.as<int64_t>());
    }
    // 2. keys
    k_keys<<<grid_for(n, 256), 256, 0, s>>>(t->x.as<double>(), t->y.as<double>(),
                                            t->z.as<double>(), n, sg, sysid,
                                            t->khi.as<uint64_t>(), t->klo.as<uint32_t>(),
                                            t->idx.as<int>(),
                                            batch ? t->khio.as<uint64_t>() : nullptr,
                                            t->m.as<double>(), t->prec.as<double4>());
    HQ_CHECK_HIP(mark(1));
    // 3. sort: by octant key; a batch then stably by system
    if ((rc = radix_sort(t))) return rc;
    if (batch && (rc = sort_by_system(t))) return rc;
    k_tie_fix<<<grid_for(n, 256), 256, 0, s>>>(t->khi.as<uint64_t>(), t->klo.as<uint32_t>(),
                                               t->idx.as<int>(), sysid, n,
                                               t->status.as<int64_t>());
    if ((rc = gather_and_emit(t, cap, mark))) return rc;
    if (gm) return build_graph_tail(t, cap, nsys);
    // the status goes to the host; the pointer kernels (over the device-held
    // node count) run while the host waits for it
    k_safe_nn<<<1, 64, 0, s>>>(t->status.as<int64_t>(), cap, nsys);
    HQ_CHECK_HIP(hipMemcpyAsync(t->hstatus, t->status.p, ST_SIZE * sizeof(int64_t),
                                hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipEventRecord(t->ev[11], s));
    launch_pointers(t, cap, nsys, t->status.as<int64_t>() + ST_SAFENN, false);
    HQ_CHECK_HIP(hipEventSynchronize(t->ev[11]));
    HQ_CHECK_HIP(hipGetLastError());
    collect_build_times(t);
    bool pointers = t->hstatus[ST_SAFENN] > 0;
    int64_t flags = t->hstatus[ST_FLAGS];
    if (flags & 1) return fail(HQ_ERR_NONFINITE, "non-finite input position");
    if (flags & 4) {
        // a run of more than TIE_RUN_MAX bodies shares the sorted bits of the
        // high word (a cube 2^-16 of the root wide): sort by the full
        // 25-level key instead -- levels 22..25 first, then 1..21 -- and
        // build the nodes again
        if ((rc = full_key_sort(t))) return rc;
        HQ_CHECK_HIP(hipMemsetAsync(t->status.p, 0, ST_SIZE * sizeof(int64_t), s));
        if ((rc = gather_and_emit(t, cap, mark))) return rc;
        if ((rc = read_status(t))) return rc;
        flags = t->hstatus[ST_FLAGS];
        pointers = false;
    }
    t->nn = t->hstatus[ST_NNODES];
    t->lost = t->hstatus[ST_LOST];
    if (flags & 2)
        return fail(HQ_ERR_CAPACITY, "INSUFFICIENT NUMBER OF NODES: %lld (needed %lld)",
                    (long long)cap, (long long)t->nn);
    t->most_depth = 0;
    for (int d = 0; d <= NDEP; ++d) t->level_start[d] = (int)t->hstatus[ST_LEVEL0 + d];
    if (t->level_start[1] != nsys)
        return fail(HQ_ERR_ARGS, "batch build produced %d roots for %d systems",
                    t->level_start[1], nsys);
    for (int d = 0; d < NDEP; ++d)
        if (t->level_start[d + 1] > t->level_start[d]) t->most_depth = d;
    // 5. parent and skip pointers (the walks' threading; the roots end every
    // walk), unless queued above
    if (!pointers) launch_pointers(t, t->nn, nsys, nullptr, false);
    // 6. moments, deepest level first
    const UpArgs ua = up_args(t);
    for (int d = t->most_depth; d >= 0; --d) {
        const int64_t a = t->level_start[d], b = t->level_start[d + 1];
        const int64_t kids = d + 2 <= NDEP ? t->level_start[d + 2] - b : 0;
        if (b > a) launch_upward(ua, a, b, kids, nullptr, d, up_staged(b - a, kids), 0, s);
    }
    HQ_CHECK_HIP(hipEventRecord(t->ev[4], s));
    HQ_CHECK_HIP(hipGetLastError());
    // no wait here: the phase times are read at the next synchronisation
    t->build_timed = true;
    t->built = true;
    return HQ_OK;
}

}  // namespace hq
