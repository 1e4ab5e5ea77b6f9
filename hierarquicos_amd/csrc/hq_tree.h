// hq_tree.h -- device-resident octree of the hierarquicos_amd library.
//
// The tree is the reference's insertion octree (fortran/octree_default.f90,
// leaf size 1, max depth 25, root = bbox centre with side 1.2 * extent),
// rebuilt data-parallel from sorted octant-path keys.  Nodes are stored in
// level order: all depth-0 nodes, then depth 1, ...; within a depth in tree
// (sorted key) order, so the children of a node are one contiguous range
// [child_start[k], child_start[k+1]).
#pragma once
#include <memory>
#include <vector>

#include "hq_internal.h"

namespace hq {
struct FmmState;  // hq_fmm.hip
void fmm_state_free(FmmState* s);
struct FmmFree {
    void operator()(FmmState* s) const { fmm_state_free(s); }
};
}  // namespace hq

struct hq_tree {
    int64_t n = 0;
    int method = 12, multipole = 8, group_size = 64, precision = HQ_FP64;
    // batch of independent systems: system s owns particles [sysoff[s], sysoff[s+1])
    // of the input order, gets its own root cube and its own tree; its root is
    // node s.  nsys == 1 is the reference's single tree.
    int nsys = 1;
    std::vector<int64_t> hsysoff;
    hq::DevBuf sysoff, sysid, ssys, khio;
    hipStream_t stream = nullptr;
    hipEvent_t ev[14] = {};
    bool have_masses = false, built = false;
    bool next_valid = false;  // the skip pointers (nnext) match the built tree

    // inputs, original order
    hq::DevBuf m, x, y, z;
    // keys and sort
    hq::DevBuf khi, klo, idx, khi2, idx2, hist, geom, l2tab, scanpart;
    hq::DevBuf prec;  // packed (x, y, z, m) per input particle
    // sorted particles (tree order)
    hq::DevBuf sx, sy, sz, sm, sklo, lcp;
    hq::DevBuf sbody;  // sorted bodies as packed (x, y, z, m) records
    // node store (level order)
    int64_t nn = 0, cap = 0;
    hq::DevBuf ndepth, ntype, nfirst, ncount, ncstart, nnext, nparent, ndigit, nmask;
    hq::DevBuf ntwig;  // per node: which children (bit i: the i-th) are twigs
    hq::DevBuf nmass, ncomx, ncomy, ncomz, nquad, noct, ncx, ncy, ncz, nkhi, nklo;
    hq::DevBuf mom32;  // float moment records for the mixed-precision walk
    hq::DevBuf wnode;  // packed per-node records for the mixed walk's traversal
    int level_start[hq::MAX_DEPTH + 2] = {};
    int most_depth = 0;
    int64_t lost = 0;
    // groups (methods 30-32) and bunches
    hq::DevBuf groups, gflag, bstart;
    int super_size = 256;  // bunching: sibling subtrees packed up to this many bodies
    int leaf = 0;          // bunched walks: cells of <= leaf bodies enter as their bodies (0: off)
    int fmm_leaf = 64;     // FMM: cells of <= fmm_leaf bodies are leaves (1..64)
    std::unique_ptr<hq::FmmState, hq::FmmFree> fmm;
    bool fused = true;     // mixed bunched walk as one fused kernel
    int split = 1;         // split evaluation of the fused walk: 0 never, 1 few bunches, 2 always
    bool build_timed = false;   // the last launched build's events await reading
    int far = 50;          // far-field factor of the fused mixed walk, in tenths (0: off)
    int far_inner = 30;    // inner far-field factor, in tenths (0: no inner tier)
    int stack_level = 0;   // the fused walk's first stack: 0 short, 1 long, 2 deep
    int walk_stack = 0;    // LDS stack entries of the last mixed walk's kernel (0: two-kernel)
    int64_t far_cells = 0; // cells of the last walk taken through the far-field expansion
    int64_t split_lcap = 0;  // log entries per bunch (from the last walk's lists; 0 unknown)
    int64_t split_walks = 0; // walks evaluated split (diagnostic)
    bool graph_split = false;
    hq::DevBuf slog, sdesc, snbat, spart;
    hq::DevBuf flist, fnum;  // far-field lists of the mixed walk: node indices and counts per bunch
    int64_t ngroups = 0;
    // walk scratch
    hq::DevBuf fx, fy, fz, counts, clist, plist, lstat;
    int64_t list_ccap = 0, list_pcap = 0;
    // integrator state: momenta in input order (positions are x, y, z above),
    // block offsets of the per-system direct sum, energy scratch
    hq::DevBuf px, py, pz, dblk, ework;
    std::vector<int> hbofs;  // host copy of dblk (per-system block prefix of the direct sum)
    bool have_state = false;
    bool graph = true;       // integrate: replay captured steps (direct and bunched mixed forces)
    // Graph mode: the build and the bunched walk take their counts (nodes,
    // level bounds, bunches) from device memory and size their grids by upper
    // bounds, so that a step can be captured.  gstate holds those counts and
    // the errors of all replayed steps (GS_* slots); gsave the state before
    // the replays, to rerun them launched if an error was flagged.
    bool graph_mode = false;
    int64_t graph_nbun = 0;  // grid of the walk in graph mode (upper bound of the bunches;
                             // only grows, so that repeated calls capture the same graph)
    int graph_lgrid[hq::MAX_DEPTH + 2] = {};  // grids of the captured upward pass (only grow;
                                              // negative: the staged form)
    // instantiated step graphs of integrate, kept for calls with the same key
    struct GraphEntry {
        uint64_t key = 0, used = 0;
        hipGraphExec_t exec = nullptr;
    };
    GraphEntry gcache[4];
    uint64_t gcache_clock = 0;
    int64_t graph_grid = 0;  // HQ_TUNE_GRAPH_GRID: that grid if > 0, else 2 x bunches + 64
    int mac = 1;             // HQ_TUNE_MAC of the bunched walks: 0 cell side, 1 side + offset
    int fmm_binw = 4096;     // HQ_TUNE_FMM_BIN_WINDOW: target groups counted in LDS per tile
    unsigned tune_user = 0;  // HQ_TUNE_* keys the caller set (bit per key): the table keeps off them
    int table_tag = -1;      // angle and precision of the last setting taken from the tuning table
    hq::DevBuf tune_buf, tune_tmp;  // autotuner: sample slots, direct forces, error; direct scratch
    hq::DevBuf gstate, gsave;
    // status block: device + host mirror
    hq::DevBuf status;
    bool groups_pending = false;  // a bunch count read-back is in flight (finish_groups)
    int64_t* hstatus = nullptr;
    // last timings / diagnostics
    float ms_keys = 0, ms_sort = 0, ms_nodes = 0, ms_upward = 0, ms_build = 0;
    float ms_walk = 0, ms_lists = 0, ms_eval = 0;
    int64_t list_cells = 0, list_parts = 0, n_p2p = 0, n_m2p = 0, exec_cells = 0, exec_parts = 0,
            max_cells = 0, max_parts = 0, max_stack = 0;
};

// Geometry buffer (doubles): the whole-input bounding box at G_* (single
// system), then one block of SG doubles per system from G_SYS: root centre,
// side and half side, and per depth d the opening size L2 = side_d^2 and the
// half side of a depth-d cube.
enum { G_SYS = 16 };
enum { S_CX = 0, S_CY, S_CZ, S_SIDE, S_HALF, S_L2 = 8, S_HALFD = 40, SG = 72 };

// status block slots (int64)
enum {
    ST_FLAGS = 0,      // bit0 non-finite input, bit1 capacity, bit2 tie run too long,
                       // bit3 list overflow, bit4 stack overflow
    ST_NNODES = 1,
    ST_LOST = 2,
    ST_LEVEL0 = 3,     // level_start[0..26] at ST_LEVEL0 + d
    ST_LISTC = 32,     // cells in lists (sum)
    ST_LISTP = 33,     // particles in lists (sum)
    ST_NEEDC = 34,     // largest cell list needed
    ST_NEEDP = 35,     // largest particle list needed
    ST_P2P = 36,
    ST_M2P = 37,
    ST_NGROUPS = 38,
    ST_EXECC = 39,     // executed target-cell interactions of the last grouped walk
    ST_EXECP = 40,     // executed target-particle interactions
    ST_MAXSTK = 41,    // deepest walk stack of the last grouped walk
    ST_LISTF = 42,     // cells of the mixed walk taken through the far-field expansion
    ST_SAFENN = 43,    // node count for kernels queued before the host checks the build
                       // (0 when the build is unusable)
    ST_SIZE = 48,
    // host-only slots after the status block (the status copy leaves them alone)
    ST_HOST_BUNCHES = 48,  // bunch count, read back while k_mom32 runs
    ST_HOST_FMM = 50,      // FMM dual walk: per-category totals (6 uint32, slots 50..52)
    ST_HOST_SIZE = 56
};

// graph-mode slots of hq_tree::gstate (int64)
enum {
    GS_STICKY = 0,     // errors of any replayed step: the GF_* bits
    GS_NN = 1,         // node count, 0 when the build overflowed
    GS_LEVEL0 = 2,     // level bounds (clamped) at GS_LEVEL0 + d, d = 0..26
    GS_NBUN = 32,      // bunches of the last walk
    GS_SIZE = 40
};
enum { GF_INPUT = 1, GF_CAPACITY = 2, GF_TIES = 4, GF_ROOTS = 8, GF_LOST = 16, GF_STACK = 32,
       GF_BUNCHES = 64 };

namespace hq {
// The mixed walk's per-node records (written by k_mom32 before each mixed walk):
//  - FP32 moments, REC floats, pre-scaled for the M2P: r[0] M; r[1..6] -3 Q
//    (xx, yy, zz, xy, xz, yz, raw second moments); r[7] -1.5 trQ; r[8..10]
//    -1.5 T, T_i = sum_j O_ijj; r[11..23] 7.5 x (O1..O9, O10, 2 O10, 2 O8,
//    2 O6) with O1..O10 in octree_default.f90 :831-843 order (xxx, xyy, xzz,
//    xxy, yyy, yzz, xxz, yyz, zzz, xyz).  Raw (not traceless) moments: with
//    softening only they give the exact Taylor series of the softened
//    potential.
//  - WalkNode: what the traversal reads of a candidate node, 48 bytes.
constexpr int REC = 24;
struct __align__(16) WalkNode {
    double cx, cy, cz, mass;   // centre of mass, mass
    int first, count, cstart;  // bodies [first, first + count), first child
    int meta;                  // children | type << 4 | depth << 12 | macq << 17
};

// Offset of a cell's centre of mass from its cube centre, |com - centre| /
// side, rounded up to a multiple of 1/32767 and returned as that multiple
// (0..32767; the offset is at most sqrt(3)/2 of the side).  Bunched walks with
// the offset acceptance test open a cell unless side + theta |com - centre|
// <= theta dist (Barnes 1994's remedy for cells whose centre of mass sits
// near a face).
__device__ __forceinline__ int mac_offset_q(double cx, double cy, double cz, double ox, double oy,
                                            double oz, double side) {
#pragma clang fp contract(off)
    const double dx = cx - ox, dy = cy - oy, dz = cz - oz;
    const double f = sqrt(dx * dx + dy * dy + dz * dz) / side;
    const double q = ceil(f * 32767.0);
    return q < 0.0 ? 0 : (q > 32767.0 ? 32767 : (int)q);
}

__device__ __forceinline__ void store_walk_records(int64_t k, int multipole, double mass,
                                                   double cx, double cy, double cz, int first,
                                                   int count, int cstart, int nch, int type,
                                                   int depth, int macq, const double* Q,
                                                   const double* O, float* __restrict__ rec,
                                                   WalkNode* __restrict__ wn) {
    // a leaf enters every walk as its body (P2P): its moment record is never
    // read, so only twigs write one
    if (type != 1) {
    float r[REC];
#pragma unroll
    for (int j = 0; j < REC; ++j) r[j] = 0.0f;
    r[0] = (float)mass;
    if (multipole >= 4) {
        for (int j = 0; j < 6; ++j) r[1 + j] = (float)(-3.0 * Q[j]);
        r[7] = (float)(-1.5 * (Q[0] + Q[1] + Q[2]));
    }
    if (multipole > 4) {
        for (int j = 0; j < 3; ++j) r[8 + j] = (float)(-1.5 * O[11 + j]);
        for (int j = 0; j < 10; ++j) r[11 + j] = (float)(7.5 * O[j]);
        r[21] = (float)(15.0 * O[9]);
        r[22] = (float)(15.0 * O[7]);
        r[23] = (float)(15.0 * O[5]);
    }
    float4* o = reinterpret_cast<float4*>(rec + k * REC);
#pragma unroll
    for (int j = 0; j < REC / 4; ++j) o[j] = make_float4(r[4 * j], r[4 * j + 1], r[4 * j + 2], r[4 * j + 3]);
    }
    WalkNode w;
    w.cx = cx;
    w.cy = cy;
    w.cz = cz;
    w.mass = mass;
    w.first = first;
    w.count = count;
    w.cstart = cstart;
    w.meta = (nch & 15) | ((type & 0xff) << 4) | ((depth & 31) << 12) | ((macq & 0x7fff) << 17);
    wn[k] = w;
}

// set while a stream capture is open: buffers must not grow then
extern thread_local bool g_capturing;
__global__ void k_fill_i64(int64_t* __restrict__ p, int64_t v, int64_t n);
__global__ void k_fill_i32(int* __restrict__ p, int v, int64_t n);
int tree_build(hq_tree* t);
int ensure_next(hq_tree* t);  // skip pointers of the last launched build
int tree_forces_bh(hq_tree* t, double theta2, double G, double eps2);
int tree_forces_groups(hq_tree* t, double theta2, double G, double eps2, int precision, bool bunched);
int tree_forces_dehnen(hq_tree* t, double theta2, double G, double eps2);
int tree_forces_fmm(hq_tree* t, int order, double theta, double G, double eps2, bool mfma, int S);
int tree_scatter_forces(hq_tree* t, double* F, int layout, int where);
int read_status(hq_tree* t);
// grids of a captured build's upward pass from the last launched build (graph mode)
void graph_grids(hq_tree* t);
// the last launched build's phase times from its events (the stream has passed them)
void collect_build_times(hq_tree* t);
int radix_sort_pairs(uint64_t* keys, int* vals, uint64_t* ktmp, int* vtmp, int64_t n, int bits,
                     DevBuf& hist, DevBuf& part, hipStream_t s);
// autotuner of the bunched walk and its table (hq_tune.hip)
struct TuneSetting {
    int bunch, super_size, leaf, split;
};
int autotune(hq_tree* t, double theta2, double G, double eps2, int precision, int reps,
             TuneSetting* best, double* best_ms);
// takes the tuning table's setting for the handle's workload class (knobs the
// caller has not set); true when it changed the handle
bool tune_from_table(hq_tree* t, double theta2, int precision);
int tune_table_save(const char* path);
int tune_table_load(const char* path);
void tune_table_clear();
// the same sort with a caller-provided count table of rs_hist_words(n) words
int radix_sort_pairs_ws(uint64_t* keys, int* vals, uint64_t* ktmp, int* vtmp, int64_t n, int bits,
                        uint32_t* hist, hipStream_t s);
size_t rs_hist_words(int64_t n);
int scan_excl_buf(uint32_t* a, int64_t len, DevBuf& part, hipStream_t s);
int direct_forces(int64_t n, const double* m, const double* x, const double* y, const double* z,
                  int64_t nt, const int64_t* targets, double G, double eps2, int precision,
                  double* F, hipStream_t s, DevBuf* scratch = nullptr);
// FP64 direct sum inside each system of the handle, input order, into t->fx/fy/fz
int direct_forces_systems(hq_tree* t, double G, double eps2);
// forces of the current positions t->x/y/z into t->fx/fy/fz; *sorted tells
// whether they are in tree order (map t->idx) or input order
int forces_now(hq_tree* t, int force, double theta2, double G, double eps2, bool* sorted);
// buffers at graph-mode sizes (before a capture)
int graph_prepare(hq_tree* t);
}  // namespace hq
