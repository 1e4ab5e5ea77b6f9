/*
 * hq.h -- C ABI of libhq.so, the MI355X (gfx950) hierarchical gravity library.
 *
 * Plain C types only.  Every entry point returns an int status (HQ_OK = 0);
 * hq_last_error() gives the message of the calling thread's last failure.
 * Arrays are float64 unless stated.  "where" says whether a pointer is host
 * memory (HQ_HOST) or device memory of the current HIP device (HQ_DEVICE).
 * Calls on one handle are serialised by the caller; different handles are
 * independent (one HIP stream each).  No call exits the process.
 *
 * Reference interfaces replaced (paths relative to Potalej/hierarquicos):
 *   f2py module `api` (fortran/api3d.f90, built by fortran/compile.sh:2)
 *   OctreeType type-bound API (fortran/octree_default.f90:145-170)
 *   compute_forces_direct (fortran/direct_forces.f90:8)
 *   morton_tree_type (fortran/morton/morton_octree.f90:49-64)
 */
#ifndef HQ_H
#define HQ_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HQ_OK 0
#define HQ_ERR_ARGS 1       /* bad argument (size, method, null pointer) */
#define HQ_ERR_CAPACITY 2   /* node capacity exceeded; the reference STOPs with
                               status 0 (octree_default.f90:457-460) */
#define HQ_ERR_LOST 3       /* particles dropped at max depth 25
                               (octree_default.f90:570-579); fatal only for the
                               grouped methods, whose reference STOPs ("body
                               lost", :1758-1760) */
#define HQ_ERR_HIP 4        /* HIP runtime error */
#define HQ_ERR_NONFINITE 5  /* NaN or Inf in the input positions */
#define HQ_ERR_DEPTH 6      /* Morton depth above 21 levels (64-bit keys) */
#define HQ_ERR_NOMEM 7      /* device allocation failed */
#define HQ_ERR_OVERFLOW 8   /* an interaction list or walk stack overflowed */

#define HQ_HOST 0
#define HQ_DEVICE 1

/* precision of the force evaluation */
#define HQ_FP64 0   /* plain FP64, the reference's arithmetic */
#define HQ_MIXED 1  /* FP32 interactions on recentred coordinates with
                       compensated sums; FP64 keys, positions, moments */

/* force layouts */
#define HQ_LAYOUT_3N 0  /* F[3*p + c], the OctreeType (3,N) layout */
#define HQ_LAYOUT_N3 1  /* F[c*N + p], the f2py (N,3) Fortran-order layout */

/* ---- f2py `api.api_mod` surface (fortran/api3d.f90) ------------------- */

/* api3d.f90:43-72 compute_forces_direct(m, x, y, z, G, eps) -> F(N,3).
 * eps is the softening length (squared inside); host arrays; F is (N,3)
 * column-major. */
int hq_compute_forces_direct(int64_t n, const double *m, const double *x, const double *y,
                             const double *z, double G, double eps, double *F);

/* api3d.f90:20-41 compute_forces(m, x, y, z, theta, G, eps, quad) -> F(N,3).
 * theta is the opening angle and eps the softening length (both squared
 * inside); method selects 10/11/12 (BH mono/quad/oct), 20/21/22 (Dehnen) or
 * 30/31/32 (grouped BH).  The legacy call passed theta unsquared into a
 * routine that expects theta^2: pass legacy_theta2 = 1 to treat theta as
 * theta^2 the way that call did. */
int hq_compute_forces(int64_t n, const double *m, const double *x, const double *y,
                      const double *z, double theta, double G, double eps, int method,
                      int legacy_theta2, double *F);

/* api3d.f90:9-18 generate_tree(m, x, y, z, quad): builds and discards. */
int hq_generate_tree(int64_t n, const double *m, const double *x, const double *y,
                     const double *z, int method);

/* ---- OctreeType handle (octree_default.f90) ---------------------------- */
typedef struct hq_tree hq_tree;

/* pre_init(m, method, group_size) (:178-282).  group_size <= 0 gives 64.
 * precision: HQ_FP64 or HQ_MIXED. */
int hq_tree_create(hq_tree **out, int64_t n, int method, int group_size, int precision);
/* A batch of nsys independent systems in one handle, for the ensemble and
 * Parareal drivers (parareal/parareal.f90:28-140 hold one OctreeType per
 * system and time slice).  System k owns counts[k] >= 1 consecutive particles
 * of every input array; each gets its own root cube and its own tree, and its
 * forces are bit-identical to those of a handle built from that system alone.
 * Every other call works on a batch handle as on a single one. */
int hq_tree_create_batch(hq_tree **out, int nsys, const int64_t *counts, int method,
                         int group_size, int precision);
int hq_tree_set_masses(hq_tree *t, const double *m, int where);
/* init(x, y, z) (:320-387): keys, radix sort, octree from sorted keys,
 * upward P2M/M2M pass, groups. */
int hq_tree_build(hq_tree *t, const double *x, const double *y, const double *z, int where);
/* Forces over all particles with theta^2, G, eps^2 as the reference takes
 * them: bh_forces_over_p (:895) for methods 10-12, dehnen_eval (:1680) for
 * 20-22, forces_groups (:1811) for 30-32.  layout: HQ_LAYOUT_3N or _N3. */
int hq_tree_forces(hq_tree *t, double theta2, double G, double eps2, double *F, int layout,
                   int where);
/* Mixed-precision bunch walk of the same tree: contiguous 64-particle
 * bunches in tree order with their tight bounding boxes under the grouped
 * (Barnes 1990) acceptance test of octree_default.f90:1927-1937, by default
 * with the cell's centre-of-mass offset added to its size (HQ_TUNE_MAC). */
int hq_tree_forces_bunched(hq_tree *t, double theta2, double G, double eps2, double *F,
                           int layout, int where, int precision);
int hq_tree_destroy(hq_tree *t);

/* Tuning knobs (no reference counterpart): HQ_TUNE_BUNCH sets the bunch size
 * of hq_tree_forces_bunched (1..64, default 64; also the group_size of
 * methods 30-32), HQ_TUNE_SUPER the size up to which sibling subtrees are
 * packed before being cut into bunches (default 256). */
#define HQ_TUNE_BUNCH 1
#define HQ_TUNE_SUPER 2
#define HQ_TUNE_FUSED 3   /* 1 (default): walk and evaluate in one kernel; 0: two kernels */
#define HQ_TUNE_LEAF 4    /* 0 (default) .. 64: in bunched walks a cell of at most this
                             many bodies enters the lists as its bodies (exact P2P),
                             the leaf-bucket size of a bucketed tree */
#define HQ_TUNE_FMM_LEAF 5 /* 1..64 (default 64): FMM cells of at most this many bodies are leaves */
#define HQ_TUNE_GRAPH 6    /* 1 (default): hq_tree_integrate with HQ_FORCE_DIRECT, or with
                              HQ_FORCE_BUNCHED at HQ_MIXED in the fused walk, captures one step
                              in a HIP graph and replays it (same results); 0: launches every
                              step */
#define HQ_TUNE_SPLIT 7    /* fused mixed walk: 1 (default) with few bunches (<= 2048) records
                              each bunch's batches and evaluates them one wavefront per batch,
                              then sums them in order (same results as the fused kernel, more
                              parallel); 0 never; 2 whenever the logs fit */
#define HQ_TUNE_FAR 8      /* mixed bunched walk, fused or split: an accepted cell whose centre
                              of mass lies at least value/10 bunch radii from the bunch centre
                              enters the bunch's local expansion (one M2L per cell, orders 5 / 3
                              / 2 for its monopole / quadrupole / octupole terms, evaluated once
                              per body) instead of one M2P per body; 0..1000, default 50; 0 off.
                              Below theta = 0.5 this factor and HQ_TUNE_FAR_INNER's grow as
                              sqrt(0.5 / theta). The two-kernel walk (HQ_TUNE_FUSED 0) has no
                              far field */
#define HQ_TUNE_FAR_INNER 9 /* inner tier of the far field: accepted cells between value/10
                              and HQ_TUNE_FAR/10 bunch radii take a local expansion of
                              higher orders (6 / 4 / 3); 0..1000, default 30; 0, or a value
                              not below HQ_TUNE_FAR, switches the tier off */
#define HQ_TUNE_GRAPH_GRID 10 /* graph replay of bunched steps: the captured walk's grid in
                              bunches; 0 (default): twice the launched first step's bunch count
                              plus 64.  A replayed step with more bunches than the grid is
                              flagged and all steps are redone launched (same results) */
#define HQ_TUNE_MAC 11     /* acceptance test of the bunched walks (fused, split, two-kernel,
                              FP64 or mixed): 0 the grouped test of octree_default.f90:1927-1937,
                              side^2 <= theta^2 dist^2 with dist from the centre of mass to the
                              bunch's box; 1 (default) the offset test
                              (side + theta |com - centre|)^2 <= theta^2 dist^2, which also
                              opens cells whose centre of mass sits near a face (at 1M Plummer
                              bodies, octupole, theta 0.5: 18% lower RMS error for 4% more
                              time).  Methods 30-32 keep the reference's test */
#define HQ_TUNE_FMM_BIN_WINDOW 12 /* matrix-core FMM: target groups whose pair counts a tile
                              of lattice pairs keeps in LDS, from the tile's smallest group
                              (1..4096, default 4096); the pairs of groups beyond it take one
                              global atomic per run.  Changes only the order of the sums */
int hq_tree_tune(hq_tree *t, int key, int64_t value);

/* Cartesian FMM of order 4 or 6 on the built tree: Dehnen's method carried to
 * higher orders (the reference's, octree_default.f90:1095-1712, stops at
 * quadrupole sources).  Expansions about the cube centres; a dual-tree walk
 * with rmax_A + rmax_B <= theta |c_A - c_B|; M2L of same-level pairs as dense
 * products with per-offset operators.  mode HQ_FMM_MFMA: that M2L in FP32 on
 * the matrix cores and the near field in FP32 (Kahan); HQ_FMM_FP64: every M2L
 * and the near field in scalar FP64.  Moments, L2L and L2P are FP64.  The far
 * field is unsoftened; eps2 softens the near field.  Single or batch handles;
 * HQ_ERR_LOST if bodies were lost at max depth. */
#define HQ_FMM_MFMA 0
#define HQ_FMM_FP64 1
int hq_tree_fmm(hq_tree *t, int order, double theta2, double G, double eps2, int mode, double *F,
                int layout, int where);
/* Autotuner (no reference counterpart): on the built tree, searches the
 * bunched walk's settings at `precision` by coordinate descent from the
 * handle's current one -- bunch size {32, 48, 64}, super-group size {128, 256,
 * 512}, leaf bucket {0, 2, 4, 8, 16, 32, 64} and evaluation tiling (fused or
 * split, HQ_TUNE_SPLIT 0 / 2) -- timing each with HIP events on the handle's
 * stream (best of `reps` runs after a warm-up).  A setting whose RMS relative
 * force error on 512 sample bodies (against the FP64 direct sum) exceeds the
 * starting setting's by more than 2% is not taken (single-system handles).
 * Keeps the fastest setting on the handle, returns it with its time in
 * milliseconds, and stores it in the process's tuning table under the
 * workload class (log2 N, systems, multipole order, theta, precision, tree
 * depth).  Handles of the same class then start from the table's setting at
 * their first bunched walk, for every knob their caller has not set with
 * hq_tree_tune.  Results of bunched walks depend on the setting (within the
 * walk's error budget), so a caller that needs bit-reproducible forces fixes
 * the knobs with hq_tree_tune. */
int hq_tree_autotune(hq_tree *t, double theta2, double G, double eps2, int precision, int reps,
                     int *bunch, int *super_size, int *leaf, double *ms);
/* The handle's current value of a tuning knob (HQ_TUNE_*). */
int hq_tree_get_tune(const hq_tree *t, int key, int64_t *value);
/* The tuning table as text, one setting per workload class; load merges a
 * file into the table (its rows win), clear empties it. */
int hq_tune_table_save(const char *path);
int hq_tune_table_load(const char *path);
int hq_tune_table_clear(void);

/* ---- Time integration around the force path ----------------------------
 * The handle holds a phase-space state in HBM, input order: positions
 * (x, y, z) and momenta (px, py, pz), FP64.  Integration updates it in place
 * and rebuilds the tree before every force evaluation, as the reference's
 * drivers call init before each force loop. */
#define HQ_SCHEME_VERLET 0 /* kick-drift-kick in momenta form:
                              benchmarks/integration_verlet/integration.f90:110-172,
                              parareal/parareal.f90:46-69 (F_til) */
#define HQ_SCHEME_EULER 1  /* symplectic Euler, kick then drift:
                              parareal/parareal.f90:8-44 (G_til) */
#define HQ_FORCE_TREE 0    /* the handle's method: bh_forces_over_p (10-12),
                              dehnen_eval (20-22), forces_groups (30-32), FP64 */
#define HQ_FORCE_BUNCHED 1 /* hq_tree_forces_bunched at the handle's precision */
#define HQ_FORCE_DIRECT 2  /* FP64 direct sum within each system
                              (parareal/api3d.f90:46-75, fortran/direct_forces.f90) */
int hq_tree_set_state(hq_tree *t, const double *x, const double *y, const double *z,
                      const double *px, const double *py, const double *pz, int where);
/* Copies the state out; a null pointer skips that component. */
int hq_tree_get_state(hq_tree *t, double *x, double *y, double *z, double *px, double *py,
                      double *pz, int where);
/* nsteps steps of dt: Verlet p += (dt/2) F, q += (dt p)/m, F(q), p += (dt/2) F
 * (integration.f90:153-171); Euler p += dt F(q), q += (dt p)/m (parareal.f90:30-36).
 * theta2 and eps2 are squared as in OctreeType%forces; G as given. */
int hq_tree_integrate(hq_tree *t, int scheme, double dt, int64_t nsteps, double theta2, double G,
                      double eps2, int force);
/* Per system k of the handle, out[7k..7k+6] = total energy (kinetic + softened
 * pair potential, integration.f90:236-262), linear momentum (3) and angular
 * momentum sum q x p (3) (parareal/api3d.f90:108-155) of the current state. */
int hq_tree_energy(hq_tree *t, double G, double eps2, double *out);

/* ---- Harness error metric ----------------------------------------------
 * evaluate_error_on_accelerations (benchmarks/test_forces.f90:66-127), with
 * a = F/m: errors[0] = sqrt((||a - a_dir||_F / a_scale)^2 / N), a_scale =
 * sum m/|q|^2 (McMillan-Aarseth); errors[1..3] = sum F; errors[4..6] = sum q x F.
 * F and Fd in the given layout; all arrays host or all device. */
int hq_errors_ma(int64_t n, const double *m, const double *x, const double *y, const double *z,
                 const double *F, const double *Fd, int layout, int where, double *errors);

/* ---- Parareal (parareal/parareal.f90:97-186) ---------------------------
 * num_windows states U(w) at t = w * tf / num_windows, w = 0..num_windows-1;
 * an initial coarse sweep, then per iteration the fine propagators of windows
 * 0..num_windows-2 (one batch handle, all windows at once) and the serial
 * correction U_k+1(w+1) = G(U_k+1(w)) + F(U_k(w)) - G(U_k(w)); stop when
 * ||U_k+1(last) - U_k(last)||_2 / (6N) <= maxerror or after itermax iterates.
 * The reference's propagators are coarse = one Euler step per window with
 * quadrupole BH forces, fine = Verlet with direct forces (qntt steps). */
#define HQ_CORR_REFERENCE 0 /* (G_new + F) - G_old, the reference's order */
#define HQ_CORR_EXACT 1     /* F + (G_new - G_old): a converged window equals the
                               sequential fine solution bit for bit */
typedef struct {
    int64_t n;             /* particles per system */
    int num_windows;       /* states per trajectory (>= 2) */
    int itermax;           /* N_itermax: iterates including the coarse sweep */
    double tf, maxerror, G, eps2;
    int coarse_force, coarse_method, coarse_precision, coarse_steps;
    double coarse_theta2;
    int fine_force, fine_method, fine_precision, fine_steps;
    double fine_theta2;
    int group_size;        /* bunch size of bunched walks (<= 0: 64) */
    int correction;        /* HQ_CORR_REFERENCE or HQ_CORR_EXACT */
    int coarse_scheme;     /* HQ_SCHEME_EULER (the reference's G_til) or HQ_SCHEME_VERLET */
} hq_parareal_params;
/* U_out: 6 * num_windows * n doubles, component-major:
 * U_out[(c * num_windows + w) * n + a], c = x, y, z, px, py, pz.
 * errors: itermax - 1 doubles (the error of iterates 2, 3, ...). */
int hq_parareal(const hq_parareal_params *prm, const double *m, const double *x0,
                const double *y0, const double *z0, const double *px0, const double *py0,
                const double *pz0, int where, double *U_out, int out_where, double *errors,
                int *iterations);

typedef struct {
    int64_t n;
    int64_t nnodes;        /* number_of_nodes */
    int64_t ngroups;       /* groups (methods 30-32) */
    int64_t lost;          /* particles dropped at max depth */
    int most_depth;        /* deepest node */
    int multipole;         /* 1, 4, 8 */
    int64_t list_cells;    /* cells in the last walk's lists (diagnostic) */
    int64_t list_parts;    /* particles in the last walk's lists (diagnostic) */
    int64_t n_p2p;         /* per-target BH counts of the last methods 10-12 walk */
    int64_t n_m2p;
    double ms_build;       /* device time of the last build (HIP events) */
    double ms_keys, ms_sort, ms_nodes, ms_upward;
    double ms_walk;        /* device time of the last force walk kernel(s) */
    double ms_lists;       /* list building part of the last grouped walk */
    double ms_eval;        /* evaluation part of the last grouped walk */
    int64_t exec_cells;    /* target-cell interactions the last grouped walk evaluated */
    int64_t exec_parts;    /* target-particle interactions (lists + in-group) */
    int64_t max_cells;     /* longest cell list of the last grouped walk */
    int64_t max_parts;     /* longest particle list */
    int64_t max_stack;     /* deepest walk stack */
    int64_t split_walks;   /* mixed walks of this handle evaluated split (HQ_TUNE_SPLIT) */
    int64_t far_cells;     /* cells of the last mixed walk taken through the far field (HQ_TUNE_FAR) */
    int64_t walk_stack;    /* LDS stack entries per wavefront of the last mixed bunched walk's
                              kernel (1152, 1536 or 4096; 0 for the two-kernel walk) */
} hq_tree_info_t;
int hq_tree_info(const hq_tree *t, hq_tree_info_t *info);

/* Node store in level order (depth-major, tree order within a depth) copied
 * to host arrays of length nnodes (child_start: nnodes+1).  Any pointer may
 * be NULL.  key_hi/key_lo hold the node's octant path, 3 bits per level with
 * digit 4*(x>=cx) + 2*(y>=cy) + (z>=cz): levels 1..21 in key_hi (level 1 in
 * bits 62..60), levels 22..25 in key_lo (level 22 in bits 11..9). */
int hq_tree_export(const hq_tree *t, int *depth, int *type, int *particle, int *child_start,
                   uint64_t *key_hi, uint32_t *key_lo, double *mass, double *com /* 3*nn */,
                   double *quad /* 6*nn */, double *oct /* 14*nn */, double *center /* 3*nn */);
/* Tree order of the particles: order[i] = original index of sorted slot i. */
int hq_tree_order(const hq_tree *t, int *order);

/* ---- direct sum (fortran/direct_forces.f90:8-64) ----------------------- */
/* Gather form over nt targets (targets == NULL: all n, in order); F is 3*nt
 * in HQ_LAYOUT_3N.  All pointers device memory; synchronous. */
int hq_direct_forces_device(int64_t n, const double *m, const double *x, const double *y,
                            const double *z, int64_t nt, const int64_t *targets, double G,
                            double eps2, int precision, double *F);

/* ---- Morton tree (fortran/morton/morton_octree.f90) -------------------- */
/* Keys of generate_nodes (:357-394): normalise, quantise, interleave.
 * L = 0 chooses the smallest depth with distinct keys (choose_depth
 * :217-244).  Outputs (host or device per `where`): keys in original
 * order, the stable sort permutation perm (sorted slot -> original index)
 * and the depth used. */
int hq_morton_keys(int64_t n, const double *m, const double *q /* (3,N) */, int L, double ssa,
                   int64_t *keys, int *perm, int *L_out, int where);
/* No counterpart in the reference: hq_morton_keys for callers that cannot
 * allocate device memory (code under HIP stream capture, serving loops).
 * m, q, keys and perm are device memory of the current device, L_out host;
 * results equal hq_morton_keys's bit for bit.  Its work arrays are one fixed
 * device array per GPU (about 12.5 MiB, held until the process ends), so n
 * is at most 1,048,576 (larger n: HQ_ERR_ARGS); keys and perm are also used
 * as sort buffers.  Runs on the null stream and returns when done; calls that
 * overlap on one GPU race on the shared array. */
int hq_morton_keys_noalloc(int64_t n, const double *m, const double *q /* (3,N) */, int L,
                           double ssa, int64_t *keys, int *perm, int *L_out);

/* ---- device memory plumbing ------------------------------------------- */
/* No counterpart in the reference (it has no device): thin wrappers of
 * hipMalloc / hipFree / hipMemcpy / hipDeviceSynchronize so that callers
 * without a HIP runtime of their own can hold inputs and forces in HBM.
 * kind: 0 host->device, 1 device->host, 2 device->device. */
int hq_dev_alloc(void **p, int64_t bytes);
int hq_dev_free(void *p);
int hq_dev_copy(void *dst, const void *src, int64_t bytes, int kind);
int hq_dev_sync(void);

/* morton_tree_type (morton_octree.f90:49-180): init(N, ms, qs, L, ssa,
 * use_quadrupole) = create + build; forces_seq/forces_par = forces.  q is
 * (3,N); L = 0 chooses the depth; forces are returned in original order. */
typedef struct hq_morton hq_morton;
int hq_morton_create(hq_morton **out, int64_t n, int L, double ssa, int use_quad);
int hq_morton_build(hq_morton *t, const double *m, const double *q, int where);
int hq_morton_forces(hq_morton *t, double G, double eps2, double theta2, int mult, double *F,
                     int layout, int where);
int hq_morton_info(const hq_morton *t, int *L, int64_t *nnodes);
/* Node store in the reference's numbering (level by level); 0-based indices,
 * first_child = -1 for leaves; any pointer may be NULL. */
int hq_morton_export(const hq_morton *t, int *perm, int64_t *prefix, int *level, int *first,
                     int *last, int *parent, int *nchild, int *first_child, int *next,
                     double *mass, double *qcm /* 3*nn */, double *quad /* 6*nn */);
int hq_morton_destroy(hq_morton *t);

const char *hq_last_error(void);
const char *hq_version(void);

#ifdef __cplusplus
}
#endif
#endif
