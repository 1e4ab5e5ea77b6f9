/*
 * hq_oracle.c -- CPU restatement of the reference's hierarchical gravity
 * force path.  TEST INFRASTRUCTURE ONLY: this file is the checker that the
 * tests, the smoke check and the CPU-baseline leg of the benchmark compare
 * the HIP library against.  No product code path links or calls it.
 *
 * Every routine restates one reference routine (cited file:line, paths
 * relative to the Potalej/hierarquicos repository) and keeps its
 * floating-point evaluation order, so that with contraction disabled
 * (-ffp-contract=off) the results are bit-identical to the reference built
 * with the same flags.  Indices are 0-based here (the reference is 1-based);
 * "-1" means "none" exactly as in the reference.
 *
 *   orc_direct_forces       fortran/direct_forces.f90:8-64   (nt == 1 branch)
 *   orc_direct_targets      gather form of the same pair force (error sample)
 *   orc_tree_*              fortran/octree_default.f90
 *       build               pre_init :178-282, init :320-387, add_node :445-479,
 *                           allocate_subnode :481-513, add_to_subnode :515-545,
 *                           add :547-580, node_size_center :424-443
 *       moments             incorporate_particles :582-628,
 *                           evaluate_multipole :744-882
 *       BH walk             bh_forces_over_p :895-1078
 *       groups              prepare_groups :1722-1809, group_worker :1833-2095
 *       Dehnen              evaluate_rmax :1095-1130,
 *                           mutual_interaction_walk_sequential :1132-1248,
 *                           evaluate_mutual_interaction_monopole :1425-1445,
 *                           evaluate_mutual_interaction_quadrupole :1447-1563,
 *                           dehnen_forces :1565-1678, dehnen_eval :1680-1712
 *   orc_morton_*            fortran/morton/morton_octree.f90
 *                           morton_key :184-213, choose_depth :217-244,
 *                           generate_nodes :328-499, argsort/mergesort :501-560,
 *                           build_next :304-326, evaluate_quadrupoles :562-625,
 *                           barnes_hut :627-766, forces_par :150-180
 *   orc_errors_ma           benchmarks/test_forces.f90:66-127
 *   orc_integrate           benchmarks/integration_verlet/integration.f90:85-172
 *                           (Verlet), parareal/parareal.f90:8-69 (G_til, F_til)
 *   orc_energy              integration.f90:236-262, parareal/api3d.f90:108-155
 *   orc_parareal            parareal/parareal.f90:97-186
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ERR_CAPACITY 1
#define ORC_ERR_ARGS 2
#define ORC_ERR_LOST 3
#define ORC_ERR_ALLOC 4

/* ------------------------------------------------------------------ */
/* direct sum: fortran/direct_forces.f90:19-38 (serial branch).        */
/* F is (3,N): F[3*a + c].                                             */
/* ------------------------------------------------------------------ */
void orc_direct_forces(int64_t n, const double *m, const double *x, const double *y,
                       const double *z, double G, double eps2, double *F)
{
    memset(F, 0, sizeof(double) * 3 * (size_t)n);
    for (int64_t a = 1; a < n; ++a) {
        for (int64_t b = 0; b < a; ++b) {
            double dx = x[b] - x[a];
            double dy = y[b] - y[a];
            double dz = z[b] - z[a];
            double dist2 = dx * dx + dy * dy + dz * dz + eps2;
            double dist = sqrt(dist2);
            double f = G * m[a] * m[b] / (dist * dist2);
            F[3 * a + 0] = F[3 * a + 0] + f * dx;
            F[3 * a + 1] = F[3 * a + 1] + f * dy;
            F[3 * a + 2] = F[3 * a + 2] + f * dz;
            F[3 * b + 0] = F[3 * b + 0] - f * dx;
            F[3 * b + 1] = F[3 * b + 1] - f * dy;
            F[3 * b + 2] = F[3 * b + 2] - f * dz;
        }
    }
}

/* Gather form of the same pair force for a subset of targets (used as the
 * FP64 direct reference on large N).  Same per-pair arithmetic as
 * direct_forces.f90:29-36; summation runs over b ascending. */
void orc_direct_targets(int64_t n, const double *m, const double *x, const double *y,
                        const double *z, int64_t nt, const int64_t *tidx, double G,
                        double eps2, double *F)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t t = 0; t < nt; ++t) {
        int64_t a = tidx[t];
        double fx = 0, fy = 0, fz = 0;
        for (int64_t b = 0; b < n; ++b) {
            if (b == a) continue;
            double dx = x[b] - x[a];
            double dy = y[b] - y[a];
            double dz = z[b] - z[a];
            double dist2 = dx * dx + dy * dy + dz * dz + eps2;
            double dist = sqrt(dist2);
            double f = G * m[a] * m[b] / (dist * dist2);
            fx += f * dx;
            fy += f * dy;
            fz += f * dz;
        }
        F[3 * t + 0] = fx;
        F[3 * t + 1] = fy;
        F[3 * t + 2] = fz;
    }
}

/* ------------------------------------------------------------------ */
/* Insertion octree (fortran/octree_default.f90)                       */
/* ------------------------------------------------------------------ */
typedef struct {
    int64_t N;
    int method, multipole, max_depth, group_size;
    double side_amp;
    int nn, cap, most_depth, status;
    double *m, *x, *y, *z;
    double *cx, *cy, *cz, *half, *L2, *mass, *qx, *qy, *qz;
    int *particle, *type, *depth, *child; /* child[8*k + r], r = 0..7 */
    uint64_t *khi;                        /* path digits, levels 1..21 */
    uint32_t *klo;                        /* path digits, levels 22..25 */
    double *quad, *oct;                   /* quad[6*k+i], oct[14*k+i] */
    /* grouped BH */
    int *count, *first, *last, *order, *groups;
    int ngroups;
    double *gx, *gy, *gz, *gm;
    /* Dehnen */
    double *rmax, *force, *hess, *third;
} orc_tree;

static void *xcalloc(size_t n, size_t s) { return calloc(n ? n : 1, s); }

void orc_tree_free(orc_tree *t)
{
    if (!t) return;
    void *ptrs[] = {t->m, t->x, t->y, t->z, t->cx, t->cy, t->cz, t->half, t->L2, t->mass,
                    t->qx, t->qy, t->qz, t->particle, t->type, t->depth, t->child, t->khi,
                    t->klo, t->quad, t->oct, t->count, t->first, t->last, t->order,
                    t->groups, t->gx, t->gy, t->gz, t->gm, t->rmax, t->force, t->hess,
                    t->third};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(t);
}

/* pre_init (:178-282): capacity 3N for N >= 1000 else 10N; method code to
 * multipole order (1, 4, 8); group size default 64. */
orc_tree *orc_tree_new(int64_t N, const double *m, int method, int group_size)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    if (!t) return NULL;
    t->N = N;
    t->method = method;
    t->max_depth = 25;
    t->side_amp = 1.2;
    t->group_size = group_size > 0 ? group_size : 64;
    switch (method) {
    case 10: case 20: case 30: t->multipole = 1; break;
    case 11: case 21: case 31: t->multipole = 4; break;
    case 12: case 22: case 32: t->multipole = 8; break;
    default: t->multipole = 1; break;
    }
    t->cap = (N >= 1000) ? (int)(3 * N) : (int)(10 * N);
    if (t->cap < 1) t->cap = 1;
    size_t c = (size_t)t->cap;
    t->m = xcalloc(N, 8); t->x = xcalloc(N, 8); t->y = xcalloc(N, 8); t->z = xcalloc(N, 8);
    t->cx = xcalloc(c, 8); t->cy = xcalloc(c, 8); t->cz = xcalloc(c, 8);
    t->half = xcalloc(c, 8); t->L2 = xcalloc(c, 8); t->mass = xcalloc(c, 8);
    t->qx = xcalloc(c, 8); t->qy = xcalloc(c, 8); t->qz = xcalloc(c, 8);
    t->particle = xcalloc(c, 4); t->type = xcalloc(c, 4); t->depth = xcalloc(c, 4);
    t->child = xcalloc(8 * c, 4); t->khi = xcalloc(c, 8); t->klo = xcalloc(c, 4);
    t->quad = xcalloc(6 * c, 8); t->oct = xcalloc(14 * c, 8);
    if (method >= 20 && method <= 22) {
        t->rmax = xcalloc(c, 8); t->force = xcalloc(3 * c, 8);
        t->hess = xcalloc(6 * c, 8); t->third = xcalloc(10 * c, 8);
    }
    if (!t->child || !t->oct) { orc_tree_free(t); return NULL; }
    memcpy(t->m, m, sizeof(double) * (size_t)N);
    return t;
}

/* clear (:284-318) */
static void tree_clear(orc_tree *t)
{
    size_t c = (size_t)t->cap;
    t->nn = 0;
    t->most_depth = 0;
    memset(t->cx, 0, c * 8); memset(t->cy, 0, c * 8); memset(t->cz, 0, c * 8);
    memset(t->half, 0, c * 8); memset(t->L2, 0, c * 8); memset(t->mass, 0, c * 8);
    memset(t->qx, 0, c * 8); memset(t->qy, 0, c * 8); memset(t->qz, 0, c * 8);
    for (size_t k = 0; k < c; ++k) { t->particle[k] = -1; t->type[k] = 0; t->depth[k] = 0; }
    memset(t->quad, 0, 6 * c * 8); memset(t->oct, 0, 14 * c * 8);
    if (t->rmax) {
        memset(t->rmax, 0, c * 8); memset(t->force, 0, 3 * c * 8);
        memset(t->hess, 0, 6 * c * 8); memset(t->third, 0, 10 * c * 8);
    }
}

/* add_node (:445-479).  khi/klo record the node's path (digits 4*xb+2*yb+zb
 * with xb = x>=cx, yb = y>=cy, zb = z>=cz) for canonical topology export. */
static int add_node(orc_tree *t, double cx, double cy, double cz, double side, int depth,
                    uint64_t khi, uint32_t klo)
{
    int idx = t->nn;
    if (idx >= t->cap) { t->status = ORC_ERR_CAPACITY; return -1; }
    t->nn = idx + 1;
    for (int r = 0; r < 8; ++r) t->child[8 * idx + r] = -1;
    t->type[idx] = 1;
    t->depth[idx] = depth;
    if (depth > t->most_depth) t->most_depth = depth;
    t->cx[idx] = cx; t->cy[idx] = cy; t->cz[idx] = cz;
    t->half[idx] = side / 2.0;
    t->L2[idx] = side * side;
    t->mass[idx] = 0.0; t->qx[idx] = 0.0; t->qy[idx] = 0.0; t->qz[idx] = 0.0;
    t->particle[idx] = -1;
    t->khi[idx] = khi; t->klo[idx] = klo;
    return idx;
}

/* allocate_subnode (:481-513); r is the 0-based reference octant
 * (reference index - 1): bit0 = x>=cx, bit1 = y<cy, bit2 = z<cz. */
static int allocate_subnode(orc_tree *t, int node, int r)
{
    double h = t->half[node];
    double h_half = h / 2.0;
    int d = t->depth[node] + 1;
    int sx = (r & 1) ? 1 : -1;
    int sy = (r & 2) ? -1 : 1;
    int sz = (r & 4) ? -1 : 1;
    double cxs = t->cx[node] + (double)sx * h_half;
    double cys = t->cy[node] + (double)sy * h_half;
    double czs = t->cz[node] + (double)sz * h_half;
    int digit = ((r & 1) ? 4 : 0) | ((r & 2) ? 0 : 2) | ((r & 4) ? 0 : 1);
    uint64_t khi = t->khi[node];
    uint32_t klo = t->klo[node];
    if (d <= 21) khi |= (uint64_t)digit << (3 * (21 - d));
    else klo |= (uint32_t)digit << (3 * (25 - d));
    int sub = add_node(t, cxs, cys, czs, h, d, khi, klo);
    if (sub < 0) return -1;
    t->child[8 * node + r] = sub;
    return sub;
}

/* add_to_subnode (:515-545): octant choice, create if missing. */
static int subnode_for(orc_tree *t, int node, int p)
{
    int ix = t->x[p] >= t->cx[node] ? 1 : 0;
    int iy = t->y[p] < t->cy[node] ? 1 : 0;
    int iz = t->z[p] < t->cz[node] ? 1 : 0;
    int r = ix + 2 * iy + 4 * iz;
    int sub = t->child[8 * node + r];
    if (sub == -1) sub = allocate_subnode(t, node, r);
    return sub;
}

/* add (:547-580), iterative: the recursion only ever continues with the
 * incoming particle p; the displaced particle always lands in a fresh leaf. */
static int tree_add(orc_tree *t, int node, int p)
{
    for (;;) {
        if (t->particle[node] == -1 && t->type[node] == 1) {
            t->particle[node] = p;
            return 0;
        }
        if (t->type[node] == 1) {
            t->type[node] = 2;
            int old = t->particle[node];
            t->particle[node] = -1;
            if (t->depth[node] < t->max_depth) {
                int s = subnode_for(t, node, old);
                if (s < 0) return -1;
                if (tree_add(t, s, old) != 0) return -1;
            }
        }
        if (t->depth[node] < t->max_depth) {
            int s = subnode_for(t, node, p);
            if (s < 0) return -1;
            node = s;
            continue;
        }
        return 0;
    }
}

/* incorporate_particles (:582-628) */
static void incorporate_particles(orc_tree *t)
{
    for (int k = t->nn - 1; k >= 0; --k) {
        if (t->type[k] == 1) {
            int p = t->particle[k];
            if (p == -1) continue;
            t->mass[k] = t->m[p]; t->qx[k] = t->x[p]; t->qy[k] = t->y[p]; t->qz[k] = t->z[p];
        } else {
            t->mass[k] = 0.0; t->qx[k] = 0.0; t->qy[k] = 0.0; t->qz[k] = 0.0;
            for (int r = 0; r < 8; ++r) {
                int c = t->child[8 * k + r];
                if (c == -1) continue;
                t->mass[k] = t->mass[k] + t->mass[c];
                t->qx[k] = t->qx[k] + t->mass[c] * t->qx[c];
                t->qy[k] = t->qy[k] + t->mass[c] * t->qy[c];
                t->qz[k] = t->qz[k] + t->mass[c] * t->qz[c];
            }
            t->qx[k] = t->qx[k] / t->mass[k];
            t->qy[k] = t->qy[k] / t->mass[k];
            t->qz[k] = t->qz[k] / t->mass[k];
        }
    }
}

/* evaluate_multipole (:744-882) */
static void evaluate_multipole(orc_tree *t)
{
    int oct = t->multipole > 4;
    for (int k = t->nn - 1; k >= 0; --k) {
        if (t->type[k] == 1) {
            int p = t->particle[k];
            if (p == -1) continue;
            t->mass[k] = t->m[p]; t->qx[k] = t->x[p]; t->qy[k] = t->y[p]; t->qz[k] = t->z[p];
            continue;
        }
        if (t->depth[k] == t->max_depth) continue;
        t->mass[k] = 0.0; t->qx[k] = 0.0; t->qy[k] = 0.0; t->qz[k] = 0.0;
        for (int r = 0; r < 8; ++r) {
            int c = t->child[8 * k + r];
            if (c == -1) continue;
            t->mass[k] = t->mass[k] + t->mass[c];
            t->qx[k] = t->qx[k] + t->mass[c] * t->qx[c];
            t->qy[k] = t->qy[k] + t->mass[c] * t->qy[c];
            t->qz[k] = t->qz[k] + t->mass[c] * t->qz[c];
        }
        t->qx[k] = t->qx[k] / t->mass[k];
        t->qy[k] = t->qy[k] / t->mass[k];
        t->qz[k] = t->qz[k] / t->mass[k];
        double *Q = t->quad + 6 * (size_t)k;
        double *O = t->oct + 14 * (size_t)k;
        for (int r = 0; r < 8; ++r) {
            int c = t->child[8 * k + r];
            if (c == -1) continue;
            double pm = t->mass[c];
            double dxi = t->qx[c] - t->qx[k];
            double dyi = t->qy[c] - t->qy[k];
            double dzi = t->qz[c] - t->qz[k];
            const double *Qc = t->quad + 6 * (size_t)c;
            const double *Oc = t->oct + 14 * (size_t)c;
            if (t->type[c] == 2) {
                for (int i = 0; i < 6; ++i) Q[i] = Q[i] + Qc[i];
                if (oct) for (int i = 0; i < 14; ++i) O[i] = O[i] + Oc[i];
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
            }
            if (oct && t->type[c] == 2) {
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
        if (oct) {
            double s = 0.0;
            for (int i = 0; i < 10; ++i) s = s + O[i];
            O[10] = s;
            O[11] = (O[0] + O[1]) + O[2];
            O[12] = (O[3] + O[4]) + O[5];
            O[13] = (O[6] + O[7]) + O[8];
        }
    }
}

/* prepare_groups (:1722-1809) */
static int prepare_groups(orc_tree *t)
{
    size_t c = (size_t)t->cap;
    if (!t->count) {
        t->count = xcalloc(c, 4); t->first = xcalloc(c, 4); t->last = xcalloc(c, 4);
        t->groups = xcalloc(c, 4); t->order = xcalloc(t->N, 4);
        t->gx = xcalloc(t->N, 8); t->gy = xcalloc(t->N, 8);
        t->gz = xcalloc(t->N, 8); t->gm = xcalloc(t->N, 8);
    }
    for (int k = t->nn - 1; k >= 0; --k) {
        if (t->type[k] == 1) {
            t->count[k] = t->particle[k] != -1 ? 1 : 0;
        } else {
            int s = 0;
            for (int r = 0; r < 8; ++r) {
                int ch = t->child[8 * k + r];
                if (ch != -1) s += t->count[ch];
            }
            t->count[k] = s;
        }
    }
    if (t->nn == 0 || t->count[0] != t->N) return ORC_ERR_LOST;
    t->first[0] = 0;
    for (int k = 0; k < t->nn; ++k) {
        int f = t->first[k];
        if (t->type[k] == 1) {
            if (t->count[k] == 1) t->order[f] = t->particle[k];
        } else {
            for (int r = 0; r < 8; ++r) {
                int ch = t->child[8 * k + r];
                if (ch == -1) continue;
                t->first[ch] = f;
                f += t->count[ch];
            }
        }
        t->last[k] = t->first[k] + t->count[k] - 1;
    }
    for (int64_t i = 0; i < t->N; ++i) {
        int p = t->order[i];
        t->gm[i] = t->m[p]; t->gx[i] = t->x[p]; t->gy[i] = t->y[p]; t->gz[i] = t->z[p];
    }
    t->ngroups = 0;
    int stack[8 * (25 + 2)];
    int top = 0;
    stack[top++] = 0;
    while (top > 0) {
        int k = stack[--top];
        if (t->count[k] == 0) continue;
        if (t->count[k] <= t->group_size || t->type[k] == 1) {
            t->groups[t->ngroups++] = k;
        } else {
            for (int r = 0; r < 8; ++r) {
                int ch = t->child[8 * k + r];
                if (ch == -1) continue;
                stack[top++] = ch;
            }
        }
    }
    return ORC_OK;
}

/* init (:320-387) */
int orc_tree_init(orc_tree *t, const double *x, const double *y, const double *z)
{
    t->status = ORC_OK;
    tree_clear(t);
    memcpy(t->x, x, sizeof(double) * (size_t)t->N);
    memcpy(t->y, y, sizeof(double) * (size_t)t->N);
    memcpy(t->z, z, sizeof(double) * (size_t)t->N);
    /* node_size_center (:424-443) */
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY,
           zmin = INFINITY, zmax = -INFINITY;
    for (int64_t p = 0; p < t->N; ++p) {
        if (x[p] < xmin) xmin = x[p];
        if (x[p] > xmax) xmax = x[p];
        if (y[p] < ymin) ymin = y[p];
        if (y[p] > ymax) ymax = y[p];
        if (z[p] < zmin) zmin = z[p];
        if (z[p] > zmax) zmax = z[p];
    }
    double c0 = 0.5 * (xmin + xmax), c1 = 0.5 * (ymin + ymax), c2 = 0.5 * (zmin + zmax);
    double ext = xmax - xmin;
    if (ymax - ymin > ext) ext = ymax - ymin;
    if (zmax - zmin > ext) ext = zmax - zmin;
    int root = add_node(t, c0, c1, c2, t->side_amp * ext, 0, 0, 0);
    if (root < 0) return t->status;
    for (int64_t p = 0; p < t->N; ++p)
        if (tree_add(t, root, (int)p) != 0) return t->status;
    if (t->multipole == 1) incorporate_particles(t);
    else evaluate_multipole(t);
    if (t->method >= 30 && t->method <= 32) {
        int s = prepare_groups(t);
        if (s) { t->status = s; return s; }
    }
    return ORC_OK;
}

int orc_tree_nnodes(const orc_tree *t) { return t->nn; }
int orc_tree_ngroups(const orc_tree *t) { return t->ngroups; }

/* Copies the node store out (caller buffers sized >= nn; child 8*nn). */
void orc_tree_export(const orc_tree *t, double *cx, double *cy, double *cz, double *half,
                     double *L2, double *mass, double *qx, double *qy, double *qz,
                     int *particle, int *type, int *depth, int *child, uint64_t *khi,
                     uint32_t *klo, double *quad, double *oct)
{
    size_t n = (size_t)t->nn;
    if (cx) memcpy(cx, t->cx, n * 8);
    if (cy) memcpy(cy, t->cy, n * 8);
    if (cz) memcpy(cz, t->cz, n * 8);
    if (half) memcpy(half, t->half, n * 8);
    if (L2) memcpy(L2, t->L2, n * 8);
    if (mass) memcpy(mass, t->mass, n * 8);
    if (qx) memcpy(qx, t->qx, n * 8);
    if (qy) memcpy(qy, t->qy, n * 8);
    if (qz) memcpy(qz, t->qz, n * 8);
    if (particle) memcpy(particle, t->particle, n * 4);
    if (type) memcpy(type, t->type, n * 4);
    if (depth) memcpy(depth, t->depth, n * 4);
    if (child) memcpy(child, t->child, 8 * n * 4);
    if (khi) memcpy(khi, t->khi, n * 8);
    if (klo) memcpy(klo, t->klo, n * 4);
    if (quad) memcpy(quad, t->quad, 6 * n * 8);
    if (oct) memcpy(oct, t->oct, 14 * n * 8);
}

/* bh_forces_over_p (:895-1078).  counts[0] += leaf / monopole interactions,
 * counts[1] += multipole (quadrupole or octupole) cell interactions. */
static void bh_forces_p(const orc_tree *t, int p, double theta2, double G, double eps2,
                        int *stack, double out[3], int64_t counts[2])
{
    double pm = t->m[p], px = t->x[p], py = t->y[p], pz = t->z[p];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    int top = 0;
    stack[top++] = 0;
    while (top > 0) {
        int k = stack[--top];
        if (t->mass[k] == 0.0) continue;
        if (t->type[k] == 1 && t->particle[k] == p) continue;
        double dx = t->qx[k] - px;
        double dy = t->qy[k] - py;
        double dz = t->qz[k] - pz;
        double dist2 = dx * dx + dy * dy + dz * dz;
        double L2 = t->L2[k];
        if (t->type[k] == 1 || L2 <= theta2 * dist2 || t->depth[k] == t->max_depth) {
            if (t->type[k] == 1 || t->multipole == 1) {
                double rinv = 1.0 / sqrt(dist2 + eps2);
                rinv = rinv * rinv * rinv;
                double f = G * pm * t->mass[k] * rinv;
                f0 = f0 + f * dx;
                f1 = f1 + f * dy;
                f2 = f2 + f * dz;
                counts[0]++;
            } else {
                double invR = 1.0 / sqrt(dist2 + eps2);
                double invR2 = 1.0 / (dist2 + eps2);
                double invR3 = invR * invR2;
                double invR5 = invR3 * invR * invR;
                double invR7 = invR5 * invR * invR;
                double invR9 = invR7 * invR * invR;
                double Mx = dx * invR3 * t->mass[k];
                double My = dy * invR3 * t->mass[k];
                double Mz = dz * invR3 * t->mass[k];
                const double *Q = t->quad + 6 * (size_t)k;
                double mxi2 = Q[0], myi2 = Q[1], mzi2 = Q[2], mxyi = Q[3], mxzi = Q[4],
                       myzi = Q[5];
                double dx2 = 15.0 * dx * dx * invR2 - 3.0;
                double dy2 = 15.0 * dy * dy * invR2 - 3.0;
                double dz2 = 15.0 * dz * dz * invR2 - 3.0;
                double dxyz = 15.0 * dx * dy * dz * invR2;
                double Qx = (mxi2 * (dx2 - 6.0) + myi2 * dy2 + mzi2 * dz2) * 0.5 * dx;
                double Qy = (myi2 * (dy2 - 6.0) + mzi2 * dz2 + mxi2 * dx2) * 0.5 * dy;
                double Qz = (mzi2 * (dz2 - 6.0) + mxi2 * dx2 + myi2 * dy2) * 0.5 * dz;
                Qx = Qx + mxyi * dy * dx2 + mxzi * dz * dx2 + myzi * dxyz;
                Qy = Qy + myzi * dz * dy2 + mxyi * dx * dy2 + mxzi * dxyz;
                Qz = Qz + mxzi * dx * dz2 + myzi * dy * dz2 + mxyi * dxyz;
                if (t->multipole == 4) {
                    f0 = f0 + G * pm * (Mx + invR5 * Qx);
                    f1 = f1 + G * pm * (My + invR5 * Qy);
                    f2 = f2 + G * pm * (Mz + invR5 * Qz);
                } else {
                    const double *O = t->oct + 14 * (size_t)k;
                    double Ox = -(1.5 * invR5 * O[11]);
                    double Oy = -(1.5 * invR5 * O[12]);
                    double Oz = -(1.5 * invR5 * O[13]);
                    double Oc = 7.5 * invR7 * (dx * O[11] + dy * O[12] + dz * O[13]);
                    Ox = Ox + dx * Oc;
                    Oy = Oy + dy * Oc;
                    Oz = Oz + dz * Oc;
                    double ex2 = dx * dx, ey2 = dy * dy, ez2 = dz * dz;
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
                counts[1]++;
            }
        } else {
            for (int r = 0; r < 8; ++r) {
                int c = t->child[8 * k + r];
                if (c != -1) stack[top++] = c;
            }
        }
    }
    out[0] = f0; out[1] = f1; out[2] = f2;
}

/* Forces over a list of targets (all when tidx == NULL); F is (3,nt).
 * counts (may be NULL) gets [n_leaf_or_mono, n_multipole] per target. */
void orc_tree_bh_forces(const orc_tree *t, int64_t nt, const int64_t *tidx, double theta2,
                        double G, double eps2, double *F, int64_t *counts)
{
    if (!tidx) nt = t->N;
#pragma omp parallel
    {
        int *stack = (int *)malloc(sizeof(int) * ((size_t)t->nn + 16));
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < nt; ++i) {
            int p = tidx ? (int)tidx[i] : (int)i;
            int64_t c[2] = {0, 0};
            bh_forces_p(t, p, theta2, G, eps2, stack, F + 3 * i, c);
            if (counts) { counts[2 * i] = c[0]; counts[2 * i + 1] = c[1]; }
        }
        free(stack);
    }
}

/* group_worker (:1833-2095), serial over groups (the per-group arithmetic and
 * order do not depend on the thread count).  F is (3,N) in original order.
 * counts (may be NULL): [list P2P entries incl. in-group, list cells] per
 * group, written to counts[2*g]. */
int orc_tree_forces_groups(const orc_tree *t, double theta2, double G, double eps2, double *F,
                           int64_t *counts)
{
    if (!t->count) return ORC_ERR_ARGS;
    int mp = t->multipole;
    int ncd = mp == 1 ? 4 : (mp == 4 ? 10 : 24);
#pragma omp parallel
    {
        size_t scap = 2048, ccap = 512;
        double *sx = malloc(scap * 8), *sy = malloc(scap * 8), *sz = malloc(scap * 8),
               *sm = malloc(scap * 8);
        double *cd = calloc(ccap * (size_t)ncd, 8);
        int stack[8 * (25 + 2)];
#pragma omp for schedule(dynamic, 1)
        for (int gi = 0; gi < t->ngroups; ++gi) {
            int g = t->groups[gi];
            int gf = t->first[g], gl = t->last[g];
            double gmin[3] = {t->cx[g] - t->half[g], t->cy[g] - t->half[g], t->cz[g] - t->half[g]};
            double gmax[3] = {t->cx[g] + t->half[g], t->cy[g] + t->half[g], t->cz[g] + t->half[g]};
            size_t ns = 0, nc = 0;
            int top = 0;
            stack[top++] = 0;
            while (top > 0) {
                int k = stack[--top];
                if (k == g) continue;
                if (t->mass[k] == 0.0) continue;
                if (t->first[k] <= gf && t->last[k] >= gl) {
                    for (int r = 0; r < 8; ++r) {
                        int c = t->child[8 * k + r];
                        if (c == -1) continue;
                        stack[top++] = c;
                    }
                    continue;
                }
                if (t->type[k] == 1) {
                    int a = t->first[k], b = t->last[k];
                    size_t cnt = (size_t)(b - a + 1);
                    while (ns + cnt > scap) {
                        scap *= 2;
                        sx = realloc(sx, scap * 8); sy = realloc(sy, scap * 8);
                        sz = realloc(sz, scap * 8); sm = realloc(sm, scap * 8);
                    }
                    for (int j = a; j <= b; ++j) {
                        sm[ns] = t->gm[j]; sx[ns] = t->gx[j]; sy[ns] = t->gy[j]; sz[ns] = t->gz[j];
                        ns++;
                    }
                    continue;
                }
                double cc[3] = {t->qx[k], t->qy[k], t->qz[k]};
                double dist2 = 0.0;
                for (int i = 0; i < 3; ++i) {
                    double dd = gmin[i] - cc[i];
                    if (dd < 0.0) dd = 0.0;
                    double e = cc[i] - gmax[i];
                    if (e > dd) dd = e;
                    dist2 = dist2 + dd * dd;
                }
                if (t->L2[k] <= theta2 * dist2 || t->depth[k] >= t->max_depth) {
                    if (nc + 1 > ccap) {
                        size_t nccap = ccap * 2;
                        double *ncdp = calloc(nccap * (size_t)ncd, 8);
                        memcpy(ncdp, cd, ccap * (size_t)ncd * 8);
                        free(cd);
                        cd = ncdp;
                        ccap = nccap;
                    }
                    double *rec = cd + (size_t)ncd * nc;
                    rec[0] = t->mass[k];
                    rec[1] = cc[0]; rec[2] = cc[1]; rec[3] = cc[2];
                    if (mp >= 4) memcpy(rec + 4, t->quad + 6 * (size_t)k, 6 * 8);
                    if (mp >= 8) memcpy(rec + 10, t->oct + 14 * (size_t)k, 14 * 8);
                    nc++;
                } else {
                    for (int r = 0; r < 8; ++r) {
                        int c = t->child[8 * k + r];
                        if (c == -1) continue;
                        stack[top++] = c;
                    }
                }
            }
            if (counts) {
                counts[2 * (size_t)gi] = (int64_t)ns + (gl - gf);
                counts[2 * (size_t)gi + 1] = (int64_t)nc;
            }
            for (int i = gf; i <= gl; ++i) {
                double xi = t->gx[i], yi = t->gy[i], zi = t->gz[i];
                double ax = 0.0, ay = 0.0, az = 0.0;
                for (int j = gf; j <= gl; ++j) {
                    if (j == i) continue;
                    double dx = t->gx[j] - xi, dy = t->gy[j] - yi, dz = t->gz[j] - zi;
                    double dist2 = dx * dx + dy * dy + dz * dz + eps2;
                    double invr = 1.0 / sqrt(dist2);
                    double invr3 = invr * invr * invr;
                    double f = t->gm[j] * invr3;
                    ax = ax + f * dx; ay = ay + f * dy; az = az + f * dz;
                }
                for (size_t j = 0; j < ns; ++j) {
                    double dx = sx[j] - xi, dy = sy[j] - yi, dz = sz[j] - zi;
                    double dist2 = dx * dx + dy * dy + dz * dz + eps2;
                    double invr = 1.0 / sqrt(dist2);
                    double invr3 = invr * invr * invr;
                    double f = sm[j] * invr3;
                    ax = ax + f * dx; ay = ay + f * dy; az = az + f * dz;
                }
                for (size_t j = 0; j < nc; ++j) {
                    const double *c = cd + (size_t)ncd * j; /* c[i] is cd(i+1, j) */
                    double dx = c[1] - xi, dy = c[2] - yi, dz = c[3] - zi;
                    double dist2 = dx * dx + dy * dy + dz * dz + eps2;
                    double invr = 1.0 / sqrt(dist2);
                    double invr3 = invr * invr * invr;
                    if (mp == 1) {
                        double f = c[0] * invr3;
                        ax = ax + f * dx; ay = ay + f * dy; az = az + f * dz;
                        continue;
                    }
                    double invr5 = invr3 * invr * invr;
                    double invr7 = invr5 * invr * invr;
                    double qx = c[4] * dx + c[7] * dy + c[8] * dz;
                    double qy = c[7] * dx + c[5] * dy + c[9] * dz;
                    double qz = c[8] * dx + c[9] * dy + c[6] * dz;
                    double dqd = dx * qx + dy * qy + dz * qz;
                    double trq = c[4] + c[5] + c[6];
                    double s = c[0] * invr3 - 1.5 * trq * invr5 + 7.5 * dqd * invr7;
                    ax = ax + s * dx - 3.0 * qx * invr5;
                    ay = ay + s * dy - 3.0 * qy * invr5;
                    az = az + s * dz - 3.0 * qz * invr5;
                    if (mp == 8) {
                        double invr9 = invr7 * invr * invr;
                        double oc = 7.5 * invr7 * (dx * c[21] + dy * c[22] + dz * c[23]);
                        double ox = -1.5 * invr5 * c[21] + dx * oc;
                        double oy = -1.5 * invr5 * c[22] + dy * oc;
                        double oz = -1.5 * invr5 * c[23] + dz * oc;
                        ox = ox + 7.5 * invr7 * (dx * dx * c[10] + dy * dy * c[11] + dz * dz * c[12]);
                        ox = ox + 15.0 * invr7 * (dx * dy * c[13] + dx * dz * c[16] + dy * dz * c[19]);
                        oy = oy + 7.5 * invr7 * (dx * dx * c[13] + dy * dy * c[14] + dz * dz * c[15]);
                        oy = oy + 15.0 * invr7 * (dx * dy * c[11] + dx * dz * c[19] + dy * dz * c[17]);
                        oz = oz + 7.5 * invr7 * (dx * dx * c[16] + dy * dy * c[17] + dz * dz * c[18]);
                        oz = oz + 15.0 * invr7 * (dx * dy * c[19] + dx * dz * c[12] + dy * dz * c[15]);
                        oc = dx * dx * dx * c[10] + dy * dy * dy * c[14] + dz * dz * dz * c[18];
                        oc = oc + 3.0 * (dx * dx * dy * c[13] + dx * dx * dz * c[16] + dx * dy * dy * c[11]);
                        oc = oc + 3.0 * (dy * dy * dz * c[17] + dx * dz * dz * c[12] + dy * dz * dz * c[15]);
                        oc = oc + 6.0 * dx * dy * dz * c[19];
                        oc = -17.5 * invr9 * oc;
                        ox = ox + dx * oc; oy = oy + dy * oc; oz = oz + dz * oc;
                        ax = ax + ox; ay = ay + oy; az = az + oz;
                    }
                }
                int k = t->order[i];
                F[3 * (size_t)k + 0] = G * t->gm[i] * ax;
                F[3 * (size_t)k + 1] = G * t->gm[i] * ay;
                F[3 * (size_t)k + 2] = G * t->gm[i] * az;
            }
        }
        free(sx); free(sy); free(sz); free(sm); free(cd);
    }
    return ORC_OK;
}

/* Group bookkeeping export: groups[ngroups] (node ids), first/last/count per node. */
void orc_tree_groups_export(const orc_tree *t, int *groups, int *first, int *last, int *count,
                            int *order)
{
    size_t n = (size_t)t->nn;
    if (groups) memcpy(groups, t->groups, (size_t)t->ngroups * 4);
    if (first) memcpy(first, t->first, n * 4);
    if (last) memcpy(last, t->last, n * 4);
    if (count) memcpy(count, t->count, n * 4);
    if (order) memcpy(order, t->order, (size_t)t->N * 4);
}

/* ---------------- Dehnen (:1095-1712) ---------------- */
static void evaluate_rmax(orc_tree *t)
{
    memset(t->rmax, 0, (size_t)t->cap * 8);
    for (int k = t->nn - 1; k >= 0; --k) {
        if (t->type[k] == 1) continue;
        double rc = 0.0;
        for (int r = 0; r < 8; ++r) {
            int c = t->child[8 * k + r];
            if (c == -1) continue;
            if (t->mass[c] == 0.0) continue;
            double dx = t->qx[c] - t->qx[k];
            double dy = t->qy[c] - t->qy[k];
            double dz = t->qz[c] - t->qz[k];
            double dist = sqrt(dx * dx + dy * dy + dz * dz);
            double v = dist + t->rmax[c];
            if (v > rc) rc = v;
        }
        t->rmax[k] = rc;
    }
}

/* ind_oct (:234-242), 0-based octupole slot for (j,k,l). */
static const int IND_OCT[3][3][3] = {
    {{0, 3, 6}, {3, 1, 9}, {6, 9, 2}},
    {{3, 1, 9}, {1, 4, 7}, {9, 7, 5}},
    {{6, 9, 2}, {9, 7, 5}, {2, 5, 8}}};

static void m2l_quadrupole(const orc_tree *t, int a, int b, double G, double eps2, double fa[3],
                           double fb[3], double h[6], double tt[10], double fc[3])
{
    double Ma = t->mass[a], Mb = t->mass[b];
    double dx = t->qx[a] - t->qx[b];
    double dy = t->qy[a] - t->qy[b];
    double dz = t->qz[a] - t->qz[b];
    double r2 = dx * dx + dy * dy + dz * dz + eps2;
    double rinv3 = 1.0 / pow(r2, 1.5);
    double rinv5 = 1.0 / pow(r2, 2.5);
    double rinv7 = 1.0 / pow(r2, 3.5);
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
    for (int i = 0; i < 6; ++i) {
        Qa[i] = t->quad[6 * (size_t)a + i] / Ma;
        Qb[i] = t->quad[6 * (size_t)b + i] / Mb;
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
    for (int i = 0; i < 3; ++i) {
        fa[i] = Mb * (D1[i] + 0.5 * CQb[i]);
        fb[i] = -(Ma * (D1[i] + 0.5 * CQa[i]));
    }
    for (int i = 0; i < 6; ++i) h[i] = D2[i];
    for (int i = 0; i < 10; ++i) tt[i] = D3[i];
    if (t->multipole >= 8) {
        double R[3] = {dx, dy, dz};
        double ci[3] = {0, 0, 0}, CRR[3] = {0, 0, 0};
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) {
                    int io = IND_OCT[j][k][l];
                    double Cjkl = Mb * t->oct[14 * (size_t)a + io] - Ma * t->oct[14 * (size_t)b + io];
                    if (k == l) ci[j] = ci[j] + Cjkl;
                    CRR[j] = CRR[j] + Cjkl * R[k] * R[l];
                }
        for (int j = 0; j < 3; ++j)
            fc[j] = 0.5 * (3.0 * G * rinv5 * ci[j] - 15.0 * G * rinv7 * CRR[j]);
    }
}

int orc_tree_dehnen(orc_tree *t, double theta2, double eps2, double G, double *F,
                    int64_t *npairs)
{
    if (!t->rmax) return ORC_ERR_ARGS;
    size_t c = (size_t)t->cap;
    memset(t->force, 0, 3 * c * 8);
    memset(t->hess, 0, 6 * c * 8);
    memset(t->third, 0, 10 * c * 8);
    evaluate_rmax(t);
    size_t scap = 36 * (size_t)t->nn + 16;
    int *sa = malloc(scap * 4), *sb = malloc(scap * 4);
    if (!sa || !sb) { free(sa); free(sb); return ORC_ERR_ALLOC; }
    int64_t pairs = 0;
    size_t top = 0;
    sa[top] = 0; sb[top] = 0; top++;
    while (top > 0) {
        top--;
        int a = sa[top], b = sb[top];
        if (t->mass[a] == 0.0) continue;
        if (t->mass[b] == 0.0) continue;
        if (a == b) {
            if (t->type[a] == 1) continue;
            for (int i = 0; i < 8; ++i) {
                int ca = t->child[8 * a + i];
                if (ca == -1) continue;
                for (int j = i; j < 8; ++j) {
                    int cb = t->child[8 * a + j];
                    if (cb == -1) continue;
                    if (top + 1 > scap) { free(sa); free(sb); return ORC_ERR_CAPACITY; }
                    sa[top] = ca; sb[top] = cb; top++;
                }
            }
            continue;
        }
        double dx = t->qx[a] - t->qx[b];
        double dy = t->qy[a] - t->qy[b];
        double dz = t->qz[a] - t->qz[b];
        double dist2 = dx * dx + dy * dy + dz * dz;
        double rs = t->rmax[a] + t->rmax[b];
        double rsum2 = rs * rs;
        if (rsum2 <= theta2 * dist2) {
            pairs++;
            if (t->multipole == 1) {
                double ex = t->qx[b] - t->qx[a];
                double ey = t->qy[b] - t->qy[a];
                double ez = t->qz[b] - t->qz[a];
                double d2 = ex * ex + ey * ey + ez * ez;
                double rinv3 = 1.0 / pow(d2 + eps2, 1.5);
                double factor = G * rinv3;
                double f[3] = {factor * ex, factor * ey, factor * ez};
                for (int i = 0; i < 3; ++i) {
                    t->force[3 * (size_t)a + i] = t->force[3 * (size_t)a + i] + t->mass[b] * f[i];
                    t->force[3 * (size_t)b + i] = t->force[3 * (size_t)b + i] - t->mass[a] * f[i];
                }
            } else {
                double fa[3], fb[3], h[6], tt[10], fc[3] = {0, 0, 0};
                m2l_quadrupole(t, a, b, G, eps2, fa, fb, h, tt, fc);
                double ma = t->mass[a], mb = t->mass[b];
                for (int i = 0; i < 3; ++i) {
                    t->force[3 * (size_t)a + i] = t->force[3 * (size_t)a + i] + fa[i] + fc[i] / ma;
                    t->force[3 * (size_t)b + i] = t->force[3 * (size_t)b + i] + fb[i] - fc[i] / mb;
                }
                for (int i = 0; i < 6; ++i) {
                    t->hess[6 * (size_t)a + i] = t->hess[6 * (size_t)a + i] + mb * h[i];
                    t->hess[6 * (size_t)b + i] = t->hess[6 * (size_t)b + i] + ma * h[i];
                }
                for (int i = 0; i < 10; ++i) {
                    t->third[10 * (size_t)a + i] = t->third[10 * (size_t)a + i] + mb * tt[i];
                    t->third[10 * (size_t)b + i] = t->third[10 * (size_t)b + i] - ma * tt[i];
                }
            }
            continue;
        }
        if (t->rmax[a] >= t->rmax[b]) {
            for (int i = 0; i < 8; ++i) {
                int ca = t->child[8 * a + i];
                if (ca == -1) continue;
                if (top + 1 > scap) { free(sa); free(sb); return ORC_ERR_CAPACITY; }
                sa[top] = ca; sb[top] = b; top++;
            }
        } else {
            for (int i = 0; i < 8; ++i) {
                int cb = t->child[8 * b + i];
                if (cb == -1) continue;
                if (top + 1 > scap) { free(sa); free(sb); return ORC_ERR_CAPACITY; }
                sa[top] = a; sb[top] = cb; top++;
            }
        }
    }
    free(sa); free(sb);
    if (npairs) *npairs = pairs;
    /* dehnen_forces (:1565-1678) */
    memset(F, 0, 3 * (size_t)t->N * 8);
    for (int k = 0; k < t->nn; ++k) {
        if (t->type[k] == 1) continue;
        const double *H = t->hess + 6 * (size_t)k;
        const double *T = t->third + 10 * (size_t)k;
        double hxx = H[0], hyy = H[1], hzz = H[2], hxy = H[3], hxz = H[4], hyz = H[5];
        double txxx = T[0], txxy = T[1], txxz = T[2], txyy = T[3], txyz = T[4], txzz = T[5],
               tyyy = T[6], tyyz = T[7], tyzz = T[8], tzzz = T[9];
        for (int r = 0; r < 8; ++r) {
            int ch = t->child[8 * k + r];
            if (ch == -1) continue;
            if (t->mass[ch] == 0.0) continue;
            double fcd[3];
            for (int i = 0; i < 3; ++i) fcd[i] = t->force[3 * (size_t)ch + i];
            for (int i = 0; i < 3; ++i) fcd[i] = fcd[i] + t->force[3 * (size_t)k + i];
            if (t->multipole >= 4) {
                double dx = t->qx[ch] - t->qx[k];
                double dy = t->qy[ch] - t->qy[k];
                double dz = t->qz[ch] - t->qz[k];
                fcd[0] = fcd[0] + hxx * dx + hxy * dy + hxz * dz + txxy * dx * dy + txxz * dx * dz +
                         txyz * dy * dz + 0.5 * (txxx * dx * dx + txyy * dy * dy + txzz * dz * dz);
                fcd[1] = fcd[1] + hxy * dx + hyy * dy + hyz * dz + txyy * dx * dy + txyz * dx * dz +
                         tyyz * dy * dz + 0.5 * (txxy * dx * dx + tyyy * dy * dy + tyzz * dz * dz);
                fcd[2] = fcd[2] + hxz * dx + hyz * dy + hzz * dz + txyz * dx * dy + txzz * dx * dz +
                         tyzz * dy * dz + 0.5 * (txxz * dx * dx + tyyz * dy * dy + tzzz * dz * dz);
                if (t->type[ch] != 1) {
                    double *Hc = t->hess + 6 * (size_t)ch;
                    Hc[0] = Hc[0] + hxx + txxx * dx + txxy * dy + txxz * dz;
                    Hc[1] = Hc[1] + hyy + txyy * dx + tyyy * dy + tyyz * dz;
                    Hc[2] = Hc[2] + hzz + txzz * dx + tyzz * dy + tzzz * dz;
                    Hc[3] = Hc[3] + hxy + txxy * dx + txyy * dy + txyz * dz;
                    Hc[4] = Hc[4] + hxz + txxz * dx + txyz * dy + txzz * dz;
                    Hc[5] = Hc[5] + hyz + txyz * dx + tyyz * dy + tyzz * dz;
                }
                double *Tc = t->third + 10 * (size_t)ch;
                for (int i = 0; i < 10; ++i) Tc[i] = Tc[i] + T[i];
            }
            for (int i = 0; i < 3; ++i) t->force[3 * (size_t)ch + i] = fcd[i];
            if (t->type[ch] == 1) {
                int p = t->particle[ch];
                if (p == -1) continue;
                for (int i = 0; i < 3; ++i) F[3 * (size_t)p + i] = t->m[p] * t->force[3 * (size_t)ch + i];
            }
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Morton octree (fortran/morton/morton_octree.f90)                    */
/* ------------------------------------------------------------------ */

/* morton_key (:184-213) */
int64_t orc_morton_key(const double q3[3], int L)
{
    int64_t scale = (int64_t)1 << L;
    int64_t ix = (int64_t)floor(q3[0] * (double)scale);
    int64_t iy = (int64_t)floor(q3[1] * (double)scale);
    int64_t iz = (int64_t)floor(q3[2] * (double)scale);
    if (ix > scale - 1) ix = scale - 1;
    if (iy > scale - 1) iy = scale - 1;
    if (iz > scale - 1) iz = scale - 1;
    int64_t key = 0;
    for (int b = L - 1; b >= 0; --b) {
        int64_t xb = (ix >> b) & 1, yb = (iy >> b) & 1, zb = (iz >> b) & 1;
        key = (key << 3) + ((xb << 2) + (yb << 1) + zb);
    }
    return key;
}

/* stable argsort (mergesort :501-560 keeps ties in input order). */
static void msort(const int64_t *keys, int *perm, int *tmp, int64_t l, int64_t r)
{
    if (l >= r) return;
    int64_t mid = (l + r) / 2;
    msort(keys, perm, tmp, l, mid);
    msort(keys, perm, tmp, mid + 1, r);
    int64_t i = l, j = mid + 1, k = l;
    while (i <= mid && j <= r) {
        if (keys[perm[i]] <= keys[perm[j]]) tmp[k++] = perm[i++];
        else tmp[k++] = perm[j++];
    }
    while (i <= mid) tmp[k++] = perm[i++];
    while (j <= r) tmp[k++] = perm[j++];
    memcpy(perm + l, tmp + l, (size_t)(r - l + 1) * sizeof(int));
}

void orc_argsort(int64_t n, const int64_t *keys, int *perm)
{
    int *tmp = malloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
    msort(keys, perm, tmp, 0, n - 1);
    free(tmp);
}

typedef struct {
    int64_t N;
    int L, nnodes;
    double ssa_par;
    int *perm;
    double *pm, *pq; /* permuted masses, positions (3,N) */
    int64_t *keys;   /* sorted keys */
    /* nodes */
    int64_t *prefix;
    int *level, *first, *last, *parent, *nchild, *child, *next;
    double *size2, *mass, *qcm, *quad;
    int *level_begin, *level_end;
} orc_morton;

void orc_morton_free(orc_morton *t)
{
    if (!t) return;
    void *p[] = {t->perm, t->pm, t->pq, t->keys, t->prefix, t->level, t->first, t->last,
                 t->parent, t->nchild, t->child, t->next, t->size2, t->mass, t->qcm, t->quad,
                 t->level_begin, t->level_end};
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); ++i) free(p[i]);
    free(t);
}

/* Normalised coordinates (generate_nodes :357-384); qn is (3,N). */
static void morton_normalise(int64_t N, const double *ms, const double *qs, double ssa_amp,
                             double *qn, double out_com[3], double *out_mass, double *out_ssa)
{
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t p = 0; p < N; ++p)
        for (int c = 0; c < 3; ++c) {
            double v = qs[3 * p + c];
            if (v < mn[c]) mn[c] = v;
            if (v > mx[c]) mx[c] = v;
        }
    double ext = mx[0] - mn[0];
    if (mx[1] - mn[1] > ext) ext = mx[1] - mn[1];
    if (mx[2] - mn[2] > ext) ext = mx[2] - mn[2];
    double ssa = ssa_amp * ext;
    double qcm[3] = {0, 0, 0}, mass = 0.0;
    for (int64_t p = 0; p < N; ++p) {
        for (int c = 0; c < 3; ++c) qcm[c] = qcm[c] + ms[p] * qs[3 * p + c];
        mass = mass + ms[p];
    }
    for (int c = 0; c < 3; ++c) qcm[c] = qcm[c] / mass;
    for (int64_t p = 0; p < N; ++p)
        for (int c = 0; c < 3; ++c) qn[3 * p + c] = qs[3 * p + c] - qcm[c];
    for (int64_t p = 0; p < 3 * N; ++p) qn[p] = qn[p] / ssa;
    double mn2[3] = {INFINITY, INFINITY, INFINITY}, mx2[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t p = 0; p < N; ++p)
        for (int c = 0; c < 3; ++c) {
            double v = qn[3 * p + c];
            if (v < mn2[c]) mn2[c] = v;
            if (v > mx2[c]) mx2[c] = v;
        }
    double cen[3];
    for (int c = 0; c < 3; ++c) cen[c] = 0.5 * (mn2[c] + mx2[c]);
    for (int64_t p = 0; p < N; ++p)
        for (int c = 0; c < 3; ++c) qn[3 * p + c] = qn[3 * p + c] + 0.5 - cen[c];
    if (out_com) memcpy(out_com, qcm, sizeof(qcm));
    if (out_mass) *out_mass = mass;
    if (out_ssa) *out_ssa = ssa;
}

/* Keys only (normalisation + morton_key), keys in original particle order.
 * With L == 0 the depth is chosen as in choose_depth (:217-244): the
 * smallest L >= 1 whose keys are all distinct. Returns the L used. */
int orc_morton_keys(int64_t N, const double *ms, const double *qs, int L, double ssa_amp,
                    int64_t *keys)
{
    double *qn = malloc(sizeof(double) * 3 * (size_t)(N ? N : 1));
    morton_normalise(N, ms, qs, ssa_amp, qn, NULL, NULL, NULL);
    if (L == 0) {
        int64_t *tmp = malloc(sizeof(int64_t) * (size_t)(N ? N : 1));
        int *perm = malloc(sizeof(int) * (size_t)(N ? N : 1));
        int dup = 1;
        for (L = 1; L <= 21; ++L) {
            for (int64_t p = 0; p < N; ++p) keys[p] = orc_morton_key(qn + 3 * p, L);
            orc_argsort(N, keys, perm);
            dup = 0;
            for (int64_t i = 1; i < N && !dup; ++i) dup = keys[perm[i]] == keys[perm[i - 1]];
            if (!dup) break;
        }
        free(tmp); free(perm);
        if (dup) { free(qn); return -21; } /* 64-bit keys hold at most 21 levels */
    } else {
        for (int64_t p = 0; p < N; ++p) keys[p] = orc_morton_key(qn + 3 * p, L);
    }
    free(qn);
    return L;
}

static int build_next_iter(orc_morton *t)
{
    /* build_next (:304-326): next(idx) = sibling; children chain; last child
     * inherits the parent's sibling.  Iterative pre-order with explicit stack. */
    int *stk = malloc(sizeof(int) * 2 * ((size_t)t->nnodes + 2));
    if (!stk) return ORC_ERR_ALLOC;
    int top = 0;
    stk[top++] = 0; stk[top++] = -1; /* node 0 (root), sibling "0" -> -1 (end) */
    while (top > 0) {
        int sib = stk[--top];
        int idx = stk[--top];
        t->next[idx] = sib;
        int nc = t->nchild[idx];
        for (int i = nc - 1; i >= 0; --i) {
            int c = t->child[8 * idx + i];
            int s = (i < nc - 1) ? t->child[8 * idx + i + 1] : sib;
            stk[top++] = c; stk[top++] = s;
        }
    }
    free(stk);
    return ORC_OK;
}

/* generate_nodes (:328-499) + evaluate_quadrupoles (:562-625).
 * qs is (3,N).  L == 0 chooses the depth.  Returns NULL on failure. */
orc_morton *orc_morton_build(int64_t N, const double *ms, const double *qs, int L, double ssa_amp,
                             int use_quad)
{
    orc_morton *t = calloc(1, sizeof(orc_morton));
    size_t n = (size_t)(N ? N : 1);
    t->N = N;
    t->keys = malloc(8 * n);
    t->perm = malloc(4 * n);
    t->pm = malloc(8 * n);
    t->pq = malloc(24 * n);
    double *qn = malloc(24 * n);
    double qcm[3], mass, ssa;
    morton_normalise(N, ms, qs, ssa_amp, qn, qcm, &mass, &ssa);
    int64_t *keys = malloc(8 * n);
    if (L == 0) {
        int dup = 1;
        for (L = 1; L <= 21; ++L) {
            for (int64_t p = 0; p < N; ++p) keys[p] = orc_morton_key(qn + 3 * p, L);
            orc_argsort(N, keys, t->perm);
            dup = 0;
            for (int64_t i = 1; i < N && !dup; ++i) dup = keys[t->perm[i]] == keys[t->perm[i - 1]];
            if (!dup) break;
        }
        if (dup) L = 21; /* 64-bit keys hold at most 21 levels; duplicates share a leaf */
    } else {
        for (int64_t p = 0; p < N; ++p) keys[p] = orc_morton_key(qn + 3 * p, L);
        orc_argsort(N, keys, t->perm);
    }
    free(qn);
    t->L = L;
    for (int64_t p = 0; p < N; ++p) {
        int o = t->perm[p];
        t->keys[p] = keys[o];
        t->pm[p] = ms[o];
        for (int c = 0; c < 3; ++c) t->pq[3 * p + c] = qs[3 * (size_t)o + c];
    }
    free(keys);
    size_t cap = (size_t)(L + 1) * n;
    t->prefix = malloc(8 * cap); t->level = malloc(4 * cap); t->first = malloc(4 * cap);
    t->last = malloc(4 * cap); t->parent = malloc(4 * cap); t->nchild = malloc(4 * cap);
    t->child = malloc(32 * cap); t->next = malloc(4 * cap); t->size2 = malloc(8 * cap);
    t->mass = malloc(8 * cap); t->qcm = malloc(24 * cap); t->quad = calloc(6 * cap, 8);
    t->level_begin = malloc(4 * (size_t)(L + 1)); t->level_end = malloc(4 * (size_t)(L + 1));
#define CREATE(idx, lev, pref, f, l, ms_, q, sz)                                  \
    do {                                                                          \
        t->level[idx] = (lev); t->prefix[idx] = (pref); t->first[idx] = (f);       \
        t->last[idx] = (l); t->mass[idx] = (ms_);                                  \
        t->qcm[3 * (idx)] = (q)[0]; t->qcm[3 * (idx) + 1] = (q)[1];                \
        t->qcm[3 * (idx) + 2] = (q)[2];                                            \
        t->size2[idx] = (sz) * (sz); t->parent[idx] = -1; t->nchild[idx] = 0;      \
        for (int cc_ = 0; cc_ < 8; ++cc_) t->child[8 * (idx) + cc_] = -1;          \
    } while (0)
    int nn = 1;
    CREATE(0, 0, N ? (t->keys[0] >> (3 * L)) : 0, 0, (int)N - 1, mass, qcm, ssa);
    t->level_begin[0] = 0;
    t->level_end[0] = 0;
    for (int level = 1; level <= L; ++level) {
        t->level_begin[level] = nn;
        double sz = ssa * pow(2.0, -level);
        for (int par = t->level_begin[level - 1]; par <= t->level_end[level - 1]; ++par) {
            int f = t->first[par], l = t->last[par];
            int64_t current = -1;
            int cf = 0;
            double m = 0, q[3] = {0, 0, 0};
            for (int i = f; i <= l; ++i) {
                int64_t pref = t->keys[i] >> (3 * (L - level));
                if (current == -1) {
                    current = pref; cf = i; m = t->pm[i];
                    for (int c = 0; c < 3; ++c) q[c] = m * t->pq[3 * (size_t)i + c];
                } else if (pref != current) {
                    for (int c = 0; c < 3; ++c) q[c] = q[c] / m;
                    CREATE(nn, level, current, cf, i - 1, m, q, sz);
                    t->parent[nn] = par;
                    t->child[8 * par + t->nchild[par]] = nn;
                    t->nchild[par]++;
                    nn++;
                    current = pref; cf = i; m = t->pm[i];
                    for (int c = 0; c < 3; ++c) q[c] = m * t->pq[3 * (size_t)i + c];
                } else {
                    m = m + t->pm[i];
                    for (int c = 0; c < 3; ++c) q[c] = q[c] + t->pm[i] * t->pq[3 * (size_t)i + c];
                }
            }
            for (int c = 0; c < 3; ++c) q[c] = q[c] / m;
            CREATE(nn, level, current, cf, l, m, q, sz);
            t->parent[nn] = par;
            t->child[8 * par + t->nchild[par]] = nn;
            t->nchild[par]++;
            nn++;
        }
        t->level_end[level] = nn - 1;
    }
#undef CREATE
    t->nnodes = nn;
    build_next_iter(t);
    if (use_quad) {
        for (int level = L; level >= 0; --level) {
            for (int k = t->level_begin[level]; k <= t->level_end[level]; ++k) {
                double *Q = t->quad + 6 * (size_t)k;
                for (int i = 0; i < 6; ++i) Q[i] = 0.0;
                const double *qc = t->qcm + 3 * (size_t)k;
                if (t->nchild[k] <= 1) {
                    for (int p = t->first[k]; p <= t->last[k]; ++p) {
                        double pm = t->pm[p];
                        double dxi = t->pq[3 * (size_t)p] - qc[0];
                        double dyi = t->pq[3 * (size_t)p + 1] - qc[1];
                        double dzi = t->pq[3 * (size_t)p + 2] - qc[2];
                        Q[0] = Q[0] + pm * (dxi * dxi);
                        Q[1] = Q[1] + pm * (dyi * dyi);
                        Q[2] = Q[2] + pm * (dzi * dzi);
                        Q[3] = Q[3] + pm * dxi * dyi;
                        Q[4] = Q[4] + pm * dxi * dzi;
                        Q[5] = Q[5] + pm * dyi * dzi;
                    }
                } else {
                    for (int cp = 0; cp < t->nchild[k]; ++cp) {
                        int c = t->child[8 * k + cp];
                        double pm = t->mass[c];
                        double dxi = t->qcm[3 * (size_t)c] - qc[0];
                        double dyi = t->qcm[3 * (size_t)c + 1] - qc[1];
                        double dzi = t->qcm[3 * (size_t)c + 2] - qc[2];
                        const double *Qc = t->quad + 6 * (size_t)c;
                        Q[0] = Q[0] + Qc[0] + pm * dxi * dxi;
                        Q[1] = Q[1] + Qc[1] + pm * dyi * dyi;
                        Q[2] = Q[2] + Qc[2] + pm * dzi * dzi;
                        Q[3] = Q[3] + Qc[3] + pm * dxi * dyi;
                        Q[4] = Q[4] + Qc[4] + pm * dxi * dzi;
                        Q[5] = Q[5] + Qc[5] + pm * dyi * dzi;
                    }
                }
            }
        }
    }
    return t;
}

int orc_morton_depth(const orc_morton *t) { return t->L; }
int orc_morton_nnodes(const orc_morton *t) { return t->nnodes; }

void orc_morton_export(const orc_morton *t, int *perm, int64_t *keys, int64_t *prefix, int *level,
                       int *first, int *last, int *parent, int *nchild, int *child, int *next,
                       double *size2, double *mass, double *qcm, double *quad)
{
    size_t n = (size_t)t->N, nn = (size_t)t->nnodes;
    if (perm) memcpy(perm, t->perm, 4 * n);
    if (keys) memcpy(keys, t->keys, 8 * n);
    if (prefix) memcpy(prefix, t->prefix, 8 * nn);
    if (level) memcpy(level, t->level, 4 * nn);
    if (first) memcpy(first, t->first, 4 * nn);
    if (last) memcpy(last, t->last, 4 * nn);
    if (parent) memcpy(parent, t->parent, 4 * nn);
    if (nchild) memcpy(nchild, t->nchild, 4 * nn);
    if (child) memcpy(child, t->child, 32 * nn);
    if (next) memcpy(next, t->next, 4 * nn);
    if (size2) memcpy(size2, t->size2, 8 * nn);
    if (mass) memcpy(mass, t->mass, 8 * nn);
    if (qcm) memcpy(qcm, t->qcm, 24 * nn);
    if (quad) memcpy(quad, t->quad, 48 * nn);
}

/* barnes_hut (:627-766) over sorted index p; F written to original order. */
static void morton_bh_p(const orc_morton *t, int p, double theta2, double G, double eps2,
                        int mult, double out[3])
{
    double mp = t->pm[p];
    const double *qp = t->pq + 3 * (size_t)p;
    double f0 = 0, f1 = 0, f2 = 0;
    int idx = 0;
    while (idx != -1) {
        int only_one = t->first[idx] == t->last[idx];
        /* (:680-683) compares last with the 1-based sentinel 0; with 0-based
         * indices that sentinel is -1, which no node has: never taken. */
        if (t->nchild[idx] == 0) {
            for (int b = t->first[idx]; b <= t->last[idx]; ++b) {
                if (p == b) continue;
                const double *qb = t->pq + 3 * (size_t)b;
                double dx = qb[0] - qp[0], dy = qb[1] - qp[1], dz = qb[2] - qp[2];
                double dist2 = dx * dx + dy * dy + dz * dz + eps2;
                double inv = 1.0 / sqrt(dist2);
                inv = inv * inv * inv;
                double f = G * mp * t->pm[b] * inv;
                f0 = f0 + f * dx; f1 = f1 + f * dy; f2 = f2 + f * dz;
            }
            idx = t->next[idx];
            continue;
        }
        if (t->first[idx] <= p && p <= t->last[idx]) {
            idx = t->child[8 * idx];
            continue;
        }
        const double *qc = t->qcm + 3 * (size_t)idx;
        double dx = qc[0] - qp[0], dy = qc[1] - qp[1], dz = qc[2] - qp[2];
        double dist2 = dx * dx + dy * dy + dz * dz;
        if (only_one || t->size2[idx] <= theta2 * dist2) {
            if (t->mass[idx] > 0) {
                dist2 = dist2 + eps2;
                if (!mult || only_one) {
                    double inv = 1.0 / sqrt(dist2);
                    inv = inv * inv * inv;
                    double f = G * mp * t->mass[idx] * inv;
                    f0 = f0 + f * dx; f1 = f1 + f * dy; f2 = f2 + f * dz;
                } else {
                    double invR = 1.0 / sqrt(dist2);
                    double invR2 = 1.0 / dist2;
                    double invR3 = invR * invR2;
                    double invR5 = invR3 * invR * invR;
                    double mx = dx * invR3 * t->mass[idx];
                    double my = dy * invR3 * t->mass[idx];
                    double mz = dz * invR3 * t->mass[idx];
                    const double *Q = t->quad + 6 * (size_t)idx;
                    double dx2 = 15.0 * dx * dx * invR2 - 3.0;
                    double dy2 = 15.0 * dy * dy * invR2 - 3.0;
                    double dz2 = 15.0 * dz * dz * invR2 - 3.0;
                    double dxyz = 15.0 * dx * dy * dz * invR2;
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
            idx = t->next[idx];
            continue;
        }
        idx = t->child[8 * idx];
    }
    out[0] = f0; out[1] = f1; out[2] = f2;
}

/* forces_par (:150-180); F is (3,N) in original particle order. */
void orc_morton_forces(const orc_morton *t, double G, double eps2, double theta2, int mult,
                       double *F)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < t->N; ++p) {
        double o[3];
        morton_bh_p(t, (int)p, theta2, G, eps2, mult, o);
        int q = t->perm[p];
        F[3 * (size_t)q] = o[0]; F[3 * (size_t)q + 1] = o[1]; F[3 * (size_t)q + 2] = o[2];
    }
}

/* ------------------------------------------------------------------ */
/* Error metrics: benchmarks/test_forces.f90:66-127 (McMillan-Aarseth  */
/* normalised RMS + sum of forces + torque), plus per-particle relative */
/* error statistics on accelerations.                                   */
/* ------------------------------------------------------------------ */
void orc_errors_ma(int64_t N, const double *ms, const double *qs, const double *F,
                   const double *Fd, double errors[7])
{
    double a_scale = 0.0, s2 = 0.0;
    for (int i = 0; i < 7; ++i) errors[i] = 0.0;
    for (int64_t p = 0; p < N; ++p)
        for (int c = 0; c < 3; ++c) errors[1 + c] = errors[1 + c] + F[3 * p + c];
    for (int64_t p = 0; p < N; ++p) {
        const double *q = qs + 3 * p, *f = F + 3 * p;
        for (int c = 0; c < 3; ++c) {
            double d = F[3 * p + c] / ms[p] - Fd[3 * p + c] / ms[p];
            s2 += d * d;
        }
        a_scale = a_scale + ms[p] / (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        errors[4] = errors[4] + (q[1] * f[2] - f[1] * q[2]);
        errors[5] = errors[5] + (-q[0] * f[2] + f[0] * q[2]);
        errors[6] = errors[6] + (q[0] * f[1] - f[0] * q[1]);
    }
    double r = sqrt(s2) / a_scale;
    errors[0] = sqrt(r * r / (double)N);
}

/* ------------------------------------------------------------------ */
/* Time integration (momenta form).                                    */
/* force: 0 = tree of `method` (10-12, bh_forces_over_p), 2 = direct   */
/* sum in the reference's pair order (parareal/api3d.f90:46-75, which   */
/* is fortran/direct_forces.f90's serial loop).                         */
/* scheme 0 = Verlet (integration.f90:153-171, parareal.f90:57-66),     */
/* scheme 1 = symplectic Euler (parareal.f90:30-36).                    */
/* The reference's benchmark drift divides by ms(1) (equal masses);     */
/* Parareal's by m(a): per-particle masses here, identical for equal.   */
/* ------------------------------------------------------------------ */
static int orc_forces_of(int64_t n, const double *m, const double *x, const double *y,
                         const double *z, int force, int method, double theta2, double G,
                         double eps2, double *F)
{
    if (force == 2) {
        orc_direct_forces(n, m, x, y, z, G, eps2, F);
        return ORC_OK;
    }
    orc_tree *t = orc_tree_new(n, m, method, 64);
    if (!t) return ORC_ERR_ALLOC;
    int st = orc_tree_init(t, x, y, z);
    if (st == ORC_OK) orc_tree_bh_forces(t, n, NULL, theta2, G, eps2, F, NULL);
    orc_tree_free(t);
    return st;
}

int orc_integrate(int64_t n, const double *m, double *x, double *y, double *z, double *px,
                  double *py, double *pz, int scheme, double dt, int64_t nsteps, int force,
                  int method, double theta2, double G, double eps2)
{
    double *F = (double *)malloc(sizeof(double) * 3 * (size_t)(n ? n : 1));
    double *q[3] = {x, y, z}, *p[3] = {px, py, pz};
    int st = ORC_OK;
    if (!F) return ORC_ERR_ALLOC;
    if (nsteps <= 0) goto done;
    if (scheme == 0) {
        double h = 0.5 * dt;
        if ((st = orc_forces_of(n, m, x, y, z, force, method, theta2, G, eps2, F))) goto done;
        for (int64_t s = 0; s < nsteps; ++s) {
            for (int64_t a = 0; a < n; ++a)
                for (int c = 0; c < 3; ++c) {
                    p[c][a] = p[c][a] + h * F[3 * a + c];
                    q[c][a] = q[c][a] + dt * p[c][a] / m[a];
                }
            if ((st = orc_forces_of(n, m, x, y, z, force, method, theta2, G, eps2, F))) goto done;
            for (int64_t a = 0; a < n; ++a)
                for (int c = 0; c < 3; ++c) p[c][a] = p[c][a] + h * F[3 * a + c];
        }
    } else {
        for (int64_t s = 0; s < nsteps; ++s) {
            if ((st = orc_forces_of(n, m, x, y, z, force, method, theta2, G, eps2, F))) goto done;
            for (int64_t a = 0; a < n; ++a)
                for (int c = 0; c < 3; ++c) {
                    p[c][a] = p[c][a] + dt * F[3 * a + c];
                    q[c][a] = q[c][a] + dt * p[c][a] / m[a];
                }
        }
    }
done:
    free(F);
    return st;
}

/* out[0] = E = T + V with the softened pair potential, out[1..3] = sum p,
 * out[4..6] = sum q x p (integration.f90:236-262; parareal/api3d.f90:108-155) */
void orc_energy(int64_t n, const double *m, const double *x, const double *y, const double *z,
                const double *px, const double *py, const double *pz, double G, double eps2,
                double *out)
{
    double T = 0, V = 0, P[3] = {0, 0, 0}, J[3] = {0, 0, 0};
    for (int64_t a = 0; a < n; ++a) {
        T = T + (px[a] * px[a] + py[a] * py[a] + pz[a] * pz[a]) / (2.0 * m[a]);
        for (int64_t b = 0; b < a; ++b) {
            double dx = x[b] - x[a], dy = y[b] - y[a], dz = z[b] - z[a];
            double dist = sqrt(dx * dx + dy * dy + dz * dz + eps2);
            V = V - G * m[a] * m[b] / dist;
        }
        P[0] += px[a]; P[1] += py[a]; P[2] += pz[a];
        J[0] += y[a] * pz[a] - z[a] * py[a];
        J[1] += z[a] * px[a] - x[a] * pz[a];
        J[2] += x[a] * py[a] - y[a] * px[a];
    }
    out[0] = T + V;
    for (int c = 0; c < 3; ++c) { out[1 + c] = P[c]; out[4 + c] = J[c]; }
}

/* integrate_parareal (parareal.f90:97-186) with the reference's correction
 * order (G_new + F) - G_old.  State of window w (0-based) is U[w*6n ...]:
 * x, y, z, px, py, pz blocks of n.  U_out gets the last iterate; errors gets
 * the error of iterates 2.. (itermax - 1 values).  cf/ff: force modes of the
 * coarse and fine propagators (0 tree of cmethod/fmethod, 2 direct). */
int orc_parareal(int64_t n, const double *m, const double *u0, double tf, int nw, int itermax,
                 double maxerror, double G, double eps2, int cf, int cmethod, double ctheta2,
                 int csteps, int ff, int fmethod, double ftheta2, int fsteps, double *U_out,
                 double *errors, int *iterations)
{
    const size_t W = 6 * (size_t)n;
    double *Up = (double *)calloc(W * nw, sizeof(double));
    double *Un = (double *)calloc(W * nw, sizeof(double));
    double *Ft = (double *)calloc(W * nw, sizeof(double));
    double *Go = (double *)calloc(W * nw, sizeof(double));
    double *g = (double *)calloc(W, sizeof(double));
    int st = ORC_OK, iters = 1;
    double window = tf / nw, dtc = window / csteps, dtf = window / fsteps;
    if (!Up || !Un || !Ft || !Go || !g) { st = ORC_ERR_ALLOC; goto out; }
#define ORC_PROP(dst, src, scheme, dt, steps, force, meth, th2)                                 \
    do {                                                                                       \
        memcpy((dst), (src), W * sizeof(double));                                              \
        st = orc_integrate(n, m, (dst), (dst) + n, (dst) + 2 * n, (dst) + 3 * n, (dst) + 4 * n, \
                           (dst) + 5 * n, (scheme), (dt), (steps), (force), (meth), (th2), G,   \
                           eps2);                                                              \
    } while (0)
    memcpy(Un, u0, W * sizeof(double));
    for (int w = 0; w + 1 < nw; ++w) {
        ORC_PROP(Un + (w + 1) * W, Un + w * W, 1, dtc, csteps, cf, cmethod, ctheta2);
        if (st) goto out;
        memcpy(Go + w * W, Un + (w + 1) * W, W * sizeof(double));
    }
    for (int k = 1; k < itermax; ++k) {
        double *tmp = Up; Up = Un; Un = tmp;
        ++iters;
        int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
        for (int w = 0; w < nw - 1; ++w) {
            double *d = Ft + w * W;
            memcpy(d, Up + w * W, W * sizeof(double));
            bad |= orc_integrate(n, m, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, 0,
                                 dtf, fsteps, ff, fmethod, ftheta2, G, eps2);
        }
        if (bad) { st = bad; goto out; }
        memcpy(Un, u0, W * sizeof(double));
        for (int w = 0; w + 1 < nw; ++w) {
            ORC_PROP(g, Un + w * W, 1, dtc, csteps, cf, cmethod, ctheta2);
            if (st) goto out;
            double *u = Un + (w + 1) * W, *f = Ft + w * W, *go = Go + w * W;
            for (size_t i = 0; i < W; ++i) {
                u[i] = g[i] + f[i];
                u[i] = u[i] - go[i];
            }
            memcpy(go, g, W * sizeof(double));
        }
        double e2 = 0.0;
        for (size_t i = 0; i < W; ++i) {
            double d = Un[(nw - 1) * W + i] - Up[(nw - 1) * W + i];
            e2 += d * d;
        }
        double err = sqrt(e2) / (6.0 * (double)n);
        if (errors) errors[k - 1] = err;
        if (err <= maxerror) break;
    }
#undef ORC_PROP
    memcpy(U_out, Un, W * nw * sizeof(double));
    if (iterations) *iterations = iters;
out:
    free(Up); free(Un); free(Ft); free(Go); free(g);
    return st;
}
