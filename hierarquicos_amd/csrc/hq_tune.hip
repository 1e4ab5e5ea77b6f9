// hq_tune.hip -- the autotuner of the bunched walk and its table of settings.
//
// No counterpart in the reference (its only knob is the group size of
// forces_groups, octree_default.f90:1722-1809).  The axes are the walk-group
// (bunch) size, the super-group size the bunches are cut from, the leaf
// bucket (cells of at most this many bodies enter the lists as their bodies,
// the leaf size of a bucketed tree: 0..64) and the evaluation tiling (the
// fused walk evaluates each bunch's 64-entry batches on its own wavefront;
// the split walk gives every batch a wavefront).  Each setting is timed with
// HIP events on the handle's stream, best of `reps` after a warm-up, and its
// RMS relative force error on a sample of bodies is measured against the
// FP64 direct sum: a setting may not raise it more than 2% above the starting
// setting's.  The search is coordinate descent from the handle's setting
// (two sweeps over the axes).
//
// The best setting is stored in a process-wide table keyed by the workload
// class: log2 N, number of systems, multipole order, theta, precision and the
// tree depth (the shape of the distribution).  A handle whose caller has not
// fixed a knob with hq_tree_tune takes it from the table at its first bunched
// walk of that class; the table can be saved to and loaded from a text file.
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "hq_tree.h"

namespace hq {

namespace {

struct TuneKey {
    int log2n, nsys_bits, multipole, theta100, precision, depth;
    bool operator<(const TuneKey& o) const {
        return std::tie(log2n, nsys_bits, multipole, theta100, precision, depth) <
               std::tie(o.log2n, o.nsys_bits, o.multipole, o.theta100, o.precision, o.depth);
    }
};

struct TuneEntry {
    TuneSetting s;
    double ms;
};

std::mutex g_mu;
std::map<TuneKey, TuneEntry> g_table;

int ilog2_round(int64_t v) { return v <= 1 ? 0 : (int)lround(log2((double)v)); }

TuneKey key_of(const hq_tree* t, double theta2, int precision) {
    return TuneKey{ilog2_round(t->n), ilog2_round(t->nsys), t->multipole,
                   (int)lround(100.0 * sqrt(theta2 > 0.0 ? theta2 : 0.0)), precision, t->most_depth};
}

constexpr unsigned TABLE_KEYS = (1u << HQ_TUNE_BUNCH) | (1u << HQ_TUNE_SUPER) |
                                (1u << HQ_TUNE_LEAF) | (1u << HQ_TUNE_SPLIT);

void apply(hq_tree* t, const TuneSetting& s, bool respect_user) {
    const unsigned u = respect_user ? t->tune_user : 0u;
    if (!(u & (1u << HQ_TUNE_BUNCH))) t->group_size = s.bunch;
    if (!(u & (1u << HQ_TUNE_SUPER))) t->super_size = s.super_size;
    if (!(u & (1u << HQ_TUNE_LEAF))) t->leaf = s.leaf;
    if (!(u & (1u << HQ_TUNE_SPLIT))) t->split = s.split;
}

TuneSetting current(const hq_tree* t) {
    return TuneSetting{t->group_size, t->super_size, t->leaf, t->split};
}

// sum over the sample of |F - Fd|^2 / |Fd|^2 (one workgroup, fixed order)
__global__ __launch_bounds__(256) void k_sample_err(const int64_t* __restrict__ slot, int64_t nt,
                                                    const double* __restrict__ fx,
                                                    const double* __restrict__ fy,
                                                    const double* __restrict__ fz,
                                                    const double* __restrict__ Fd,
                                                    double* __restrict__ out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < nt; i += 256) {
        const int64_t p = slot[i];
        const double d0 = fx[p] - Fd[3 * i], d1 = fy[p] - Fd[3 * i + 1], d2 = fz[p] - Fd[3 * i + 2];
        const double r2 = Fd[3 * i] * Fd[3 * i] + Fd[3 * i + 1] * Fd[3 * i + 1] +
                          Fd[3 * i + 2] * Fd[3 * i + 2];
        acc += r2 > 0.0 ? (d0 * d0 + d1 * d1 + d2 * d2) / r2 : 0.0;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

bool tune_from_table(hq_tree* t, double theta2, int precision) {
    if (t->graph_mode || capturing() || (t->tune_user & TABLE_KEYS) == TABLE_KEYS) return false;
    const TuneKey k = key_of(t, theta2, precision);
    const int tag = k.theta100 * 2 + precision;
    if (t->table_tag == tag) return false;  // once per angle and precision
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(k);
    if (it == g_table.end()) return false;
    apply(t, it->second.s, true);
    t->table_tag = tag;
    return true;
}

static int tune_search(hq_tree* t, double theta2, double G, double eps2, int precision, int reps,
                       const TuneSetting& start, TuneSetting* best_out, double* best_ms);

int autotune(hq_tree* t, double theta2, double G, double eps2, int precision, int reps,
             TuneSetting* best_out, double* best_ms) {
    const TuneSetting start = current(t);
    const unsigned user0 = t->tune_user;
    t->tune_user |= TABLE_KEYS;  // the walks below run the candidates, not the table's setting
    int rc = tune_search(t, theta2, G, eps2, precision, reps, start, best_out, best_ms);
    if (rc) {
        apply(t, start, false);
        t->tune_user = user0;
    }
    return rc;
}

static int tune_search(hq_tree* t, double theta2, double G, double eps2, int precision, int reps,
                       const TuneSetting& start, TuneSetting* best_out, double* best_ms) {
    hipStream_t s = t->stream;
    const int64_t n = t->n;
    int rc;
    // the sample: evenly spaced tree-order slots, the FP64 direct sum over the
    // sorted bodies at each
    const int64_t nt = std::min<int64_t>(n, 512);
    std::vector<int64_t> hslot(nt);
    for (int64_t i = 0; i < nt; ++i) hslot[i] = (i * n) / nt;
    if ((rc = t->tune_buf.ensure((size_t)(nt + 3 * nt + 1) * sizeof(double)))) return rc;
    int64_t* slot = t->tune_buf.as<int64_t>();
    double* Fd = t->tune_buf.as<double>() + nt;
    double* err = Fd + 3 * nt;
    HQ_CHECK_HIP(hipMemcpyAsync(slot, hslot.data(), nt * sizeof(int64_t), hipMemcpyHostToDevice, s));
    if ((rc = direct_forces(n, t->sm.as<double>(), t->sx.as<double>(), t->sy.as<double>(),
                            t->sz.as<double>(), nt, slot, G, eps2, HQ_FP64, Fd, s, &t->tune_tmp)))
        return rc;
    // time (best of reps after a warm-up) and sample error of one setting
    auto eval = [&](const TuneSetting& st, double* ms, double* e2) -> int {
        apply(t, st, false);
        int r;
        double tmin = 1e300;
        for (int k = 0; k <= reps; ++k) {
            HQ_CHECK_HIP(hipEventRecord(t->ev[9], s));
            if ((r = tree_forces_groups(t, theta2, G, eps2, precision, true))) return r;
            HQ_CHECK_HIP(hipEventRecord(t->ev[10], s));
            HQ_CHECK_HIP(hipEventSynchronize(t->ev[10]));
            float el = 0.0f;
            HQ_CHECK_HIP(hipEventElapsedTime(&el, t->ev[9], t->ev[10]));
            if (k > 0) tmin = std::min(tmin, (double)el);
        }
        k_sample_err<<<1, 256, 0, s>>>(slot, nt, t->fx.as<double>(), t->fy.as<double>(),
                                       t->fz.as<double>(), Fd, err);
        double h = 0.0;
        HQ_CHECK_HIP(hipMemcpyAsync(&h, err, sizeof(double), hipMemcpyDeviceToHost, s));
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        *ms = tmin;
        *e2 = h;
        return HQ_OK;
    };
    TuneSetting best = start;
    double bms = 0.0, e0 = 0.0;
    if ((rc = eval(start, &bms, &e0))) return rc;
    // mean squares: RMS within 2%.  A batch's sample mixes systems, so its
    // direct sum is not the walk's reference: no guard there
    const double e_cap = t->nsys > 1 ? INFINITY : 1.02 * 1.02 * e0 + 1e-300;
    static const int kB[] = {32, 48, 64}, kS[] = {128, 256, 512},
                     kL[] = {0, 2, 4, 8, 16, 32, 64}, kP[] = {0, 2};
    for (int sweep = 0; sweep < 2; ++sweep) {
        const TuneSetting before = best;
        for (int axis = 0; axis < 4; ++axis) {
            const int* vals = axis == 0 ? kL : axis == 1 ? kS : axis == 2 ? kB : kP;
            const int nv = axis == 0 ? 7 : axis == 3 ? 2 : 3;
            for (int v = 0; v < nv; ++v) {
                TuneSetting c = best;
                int& f = axis == 0 ? c.leaf : axis == 1 ? c.super_size : axis == 2 ? c.bunch : c.split;
                if (f == vals[v]) continue;
                f = vals[v];
                if (c.super_size < c.bunch) continue;
                double ms = 0.0, e2 = 0.0;
                if ((rc = eval(c, &ms, &e2))) return rc;
                if (e2 <= e_cap && ms < bms) best = c, bms = ms;
            }
        }
        if (best.bunch == before.bunch && best.super_size == before.super_size &&
            best.leaf == before.leaf && best.split == before.split)
            break;
    }
    // the chosen setting, once more, so that the handle's derived state (stack
    // level, split log size) is the one it leaves
    apply(t, best, false);
    if ((rc = tree_forces_groups(t, theta2, G, eps2, precision, true))) return rc;
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    // the handle keeps its tuned setting (TABLE_KEYS stay marked as set)
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_table[key_of(t, theta2, precision)] = TuneEntry{best, bms};
    }
    *best_out = best;
    *best_ms = bms;
    return HQ_OK;
}

int tune_table_save(const char* path) {
    FILE* f = fopen(path, "w");
    if (!f) return fail(HQ_ERR_ARGS, "cannot write %s", path);
    fprintf(f, "# hierarquicos_amd tuning table: log2n nsys_bits multipole theta100 precision depth "
               "bunch super leaf split ms\n");
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& kv : g_table) {
        const TuneKey& k = kv.first;
        const TuneEntry& e = kv.second;
        fprintf(f, "%d %d %d %d %d %d %d %d %d %d %.6f\n", k.log2n, k.nsys_bits, k.multipole,
                k.theta100, k.precision, k.depth, e.s.bunch, e.s.super_size, e.s.leaf, e.s.split,
                e.ms);
    }
    const bool ok = fclose(f) == 0;
    return ok ? HQ_OK : fail(HQ_ERR_ARGS, "cannot write %s", path);
}

int tune_table_load(const char* path) {
    FILE* f = fopen(path, "r");
    if (!f) return fail(HQ_ERR_ARGS, "cannot read %s", path);
    char line[512];
    int row = 0;
    std::vector<std::pair<TuneKey, TuneEntry>> rows;
    while (fgets(line, sizeof(line), f)) {
        ++row;
        if (line[0] == '#' || line[0] == '\n') continue;
        TuneKey k;
        TuneEntry e;
        if (sscanf(line, "%d %d %d %d %d %d %d %d %d %d %lf", &k.log2n, &k.nsys_bits, &k.multipole,
                   &k.theta100, &k.precision, &k.depth, &e.s.bunch, &e.s.super_size, &e.s.leaf,
                   &e.s.split, &e.ms) != 11 ||
            e.s.bunch < 1 || e.s.bunch > 64 || e.s.super_size < 1 || e.s.leaf < 0 ||
            e.s.leaf > 64 || e.s.split < 0 || e.s.split > 2) {
            fclose(f);
            return fail(HQ_ERR_ARGS, "%s:%d: not a tuning table row", path, row);
        }
        rows.emplace_back(k, e);
    }
    fclose(f);
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& r : rows) g_table[r.first] = r.second;
    return HQ_OK;
}

void tune_table_clear() {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
}

}  // namespace hq
