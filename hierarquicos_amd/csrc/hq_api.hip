// hq_api.hip -- the C ABI (include/hq.h) over the tree, walk, direct and
// Morton kernels.  Host-side orchestration only; every force path runs on
// the GPU.  Errors are returned as status codes with a per-thread message.
#include <math.h>
#include <stdarg.h>
#include <string.h>

#include <algorithm>
#include <memory>

#include "hq_tree.h"

namespace hq {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

thread_local bool g_capturing = false;
bool capturing() { return g_capturing; }

__global__ void k_fill_i64(int64_t* __restrict__ p, int64_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_i32(int* __restrict__ p, int v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

int morton_keys(int64_t n, const double* m, const double* q, int L, double ssa, int64_t* keys,
                int* perm, int* L_out, hipStream_t s, double* par_out = nullptr);

static int multipole_of(int method) {
    switch (method) {
        case 10: case 20: case 30: return 1;
        case 11: case 21: case 31: return 4;
        case 12: case 22: case 32: return 8;
        default: return 0;
    }
}

// copy n doubles from host or device into a device buffer
static int stage(DevBuf& dst, const double* src, int64_t n, int where, hipStream_t s) {
    int rc;
    if ((rc = dst.ensure(n * sizeof(double)))) return rc;
    HQ_CHECK_HIP(hipMemcpyAsync(dst.p, src, n * sizeof(double),
                                where == HQ_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                                s));
    return HQ_OK;
}

}  // namespace hq

using namespace hq;

extern "C" {

const char* hq_last_error(void) { return g_err; }

int hq_dev_alloc(void** p, int64_t bytes) {
    if (!p || bytes < 0) return fail(HQ_ERR_ARGS, "bad argument");
    *p = nullptr;
    hipError_t e = hipMalloc(p, bytes > 0 ? (size_t)bytes : 1);
    if (e != hipSuccess) return fail(HQ_ERR_NOMEM, "hipMalloc(%lld): %s", (long long)bytes, hipGetErrorString(e));
    return HQ_OK;
}

int hq_dev_free(void* p) {
    if (p) HQ_CHECK_HIP(hipFree(p));
    return HQ_OK;
}

int hq_dev_copy(void* dst, const void* src, int64_t bytes, int kind) {
    if (bytes <= 0) return HQ_OK;
    if (!dst || !src || kind < 0 || kind > 2) return fail(HQ_ERR_ARGS, "bad argument");
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                      : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    HQ_CHECK_HIP(hipMemcpy(dst, src, (size_t)bytes, k));
    return HQ_OK;
}

int hq_dev_sync(void) {
    HQ_CHECK_HIP(hipDeviceSynchronize());
    return HQ_OK;
}
const char* hq_version(void) { return "hierarquicos_amd 0.1 (gfx950)"; }

int hq_tree_create(hq_tree** out, int64_t n, int method, int group_size, int precision) {
    if (!out) return fail(HQ_ERR_ARGS, "null handle pointer");
    *out = nullptr;
    if (n < 1 || n > (int64_t)1 << 30) return fail(HQ_ERR_ARGS, "particle count %lld out of range", (long long)n);
    int mp = multipole_of(method);
    if (!mp) return fail(HQ_ERR_ARGS, "method not identified! (%d)", method);
    if (precision != HQ_FP64 && precision != HQ_MIXED) return fail(HQ_ERR_ARGS, "bad precision %d", precision);
    std::unique_ptr<hq_tree> t(new hq_tree());
    t->n = n;
    t->method = method;
    t->multipole = mp;
    t->group_size = group_size > 0 ? group_size : 64;
    t->precision = precision;
    HQ_CHECK_HIP(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    for (auto& e : t->ev) HQ_CHECK_HIP(hipEventCreate(&e));
    int rc;
    if ((rc = t->status.ensure(ST_SIZE * sizeof(int64_t)))) return rc;
    // the status block, then host slots of other read-backs (ST_HOST_BUNCHES)
    HQ_CHECK_HIP(hipHostMalloc((void**)&t->hstatus, ST_HOST_SIZE * sizeof(int64_t), hipHostMallocDefault));
    // one system [0, n): a batch handle overwrites these
    t->hsysoff = {0, n};
    if ((rc = t->sysoff.ensure(2 * sizeof(int64_t)))) return rc;
    HQ_CHECK_HIP(hipMemcpy(t->sysoff.p, t->hsysoff.data(), 2 * sizeof(int64_t), hipMemcpyHostToDevice));
    *out = t.release();
    return HQ_OK;
}

int hq_tree_create_batch(hq_tree** out, int nsys, const int64_t* counts, int method,
                         int group_size, int precision) {
    if (!out) return fail(HQ_ERR_ARGS, "null handle pointer");
    *out = nullptr;
    if (nsys < 1 || !counts) return fail(HQ_ERR_ARGS, "batch needs at least one system");
    std::vector<int64_t> off(nsys + 1, 0);
    for (int k = 0; k < nsys; ++k) {
        if (counts[k] < 1)
            return fail(HQ_ERR_ARGS, "system %d of the batch has %lld particles", k, (long long)counts[k]);
        off[k + 1] = off[k] + counts[k];
        if (off[k + 1] > (int64_t)1 << 30)
            return fail(HQ_ERR_ARGS, "batch holds more than 2^30 particles");
    }
    hq_tree* t = nullptr;
    int rc = hq_tree_create(&t, off[nsys], method, group_size, precision);
    if (rc) return rc;
    std::unique_ptr<hq_tree, int (*)(hq_tree*)> guard(t, hq_tree_destroy);
    t->nsys = nsys;
    t->hsysoff = off;
    if ((rc = t->sysoff.ensure((nsys + 1) * sizeof(int64_t)))) return rc;
    HQ_CHECK_HIP(hipMemcpyAsync(t->sysoff.p, off.data(), (nsys + 1) * sizeof(int64_t),
                                hipMemcpyHostToDevice, t->stream));
    HQ_CHECK_HIP(hipStreamSynchronize(t->stream));
    *out = guard.release();
    return HQ_OK;
}

int hq_tree_destroy(hq_tree* t) {
    if (!t) return HQ_OK;
    if (t->stream) hipStreamSynchronize(t->stream);
    for (auto& g : t->gcache)
        if (g.exec) hipGraphExecDestroy(g.exec);
    for (auto& e : t->ev)
        if (e) hipEventDestroy(e);
    if (t->hstatus) hipHostFree(t->hstatus);
    hipStream_t s = t->stream;
    delete t;  // DevBuf members free their device memory
    if (s) hipStreamDestroy(s);
    return HQ_OK;
}

int hq_tree_tune(hq_tree* t, int key, int64_t value) {
    if (!t) return fail(HQ_ERR_ARGS, "null handle");
    if (key == HQ_TUNE_BUNCH) {
        if (value < 1 || value > 64) return fail(HQ_ERR_ARGS, "bunch size %lld outside 1..64", (long long)value);
        t->group_size = (int)value;
    } else if (key == HQ_TUNE_SUPER) {
        if (value < 1 || value > (1 << 20)) return fail(HQ_ERR_ARGS, "super-group size %lld out of range", (long long)value);
        t->super_size = (int)value;
    } else if (key == HQ_TUNE_FUSED) {
        t->fused = value != 0;
    } else if (key == HQ_TUNE_FMM_LEAF) {
        if (value < 1 || value > 64) return fail(HQ_ERR_ARGS, "FMM leaf size %lld outside 1..64", (long long)value);
        t->fmm_leaf = (int)value;
    } else if (key == HQ_TUNE_GRAPH) {
        t->graph = value != 0;
    } else if (key == HQ_TUNE_SPLIT) {
        if (value < 0 || value > 2) return fail(HQ_ERR_ARGS, "split mode %lld outside 0..2", (long long)value);
        t->split = (int)value;
    } else if (key == HQ_TUNE_FAR) {
        if (value < 0 || value > 1000) return fail(HQ_ERR_ARGS, "far factor %lld outside 0..1000", (long long)value);
        t->far = (int)value;
    } else if (key == HQ_TUNE_FAR_INNER) {
        if (value < 0 || value > 1000) return fail(HQ_ERR_ARGS, "inner far factor %lld outside 0..1000", (long long)value);
        t->far_inner = (int)value;
    } else if (key == HQ_TUNE_MAC) {
        if (value < 0 || value > 1) return fail(HQ_ERR_ARGS, "acceptance test %lld outside 0..1", (long long)value);
        t->mac = (int)value;
    } else if (key == HQ_TUNE_FMM_BIN_WINDOW) {
        if (value < 1 || value > 4096) return fail(HQ_ERR_ARGS, "FMM bin window %lld outside 1..4096", (long long)value);
        t->fmm_binw = (int)value;
    } else if (key == HQ_TUNE_GRAPH_GRID) {
        if (value < 0 || value > ((int64_t)1 << 30)) return fail(HQ_ERR_ARGS, "graph grid %lld out of range", (long long)value);
        t->graph_grid = value;
    } else if (key == HQ_TUNE_LEAF) {
        if (value < 0 || value > 64) return fail(HQ_ERR_ARGS, "leaf bucket %lld outside 0..64", (long long)value);
        t->leaf = (int)value;
    } else {
        return fail(HQ_ERR_ARGS, "unknown tuning key %d", key);
    }
    t->tune_user |= 1u << key;  // set by the caller: the tuning table leaves it alone
    return HQ_OK;
}

int hq_tree_set_masses(hq_tree* t, const double* m, int where) {
    if (!t || !m) return fail(HQ_ERR_ARGS, "null argument");
    int rc = stage(t->m, m, t->n, where, t->stream);
    if (rc) return rc;
    HQ_CHECK_HIP(hipStreamSynchronize(t->stream));
    t->have_masses = true;
    return HQ_OK;
}

int hq_tree_build(hq_tree* t, const double* x, const double* y, const double* z, int where) {
    if (!t || !x || !y || !z) return fail(HQ_ERR_ARGS, "null argument");
    if (!t->have_masses) return fail(HQ_ERR_ARGS, "Not allocated! Run pre_init first!");
    int rc;
    t->built = false;
    if ((rc = stage(t->x, x, t->n, where, t->stream)) || (rc = stage(t->y, y, t->n, where, t->stream)) ||
        (rc = stage(t->z, z, t->n, where, t->stream)))
        return rc;
    return tree_build(t);
}

int hq_tree_forces(hq_tree* t, double theta2, double G, double eps2, double* F, int layout,
                   int where) {
    if (!t || !F) return fail(HQ_ERR_ARGS, "null argument");
    if (!t->built) return fail(HQ_ERR_ARGS, "tree not built (call hq_tree_build)");
    if (layout != HQ_LAYOUT_3N && layout != HQ_LAYOUT_N3) return fail(HQ_ERR_ARGS, "bad layout");
    int rc;
    if (t->method < 20) rc = tree_forces_bh(t, theta2, G, eps2);
    else if (t->method < 30) rc = tree_forces_dehnen(t, theta2, G, eps2);
    else rc = tree_forces_groups(t, theta2, G, eps2, HQ_FP64, false);
    if (rc) return rc;
    return tree_scatter_forces(t, F, layout, where);
}

int hq_tree_forces_bunched(hq_tree* t, double theta2, double G, double eps2, double* F,
                           int layout, int where, int precision) {
    if (!t || !F) return fail(HQ_ERR_ARGS, "null argument");
    if (!t->built) return fail(HQ_ERR_ARGS, "tree not built (call hq_tree_build)");
    if (layout != HQ_LAYOUT_3N && layout != HQ_LAYOUT_N3) return fail(HQ_ERR_ARGS, "bad layout");
    if (precision != HQ_FP64 && precision != HQ_MIXED) return fail(HQ_ERR_ARGS, "bad precision");
    int rc = tree_forces_groups(t, theta2, G, eps2, precision, true);
    if (rc) return rc;
    return tree_scatter_forces(t, F, layout, where);
}

int hq_tree_autotune(hq_tree* t, double theta2, double G, double eps2, int precision, int reps,
                     int* bunch, int* super_size, int* leaf, double* ms) {
    if (!t) return fail(HQ_ERR_ARGS, "null handle");
    if (!t->built) return fail(HQ_ERR_ARGS, "tree not built (call hq_tree_build)");
    if (precision != HQ_FP64 && precision != HQ_MIXED) return fail(HQ_ERR_ARGS, "bad precision");
    TuneSetting best;
    double best_ms = 0.0;
    const int rc = autotune(t, theta2, G, eps2, precision, reps < 1 ? 1 : reps, &best, &best_ms);
    if (rc) return rc;
    if (bunch) *bunch = best.bunch;
    if (super_size) *super_size = best.super_size;
    if (leaf) *leaf = best.leaf;
    if (ms) *ms = best_ms;
    return HQ_OK;
}

int hq_tree_get_tune(const hq_tree* t, int key, int64_t* value) {
    if (!t || !value) return fail(HQ_ERR_ARGS, "null argument");
    switch (key) {
        case HQ_TUNE_BUNCH: *value = t->group_size; break;
        case HQ_TUNE_SUPER: *value = t->super_size; break;
        case HQ_TUNE_FUSED: *value = t->fused ? 1 : 0; break;
        case HQ_TUNE_LEAF: *value = t->leaf; break;
        case HQ_TUNE_FMM_LEAF: *value = t->fmm_leaf; break;
        case HQ_TUNE_GRAPH: *value = t->graph ? 1 : 0; break;
        case HQ_TUNE_SPLIT: *value = t->split; break;
        case HQ_TUNE_FAR: *value = t->far; break;
        case HQ_TUNE_FAR_INNER: *value = t->far_inner; break;
        case HQ_TUNE_GRAPH_GRID: *value = t->graph_grid; break;
        case HQ_TUNE_MAC: *value = t->mac; break;
        case HQ_TUNE_FMM_BIN_WINDOW: *value = t->fmm_binw; break;
        default: return fail(HQ_ERR_ARGS, "unknown tuning key %d", key);
    }
    return HQ_OK;
}

int hq_tune_table_save(const char* path) {
    if (!path) return fail(HQ_ERR_ARGS, "null path");
    return tune_table_save(path);
}

int hq_tune_table_load(const char* path) {
    if (!path) return fail(HQ_ERR_ARGS, "null path");
    return tune_table_load(path);
}

int hq_tune_table_clear(void) {
    tune_table_clear();
    return HQ_OK;
}

int hq_tree_fmm(hq_tree* t, int order, double theta2, double G, double eps2, int mode, double* F,
                int layout, int where) {
    if (!t || !F) return fail(HQ_ERR_ARGS, "null argument");
    if (!t->built) return fail(HQ_ERR_ARGS, "tree not built (call hq_tree_build)");
    if (order != 4 && order != 6) return fail(HQ_ERR_ARGS, "FMM order %d (4 or 6)", order);
    if (mode != HQ_FMM_MFMA && mode != HQ_FMM_FP64) return fail(HQ_ERR_ARGS, "bad FMM mode %d", mode);
    if (!(theta2 > 0.0) || theta2 >= 1.0) return fail(HQ_ERR_ARGS, "FMM needs 0 < theta^2 < 1");
    if (layout != HQ_LAYOUT_3N && layout != HQ_LAYOUT_N3) return fail(HQ_ERR_ARGS, "bad layout");
    if (t->lost > 0) return fail(HQ_ERR_LOST, "body lost (%lld particles at max depth)", (long long)t->lost);
    int rc = tree_forces_fmm(t, order, sqrt(theta2), G, eps2, mode == HQ_FMM_MFMA, t->fmm_leaf);
    if (rc) return rc;
    return tree_scatter_forces(t, F, layout, where);
}

int hq_tree_info(const hq_tree* t, hq_tree_info_t* info) {
    if (!t || !info) return fail(HQ_ERR_ARGS, "null argument");
    collect_build_times(const_cast<hq_tree*>(t));  // the last build's phase times, if pending
    memset(info, 0, sizeof(*info));
    info->n = t->n;
    info->nnodes = t->nn;
    info->ngroups = t->ngroups;
    info->lost = t->lost;
    info->most_depth = t->most_depth;
    info->multipole = t->multipole;
    info->list_cells = t->list_cells;
    info->list_parts = t->list_parts;
    info->far_cells = t->far_cells;
    info->n_p2p = t->n_p2p;
    info->n_m2p = t->n_m2p;
    info->ms_build = t->ms_build;
    info->ms_keys = t->ms_keys;
    info->ms_sort = t->ms_sort;
    info->ms_nodes = t->ms_nodes;
    info->ms_upward = t->ms_upward;
    info->ms_walk = t->ms_walk;
    info->ms_lists = t->ms_lists;
    info->ms_eval = t->ms_eval;
    info->exec_cells = t->exec_cells;
    info->exec_parts = t->exec_parts;
    info->max_cells = t->max_cells;
    info->max_parts = t->max_parts;
    info->max_stack = t->max_stack;
    info->split_walks = t->split_walks;
    info->walk_stack = t->walk_stack;
    return HQ_OK;
}

int hq_tree_export(const hq_tree* t, int* depth, int* type, int* particle, int* child_start,
                   uint64_t* key_hi, uint32_t* key_lo, double* mass, double* com, double* quad,
                   double* oct, double* center) {
    if (!t || !t->built) return fail(HQ_ERR_ARGS, "tree not built");
    const int64_t nn = t->nn, S = t->cap;
    hipStream_t s = t->stream;
    auto d2h = [&](void* dst, const void* src, size_t bytes) -> int {
        HQ_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        return HQ_OK;
    };
    std::unique_ptr<uint8_t[]> u8(new uint8_t[nn]);
    std::unique_ptr<int8_t[]> i8(new int8_t[nn]);
    int rc;
    if (depth) {
        if ((rc = d2h(u8.get(), t->ndepth.p, nn))) return rc;
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        for (int64_t k = 0; k < nn; ++k) depth[k] = u8[k];
    }
    if (type || particle) {
        if ((rc = d2h(i8.get(), t->ntype.p, nn))) return rc;
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        if (type)
            for (int64_t k = 0; k < nn; ++k) type[k] = i8[k];
    }
    if (particle) {
        // leaf -> original index of its particle; twig -> -1
        std::unique_ptr<int[]> first(new int[nn]);
        std::unique_ptr<int[]> order(new int[t->n]);
        if ((rc = d2h(first.get(), t->nfirst.p, nn * 4)) || (rc = d2h(order.get(), t->idx.p, t->n * 4)))
            return rc;
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        for (int64_t k = 0; k < nn; ++k) particle[k] = i8[k] == 1 ? order[first[k]] : -1;
    }
    if (child_start && (rc = d2h(child_start, t->ncstart.p, (nn + 1) * 4))) return rc;
    if (key_hi && (rc = d2h(key_hi, t->nkhi.p, nn * 8))) return rc;
    if (key_lo && (rc = d2h(key_lo, t->nklo.p, nn * 4))) return rc;
    if (mass && (rc = d2h(mass, t->nmass.p, nn * 8))) return rc;
    if (com) {
        std::unique_ptr<double[]> tmp(new double[3 * nn]);
        if ((rc = d2h(tmp.get(), t->ncomx.p, nn * 8)) || (rc = d2h(tmp.get() + nn, t->ncomy.p, nn * 8)) ||
            (rc = d2h(tmp.get() + 2 * nn, t->ncomz.p, nn * 8)))
            return rc;
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        for (int64_t k = 0; k < nn; ++k)
            for (int c = 0; c < 3; ++c) com[3 * k + c] = tmp[c * nn + k];
    }
    if (center) {
        std::unique_ptr<double[]> tmp(new double[3 * nn]);
        if ((rc = d2h(tmp.get(), t->ncx.p, nn * 8)) || (rc = d2h(tmp.get() + nn, t->ncy.p, nn * 8)) ||
            (rc = d2h(tmp.get() + 2 * nn, t->ncz.p, nn * 8)))
            return rc;
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        for (int64_t k = 0; k < nn; ++k)
            for (int c = 0; c < 3; ++c) center[3 * k + c] = tmp[c * nn + k];
    }
    auto soa = [&](double* dst, const DevBuf& src, int ncomp) -> int {
        std::unique_ptr<double[]> tmp(new double[(size_t)ncomp * nn]);
        for (int j = 0; j < ncomp; ++j)
            HQ_CHECK_HIP(hipMemcpyAsync(tmp.get() + (size_t)j * nn, src.as<double>() + (size_t)j * S,
                                        nn * 8, hipMemcpyDeviceToHost, s));
        HQ_CHECK_HIP(hipStreamSynchronize(s));
        for (int64_t k = 0; k < nn; ++k)
            for (int j = 0; j < ncomp; ++j) dst[(size_t)ncomp * k + j] = tmp[(size_t)j * nn + k];
        return HQ_OK;
    };
    if (quad) {
        if (t->multipole >= 4) {
            if ((rc = soa(quad, t->nquad, 6))) return rc;
        } else {
            memset(quad, 0, nn * 6 * 8);
        }
    }
    if (oct) {
        if (t->multipole > 4) {
            if ((rc = soa(oct, t->noct, 14))) return rc;
        } else {
            memset(oct, 0, nn * 14 * 8);
        }
    }
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    return HQ_OK;
}

int hq_tree_order(const hq_tree* t, int* order) {
    if (!t || !t->built || !order) return fail(HQ_ERR_ARGS, "tree not built");
    HQ_CHECK_HIP(hipMemcpyAsync(order, t->idx.p, t->n * 4, hipMemcpyDeviceToHost, t->stream));
    HQ_CHECK_HIP(hipStreamSynchronize(t->stream));
    return HQ_OK;
}

int hq_direct_forces_device(int64_t n, const double* m, const double* x, const double* y,
                            const double* z, int64_t nt, const int64_t* targets, double G,
                            double eps2, int precision, double* F) {
    if (n < 1 || !m || !x || !y || !z || !F) return fail(HQ_ERR_ARGS, "bad argument");
    if (!targets) nt = n;
    int rc = direct_forces(n, m, x, y, z, nt, targets, G, eps2, precision, F, 0);
    if (rc) return rc;
    HQ_CHECK_HIP(hipStreamSynchronize(0));
    return HQ_OK;
}

int hq_compute_forces_direct(int64_t n, const double* m, const double* x, const double* y,
                             const double* z, double G, double eps, double* F) {
    if (n < 1 || !m || !x || !y || !z || !F) return fail(HQ_ERR_ARGS, "bad argument");
    DevBuf dm, dx, dy, dz, dF, dN3;
    int rc;
    hipStream_t s = 0;
    if ((rc = stage(dm, m, n, HQ_HOST, s)) || (rc = stage(dx, x, n, HQ_HOST, s)) ||
        (rc = stage(dy, y, n, HQ_HOST, s)) || (rc = stage(dz, z, n, HQ_HOST, s)) ||
        (rc = dF.ensure(3 * n * 8)))
        return rc;
    if ((rc = direct_forces(n, dm.as<double>(), dx.as<double>(), dy.as<double>(), dz.as<double>(),
                            n, nullptr, G, eps * eps, HQ_FP64, dF.as<double>(), s)))
        return rc;
    std::unique_ptr<double[]> tmp(new double[3 * n]);
    HQ_CHECK_HIP(hipMemcpyAsync(tmp.get(), dF.p, 3 * n * 8, hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    for (int64_t p = 0; p < n; ++p)
        for (int c = 0; c < 3; ++c) F[c * n + p] = tmp[3 * p + c];
    return HQ_OK;
}

int hq_compute_forces(int64_t n, const double* m, const double* x, const double* y,
                      const double* z, double theta, double G, double eps, int method,
                      int legacy_theta2, double* F) {
    hq_tree* t = nullptr;
    int rc = hq_tree_create(&t, n, method, 64, HQ_FP64);
    if (rc) return rc;
    double theta2 = legacy_theta2 ? theta : theta * theta;
    if (!(rc = hq_tree_set_masses(t, m, HQ_HOST)) && !(rc = hq_tree_build(t, x, y, z, HQ_HOST)))
        rc = hq_tree_forces(t, theta2, G, eps * eps, F, HQ_LAYOUT_N3, HQ_HOST);
    hq_tree_destroy(t);
    return rc;
}

int hq_generate_tree(int64_t n, const double* m, const double* x, const double* y,
                     const double* z, int method) {
    hq_tree* t = nullptr;
    int rc = hq_tree_create(&t, n, method, 64, HQ_FP64);
    if (rc) return rc;
    if (!(rc = hq_tree_set_masses(t, m, HQ_HOST))) rc = hq_tree_build(t, x, y, z, HQ_HOST);
    hq_tree_destroy(t);
    return rc;
}

int hq_morton_keys(int64_t n, const double* m, const double* q, int L, double ssa, int64_t* keys,
                   int* perm, int* L_out, int where) {
    if (n < 1 || !m || !q || !keys || !perm || !L_out) return fail(HQ_ERR_ARGS, "bad argument");
    hipStream_t s = 0;
    int rc;
    if (where == HQ_DEVICE) {
        rc = morton_keys(n, m, q, L, ssa, keys, perm, L_out, s);
        return rc;
    }
    DevBuf dm, dq, dk, dp;
    if ((rc = stage(dm, m, n, HQ_HOST, s)) || (rc = stage(dq, q, 3 * n, HQ_HOST, s)) ||
        (rc = dk.ensure(n * 8)) || (rc = dp.ensure(n * 4)))
        return rc;
    if ((rc = morton_keys(n, dm.as<double>(), dq.as<double>(), L, ssa, dk.as<int64_t>(),
                          dp.as<int>(), L_out, s)))
        return rc;
    HQ_CHECK_HIP(hipMemcpyAsync(keys, dk.p, n * 8, hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipMemcpyAsync(perm, dp.p, n * 4, hipMemcpyDeviceToHost, s));
    HQ_CHECK_HIP(hipStreamSynchronize(s));
    return HQ_OK;
}

}  // extern "C"
