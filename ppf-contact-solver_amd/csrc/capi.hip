// capi.hip -- extern "C" entry points declared in include/ppf_mi355x.h.
//
// Plain pointers and sizes only; no torch or C++ types cross the boundary.
#include "common.hpp"
#include "eig3.hpp"
#include "elastic.hpp"
#include "isect.hpp"
#include "assemble.hpp"
#include "learned.hpp"
#include "linesearch.hpp"
#include "narrow.hpp"
#include "newton.hpp"
#include "pcg.hpp"
#include <atomic>
#include <cstdarg>
#include <cstring>

namespace ppf {

static thread_local char g_err[1024];
static std::atomic<uint64_t> g_allocs{0}, g_frees{0};

void count_alloc() { g_allocs.fetch_add(1, std::memory_order_relaxed); }

namespace {
// walk statistics: 4 kinds x 5 counters on the device, kernel time per kind
struct WalkStatsState {
    unsigned long long *dev = nullptr;
    bool on = false;
    double ms[4] = {};
    hipEvent_t ev[2] = {nullptr, nullptr};
} g_walk;
} // namespace

unsigned long long *walk_stats_device() { return g_walk.on ? g_walk.dev : nullptr; }

WalkClock::WalkClock(int kind, hipStream_t s) : kind_(kind), s_(s), on_(g_walk.on) {
    if (on_) {
        (void)hipEventRecord(g_walk.ev[0], s_);
    }
}

void WalkClock::stop() {
    if (!on_) {
        return;
    }
    on_ = false;
    float ms = 0.0f;
    if (hipEventRecord(g_walk.ev[1], s_) == hipSuccess &&
        hipEventSynchronize(g_walk.ev[1]) == hipSuccess &&
        hipEventElapsedTime(&ms, g_walk.ev[0], g_walk.ev[1]) == hipSuccess) {
        g_walk.ms[kind_] += ms;
    }
}
void count_free() { g_frees.fetch_add(1, std::memory_order_relaxed); }
void alloc_totals(uint64_t *allocs, uint64_t *frees) {
    *allocs = g_allocs.load(std::memory_order_relaxed);
    *frees = g_frees.load(std::memory_order_relaxed);
}

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

namespace {

__global__ void k_bj_invert(uint32_t n, const float *__restrict__ in,
                            float *__restrict__ out, uint32_t *__restrict__ bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    float a[9], o[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        a[e] = in[9 * (size_t)i + e];
    }
    const bool ok = bj_invert(a, o);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        out[9 * (size_t)i + e] = o[e];
    }
    if (!ok && bad) {
        atomicAdd(bad, 1u);
    }
}

__global__ void k_symm3x3(uint32_t n, const float *__restrict__ in,
                          float *__restrict__ val, float *__restrict__ vec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    m3 a;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        a.m[e] = in[9 * (size_t)i + e];
    }
    f3 l;
    m3 q;
    symm3x3(a, l, q);
    store3(val + 3 * (size_t)i, l);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        vec[9 * (size_t)i + e] = q.m[e];
    }
}

} // namespace
} // namespace ppf

using namespace ppf;

struct ppf_pcg_plan {
    PcgPlanImpl *impl;
};

struct ppf_narrow {
    Narrow impl;
    IntersectionCheck isect;
};

struct ppf_contact {
    ContactAssembler impl;
};

struct ppf_ccd {
    LineSearch impl;
};

struct ppf_elastic {
    ElasticAssembler impl;
};

struct ppf_stepper {
    Stepper impl;
};

struct ppf_lpc {
    Lpc impl;
};

extern "C" {

const char *ppf_last_error(void) { return g_err; }

int ppf_device_alloc(void **ptr, size_t bytes) {
    if (!ptr) {
        set_error("ppf_device_alloc: null pointer");
        return PPF_ERR_ARG;
    }
    PPF_HIP(ppf::dev_alloc(ptr, bytes ? bytes : 1));
    return PPF_OK;
}

void ppf_device_free(void *ptr) { (void)ppf::dev_free(ptr); }

int ppf_walk_stats_enable(int on) {
    if (on) {
        if (!g_walk.dev) {
            PPF_HIP(dev_alloc(&g_walk.dev, 4 * 5 * sizeof(unsigned long long)));
            for (hipEvent_t &e : g_walk.ev) {
                PPF_HIP(hipEventCreate(&e));
            }
        }
        PPF_HIP(hipMemset(g_walk.dev, 0, 4 * 5 * sizeof(unsigned long long)));
        for (double &m : g_walk.ms) {
            m = 0.0;
        }
    }
    g_walk.on = on != 0;
    return PPF_OK;
}

int ppf_walk_stats_read(ppf_walk_stats *out) {
    if (!out) {
        set_error("ppf_walk_stats_read: null output");
        return PPF_ERR_ARG;
    }
    unsigned long long h[20] = {};
    if (g_walk.dev) {
        PPF_HIP(hipDeviceSynchronize());
        PPF_HIP(hipMemcpy(h, g_walk.dev, sizeof(h), hipMemcpyDeviceToHost));
    }
    for (int k = 0; k < 4; ++k) {
        out[k].launches = h[5 * k];
        out[k].queries = h[5 * k + 1];
        out[k].nodes = h[5 * k + 2];
        out[k].leaves = h[5 * k + 3];
        out[k].sweeps = h[5 * k + 4];
        out[k].ms = g_walk.ms[k];
    }
    return PPF_OK;
}

void ppf_alloc_counts(uint64_t *allocs, uint64_t *frees) {
    uint64_t a = 0, f = 0;
    ppf::alloc_totals(&a, &f);
    if (allocs) *allocs = a;
    if (frees) *frees = f;
}

int ppf_device_info(char *arch, uint32_t arch_len, uint32_t *cu_count) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return PPF_ERR_NODEV;
    }
    hipDeviceProp_t prop;
    PPF_HIP(hipGetDeviceProperties(&prop, 0));
    if (arch && arch_len) {
        snprintf(arch, arch_len, "%s", prop.gcnArchName);
    }
    if (cu_count) {
        *cu_count = (uint32_t)prop.multiProcessorCount;
    }
    return PPF_OK;
}

int ppf_pcg_plan_create(uint32_t n, uint32_t nnz, const uint32_t *d_row_offset,
                        const uint32_t *d_col_index, void *stream,
                        ppf_pcg_plan **out) {
    if (!out) {
        set_error("ppf_pcg_plan_create: out is null");
        return PPF_ERR_ARG;
    }
    PcgPlanImpl *impl = nullptr;
    int rc = pcg_plan_create(n, nnz, d_row_offset, d_col_index,
                             (hipStream_t)stream, &impl);
    if (rc != PPF_OK) {
        return rc;
    }
    *out = new ppf_pcg_plan{impl};
    return PPF_OK;
}

void ppf_pcg_plan_destroy(ppf_pcg_plan *plan) {
    if (plan) {
        pcg_plan_destroy(plan->impl);
        delete plan;
    }
}

int ppf_pcg_plan_repattern(ppf_pcg_plan *plan, uint32_t nnz, const uint32_t *d_row_offset,
                           const uint32_t *d_col_index, void *stream) {
    if (!plan) {
        set_error("ppf_pcg_plan_repattern: plan is null");
        return PPF_ERR_ARG;
    }
    return pcg_plan_repattern(plan->impl, nnz, d_row_offset, d_col_index,
                              (hipStream_t)stream);
}

uint32_t ppf_lpc_weight_count(void) { return LpcLayout::count; }

int ppf_lpc_create(ppf_lpc **out) {
    if (!out) {
        set_error("ppf_lpc_create: out is null");
        return PPF_ERR_ARG;
    }
    *out = new ppf_lpc();
    return PPF_OK;
}

void ppf_lpc_destroy(ppf_lpc *h) { delete h; }

int ppf_lpc_set_weights(ppf_lpc *h, const float *host_weights, uint32_t count, void *stream) {
    if (!h || !host_weights) {
        set_error("ppf_lpc_set_weights: invalid argument");
        return PPF_ERR_ARG;
    }
    return h->impl.set_weights(host_weights, count, (hipStream_t)stream);
}

int ppf_lpc_get_weights(ppf_lpc *h, float *host_weights, uint32_t count, void *stream) {
    if (!h || !host_weights) {
        set_error("ppf_lpc_get_weights: invalid argument");
        return PPF_ERR_ARG;
    }
    return h->impl.get_weights(host_weights, count, (hipStream_t)stream);
}

int ppf_lpc_factor(ppf_lpc *h, ppf_pcg_plan *plan, const float *d_value,
                   const float *d_cdiag, const float *d_b, void *stream) {
    if (!h || !plan || !d_value || !d_cdiag || !d_b) {
        set_error("ppf_lpc_factor: invalid argument");
        return PPF_ERR_ARG;
    }
    const hipStream_t s = (hipStream_t)stream;
    // the factor lives in the plan's row order
    const float *ov = nullptr, *oc = nullptr, *ob = nullptr;
    int rc = pcg_order_system(plan->impl, d_value, d_cdiag, d_b, &ov, &oc, &ob, s);
    if (rc == PPF_OK) {
        rc = h->impl.prepare(pcg_pattern(plan->impl), ov, oc, ob, s);
    }
    return rc == PPF_OK ? h->impl.forward(s) : rc;
}

int ppf_lpc_outputs(ppf_lpc *h, float *d_t, float *d_u, float *d_gd, float *d_gl,
                    void *stream) {
    if (!h || !h->impl.Gd()) {
        set_error("ppf_lpc_outputs: no factor computed");
        return PPF_ERR_ARG;
    }
    const hipStream_t s = (hipStream_t)stream;
    const size_t n = h->impl.n(), z = h->impl.nnz();
    if (d_t) PPF_HIP(hipMemcpyAsync(d_t, h->impl.t(), 6 * n * 4, hipMemcpyDeviceToDevice, s));
    if (d_u) PPF_HIP(hipMemcpyAsync(d_u, h->impl.u(), 9 * z * 4, hipMemcpyDeviceToDevice, s));
    if (d_gd) PPF_HIP(hipMemcpyAsync(d_gd, h->impl.Gd(), 9 * n * 4, hipMemcpyDeviceToDevice, s));
    if (d_gl) PPF_HIP(hipMemcpyAsync(d_gl, h->impl.Gl(), 9 * z * 4, hipMemcpyDeviceToDevice, s));
    return PPF_OK;
}

int ppf_lpc_gradient(ppf_lpc *h, ppf_pcg_plan *plan, const float *d_value,
                     const float *d_cdiag, const float *d_probes, uint32_t nprobe,
                     float *loss, float *host_grad, void *stream) {
    if (!h || !plan || !d_value || !d_cdiag || !d_probes || !loss || !h->impl.Gd()) {
        set_error("ppf_lpc_gradient: invalid argument (or no factor computed)");
        return PPF_ERR_ARG;
    }
    if (pcg_plan_ordered(plan->impl)) {
        set_error("ppf_lpc_gradient: training runs on plans in the original row order");
        return PPF_ERR_ARG;
    }
    const hipStream_t s = (hipStream_t)stream;
    int rc = h->impl.gradient(plan->impl, d_value, d_cdiag, d_probes, nprobe, loss, s);
    if (rc == PPF_OK && host_grad) {
        PPF_HIP(hipMemcpyAsync(host_grad, h->impl.grad(), LpcLayout::count * sizeof(float),
                               hipMemcpyDeviceToHost, s));
        PPF_HIP(hipStreamSynchronize(s));
    }
    return rc;
}

int ppf_lpc_adam(ppf_lpc *h, float lr, uint32_t step, void *stream) {
    if (!h || step == 0) {
        set_error("ppf_lpc_adam: invalid argument");
        return PPF_ERR_ARG;
    }
    return h->impl.adam(lr, step, (hipStream_t)stream);
}

int ppf_pcg_solve_lpc(ppf_pcg_plan *plan, ppf_lpc *h, const float *d_value,
                      const float *d_cdiag, const float *d_b, float *d_x, float tol,
                      uint32_t max_iter, ppf_pcg_result *out, void *stream) {
    if (!plan || !h || !h->impl.Gd() || !out) {
        set_error("ppf_pcg_solve_lpc: invalid argument (or no factor computed)");
        return PPF_ERR_ARG;
    }
    if (h->impl.n() != pcg_pattern(plan->impl).n || h->impl.nnz() != pcg_pattern(plan->impl).nnz) {
        set_error("ppf_pcg_solve_lpc: the factor was computed for another system");
        return PPF_ERR_ARG;
    }
    const PcgFactor fac = h->impl.factor();
    return pcg_solve(plan->impl, d_value, d_cdiag, d_b, d_x, tol, max_iter, nullptr, out,
                     (hipStream_t)stream, &fac);
}

namespace {

// The one-result entry points take one-scene plans only.
bool one_scene(const ppf_pcg_plan *plan, const char *who) {
    if (!plan) {
        set_error("%s: plan is null", who);
        return false;
    }
    if (pcg_plan_scenes(plan->impl) != 1) {
        set_error("%s: the plan holds %u scenes (ppf_pcg_solve_batch solves them)", who,
                  pcg_plan_scenes(plan->impl));
        return false;
    }
    return true;
}

} // namespace

int ppf_pcg_plan_create_batch(uint32_t n, uint32_t nnz, const uint32_t *d_row_offset,
                              const uint32_t *d_col_index, uint32_t nscene,
                              const uint32_t *h_scene_row, void *stream, ppf_pcg_plan **out) {
    if (!out || !h_scene_row || nscene == 0) {
        set_error("ppf_pcg_plan_create_batch: invalid argument");
        return PPF_ERR_ARG;
    }
    PcgPlanImpl *impl = nullptr;
    int rc = pcg_plan_create_scenes(n, nnz, d_row_offset, d_col_index, nscene, h_scene_row,
                                    (hipStream_t)stream, &impl);
    if (rc != PPF_OK) {
        return rc;
    }
    *out = new ppf_pcg_plan{impl};
    return PPF_OK;
}

int ppf_pcg_plan_set_order(ppf_pcg_plan *plan, const uint32_t *d_order, void *stream) {
    if (!plan) {
        set_error("ppf_pcg_plan_set_order: plan is null");
        return PPF_ERR_ARG;
    }
    return pcg_plan_set_order(plan->impl, d_order, (hipStream_t)stream);
}

int ppf_pcg_plan_set_tiles(ppf_pcg_plan *plan, uint32_t spmv_rows, uint32_t vector_rows,
                           void *stream) {
    if (!plan) {
        set_error("ppf_pcg_plan_set_tiles: plan is null");
        return PPF_ERR_ARG;
    }
    return pcg_plan_set_tiles(plan->impl, spmv_rows, vector_rows, (hipStream_t)stream);
}

int ppf_pcg_plan_set_spmv_form(ppf_pcg_plan *plan, int form, int *form_in_use, void *stream) {
    if (!plan) {
        set_error("ppf_pcg_plan_set_spmv_form: plan is null");
        return PPF_ERR_ARG;
    }
    const int rc = pcg_plan_set_spmv_form(plan->impl, form, (hipStream_t)stream);
    if (form_in_use) {
        *form_in_use = pcg_plan_spmv_form(plan->impl);
    }
    return rc;
}

int ppf_pcg_plan_autotune(ppf_pcg_plan *plan, const float *d_value, const float *d_cdiag,
                          const float *d_b, float tol, uint32_t *h_tile_rows, float *h_ms,
                          void *stream) {
    if (!plan || !h_tile_rows) {
        set_error("ppf_pcg_plan_autotune: invalid argument");
        return PPF_ERR_ARG;
    }
    return pcg_plan_autotune(plan->impl, d_value, d_cdiag, d_b, tol, h_tile_rows,
                             h_tile_rows + 1, h_ms, (hipStream_t)stream);
}

int ppf_pcg_solve_batch(ppf_pcg_plan *plan, const float *d_value, const float *d_cdiag,
                        const float *d_b, float *d_x, float tol, uint32_t max_iter,
                        ppf_pcg_result *h_out, void *stream) {
    if (!plan) {
        set_error("ppf_pcg_solve_batch: plan is null");
        return PPF_ERR_ARG;
    }
    return pcg_solve(plan->impl, d_value, d_cdiag, d_b, d_x, tol, max_iter, nullptr, h_out,
                     (hipStream_t)stream);
}

int ppf_pcg_solve(ppf_pcg_plan *plan, const float *d_value,
                  const float *d_cdiag, const float *d_b, float *d_x, float tol,
                  uint32_t max_iter, ppf_pcg_result *out, void *stream) {
    if (!one_scene(plan, "ppf_pcg_solve")) {
        return PPF_ERR_ARG;
    }
    return pcg_solve(plan->impl, d_value, d_cdiag, d_b, d_x, tol, max_iter,
                     nullptr, out, (hipStream_t)stream);
}

int ppf_pcg_solve_hist(ppf_pcg_plan *plan, const float *d_value,
                       const float *d_cdiag, const float *d_b, float *d_x,
                       float tol, uint32_t max_iter, float *d_err_hist,
                       ppf_pcg_result *out, void *stream) {
    if (!one_scene(plan, "ppf_pcg_solve_hist")) {
        return PPF_ERR_ARG;
    }
    return pcg_solve(plan->impl, d_value, d_cdiag, d_b, d_x, tol, max_iter,
                     d_err_hist, out, (hipStream_t)stream);
}

int ppf_pcg_spmv(ppf_pcg_plan *plan, const float *d_value, const float *d_cdiag,
                 const float *d_x, float *d_y, void *stream) {
    if (!plan || !d_value || !d_cdiag || !d_x || !d_y) {
        set_error("ppf_pcg_spmv: invalid argument");
        return PPF_ERR_ARG;
    }
    return pcg_spmv(plan->impl, d_value, d_cdiag, d_x, d_y, (hipStream_t)stream);
}

int ppf_pcg_time_spmv(ppf_pcg_plan *plan, const float *d_value,
                      const float *d_cdiag, const float *d_x, float *d_y,
                      uint32_t reps, float *avg_ms, void *stream) {
    if (!plan) {
        set_error("ppf_pcg_time_spmv: plan is null");
        return PPF_ERR_ARG;
    }
    return pcg_time_spmv(plan->impl, d_value, d_cdiag, d_x, d_y, reps, avg_ms,
                         (hipStream_t)stream);
}

int ppf_pcg_solve_noalloc(uint32_t n, uint32_t nnz, const uint32_t *d_row_offset,
                          const uint32_t *d_col_index, const float *d_value,
                          const float *d_cdiag, const float *d_b, float *d_x, float tol,
                          uint32_t max_iter, ppf_pcg_result *out, void *stream) {
    return pcg_solve_noalloc(n, nnz, d_row_offset, d_col_index, d_value, d_cdiag, d_b, d_x,
                             tol, max_iter, out, (hipStream_t)stream);
}

int ppf_bj_invert(uint32_t n, const float *d_in, float *d_out, uint32_t *d_bad,
                  void *stream) {
    if (n == 0) {
        return PPF_OK;
    }
    if (!d_in || !d_out) {
        set_error("ppf_bj_invert: null pointer");
        return PPF_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    k_bj_invert<<<(n + 255) / 256, 256, 0, st>>>(n, d_in, d_out, d_bad);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipStreamSynchronize(st));
    return PPF_OK;
}

int ppf_symm3x3(uint32_t n, const float *d_in, float *d_val, float *d_vec,
                void *stream) {
    if (n == 0) {
        return PPF_OK;
    }
    if (!d_in || !d_val || !d_vec) {
        set_error("ppf_symm3x3: null pointer");
        return PPF_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    k_symm3x3<<<(n + 255) / 256, 256, 0, st>>>(n, d_in, d_val, d_vec);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipStreamSynchronize(st));
    return PPF_OK;
}

int ppf_narrow_create(ppf_narrow **out) {
    if (!out) {
        set_error("ppf_narrow_create: out is null");
        return PPF_ERR_ARG;
    }
    *out = new ppf_narrow();
    return PPF_OK;
}

void ppf_narrow_destroy(ppf_narrow *h) { delete h; }

int ppf_narrow_build(ppf_narrow *h, const ppf_contact_mesh *m, void *stream) {
    if (!h || !m || !m->x) {
        set_error("ppf_narrow_build: invalid argument");
        return PPF_ERR_ARG;
    }
    return h->impl.build_trees(*m, (hipStream_t)stream);
}

int ppf_narrow_tree(ppf_narrow *h, int kind, uint32_t *d_nodes, float *d_boxes,
                    uint32_t *n, uint32_t *root, void *stream) {
    if (!h || kind < 0 || kind > 2 || !n || !root) {
        set_error("ppf_narrow_tree: invalid argument");
        return PPF_ERR_ARG;
    }
    Lbvh &t = h->impl.tree(kind);
    const BvhView v = t.view();
    *n = v.n;
    *root = v.root;
    if (v.n == 0) {
        return PPF_OK;
    }
    const size_t nodes = 2 * (size_t)v.n - 1;
    hipStream_t st = (hipStream_t)stream;
    if (d_nodes) {
        PPF_HIP(hipMemcpyAsync(d_nodes, v.node, 2 * nodes * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, st));
    }
    if (d_boxes) {
        PPF_HIP(hipMemcpyAsync(d_boxes, v.box, nodes * sizeof(Aabb),
                               hipMemcpyDeviceToDevice, st));
    }
    PPF_HIP(hipStreamSynchronize(st));
    return PPF_OK;
}

int ppf_narrow_proximity(ppf_narrow *h, const ppf_contact_mesh *m,
                         ppf_pair *d_out, uint64_t cap, uint64_t *count,
                         void *stream) {
    if (!h || !m || !count) {
        set_error("ppf_narrow_proximity: invalid argument");
        return PPF_ERR_ARG;
    }
    if (h->impl.tree(0).view().n != m->nf || h->impl.tree(1).view().n != m->ne ||
        h->impl.tree(2).view().n != m->nv) {
        set_error("ppf_narrow_proximity: trees were built for another mesh");
        return PPF_ERR_ARG;
    }
    return h->impl.proximity(*m, d_out, cap, count, (hipStream_t)stream);
}

int ppf_narrow_static(ppf_narrow *h, const ppf_contact_mesh *m,
                      const struct ppf_static_mesh *sm, const float *d_vmass,
                      const float *d_emass, const float *d_fmass, ppf_pair *d_out,
                      uint64_t cap, uint64_t *count, void *stream) {
    if (!h || !m || !count || !d_vmass || !d_emass || !d_fmass) {
        set_error("ppf_narrow_static: invalid argument");
        return PPF_ERR_ARG;
    }
    if (h->impl.tree(0).view().n != m->nf || h->impl.tree(1).view().n != m->ne ||
        h->impl.tree(2).view().n != m->nv) {
        set_error("ppf_narrow_static: trees were built for another mesh");
        return PPF_ERR_ARG;
    }
    if (!sm || sm->nv == 0) {
        *count = 0;
        return PPF_OK;
    }
    if (!sm->x || (sm->nf && !sm->face) || (sm->ne && !sm->edge)) {
        set_error("ppf_narrow_static: static mesh arrays missing");
        return PPF_ERR_ARG;
    }
    return h->impl.static_proximity(*m, *sm, d_vmass, d_emass, d_fmass, d_out, cap, count,
                                    (hipStream_t)stream);
}

int ppf_intersection_check(ppf_narrow *h, const ppf_contact_mesh *m,
                           const struct ppf_static_mesh *sm, const float *d_vmass,
                           const float *d_emass, const float *d_fmass,
                           uint32_t surface_vert_count, ppf_isect_record *records,
                           uint32_t cap, uint32_t *count, void *stream) {
    if (!h || !m || !m->x || !count || !d_vmass || !d_emass || !d_fmass) {
        set_error("ppf_intersection_check: invalid argument");
        return PPF_ERR_ARG;
    }
    if (h->impl.tree(0).view().n != m->nf || h->impl.tree(1).view().n != m->ne ||
        h->impl.tree(2).view().n != m->nv) {
        set_error("ppf_intersection_check: trees were built for another mesh");
        return PPF_ERR_ARG;
    }
    if (sm && sm->nv && (!sm->x || (sm->nf && !sm->face))) {
        set_error("ppf_intersection_check: static mesh arrays missing");
        return PPF_ERR_ARG;
    }
    return h->isect.run(h->impl.tree(0).view(), h->impl.tree(1).view(),
                        h->impl.tree(2).view(), *m, sm, d_vmass, d_emass, d_fmass,
                        surface_vert_count, records, cap, count, (hipStream_t)stream);
}

int ppf_contact_create(ppf_contact **out) {
    if (!out) {
        set_error("ppf_contact_create: out is null");
        return PPF_ERR_ARG;
    }
    *out = new ppf_contact();
    return PPF_OK;
}

void ppf_contact_destroy(ppf_contact *h) { delete h; }

int ppf_contact_assemble(ppf_contact *h, const ppf_contact_mesh *m,
                         const ppf_contact_params *p, const ppf_pair *d_pairs,
                         uint64_t npairs, float *d_grad, uint64_t *nnz,
                         void *stream) {
    if (!h || !m || !p || !d_grad || !nnz || (npairs && !d_pairs) || !p->x0 ||
        !p->mass || !p->v_fric || !p->e_fric || !p->f_fric) {
        set_error("ppf_contact_assemble: invalid argument");
        return PPF_ERR_ARG;
    }
    if (p->sm && p->sm->nv &&
        (!p->sm->x || !p->sm->v_fric || (p->sm->ne && (!p->sm->edge || !p->sm->e_fric)) ||
         (p->sm->nf && (!p->sm->face || !p->sm->f_fric)))) {
        set_error("ppf_contact_assemble: collision mesh arrays missing");
        return PPF_ERR_ARG;
    }
    int rc = h->impl.assemble(*m, *p, d_pairs, npairs, d_grad, (hipStream_t)stream);
    *nnz = h->impl.nnz();
    if (rc == PPF_OK) {
        PPF_HIP(hipStreamSynchronize((hipStream_t)stream));
    }
    return rc;
}

int ppf_contact_blocks(ppf_contact *h, const uint32_t **d_row_offset,
                       const uint32_t **d_rows, const uint32_t **d_cols,
                       const float **d_vals, uint64_t *nnz) {
    if (!h) {
        set_error("ppf_contact_blocks: null handle");
        return PPF_ERR_ARG;
    }
    if (d_row_offset) *d_row_offset = h->impl.row_offset();
    if (d_rows) *d_rows = h->impl.rows();
    if (d_cols) *d_cols = h->impl.cols();
    if (d_vals) *d_vals = h->impl.values();
    if (nnz) *nnz = h->impl.nnz();
    return PPF_OK;
}

int ppf_ccd_create(ppf_ccd **out) {
    if (!out) {
        set_error("ppf_ccd_create: out is null");
        return PPF_ERR_ARG;
    }
    *out = new ppf_ccd();
    return PPF_OK;
}

void ppf_ccd_destroy(ppf_ccd *h) { delete h; }

int ppf_line_search(ppf_ccd *h, const ppf_contact_mesh *m, const ppf_static_mesh *sm,
                    const ppf_ls_params *p, int rebuild, float *toi, int *overlap,
                    void *stream) {
    if (!h || !m || !p || !toi || !overlap || !m->x || !p->x1 || !p->vmass ||
        !p->emass || !p->fmass || !(p->max_t > 0.0f)) {
        set_error("ppf_line_search: invalid argument");
        return PPF_ERR_ARG;
    }
    if (sm && sm->nv && (!sm->x || (sm->ne && !sm->edge) || (sm->nf && !sm->face))) {
        set_error("ppf_line_search: static mesh arrays missing");
        return PPF_ERR_ARG;
    }
    return h->impl.run(*m, sm, *p, rebuild != 0, toi, overlap, (hipStream_t)stream);
}

int ppf_ccd_batch(int kind, uint32_t n, const float *d_x0, const float *d_x1,
                  const float *d_offset, const float *d_ghat, float max_t,
                  float ccd_eps, float *d_toi, int32_t *d_overlap, void *stream) {
    if (kind < 0 || kind > 3 || (n && (!d_x0 || !d_x1 || !d_offset || !d_ghat || !d_toi)) ||
        !(max_t > 0.0f)) {
        set_error("ppf_ccd_batch: invalid argument");
        return PPF_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    int rc = ccd_batch(kind, n, d_x0, d_x1, d_offset, d_ghat, max_t, ccd_eps, d_toi,
                       d_overlap, st);
    if (rc == PPF_OK) {
        PPF_HIP(hipStreamSynchronize(st));
    }
    return rc;
}

int ppf_elastic_create(const ppf_shell *topology, void *stream, ppf_elastic **out) {
    if (!topology || !out || (topology->nf && !topology->face) ||
        (topology->nh && !topology->hinge)) {
        set_error("ppf_elastic_create: invalid argument");
        return PPF_ERR_ARG;
    }
    ppf_elastic *h = new ppf_elastic();
    int rc = h->impl.create(*topology, (hipStream_t)stream);
    if (rc != PPF_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return PPF_OK;
}

void ppf_elastic_destroy(ppf_elastic *h) { delete h; }

int ppf_elastic_pattern(ppf_elastic *h, const uint32_t **d_row_offset,
                        const uint32_t **d_col_index, uint32_t *nnz) {
    if (!h) {
        set_error("ppf_elastic_pattern: null handle");
        return PPF_ERR_ARG;
    }
    if (d_row_offset) *d_row_offset = h->impl.row_offset();
    if (d_col_index) *d_col_index = h->impl.col_index();
    if (nnz) *nnz = h->impl.nnz();
    return PPF_OK;
}

int ppf_elastic_assemble(ppf_elastic *h, const ppf_shell *sh, const ppf_elastic_params *p,
                         float *d_force, float *d_cdiag, float *d_value, void *stream) {
    if (!h || !sh || !p || !d_force || !d_cdiag || (h->impl.nnz() && !d_value) || !sh->x ||
        !sh->x0 || !sh->target || !sh->vmass || !sh->varea || !sh->vflag || !sh->v2f_off ||
        (sh->nf && (!sh->fmass || !sh->inv_rest || !sh->f_mu || !sh->f_lambda || !sh->fflag)) ||
        (sh->nh && (!sh->h_rest || !sh->h_length || !sh->h_area || !sh->h_bend || !sh->hflag)) ||
        !(p->dt > 0.0f)) {
        set_error("ppf_elastic_assemble: invalid argument");
        return PPF_ERR_ARG;
    }
    if (sh->nv != h->impl.nv() || sh->nf != h->impl.nf() || sh->nh != h->impl.nh()) {
        set_error("ppf_elastic_assemble: shell sizes differ from the pattern's");
        return PPF_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    int rc = h->impl.assemble(*sh, *p, d_force, d_cdiag, d_value, st);
    if (rc == PPF_OK) {
        PPF_HIP(hipStreamSynchronize(st));
    }
    return rc;
}

static bool strain_limit_shell_ok(const ppf_elastic *h, const ppf_shell *sh) {
    return sh->nv == h->impl.nv() && sh->nf == h->impl.nf() &&
           (!sh->nf || (sh->face && sh->inv_rest && sh->fflag));
}

int ppf_elastic_strain_limit(ppf_elastic *h, const ppf_shell *sh, float eiganalysis_eps,
                             float *d_force, float *d_value, void *stream) {
    if (!h || !sh || !d_force || (h->impl.nnz() && !d_value) || !sh->x || !sh->x0 ||
        (sh->nf && !sh->fmass) || !strain_limit_shell_ok(h, sh) || !(eiganalysis_eps >= 0.0f)) {
        set_error("ppf_elastic_strain_limit: invalid argument");
        return PPF_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    int rc = h->impl.strain_limit(*sh, eiganalysis_eps, d_force, d_value, st);
    if (rc == PPF_OK) {
        PPF_HIP(hipStreamSynchronize(st));
    }
    return rc;
}

int ppf_elastic_plasticity(ppf_elastic *h, const ppf_shell *sh, float dt, float *d_inv_rest,
                           float *d_h_rest, void *stream) {
    if (!h || !sh || !sh->x || !(dt > 0.0f) || sh->nv != h->impl.nv() ||
        sh->nf != h->impl.nf() || sh->nh != h->impl.nh() ||
        (sh->f_plasticity && sh->nf && (!sh->f_plastic_threshold || !sh->face || !d_inv_rest)) ||
        (sh->h_plasticity && sh->nh &&
         (!sh->h_plastic_threshold || !sh->hinge || !sh->hflag || !d_h_rest))) {
        set_error("ppf_elastic_plasticity: invalid argument");
        return PPF_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    int rc = h->impl.plasticity(*sh, dt, nullptr, nullptr, d_inv_rest, d_h_rest, st);
    if (rc == PPF_OK) {
        PPF_HIP(hipStreamSynchronize(st));
    }
    return rc;
}

int ppf_elastic_strain_limit_toi(ppf_elastic *h, const ppf_shell *sh, const float *d_x0,
                                 const float *d_x1, float line_search_max_t, float *toi,
                                 void *stream) {
    if (!h || !sh || !d_x0 || !d_x1 || !toi || !strain_limit_shell_ok(h, sh) ||
        !(line_search_max_t > 0.0f)) {
        set_error("ppf_elastic_strain_limit_toi: invalid argument");
        return PPF_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    uint32_t *word = nullptr;
    int rc = h->impl.toi_word(&word);
    if (rc != PPF_OK) {
        return rc;
    }
    const uint32_t init = __builtin_bit_cast(uint32_t, line_search_max_t);
    PPF_HIP(hipMemcpyAsync(word, &init, sizeof init, hipMemcpyHostToDevice, st));
    rc = h->impl.strain_limit_toi(*sh, d_x0, d_x1, line_search_max_t, nullptr, word, st);
    if (rc != PPF_OK) {
        return rc;
    }
    uint32_t bits = init;
    PPF_HIP(hipMemcpyAsync(&bits, word, sizeof bits, hipMemcpyDeviceToHost, st));
    PPF_HIP(hipStreamSynchronize(st));
    *toi = __builtin_bit_cast(float, bits) / line_search_max_t;
    return PPF_OK;
}

int ppf_stepper_create(const ppf_scene *scene, void *stream, ppf_stepper **out) {
    if (!scene || !out || !scene->mesh.x || !scene->prev || !scene->emass ||
        !scene->v_fric || !scene->e_fric || !scene->f_fric) {
        set_error("ppf_stepper_create: invalid argument");
        return PPF_ERR_ARG;
    }
    ppf_stepper *h = new ppf_stepper();
    int rc = h->impl.create(*scene, (hipStream_t)stream);
    if (rc != PPF_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return PPF_OK;
}

void ppf_stepper_destroy(ppf_stepper *h) { delete h; }

int ppf_stepper_advance(ppf_stepper *h, const ppf_scene *scene, const ppf_step_params *p,
                        ppf_step_result *out, void *stream) {
    if (!h || !scene || !p || !out || !(p->dt > 0.0f) || !(p->line_search_max_t > 0.0f) ||
        p->cg_max_iter == 0) {
        set_error("ppf_stepper_advance: invalid argument");
        return PPF_ERR_ARG;
    }
    return h->impl.advance(*scene, *p, out, (hipStream_t)stream);
}

int ppf_stepper_advance_scenes(ppf_stepper *h, const ppf_scene *scene,
                               const ppf_step_params *p, const float *prev_dt,
                               ppf_step_result *out, void *stream) {
    if (!h || !scene || !p || !prev_dt || !out || !(p->dt > 0.0f) ||
        !(p->line_search_max_t > 0.0f) || p->cg_max_iter == 0) {
        set_error("ppf_stepper_advance_scenes: invalid argument");
        return PPF_ERR_ARG;
    }
    return h->impl.advance_scenes(*scene, *p, prev_dt, out, (hipStream_t)stream);
}

int ppf_stepper_set_preconditioner(ppf_stepper *s, ppf_lpc *h) {
    if (!s) {
        set_error("ppf_stepper_set_preconditioner: stepper is null");
        return PPF_ERR_ARG;
    }
    s->impl.set_lpc(h ? &h->impl : nullptr);
    return PPF_OK;
}

int ppf_stepper_update_pins(ppf_stepper *h, const ppf_scene *scene, void *stream) {
    if (!h || !scene) {
        set_error("ppf_stepper_update_pins: invalid argument");
        return PPF_ERR_ARG;
    }
    return h->impl.refresh_pins(*scene, (hipStream_t)stream);
}

int ppf_stepper_system(ppf_stepper *h, uint32_t *n, uint32_t *nnz, uint32_t *d_row_offset,
                       uint32_t *d_col_index, float *d_value, float *d_cdiag, float *d_rhs,
                       void *stream) {
    if (!h || !n || !nnz) {
        set_error("ppf_stepper_system: invalid argument");
        return PPF_ERR_ARG;
    }
    const uint32_t *rp, *ci;
    const float *val, *cdiag, *rhs;
    h->impl.system(nnz, &rp, &ci, &val, &cdiag, &rhs);
    *n = h->impl.nv();
    if (*nnz == 0) {
        set_error("ppf_stepper_system: no system solved yet");
        return PPF_ERR_ARG;
    }
    const hipStream_t s = (hipStream_t)stream;
    const size_t nr = *n, nz = *nnz;
    auto cp = [&](void *dst, const void *src, size_t bytes) -> int {
        if (dst) {
            PPF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
        }
        return PPF_OK;
    };
    if (cp(d_row_offset, rp, (nr + 1) * 4) || cp(d_col_index, ci, nz * 4) ||
        cp(d_value, val, 9 * nz * 4) || cp(d_cdiag, cdiag, 9 * nr * 4) ||
        cp(d_rhs, rhs, 3 * nr * 4)) {
        return PPF_ERR_HIP;
    }
    return PPF_OK;
}

int ppf_stepper_solve_order(ppf_stepper *h, uint32_t *d_order, void *stream) {
    if (!h || !d_order || !h->impl.solve_order()) {
        set_error("ppf_stepper_solve_order: invalid argument");
        return PPF_ERR_ARG;
    }
    PPF_HIP(hipMemcpyAsync(d_order, h->impl.solve_order(), h->impl.nv() * sizeof(uint32_t),
                           hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return PPF_OK;
}

uint32_t ppf_stepper_intersection_records(ppf_stepper *h, ppf_isect_record *out,
                                          uint32_t cap) {
    return h ? h->impl.records(out, cap) : 0;
}

int ppf_stepper_contact_pattern(ppf_stepper *h, uint32_t *nnz, uint32_t *d_row_offset,
                                uint32_t *d_col_index, float *d_value, void *stream) {
    if (!h || !nnz) {
        set_error("ppf_stepper_contact_pattern: invalid argument");
        return PPF_ERR_ARG;
    }
    const ContactAssembler &c = h->impl.contact();
    const hipStream_t s = (hipStream_t)stream;
    const uint64_t k = c.row_offset() ? c.nnz() : 0;
    *nnz = (uint32_t)k;
    const uint32_t nv = h->impl.nv();
    if (d_row_offset) {
        if (c.row_offset()) {
            PPF_HIP(hipMemcpyAsync(d_row_offset, c.row_offset(), (nv + 1) * sizeof(uint32_t),
                                   hipMemcpyDeviceToDevice, s));
        } else {
            PPF_HIP(hipMemsetAsync(d_row_offset, 0, (nv + 1) * sizeof(uint32_t), s));
        }
    }
    if (k && d_col_index) {
        PPF_HIP(hipMemcpyAsync(d_col_index, c.cols(), k * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, s));
    }
    if (k && d_value) {
        PPF_HIP(hipMemcpyAsync(d_value, c.values(), 9 * k * sizeof(float),
                               hipMemcpyDeviceToDevice, s));
    }
    return PPF_OK;
}

} // extern "C"
