// ensemble.hip -- many independent IMPEC problems of different sizes advanced
// together.  Every pass of an IMPEC step (Impec2DSolver.cpp:339-388) is one
// batched launch over all members; the pressure solve's loop state lives per
// member on the device, so a step is four replays of captured graphs (set-up,
// a block of CG iterations repeated until no member iterates, the end of the
// solve, velocities and transport) with one small host read after each
// iteration block.  Members are ordinary contexts on the ensemble's stream:
// set_* / get_* / diagnostics calls on them work as on any context, and each
// member's fields after a step are bitwise those its own impec_step computes.
#include "impec_device.h"
#include <math.h>
#include <new>
#include <stdarg.h>
#include <stdlib.h>
#include <vector>

namespace impec {

struct EnsSnap {
    const void *K11, *K22, *src_idx, *src_q;
    int nsrc, mg_rows;
    double bnorm, bsum, q_pin;
    int symmetric;
};

static EnsSnap snap_of(const impec_ctx *c) {
    EnsSnap s;
    s.K11 = c->K11; s.K22 = c->K22; s.src_idx = c->src_idx; s.src_q = c->src_q;
    s.nsrc = c->nsrc; s.mg_rows = c->P.mg_rows;
    s.bnorm = c->bnorm; s.bsum = c->bsum; s.q_pin = c->q_pin;
    s.symmetric = c->symmetric;
    return s;
}

static bool snap_equal(const EnsSnap &a, const EnsSnap &b) {
    return a.K11 == b.K11 && a.K22 == b.K22 && a.src_idx == b.src_idx && a.src_q == b.src_q && a.nsrc == b.nsrc &&
           a.mg_rows == b.mg_rows && a.bnorm == b.bnorm && a.bsum == b.bsum && a.q_pin == b.q_pin &&
           a.symmetric == b.symmetric;
}

#define ENS_CHECK(e, call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ens_set_error((e), IMPEC_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,   \
                                 hipGetErrorString(e_));                                       \
    } while (0)

// members whose solve is still running
__global__ void kb_count_active(double *const *__restrict__ S, int n, int *__restrict__ out) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int mine = 0;
    for (int m = threadIdx.x; m < n; m += blockDim.x) mine += S[m][S_EST] != (double)EST_DONE;
    if (mine) atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) out[0] = cnt;
}

// per member: iterations, status, ||b - A p||^2, time step, || |b| + |A| |p| ||^2
constexpr int GATHER = 5;
__global__ void kb_gather(double *const *__restrict__ S, int n, double *__restrict__ out) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const double *s = S[m];
    out[GATHER * m + 0] = s[S_EITS];
    out[GATHER * m + 1] = s[S_ESTATUS];
    out[GATHER * m + 2] = s[S_TT];
    out[GATHER * m + 3] = s[S_DT];
    out[GATHER * m + 4] = s[S_FLOOR];
}

static void drop_graphs(impec_ensemble *e) {
    hipGraphExec_t *g[4] = {&e->g_pre, &e->g_loop, &e->g_end, &e->g_post};
    for (int k = 0; k < 4; k++)
        if (*g[k]) {
            (void)hipGraphExecDestroy(*g[k]);
            *g[k] = nullptr;
        }
    e->have_graphs = 0;
}

static int stage_pre(impec_ensemble *e) {
    int st;
    if ((st = ens_fill_c(e)) || (st = ens_mobility(e)) || (st = ens_solve_begin(e))) return st;
    return IMPEC_OK;
}

static int stage_loop(impec_ensemble *e) {
    for (int k = 0; k < e->slots; k++) {
        int st = ens_cg_slot(e);
        if (st) return st;
    }
    hipLaunchKernelGGL(kb_count_active, dim3(1), dim3(256), 0, e->stream, e->d_S, e->n, e->d_active);
    ENS_CHECK(e, hipGetLastError());
    return IMPEC_OK;
}

static int stage_end(impec_ensemble *e) {
    int st = ens_solve_end(e);
    if (st) return st;
    hipLaunchKernelGGL(kb_gather, dim3((e->n + 63) / 64), dim3(64), 0, e->stream, e->d_S, e->n, e->d_out);
    ENS_CHECK(e, hipGetLastError());
    return IMPEC_OK;
}

// velocities, time step and impec_hyperbolic_solve without dispersion
static int stage_post(impec_ensemble *e) {
    int st;
    if ((st = ens_velocity_dt(e)) || (st = ens_tr(e)) || (st = ens_fill_c(e)) || (st = ens_advect_parallel(e)) ||
        (st = ens_tr(e)) || (st = ens_fill_c(e)))
        return st;
    hipLaunchKernelGGL(kb_gather, dim3((e->n + 63) / 64), dim3(64), 0, e->stream, e->d_S, e->n, e->d_out);
    ENS_CHECK(e, hipGetLastError());
    return IMPEC_OK;
}

// Every node of a captured stage must be a kernel launch with a non-empty
// grid, a workgroup of 1 to 1024 threads and at most 64 KiB of dynamic LDS:
// the stages record no copies, fills or host nodes.  Checked once per
// capture, before instantiation.
static int check_graph(impec_ensemble *e, hipGraph_t g, const char *name) {
    size_t n = 0;
    ENS_CHECK(e, hipGraphGetNodes(g, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    if (n) ENS_CHECK(e, hipGraphGetNodes(g, nodes.data(), &n));
    for (size_t k = 0; k < n; k++) {
        hipGraphNodeType type;
        ENS_CHECK(e, hipGraphNodeGetType(nodes[k], &type));
        if (type != hipGraphNodeTypeKernel)
            return ens_set_error(e, IMPEC_EHIP, "graph %s: node %zu has type %d, not a kernel launch", name, k,
                                 (int)type);
        hipKernelNodeParams kp;
        ENS_CHECK(e, hipGraphKernelNodeGetParams(nodes[k], &kp));
        const long long grid = (long long)kp.gridDim.x * kp.gridDim.y * kp.gridDim.z;
        const long long block = (long long)kp.blockDim.x * kp.blockDim.y * kp.blockDim.z;
        if (grid < 1 || block < 1 || block > 1024 || kp.sharedMemBytes > 65536 || !kp.func)
            return ens_set_error(e, IMPEC_EHIP, "graph %s: kernel node %zu with grid %lld, block %lld, LDS %u",
                                 name, k, grid, block, (unsigned)kp.sharedMemBytes);
    }
    e->graph_nodes += (long long)n;
    return IMPEC_OK;
}

static int capture(impec_ensemble *e, int (*stage)(impec_ensemble *), hipGraphExec_t *out, const char *name) {
    hipGraph_t g = nullptr;
    ENS_CHECK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int st = stage(e);
    const hipError_t he = hipStreamEndCapture(e->stream, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    ENS_CHECK(e, he);
    const int sc = check_graph(e, g, name);
    if (sc) {
        (void)hipGraphDestroy(g);
        return sc;
    }
    const hipError_t hi = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    ENS_CHECK(e, hi);
    return IMPEC_OK;
}

static int run(impec_ensemble *e, int (*stage)(impec_ensemble *), hipGraphExec_t g) {
    if (!e->use_graphs) return stage(e);
    ENS_CHECK(e, hipGraphLaunch(g, e->stream));
    return IMPEC_OK;
}

// (re)builds the launch tables when a member's inputs changed since the last
// build (sources, permeability, chunk heights)
static int prepare(impec_ensemble *e) {
    bool same = e->ready;
    for (int k = 0; k < e->n && same; k++) same = snap_equal(e->snap[k], snap_of(e->m[k]));
    if (same) return IMPEC_OK;
    drop_graphs(e);
    e->ready = 0;
    long long maxit = 1;
    for (int k = 0; k < e->n; k++) {
        impec_ctx *c = e->m[k];
        if (!c->symmetric)
            return ens_set_error(e, IMPEC_EINVAL, "member %d: the ensemble needs the symmetric 5-point operator", k);
        if (c->K22 && c->K22 != c->K11)
            return ens_set_error(e, IMPEC_EINVAL, "member %d: the ensemble needs isotropic permeability", k);
        if (c->nsrc == 0 || c->bnorm == 0.)
            return ens_set_error(e, IMPEC_EINVAL, "member %d: no source set", k);
        const long long mi = c->P.max_iterations > 0 ? c->P.max_iterations : 10000;
        if (mi > maxit) maxit = mi;
    }
    // every member stops within its iteration limit plus 4 residual rechecks
    e->max_loop = (maxit + 8) / e->slots + 2;
    int st;
    if ((st = ens_field_build(e)) || (st = ens_press_build(e)) || (st = ens_trans_build(e))) return st;
    for (int k = 0; k < e->n; k++) e->snap[k] = snap_of(e->m[k]);
    e->ready = 1;
    return IMPEC_OK;
}

}  // namespace impec

using namespace impec;

extern "C" {

int impec_ensemble_create(const impec_params *members, int count, impec_ensemble **out) {
    if (!members || !out || count < 1) return IMPEC_EINVAL;
    *out = nullptr;
    const impec_params &P0 = members[0];
    for (int k = 0; k < count; k++) {
        const impec_params &P = members[k];
        if (P.nu != 1. || P.diffusion || P.is_test || P.dim_splitting != IMPEC_SPLIT_PARALLEL) return IMPEC_EINVAL;
        if (P.initial_guess != IMPEC_GUESS_ZERO && P.initial_guess != IMPEC_GUESS_PREVIOUS) return IMPEC_EINVAL;
        if (P.precision == IMPEC_PREC_MIXED32 || P.transport32) return IMPEC_EINVAL;   // single contexts only
        if (P.precision != P0.precision || P.hyper_method != P0.hyper_method ||
            (P.porosity == 1.) != (P0.porosity == 1.))
            return IMPEC_EINVAL;
        // one or two red-black pairs per leg, the same for every member
        const bool one = P.mg_pre <= 1 && P.mg_post <= 1, two = P.mg_pre == 2 && P.mg_post == 2;
        const bool one0 = P0.mg_pre <= 1 && P0.mg_post <= 1;
        if (!(one || two) || one != one0 || P.mg_variant != 0) return IMPEC_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IMPEC_ENODEV;
    impec_ensemble *e = new (std::nothrow) impec_ensemble();
    if (!e) return IMPEC_ENOMEM;
    e->n = count;
    e->slots = 4;
    e->use_graphs = 1;
    e->m = (impec_ctx **)calloc(count, sizeof(impec_ctx *));
    e->snap = new (std::nothrow) EnsSnap[count];
    int st = (e->m && e->snap) ? IMPEC_OK : IMPEC_ENOMEM;
    if (!st && hipStreamCreate(&e->stream) != hipSuccess) st = IMPEC_EHIP;
    for (int k = 0; k < count && !st; k++) {
        st = ctx_create(&members[k], e->stream, &e->m[k]);
        if (st) ens_set_error(e, st, "member %d: %s", k, impec_strerror(st));
        else e->m[k]->fixed_buffers = 1;   // the batched records hold its c and t1
    }
    if (!st) {
        std::vector<double *> S(count);
        for (int k = 0; k < count; k++) S[k] = e->m[k]->scal;
        if (hipMalloc(&e->d_S, sizeof(double *) * count) != hipSuccess ||
            hipMemcpy(e->d_S, S.data(), sizeof(double *) * count, hipMemcpyHostToDevice) != hipSuccess ||
            hipMalloc(&e->d_out, sizeof(double) * GATHER * count) != hipSuccess ||
            hipMalloc(&e->d_active, sizeof(int)) != hipSuccess ||
            hipHostMalloc((void **)&e->h_out, sizeof(double) * GATHER * count, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&e->h_active, sizeof(int), hipHostMallocDefault) != hipSuccess)
            st = IMPEC_ENOMEM;
    }
    if (st) {
        impec_ensemble_destroy(e);
        return st;
    }
    *out = e;
    return IMPEC_OK;
}

int impec_ensemble_destroy(impec_ensemble *e) {
    if (!e) return IMPEC_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    ens_field_free(e);
    ens_press_free(e);
    ens_trans_free(e);
    if (e->m)
        for (int k = 0; k < e->n; k++) impec_destroy(e->m[k]);
    free(e->m);
    delete[] e->snap;
    if (e->d_S) (void)hipFree(e->d_S);
    if (e->d_out) (void)hipFree(e->d_out);
    if (e->d_active) (void)hipFree(e->d_active);
    if (e->h_out) (void)hipHostFree(e->h_out);
    if (e->h_active) (void)hipHostFree(e->h_active);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return IMPEC_OK;
}

const char *impec_ensemble_last_error(const impec_ensemble *e) { return e ? e->err : "null ensemble"; }

int impec_ensemble_member(impec_ensemble *e, int k, impec_ctx **ctx) {
    if (!e || !ctx || k < 0 || k >= e->n) return IMPEC_EINVAL;
    *ctx = e->m[k];
    return IMPEC_OK;
}

int impec_ensemble_set_graphs(impec_ensemble *e, int on, int slots) {
    if (!e || slots < 0 || slots > 64) return IMPEC_EINVAL;
    drop_graphs(e);
    e->use_graphs = on != 0;
    if (slots > 0) e->slots = slots;
    e->ready = 0;
    return IMPEC_OK;
}

int impec_ensemble_autotune(impec_ensemble *e, int reps, int *rows) {
    if (!e || reps < 1) return IMPEC_EINVAL;
    int st = prepare(e);
    if (st) return st;
    static const int cand[] = {8, 16, 32, 64};
    // members grouped by grid size; coordinate descent over the groups
    std::vector<int> done(e->n, 0);
    for (int k = 0; k < e->n; k++) {
        if (done[k]) continue;
        std::vector<int> grp;
        for (int q = k; q < e->n; q++)
            if (!done[q] && e->m[q]->N1 == e->m[k]->N1 && e->m[q]->N2 == e->m[k]->N2) {
                grp.push_back(q);
                done[q] = 1;
            }
        int best = e->m[k]->P.mg_rows;
        float best_ms = 1e30f;
        for (int c : cand) {
            for (int q : grp) e->m[q]->P.mg_rows = c;
            if ((st = ens_press_build(e))) return st;
            float ms = 0.f;
            if ((st = ens_vcycle_time(e, reps, &ms))) return st;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        for (int q : grp) e->m[q]->P.mg_rows = best;
    }
    e->ready = 0;   // rebuild with the chosen heights
    if ((st = prepare(e))) return st;
    if (rows)
        for (int k = 0; k < e->n; k++) rows[k] = e->m[k]->P.mg_rows;
    return IMPEC_OK;
}

int impec_ensemble_step(impec_ensemble *e, double *dt, int *its) {
    if (!e) return IMPEC_EINVAL;
    int st = prepare(e);
    if (st) return st;
    if (e->use_graphs && !e->have_graphs) {
        e->graph_nodes = 0;
        if ((st = capture(e, stage_pre, &e->g_pre, "set-up")) || (st = capture(e, stage_loop, &e->g_loop, "iterations")) ||
            (st = capture(e, stage_end, &e->g_end, "end of solve")) ||
            (st = capture(e, stage_post, &e->g_post, "transport"))) {
            drop_graphs(e);
            return st;
        }
        e->have_graphs = 1;
    }
    if ((st = run(e, stage_pre, e->g_pre))) return st;
    long long loops = 0;
    for (;;) {
        if ((st = run(e, stage_loop, e->g_loop))) return st;
        ENS_CHECK(e, hipMemcpyAsync(e->h_active, e->d_active, sizeof(int), hipMemcpyDeviceToHost, e->stream));
        ENS_CHECK(e, hipStreamSynchronize(e->stream));
        if (e->h_active[0] == 0) break;
        if (++loops > e->max_loop) return ens_set_error(e, IMPEC_ENOCONV, "ensemble pressure solve did not finish");
    }
    if ((st = run(e, stage_end, e->g_end))) return st;
    ENS_CHECK(e, hipMemcpyAsync(e->h_out, e->d_out, sizeof(double) * GATHER * e->n, hipMemcpyDeviceToHost,
                                e->stream));
    ENS_CHECK(e, hipStreamSynchronize(e->stream));
    for (int k = 0; k < e->n; k++) {
        impec_ctx *c = e->m[k];
        int mits = (int)e->h_out[GATHER * k + 0];
        const int status = (int)e->h_out[GATHER * k + 1];
        double relres = c->bnorm > 0 ? sqrt(e->h_out[GATHER * k + 2]) / c->bnorm : 0.;
        double floor_rel = c->bnorm > 0 ? IMPEC_FLOOR_C * 0x1p-53 * sqrt(e->h_out[GATHER * k + 4]) / c->bnorm : 0.;
        if (status == 1)
            return ens_set_error(e, IMPEC_ENOCONV, "member %d: no convergence after %d iterations", k, mits);
        if (status == 2) return ens_set_error(e, IMPEC_ENAN, "member %d: non-finite residual", k);
        if (relres > c->P.eps && relres > floor_rel) {
            // the shifted iterate misses the pinned system's tolerance: the
            // member's own solve (consistent phase, then the eliminated one)
            int its1 = 0;
            c->have_prev_p = 0;
            if ((st = pressure_solve(c, &its1, &relres)))
                return ens_set_error(e, st, "member %d: %s", k, c->err);
            mits = its1;
            floor_rel = c->stats.last_relres_floor;
            c->stats.total_its -= its1;   // counted once below
            c->stats.calls -= 1;
        }
        c->have_prev_p = 1;
        c->stats.last_its = mits;
        c->stats.last_vcycles = mits;
        c->stats.last_relres = relres;
        c->stats.last_relres_floor = floor_rel;
        c->stats.total_its += mits;
        c->stats.calls += 1;
        if (its) its[k] = mits;
    }
    if ((st = run(e, stage_post, e->g_post))) return st;
    if (dt) {
        ENS_CHECK(e, hipMemcpyAsync(e->h_out, e->d_out, sizeof(double) * GATHER * e->n, hipMemcpyDeviceToHost,
                                    e->stream));
        ENS_CHECK(e, hipStreamSynchronize(e->stream));
        for (int k = 0; k < e->n; k++) dt[k] = e->h_out[GATHER * k + 3];
    }
    return IMPEC_OK;
}

}  // extern "C"
