// capi.hip -- the extern "C" boundary of libimpec_hip.so (include/impec_hip.h).
#include "impec_internal.h"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

namespace impec {

// plain copy used as the reference point of HBM byte counters
__global__ void k_calib_copy(const double *__restrict__ a, double *__restrict__ b, long n) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) b[k] = a[k];
}

int set_error(impec_ctx *ctx, int code, const char *fmt, ...) {
    if (ctx) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

void prof_begin(impec_ctx *ctx) {
    if (ctx->prof_on) hipEventRecord(ctx->ev0, ctx->stream);
}

void prof_end(impec_ctx *ctx, const char *name) {
    if (!ctx->prof_on) return;
    hipEventRecord(ctx->ev1, ctx->stream);
    hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    for (int k = 0; k < ctx->nprof; k++)
        if (strcmp(ctx->prof[k].name, name) == 0) {
            ctx->prof[k].launches++;
            ctx->prof[k].ms += ms;
            return;
        }
    if (ctx->nprof < 32) {
        ctx->prof[ctx->nprof].name = name;
        ctx->prof[ctx->nprof].launches = 1;
        ctx->prof[ctx->nprof].ms = ms;
        ctx->nprof++;
    }
}

static int round_up(int x, int m) { return (x + m - 1) / m * m; }

static int dalloc(impec_ctx *ctx, double **p, size_t n) {
    if (hipMalloc((void **)p, n * sizeof(double)) != hipSuccess) {
        *p = nullptr;
        return set_error(ctx, IMPEC_ENOMEM, "device allocation of %zu doubles failed", n);
    }
    if (hipMemset(*p, 0, n * sizeof(double)) != hipSuccess) return set_error(ctx, IMPEC_EHIP, "hipMemset failed");
    return IMPEC_OK;
}

static size_t scalar_elems(const impec_ctx *ctx) { return (size_t)(ctx->N1 + 2 * ctx->B) * ctx->lds; }
static size_t vel_elems(const impec_ctx *ctx) { return (size_t)(ctx->N1 + 1 + 2 * ctx->B) * ctx->ldv; }

static int ensure_cn(impec_ctx *ctx) {
    if (ctx->cn_rhs) return IMPEC_OK;
    const size_t n = (size_t)ctx->N1 * ctx->N2;
    int st;
    double **v[8] = {&ctx->cn_rhs, &ctx->cn_r, &ctx->cn_r0, &ctx->cn_p, &ctx->cn_v,
                     &ctx->cn_r2, &ctx->cn_y, &ctx->cn_z};
    for (int k = 0; k < 8; k++)
        if ((st = dalloc(ctx, v[k], n))) return st;
    if (hipMalloc((void **)&ctx->cn_dg, n * sizeof(float)) != hipSuccess) {
        ctx->cn_dg = nullptr;
        return set_error(ctx, IMPEC_ENOMEM, "device allocation of %zu floats failed", n);
    }
    return IMPEC_OK;
}

// symmetric folded Neumann matrix => shift formulation of the pinned solve
static void update_symmetry(impec_ctx *ctx) {
    const bool aniso = ctx->K22 && ctx->K22 != ctx->K11;
    ctx->symmetric = ctx->P.nu == 1. || (ctx->dx == ctx->dy && !aniso);
}

}  // namespace impec

using namespace impec;

extern "C" {

const char *impec_strerror(int code) {
    switch (code) {
    case IMPEC_OK: return "ok";
    case IMPEC_EINVAL: return "invalid argument";
    case IMPEC_EHIP: return "HIP runtime error";
    case IMPEC_ENOCONV: return "linear solver did not converge";
    case IMPEC_ENAN: return "non-finite value";
    case IMPEC_ENOMEM: return "out of device memory";
    case IMPEC_ENODEV: return "no GPU";
    default: return "unknown error";
    }
}

// NULL: the message of the last context-free call on this thread
const char *impec_last_error(const impec_ctx *ctx) { return ctx ? ctx->err : noalloc_last_error(); }

int impec_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int impec_params_default(impec_params *p) {
    if (!p) return IMPEC_EINVAL;
    memset(p, 0, sizeof(*p));
    // examples/parallel.cfg and examples/diagonal_dispersion.cfg
    p->N1 = 64; p->N2 = 64; p->num_ghostcells = 2;
    p->source_type = IMPEC_SRC_DIAGONAL;
    p->average_procedure = IMPEC_AVG_HARMONIC;
    p->dim_splitting = IMPEC_SPLIT_PARALLEL;
    p->hyper_method = IMPEC_HYPER_MUSCL;
    p->diffusion = 0;
    p->para_method = IMPEC_PARA_CN;
    p->is_test = 0;
    p->alpha_absolute = 0;
    p->max_iterations = 10000;   // MGSolver.cpp:89
    p->precision = IMPEC_PREC_FP64;
    p->initial_guess = IMPEC_GUESS_ZERO;
    p->mg_pre = 2; p->mg_post = 2;
    p->x_lb = 0.; p->x_ub = 1.; p->y_lb = 0.; p->y_ub = 1.;
    p->nu = 2. / 3.;
    p->mu_res = 41.; p->mu_inj = 1.;
    p->porosity = 1.; p->cfl = 0.5; p->inj_fluid_conc = 1.;
    p->alpha_l = 1.; p->alpha_t = 1.; p->dt_diff_limiter = 0.9;
    p->d_m = 0.;
    p->eps = 1e-14; p->para_eps = 1e-14;
    p->I = 1.; p->tstart = 0.;
    p->inner_tol = 1e-3;
    return IMPEC_OK;
}

int impec_create(const impec_params *P, impec_ctx **out) { return impec::ctx_create(P, nullptr, out); }

}  // extern "C"

// a context on its own stream, or on `ext` (not owned) for ensemble members
int impec::ctx_create(const impec_params *P, hipStream_t ext, impec_ctx **out) {
    if (!P || !out) return IMPEC_EINVAL;
    *out = nullptr;
    if (P->N1 < 2 || P->N2 < 2) return IMPEC_EINVAL;
    if ((long long)P->N1 * P->N2 >= (1LL << 31)) return IMPEC_EINVAL;   // cell indices are 32-bit
    if (P->num_ghostcells != GHOST) return IMPEC_EINVAL;
    if (P->source_type != IMPEC_SRC_PARALLEL && P->source_type != IMPEC_SRC_DIAGONAL) return IMPEC_EINVAL;
    if (!(P->nu >= 0. && P->nu <= 1.)) return IMPEC_EINVAL;   // EllipticSolver::SetNu
    if (!(P->x_ub > P->x_lb && P->y_ub > P->y_lb)) return IMPEC_EINVAL;
    if (P->dim_splitting < 0 || P->dim_splitting > 4) return IMPEC_EINVAL;
    if (P->precision != IMPEC_PREC_FP64 && P->precision != IMPEC_PREC_MIXED && P->precision != IMPEC_PREC_MIXED32)
        return IMPEC_EINVAL;
    if (P->mg_rows < 0 || P->mg_rows > 4096) return IMPEC_EINVAL;
    if (P->mg_variant < 0 || P->mg_variant > 2) return IMPEC_EINVAL;
    if (P->cn_guess != 0 && P->cn_guess != 1) return IMPEC_EINVAL;
    if (P->cn_ahead != P->cn_ahead || P->cg_ahead != P->cg_ahead) return IMPEC_EINVAL;   // NaN
    // FP32 transport: the one-pass parallel splitting only
    if (P->transport32 != 0 && (P->transport32 != 1 || P->dim_splitting != IMPEC_SPLIT_PARALLEL)) return IMPEC_EINVAL;
    if (P->initial_guess < IMPEC_GUESS_ZERO || P->initial_guess > IMPEC_GUESS_EXTRAPOLATE) return IMPEC_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return IMPEC_ENODEV;

    impec_ctx *ctx = (impec_ctx *)calloc(1, sizeof(impec_ctx));
    if (!ctx) return IMPEC_ENOMEM;
    ctx->P = *P;
    ctx->N1 = P->N1; ctx->N2 = P->N2; ctx->B = GHOST;
    hipGetDevice(&ctx->device);
    // SimParams.cpp:449-450 and EllipticSolver.cpp:68-93
    ctx->dx = (P->x_ub - P->x_lb) / P->N1;
    ctx->dy = (P->y_ub - P->y_lb) / P->N2;
    ctx->dA = ctx->dx * ctx->dy;
    ctx->alpha = atan(ctx->dx / ctx->dy);
    ctx->theta = atan(ctx->dy / ctx->dx);
    ctx->dxi = ctx->dx / cos(ctx->theta);
    ctx->deta = ctx->dy / cos(ctx->alpha);
    const double secamt = 1. / cos(ctx->alpha - ctx->theta);
    ctx->secamt2 = secamt * secamt;
    ctx->cosa2 = cos(ctx->alpha) * cos(ctx->alpha);
    ctx->sina2 = sin(ctx->alpha) * sin(ctx->alpha);
    ctx->costsina = cos(ctx->theta) * sin(ctx->alpha);
    ctx->cosasint = cos(ctx->alpha) * sin(ctx->theta);
    ctx->cost2 = cos(ctx->theta) * cos(ctx->theta);
    ctx->sint2 = sin(ctx->theta) * sin(ctx->theta);
    ctx->cosa = cos(ctx->alpha);
    ctx->sint = sin(ctx->theta);
    ctx->cost = cos(ctx->theta);
    ctx->sina = sin(ctx->alpha);
    ctx->mob1over4 = pow(P->mu_res / P->mu_inj, 0.25);
    ctx->kappa = 1.;
    ctx->pin_i = 1; ctx->pin_j = 1;   // MGSolver.cpp:109-118
    ctx->force_unfused = P->mg_variant == 1;
    ctx->cn_ahead = P->cn_ahead == 0. ? IMPEC_CN_AHEAD : P->cn_ahead;
    ctx->cg_ahead = P->cg_ahead == 0. ? IMPEC_CG_AHEAD : P->cg_ahead;
    ctx->prev_first_sweep = 1;        // HyperbolicSolver.cpp:84
    ctx->lds = round_up(ctx->N2 + 2 * ctx->B, 16);
    ctx->ldv = round_up(ctx->N2 + 1 + 2 * ctx->B, 16);

    int st = IMPEC_OK;
    if (ext) ctx->stream = ext;
    else if (hipStreamCreate(&ctx->stream) != hipSuccess) st = set_error(ctx, IMPEC_EHIP, "stream");
    else ctx->own_stream = 1;
    if (!st && (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess))
        st = set_error(ctx, IMPEC_EHIP, "events");
    const size_t ns = scalar_elems(ctx), nv = vel_elems(ctx), nq = (size_t)ctx->N1 * ctx->N2;
    if (!st) st = dalloc(ctx, &ctx->c, ns);
    if (!st) st = dalloc(ctx, &ctx->p, ns);
    if (!st) st = dalloc(ctx, &ctx->w, ns);
    if (!st) st = dalloc(ctx, &ctx->t1, ns);
    if (!st) st = dalloc(ctx, &ctx->t2, ns);
    if (!st) st = dalloc(ctx, &ctx->t3, ns);
    if (!st) st = dalloc(ctx, &ctx->u, nv);
    if (!st) st = dalloc(ctx, &ctx->v, nv);
    if (!st) st = dalloc(ctx, &ctx->q, nq);
    if (!st) st = dalloc(ctx, &ctx->scal, S_COUNT);
    if (!st) st = dalloc(ctx, &ctx->part, (size_t)RED_BLOCKS * PART_VECS);
    if (!st) st = dalloc(ctx, &ctx->part_mid, (size_t)FIN_THREADS * PART_VECS);
    // mapped, coherent host memory: a kernel writes the scalars and then a
    // sequence number (slot S_COUNT) into it, which the host polls
    if (!st && hipHostMalloc((void **)&ctx->h_scal, sizeof(double) * (S_COUNT + 1),
                             hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        st = set_error(ctx, IMPEC_ENOMEM, "pinned host scalars");
    if (!st) {
        memset(ctx->h_scal, 0, sizeof(double) * (S_COUNT + 1));
        if (hipHostGetDevicePointer((void **)&ctx->d_hscal, ctx->h_scal, 0) != hipSuccess)
            st = set_error(ctx, IMPEC_EHIP, "device address of the host scalars");
    }
    if (!st) st = pressure_setup_levels(ctx);
    if (!st && P->initial_guess == IMPEC_GUESS_EXTRAPOLATE) {   // the two previous pressures
        st = dalloc(ctx, &ctx->p_prev, ns);
        if (!st) st = dalloc(ctx, &ctx->p_old, ns);
    }
    if (!st && P->diffusion) st = ensure_cn(ctx);
    if (!st) update_symmetry(ctx);
    if (!st && hipDeviceSynchronize() != hipSuccess) st = set_error(ctx, IMPEC_EHIP, "initial synchronisation");
    if (st) {
        impec_destroy(ctx);
        return st;
    }
    *out = ctx;
    return IMPEC_OK;
}

extern "C" {

int impec_destroy(impec_ctx *ctx) {
    if (!ctx) return IMPEC_OK;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    double *bufs[] = {ctx->c, ctx->p, ctx->u, ctx->v, ctx->q, ctx->K11, ctx->K22 != ctx->K11 ? ctx->K22 : nullptr,
                      ctx->w, ctx->w22, ctx->t1, ctx->t2, ctx->t3, ctx->scal, ctx->part, ctx->part_mid,
                      ctx->cn_rhs, ctx->cn_r, ctx->cn_r0, ctx->cn_p, ctx->cn_v,
                      ctx->cn_r2, ctx->cn_y, ctx->cn_z, ctx->src_q, ctx->tr_md, ctx->p_prev,
                      ctx->p_old};
    for (double *b : bufs)
        if (b) hipFree(b);
    if (ctx->cn_dg) hipFree(ctx->cn_dg);
    if (ctx->src_idx) hipFree(ctx->src_idx);
    if (ctx->nlev) pressure_free_levels(ctx);
    if (ctx->h_scal) hipHostFree(ctx->h_scal);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->stream && ctx->own_stream) hipStreamDestroy(ctx->stream);
    free(ctx);
    return IMPEC_OK;
}

int impec_synchronize(impec_ctx *ctx) {
    if (!ctx) return IMPEC_EINVAL;
    IMPEC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return IMPEC_OK;
}

// ---- transfers -------------------------------------------------------------
static int put2d(impec_ctx *ctx, double *dst, int ld_dst, const double *src, int rows, int cols) {
    IMPEC_HIP_CHECK(ctx, hipMemcpy2DAsync(dst, sizeof(double) * ld_dst, src, sizeof(double) * cols, sizeof(double) * cols,
                                          rows, hipMemcpyHostToDevice, ctx->stream));
    IMPEC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return IMPEC_OK;
}
static int get2d(impec_ctx *ctx, double *dst, const double *src, int ld_src, int rows, int cols) {
    IMPEC_HIP_CHECK(ctx, hipMemcpy2DAsync(dst, sizeof(double) * cols, src, sizeof(double) * ld_src, sizeof(double) * cols,
                                          rows, hipMemcpyDeviceToHost, ctx->stream));
    IMPEC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return IMPEC_OK;
}

int impec_set_concentration(impec_ctx *ctx, const double *c) {
    if (!ctx || !c) return IMPEC_EINVAL;
    return put2d(ctx, ctx->c, ctx->lds, c, ctx->N1 + 2 * ctx->B, ctx->N2 + 2 * ctx->B);
}
int impec_get_concentration(impec_ctx *ctx, double *c) {
    if (!ctx || !c) return IMPEC_EINVAL;
    return get2d(ctx, c, ctx->c, ctx->lds, ctx->N1 + 2 * ctx->B, ctx->N2 + 2 * ctx->B);
}
int impec_set_pressure(impec_ctx *ctx, const double *p) {
    if (!ctx || !p) return IMPEC_EINVAL;
    ctx->have_prev_p = 1;
    ctx->n_solves = 1;   // no earlier pressure to extrapolate from
    return put2d(ctx, ctx->p, ctx->lds, p, ctx->N1 + 2 * ctx->B, ctx->N2 + 2 * ctx->B);
}
int impec_get_pressure(impec_ctx *ctx, double *p) {
    if (!ctx || !p) return IMPEC_EINVAL;
    return get2d(ctx, p, ctx->p, ctx->lds, ctx->N1 + 2 * ctx->B, ctx->N2 + 2 * ctx->B);
}

// interleaved host velocity <-> two device component arrays
int impec_set_velocity(impec_ctx *ctx, const double *v) {
    if (!ctx || !v) return IMPEC_EINVAL;
    const int rows = ctx->N1 + 1 + 2 * ctx->B, cols = ctx->N2 + 1 + 2 * ctx->B;
    std::vector<double> a((size_t)rows * cols), b((size_t)rows * cols);
    for (size_t k = 0; k < a.size(); k++) {
        a[k] = v[2 * k];
        b[k] = v[2 * k + 1];
    }
    int st = put2d(ctx, ctx->u, ctx->ldv, a.data(), rows, cols);
    if (!st) st = put2d(ctx, ctx->v, ctx->ldv, b.data(), rows, cols);
    return st;
}
int impec_get_velocity(impec_ctx *ctx, double *v) {
    if (!ctx || !v) return IMPEC_EINVAL;
    const int rows = ctx->N1 + 1 + 2 * ctx->B, cols = ctx->N2 + 1 + 2 * ctx->B;
    std::vector<double> a((size_t)rows * cols), b((size_t)rows * cols);
    int st = get2d(ctx, a.data(), ctx->u, ctx->ldv, rows, cols);
    if (!st) st = get2d(ctx, b.data(), ctx->v, ctx->ldv, rows, cols);
    if (st) return st;
    for (size_t k = 0; k < a.size(); k++) {
        v[2 * k] = a[k];
        v[2 * k + 1] = b[k];
    }
    return IMPEC_OK;
}

int impec_set_source(impec_ctx *ctx, const double *q) {
    if (!ctx || !q) return IMPEC_EINVAL;
    const size_t n = (size_t)ctx->N1 * ctx->N2;
    IMPEC_HIP_CHECK(ctx, hipMemcpyAsync(ctx->q, q, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    std::vector<int> idx;
    std::vector<double> qv;
    double bb = 0., bs = 0.;
    for (size_t k = 0; k < n; k++) {
        const double b = -q[k];
        bb += b * b;
        bs += b;
        if (q[k] != 0.) {
            idx.push_back((int)k);
            qv.push_back(q[k]);
        }
    }
    ctx->bnorm = sqrt(bb);
    ctx->bsum = bs;
    ctx->q_pin = q[(size_t)ctx->pin_i * ctx->N2 + ctx->pin_j];
    if (ctx->src_idx) hipFree(ctx->src_idx);
    if (ctx->src_q) hipFree(ctx->src_q);
    if (ctx->tr_md) hipFree(ctx->tr_md);
    ctx->src_idx = nullptr; ctx->src_q = nullptr; ctx->tr_md = nullptr;
    ctx->nsrc = (int)idx.size();
    if (ctx->nsrc) {
        IMPEC_HIP_CHECK(ctx, hipMalloc(&ctx->src_idx, sizeof(int) * ctx->nsrc));
        IMPEC_HIP_CHECK(ctx, hipMalloc(&ctx->src_q, sizeof(double) * ctx->nsrc));
        IMPEC_HIP_CHECK(ctx, hipMalloc(&ctx->tr_md, sizeof(double) * ctx->nsrc));
        IMPEC_HIP_CHECK(ctx, hipMemcpyAsync(ctx->src_idx, idx.data(), sizeof(int) * ctx->nsrc, hipMemcpyHostToDevice,
                                            ctx->stream));
        IMPEC_HIP_CHECK(ctx, hipMemcpyAsync(ctx->src_q, qv.data(), sizeof(double) * ctx->nsrc, hipMemcpyHostToDevice,
                                            ctx->stream));
    }
    IMPEC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return IMPEC_OK;
}

int impec_set_permeability(impec_ctx *ctx, const double *K11, const double *K22) {
    if (!ctx) return IMPEC_EINVAL;
    const int rows = ctx->N1 + 2 * ctx->B, cols = ctx->N2 + 2 * ctx->B;
    const size_t ns = scalar_elems(ctx);
    if (ctx->K22 && ctx->K22 != ctx->K11) hipFree(ctx->K22);
    if (ctx->K11) hipFree(ctx->K11);
    ctx->K11 = ctx->K22 = nullptr;
    int st = IMPEC_OK;
    if (K11) {
        if ((st = dalloc(ctx, &ctx->K11, ns))) return st;
        if ((st = put2d(ctx, ctx->K11, ctx->lds, K11, rows, cols))) return st;
        ctx->K22 = ctx->K11;
        if (K22 && K22 != K11) {
            if ((st = dalloc(ctx, &ctx->K22, ns))) return st;
            if ((st = put2d(ctx, ctx->K22, ctx->lds, K22, rows, cols))) return st;
            if (!ctx->w22 && (st = dalloc(ctx, &ctx->w22, ns))) return st;
        }
    }
    update_symmetry(ctx);
    return IMPEC_OK;
}

int impec_set_pin_kappa(impec_ctx *ctx, double kappa) {
    if (!ctx || !(kappa != 0.)) return IMPEC_EINVAL;
    ctx->kappa = kappa;
    return IMPEC_OK;
}

// ---- stages ----------------------------------------------------------------
int impec_fill_boundaries(impec_ctx *ctx) {
    if (!ctx) return IMPEC_EINVAL;
    return launch_fill_scalar(ctx, ctx->c, ctx->P.is_test);
}

int impec_pressure_solve(impec_ctx *ctx, int *its, double *relres) {
    if (!ctx) return IMPEC_EINVAL;
    return pressure_solve(ctx, its, relres);
}

int impec_velocities(impec_ctx *ctx) {
    if (!ctx) return IMPEC_EINVAL;
    int st = launch_mobility(ctx);
    if (st) return st;
    return launch_velocity(ctx);
}

int impec_max_stable_timestep(impec_ctx *ctx, double *dt) {
    if (!ctx || !dt) return IMPEC_EINVAL;
    return launch_cfl(ctx, dt);
}

int impec_source_split(impec_ctx *ctx, double dt) {
    if (!ctx) return IMPEC_EINVAL;
    return launch_tr(ctx, dt);
}

int impec_advect_split(impec_ctx *ctx, double dt) {
    if (!ctx) return IMPEC_EINVAL;
    return advect_split(ctx, dt);
}

int impec_advect_x(impec_ctx *ctx, double dt) {
    if (!ctx) return IMPEC_EINVAL;
    int st = advect_x(ctx, ctx->c, ctx->t1, dt);
    if (!st) st = copy_scalar(ctx, ctx->t1, ctx->c);
    return st;
}

int impec_advect_y(impec_ctx *ctx, double dt) {
    if (!ctx) return IMPEC_EINVAL;
    int st = advect_y(ctx, ctx->c, ctx->t1, dt);
    if (!st) st = copy_scalar(ctx, ctx->t1, ctx->c);
    return st;
}

int impec_intercell_flux(impec_ctx *ctx, int dir, double dt, double *F) {
    if (!ctx || !F || (dir != 0 && dir != 1)) return IMPEC_EINVAL;
    const size_t n = dir == 0 ? (size_t)(ctx->N1 + 1) * (ctx->N2 + 2 * ctx->B)
                              : (size_t)(ctx->N1 + 2 * ctx->B) * (ctx->N2 + 1);
    double *d = nullptr;
    IMPEC_HIP_CHECK(ctx, hipMalloc(&d, n * sizeof(double)));
    int st = intercell_flux(ctx, dir, dt, d);
    hipError_t e = hipSuccess;
    if (!st) e = hipMemcpyAsync(F, d, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (!st && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d);
    if (st) return st;
    if (e != hipSuccess) return set_error(ctx, IMPEC_EHIP, "intercell flux: %s", hipGetErrorString(e));
    return IMPEC_OK;
}

int impec_dispersion_cn(impec_ctx *ctx, double dt, int *its) {
    if (!ctx) return IMPEC_EINVAL;
    int st = ensure_cn(ctx);
    if (st) return st;
    return cn_solve(ctx, dt, its);
}

// HyperbolicSplitSolver::Solve with SPLIT2ND (HyperbolicSplitSolver.cpp:162-198)
int impec_hyperbolic_solve(impec_ctx *ctx, double dt) {
    if (!ctx) return IMPEC_EINVAL;
    int st;
    const int periodic = ctx->P.is_test;
    if ((st = launch_tr(ctx, 0.5 * dt))) return st;
    if ((st = launch_fill_scalar(ctx, ctx->c, periodic))) return st;
    if (ctx->P.diffusion && (st = impec_dispersion_cn(ctx, 0.5 * dt, nullptr))) return st;
    if ((st = advect_split(ctx, dt))) return st;
    if ((st = launch_tr(ctx, 0.5 * dt))) return st;
    if ((st = launch_fill_scalar(ctx, ctx->c, periodic))) return st;
    if (ctx->P.diffusion && (st = impec_dispersion_cn(ctx, 0.5 * dt, nullptr))) return st;
    return IMPEC_OK;
}

// body of the Impec2DSolver::Solve loop (Impec2DSolver.cpp:339-388)
int impec_step(impec_ctx *ctx, double t, double t_next, double *dt_out) {
    if (!ctx) return IMPEC_EINVAL;
    int st;
    if ((st = launch_fill_scalar(ctx, ctx->c, ctx->P.is_test))) return st;
    if ((st = pressure_solve(ctx, nullptr, nullptr))) return st;
    if ((st = launch_velocity(ctx))) return st;
    double dt;
    if ((st = launch_cfl(ctx, &dt))) return st;
    if (t_next > t && t + dt >= t_next) dt = t_next - t;
    if ((st = impec_hyperbolic_solve(ctx, dt))) return st;
    if (dt_out) *dt_out = dt;
    return IMPEC_OK;
}

// ---- diagnostics ----------------------------------------------------------
int impec_total_mass(impec_ctx *ctx, double *mass) {
    if (!ctx || !mass) return IMPEC_EINVAL;
    int st = launch_diagnostics(ctx, 0);
    if (st) return st;
    *mass = ctx->h_scal[S_MASS];
    return IMPEC_OK;
}

int impec_mass_error(impec_ctx *ctx, double t, double *err) {
    if (!ctx || !err) return IMPEC_EINVAL;
    int st = launch_diagnostics(ctx, 0);
    if (st) return st;
    const double mass = ctx->h_scal[S_MASS];
    const double outflux = ctx->h_scal[S_OUTFLUX];
    const double ncpv = 1. / (ctx->dx * ctx->dy);
    double influx = (t - ctx->P.tstart) * ctx->P.I * ncpv;
    if (ctx->P.source_type == IMPEC_SRC_PARALLEL) influx *= 2.;
    *err = fabs((mass + outflux - influx) / influx);
    return IMPEC_OK;
}

int impec_divergence_error(impec_ctx *ctx, double *err) {
    if (!ctx || !err) return IMPEC_EINVAL;
    int st = launch_diagnostics(ctx, 0);
    if (st) return st;
    *err = ctx->h_scal[S_DIV] * ctx->dx * ctx->dx;
    return IMPEC_OK;
}

int impec_velocity_boundary_error(impec_ctx *ctx, double *err) {
    if (!ctx || !err) return IMPEC_EINVAL;
    int st = launch_diagnostics(ctx, 0);
    if (st) return st;
    *err = ctx->h_scal[S_NOFLOW];
    return IMPEC_OK;
}

int impec_minmax_concentration(impec_ctx *ctx, double *cmax, double *cmin) {
    if (!ctx || !cmax || !cmin) return IMPEC_EINVAL;
    return launch_minmax(ctx, cmax, cmin);
}

int impec_get_stats(impec_ctx *ctx, impec_stats *st) {
    if (!ctx || !st) return IMPEC_EINVAL;
    if (int e = sync_scalars(ctx)) return e;
    ctx->stats.total_mass_outflux = ctx->h_scal[S_OUTFLUX];
    *st = ctx->stats;
    return IMPEC_OK;
}

int impec_pressure_operator(impec_ctx *ctx, double *A9) {
    if (!ctx || !A9) return IMPEC_EINVAL;
    return pressure_operator_dump(ctx, A9);
}

int impec_parabolic_operator(impec_ctx *ctx, double dt, double *A9, double *rhs) {
    if (!ctx) return IMPEC_EINVAL;
    int st = ensure_cn(ctx);
    if (st) return st;
    return cn_operator_dump(ctx, dt, A9, rhs);
}

// ---- measurement ----------------------------------------------------------
int impec_profile_enable(impec_ctx *ctx, int on) {
    if (!ctx) return IMPEC_EINVAL;
    ctx->prof_on = on != 0;
    return IMPEC_OK;
}

int impec_profile_read(impec_ctx *ctx, const char *name, long long *launches, double *ms) {
    if (!ctx || !name) return IMPEC_EINVAL;
    for (int k = 0; k < ctx->nprof; k++)
        if (strcmp(ctx->prof[k].name, name) == 0) {
            if (launches) *launches = ctx->prof[k].launches;
            if (ms) *ms = ctx->prof[k].ms;
            return IMPEC_OK;
        }
    if (launches) *launches = 0;
    if (ms) *ms = 0.;
    return IMPEC_EINVAL;
}

int impec_profile_reset(impec_ctx *ctx) {
    if (!ctx) return IMPEC_EINVAL;
    ctx->nprof = 0;
    return IMPEC_OK;
}

int impec_copy_calibration(long long nbytes, int reps, double *ms) {
    if (nbytes < 8 || reps < 1 || !ms) return IMPEC_EINVAL;
    const long n = (long)(nbytes / 8);
    double *a = nullptr, *b = nullptr;
    hipStream_t s;
    hipEvent_t e0, e1;
    if (hipMalloc(&a, n * 8) != hipSuccess) return IMPEC_ENOMEM;
    if (hipMalloc(&b, n * 8) != hipSuccess) {
        hipFree(a);
        return IMPEC_ENOMEM;
    }
    int st = IMPEC_OK;
    if (hipStreamCreate(&s) != hipSuccess) st = IMPEC_EHIP;
    if (!st && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) st = IMPEC_EHIP;
    if (!st) {
        hipMemsetAsync(a, 0, n * 8, s);
        int blocks = (int)((n + 255) / 256);
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(k_calib_copy, dim3(blocks), dim3(256), 0, s, a, b, n);  // warm-up
        hipEventRecord(e0, s);
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_calib_copy, dim3(blocks), dim3(256), 0, s, a, b, n);
        hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) st = IMPEC_EHIP;
        float t = 0.f;
        hipEventElapsedTime(&t, e0, e1);
        *ms = t / reps;
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        hipStreamDestroy(s);
    }
    hipFree(a);
    hipFree(b);
    return st;
}

}  // extern "C"
