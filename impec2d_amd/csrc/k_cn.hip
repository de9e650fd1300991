// k_cn.hip -- Crank-Nicolson velocity-dependent dispersion
// (HyperbolicSolver::DiffusionSplitCN -> MGSolver::Solve_Parabolic_cc_to_c,
// MGSolver.cpp:846-1392): face dispersion tensors, the 9-point
// (I - dt/2 L_D) operator with Neumann folding, the explicit right-hand side,
// and a Jacobi-preconditioned BiCGStab solve (the operator is not symmetric;
// the reference runs HYPRE PCG+SMG on it).  FP64.
#include "impec_device.h"
#include <math.h>

namespace impec {

enum { ST_C = 0, ST_W, ST_SW, ST_S, ST_SE, ST_E, ST_NE, ST_N, ST_NW };

__device__ __forceinline__ void cn_fold(int i, int j, int n1, int n2, double a[9]) {
    const bool bi0 = i == 0, bi1 = i == n1 - 1, bj0 = j == 0, bj1 = j == n2 - 1;
    if (!(bi0 | bi1 | bj0 | bj1)) return;
    const double C = a[ST_C], W = a[ST_W], SW = a[ST_SW], S = a[ST_S], SE = a[ST_SE], E = a[ST_E], NE = a[ST_NE],
                 N = a[ST_N], NW = a[ST_NW];
    double o[9];
    // MGSolver.cpp:1033-1296: same folding pattern as the elliptic matrix
    if (bi0 && bj0) {
        o[0] = C + W + S + SW; o[1] = 0; o[2] = 0; o[3] = 0; o[4] = 0; o[5] = E + SE; o[6] = NE; o[7] = N + NW; o[8] = 0;
    } else if (bi0 && bj1) {
        o[0] = C + W + N + NW; o[1] = 0; o[2] = 0; o[3] = S + SW; o[4] = SE; o[5] = E + NE; o[6] = 0; o[7] = 0; o[8] = 0;
    } else if (bi1 && bj1) {
        o[0] = C + N + NE + E; o[1] = W + NW; o[2] = SW; o[3] = S + SE; o[4] = 0; o[5] = 0; o[6] = 0; o[7] = 0; o[8] = 0;
    } else if (bi1 && bj0) {
        o[0] = C + S + E + SE; o[1] = W + SW; o[2] = 0; o[3] = 0; o[4] = 0; o[5] = 0; o[6] = 0; o[7] = N + NE; o[8] = NW;
    } else if (bi1) {
        o[0] = C + E; o[1] = W; o[2] = SW; o[3] = S + SE; o[4] = 0; o[5] = 0; o[6] = 0; o[7] = N + NE; o[8] = NW;
    } else if (bi0) {
        o[0] = C + W; o[1] = 0; o[2] = 0; o[3] = S + SW; o[4] = SE; o[5] = E; o[6] = NE; o[7] = N + NW; o[8] = 0;
    } else if (bj1) {
        o[0] = C + N; o[1] = W + NW; o[2] = SW; o[3] = S; o[4] = SE; o[5] = E + NE; o[6] = 0; o[7] = 0; o[8] = 0;
    } else {
        o[0] = C + S; o[1] = W + SW; o[2] = 0; o[3] = 0; o[4] = 0; o[5] = E + SE; o[6] = NE; o[7] = N; o[8] = NW;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = o[k];
}

struct CnConst {
    int N1, N2, B, ldv;
    double phal, phat;      // phi * alpha_l, phi * alpha_t (alpha scaled as the reference does)
    double phdm;            // phi * d_m (extension)
    int guard;              // extension: zero dispersion where |u| == 0
    // matrix and rhs constants (MGSolver.cpp:937-946, 1349-1360)
    double cA, mcA;         // dt/(8.*dx*dy), -dt/(8.*dx*dy)
    double mdt8, dt8;       // -dt/8., dt/8.
    double c4dx2, c4dy2, c1dxdy;  // 4./(dx*dx), 4./(dy*dy), 1./(dx*dy)
    double cx2, cy2;        // dt/(2.*dx*dx), dt/(2.*dy*dy)
};

// SetupDiffusionTensorF / SetupDiffusionTensorG (MGSolver.cpp:846-919) on
// every x-face (i in [0,N1], j in [0,N2)) and y-face (i in [0,N1), j in [0,N2])
__global__ void k_cn_tensor(CnConst k, const double *__restrict__ u, const double *__restrict__ v,
                            double *__restrict__ d11, double *__restrict__ d12, double *__restrict__ d22,
                            double *__restrict__ d21) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > k.N1 || j > k.N2) return;
    const long L = k.ldv;
    const long o = (long)(i + k.B) * L + (j + k.B);
    const long t = (long)i * L + j;
    if (j < k.N2) {
        const double uu = u[o];
        const double vb = 1. / 4 * (v[o - L + 1] + v[o + 1] + v[o - L] + v[o]);
        const double u2 = uu * uu, v2 = vb * vb;
        const double s = sqrt(u2 + v2);
        double fl = k.phal / s, ft = k.phat / s;
        if (k.guard && !(s > 0)) { fl = 0.; ft = 0.; }
        const double a11 = fl * u2, a12 = fl * (uu * vb);
        const double b11 = ft * v2, b12 = ft * (-uu * vb);
        double x11 = a11 + b11;
        if (k.phdm != 0.) x11 = x11 + k.phdm;
        d11[t] = x11;
        d12[t] = a12 + b12;
    }
    if (i < k.N1) {
        const double ub = 1. / 4 * (u[o + L - 1] + u[o + L] + u[o - 1] + u[o]);
        const double vv = v[o];
        const double u2 = ub * ub, v2 = vv * vv;
        const double s = sqrt(u2 + v2);
        double fl = k.phal / s, ft = k.phat / s;
        if (k.guard && !(s > 0)) { fl = 0.; ft = 0.; }
        const double a21 = fl * (ub * vv), a22 = fl * v2;
        const double b21 = ft * (-ub * vv), b22 = ft * u2;
        d21[t] = a21 + b21;
        double x22 = a22 + b22;
        if (k.phdm != 0.) x22 = x22 + k.phdm;
        d22[t] = x22;
    }
}

// GetCoefficients_Parabolic (MGSolver.cpp:923-948) + folding
__device__ __forceinline__ void cn_row(const CnConst &k, const double *__restrict__ d11, const double *__restrict__ d12,
                                       const double *__restrict__ d22, const double *__restrict__ d21, int i, int j,
                                       double a[9]) {
    const long L = k.ldv, t = (long)i * L + j;
    const double d11m = d11[t], d12m = d12[t], d11p = d11[t + L], d12p = d12[t + L];
    const double d22m = d22[t], d21m = d21[t], d22p = d22[t + 1], d21p = d21[t + 1];
    a[ST_NW] = k.cA * (d12m + d21p);
    a[ST_NE] = k.mcA * (d12p + d21p);
    a[ST_N] = k.mdt8 * (k.c4dy2 * d22p + k.c1dxdy * (-d12m + d12p));
    a[ST_SW] = k.mcA * (d12m + d21m);
    a[ST_SE] = k.cA * (d12p + d21m);
    a[ST_S] = k.mdt8 * (k.c4dy2 * d22m + k.c1dxdy * (-d12p + d12m));
    a[ST_W] = k.mdt8 * (k.c4dx2 * d11m + k.c1dxdy * (-d21p + d21m));
    a[ST_C] = 1. + k.cx2 * (d11p + d11m) + k.cy2 * (d22p + d22m);
    a[ST_E] = k.mdt8 * (k.c4dx2 * d11p + k.c1dxdy * (-d21m + d21p));
    cn_fold(i, j, k.N1, k.N2, a);
}

// SetRHS_Parabolic (MGSolver.cpp:1324-1370); c ghost-filled
__global__ void k_cn_rhs(CnConst k, const double *__restrict__ c, int lds, const double *__restrict__ d11,
                         const double *__restrict__ d12, const double *__restrict__ d22,
                         const double *__restrict__ d21, double *__restrict__ rhs, double *__restrict__ diag) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= k.N1 || j >= k.N2) return;
    const long L = k.ldv, t = (long)i * L + j;
    const double d11m = d11[t], d12m = d12[t], d11p = d11[t + L], d12p = d12[t + L];
    const double d22m = d22[t], d21m = d21[t], d22p = d22[t + 1], d21p = d21[t + 1];
    const double *cc = c + (long)(i + k.B) * lds + (j + k.B);
#define C_(a, b) cc[(long)(a) * lds + (b)]
    const double r = C_(0, 0) + k.dt8 * (
        k.c4dx2 * (d11p * (C_(1, 0) - C_(0, 0)) - d11m * (C_(0, 0) - C_(-1, 0)))
        + k.c1dxdy * (d12p * (C_(0, 1) + C_(1, 1) - C_(0, -1) - C_(1, -1))
                      - d12m * (C_(0, 1) + C_(-1, 1) - C_(-1, -1) - C_(0, -1)))
        + k.c1dxdy * (d21p * (C_(1, 0) + C_(1, 1) - C_(-1, 1) - C_(-1, 0))
                      - d21m * (C_(1, 0) + C_(1, -1) - C_(-1, 0) - C_(-1, -1)))
        + k.c4dy2 * (d22p * (C_(0, 1) - C_(0, 0)) - d22m * (C_(0, 0) - C_(0, -1))));
#undef C_
    rhs[(long)i * k.N2 + j] = r;
    double a[9];
    cn_row(k, d11, d12, d22, d21, i, j, a);
    diag[(long)i * k.N2 + j] = a[ST_C];
}

// y = A x on the flat N1*N2 layout, with optional partial dot products
// (y.s and y.y) for BiCGStab
template <int MODE>
__global__ void k_cn_apply(CnConst k, const double *__restrict__ d11, const double *__restrict__ d12,
                           const double *__restrict__ d22, const double *__restrict__ d21, const double *__restrict__ x,
                           double *__restrict__ y, const double *__restrict__ w, double *__restrict__ part) {
    const long n = (long)k.N1 * k.N2;
    double s[2] = {0., 0.};
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const int i = (int)(id / k.N2), j = (int)(id - (long)i * k.N2);
        double a[9];
        cn_row(k, d11, d12, d22, d21, i, j, a);
        const double *xc = x + id;
        const long ld = k.N2;
        double acc = a[ST_C] * xc[0];
        const bool we = i > 0, ea = i < k.N1 - 1, so = j > 0, no = j < k.N2 - 1;
        if (we) acc += a[ST_W] * xc[-ld];
        if (we && so) acc += a[ST_SW] * xc[-ld - 1];
        if (so) acc += a[ST_S] * xc[-1];
        if (ea && so) acc += a[ST_SE] * xc[ld - 1];
        if (ea) acc += a[ST_E] * xc[ld];
        if (ea && no) acc += a[ST_NE] * xc[ld + 1];
        if (no) acc += a[ST_N] * xc[1];
        if (we && no) acc += a[ST_NW] * xc[-ld + 1];
        y[id] = acc;
        if (MODE == 1) s[0] += w[id] * acc;            // r0 . v
        if (MODE == 2) { s[0] += acc * w[id]; s[1] += acc * acc; }  // t . s, t . t
    }
    if (MODE) {
        block_reduce<RED_SUM, 2>(s);
        if (tid_linear() == 0) {
            part[blockIdx.x] = s[0];
            part[RED_BLOCKS + blockIdx.x] = s[1];
        }
    }
}

// BiCGStab vector steps; the scalars rho, r0.r of the next step (rho_new),
// alpha and omega live in device memory
constexpr int S_RHO_NEW = 13;

__global__ void k_bicg_init(long n, const double *__restrict__ b, double *__restrict__ r, double *__restrict__ r0,
                            double *__restrict__ p, double *__restrict__ vv, double *__restrict__ x,
                            double *__restrict__ part) {
    double s[2] = {0., 0.};
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const double bv = b[id];
        r[id] = bv;
        r0[id] = bv;
        p[id] = 0.;
        vv[id] = 0.;
        x[id] = 0.;
        s[0] += bv * bv;
        s[1] += bv * bv;
    }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        part[blockIdx.x] = s[0];
        part[RED_BLOCKS + blockIdx.x] = s[1];
    }
}

__global__ void k_bicg_scal_init(double *__restrict__ S) {
    S[S_RHO] = 1.; S[S_ALPHA] = 1.; S[S_OMEGA] = 1.;
}

// p = r + beta (p - omega v), y = p / diag; then rho <- rho_new
__global__ void k_bicg_p(long n, const double *__restrict__ r, double *__restrict__ p, const double *__restrict__ vv,
                         const double *__restrict__ diag, double *__restrict__ y, const double *__restrict__ S) {
    const double beta = (S[S_RHO_NEW] / S[S_RHO]) * (S[S_ALPHA] / S[S_OMEGA]);
    const double om = S[S_OMEGA];
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const double pv = r[id] + beta * (p[id] - om * vv[id]);
        p[id] = pv;
        y[id] = pv / diag[id];
    }
}

__global__ void k_bicg_rho(double *__restrict__ S) { S[S_RHO] = S[S_RHO_NEW]; }
__global__ void k_bicg_alpha(double *__restrict__ S) { S[S_ALPHA] = S[S_RHO] / S[S_R0V]; }
__global__ void k_bicg_omega(double *__restrict__ S) { S[S_OMEGA] = S[S_TT] > 0 ? S[S_TS] / S[S_TT] : 0.; }

// s = r - alpha v, z = s / diag
__global__ void k_bicg_s(long n, const double *__restrict__ r, const double *__restrict__ vv,
                         const double *__restrict__ diag, double *__restrict__ sv, double *__restrict__ z,
                         const double *__restrict__ S) {
    const double alpha = S[S_ALPHA];
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const double s = r[id] - alpha * vv[id];
        sv[id] = s;
        z[id] = s / diag[id];
    }
}

// x += alpha y + omega z, r = s - omega t; partials r.r, r0.r
__global__ void k_bicg_update(long n, double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                              double *__restrict__ r, const double *__restrict__ sv, const double *__restrict__ t,
                              const double *__restrict__ r0, const double *__restrict__ S, double *__restrict__ part) {
    const double alpha = S[S_ALPHA], om = S[S_OMEGA];
    double s[2] = {0., 0.};
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        x[id] += alpha * y[id] + om * z[id];
        const double rv = sv[id] - om * t[id];
        r[id] = rv;
        s[0] += rv * rv;
        s[1] += r0[id] * rv;
    }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        part[blockIdx.x] = s[0];
        part[RED_BLOCKS + blockIdx.x] = s[1];
    }
}

// solution into the interior of c
__global__ void k_cn_store(const double *__restrict__ x, double *__restrict__ c, int N1, int N2, int B, int lds) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= N1 || j >= N2) return;
    c[(long)(i + B) * lds + (j + B)] = x[(long)i * N2 + j];
}

// matrix dump: nine N1*N2 arrays
__global__ void k_cn_dump(CnConst k, const double *__restrict__ d11, const double *__restrict__ d12,
                          const double *__restrict__ d22, const double *__restrict__ d21, double *__restrict__ A) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= k.N1 || j >= k.N2) return;
    double a[9];
    cn_row(k, d11, d12, d22, d21, i, j, a);
    const long M = (long)k.N1 * k.N2;
    for (int q = 0; q < 9; q++) A[q * M + (long)i * k.N2 + j] = a[q];
}

static CnConst cn_const(const impec_ctx *ctx, double dt) {
    CnConst k;
    k.N1 = ctx->N1; k.N2 = ctx->N2; k.B = ctx->B; k.ldv = ctx->ldv;
    const double dx = ctx->dx, dy = ctx->dy;
    const double phi = ctx->P.porosity;
    const double maxdx = dx < dy ? dy : dx;  // std::max(dx, dy)
    const double al = ctx->P.alpha_absolute ? ctx->P.alpha_l : ctx->P.alpha_l * maxdx;
    const double at = ctx->P.alpha_absolute ? ctx->P.alpha_t : ctx->P.alpha_t * maxdx;
    k.phal = phi * al;
    k.phat = phi * at;
    k.phdm = phi * ctx->P.d_m;
    k.guard = ctx->P.alpha_absolute;
    k.cA = dt / (8. * dx * dy);
    k.mcA = -dt / (8. * dx * dy);
    k.mdt8 = -dt / 8.;
    k.dt8 = dt / 8.;
    k.c4dx2 = 4. / (dx * dx);
    k.c4dy2 = 4. / (dy * dy);
    k.c1dxdy = 1. / (dx * dy);
    k.cx2 = dt / (2. * dx * dx);
    k.cy2 = dt / (2. * dy * dy);
    return k;
}

static int red_grid(long n) {
    long b = (n + 1023) / 1024;
    if (b < 1) b = 1;
    if (b > RED_BLOCKS) b = RED_BLOCKS;
    return (int)b;
}

__global__ void k_copy_scal_cn(double *__restrict__ S, int dst, int src) { S[dst] = S[src]; }

// face tensors, right-hand side (into cn_rhs) and folded diagonal (into cn_dg)
static int tensors_and_rhs(impec_ctx *ctx, const CnConst &k) {
    dim3 blk(64, 4);
    dim3 gt((ctx->N2 + 1 + 63) / 64, (ctx->N1 + 1 + 3) / 4);
    hipLaunchKernelGGL(k_cn_tensor, gt, blk, 0, ctx->stream, k, ctx->u, ctx->v, ctx->d11, ctx->d12, ctx->d22,
                       ctx->d21);
    dim3 gc((ctx->N2 + 63) / 64, (ctx->N1 + 3) / 4);
    hipLaunchKernelGGL(k_cn_rhs, gc, blk, 0, ctx->stream, k, ctx->c, ctx->lds, ctx->d11, ctx->d12, ctx->d22, ctx->d21,
                       ctx->cn_rhs, ctx->cn_dg);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

static int fetch(impec_ctx *ctx) {
    IMPEC_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_scal, ctx->scal, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost,
                                        ctx->stream));
    IMPEC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return IMPEC_OK;
}

// Solve_Parabolic_cc_to_c: x0 = 0 (AssignRHS_Parabolic, MGSolver.cpp:1375-1392),
// stop on ||rhs - A x||_2 <= para_eps ||rhs||_2, solution into c's interior.
int cn_solve(impec_ctx *ctx, double dt, int *its_out) {
    const CnConst k = cn_const(ctx, dt);
    const long n = (long)ctx->N1 * ctx->N2;
    const int nb = red_grid(n);
    hipStream_t s = ctx->stream;
    double *b = ctx->cn_rhs, *x = ctx->cn_x, *r = ctx->cn_r, *r0 = ctx->cn_r0, *p = ctx->cn_p, *vv = ctx->cn_v;
    double *sv = ctx->cn_s, *t = ctx->cn_t, *y = ctx->cn_y, *dg = ctx->cn_dg, *z = ctx->cn_z;
    prof_begin(ctx);
    int st = tensors_and_rhs(ctx, k);
    prof_end(ctx, "cn_tensor_rhs");
    if (st) return st;
    hipLaunchKernelGGL(k_bicg_init, dim3(nb), dim3(256), 0, s, n, b, r, r0, p, vv, x, ctx->part);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    if ((st = finish_sum(ctx, nb, 2, S_RR))) return st;   // S_RR = b.b, S_RSUM = r0.r = b.b
    hipLaunchKernelGGL(k_bicg_scal_init, dim3(1), dim3(1), 0, s, ctx->scal);
    hipLaunchKernelGGL(k_copy_scal_cn, dim3(1), dim3(1), 0, s, ctx->scal, S_RHO_NEW, (int)S_RSUM);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    if ((st = fetch(ctx))) return st;
    const double bb = ctx->h_scal[S_RR];
    double rr = bb;
    const double tol = ctx->P.para_eps;
    const int maxit = ctx->P.max_iterations > 0 ? ctx->P.max_iterations : 10000;
    int its = 0, status = IMPEC_OK;
    while (!(rr <= tol * tol * bb)) {
        if (its >= maxit) { status = IMPEC_ENOCONV; break; }
        if (!(rr == rr)) { status = IMPEC_ENAN; break; }
        hipLaunchKernelGGL(k_bicg_p, dim3(nb), dim3(256), 0, s, n, r, p, vv, dg, y, ctx->scal);
        hipLaunchKernelGGL(k_bicg_rho, dim3(1), dim3(1), 0, s, ctx->scal);
        prof_begin(ctx);
        hipLaunchKernelGGL(k_cn_apply<1>, dim3(nb), dim3(256), 0, s, k, ctx->d11, ctx->d12, ctx->d22, ctx->d21, y, vv,
                           r0, ctx->part);
        prof_end(ctx, "cn_apply");
        if ((st = finish_sum(ctx, nb, 1, S_R0V))) return st;
        hipLaunchKernelGGL(k_bicg_alpha, dim3(1), dim3(1), 0, s, ctx->scal);
        hipLaunchKernelGGL(k_bicg_s, dim3(nb), dim3(256), 0, s, n, r, vv, dg, sv, z, ctx->scal);
        prof_begin(ctx);
        hipLaunchKernelGGL(k_cn_apply<2>, dim3(nb), dim3(256), 0, s, k, ctx->d11, ctx->d12, ctx->d22, ctx->d21, z, t,
                           sv, ctx->part);
        prof_end(ctx, "cn_apply");
        if ((st = finish_sum(ctx, nb, 2, S_TS))) return st;   // S_TS = t.s, S_TT = t.t
        hipLaunchKernelGGL(k_bicg_omega, dim3(1), dim3(1), 0, s, ctx->scal);
        hipLaunchKernelGGL(k_bicg_update, dim3(nb), dim3(256), 0, s, n, x, y, z, r, sv, t, r0, ctx->scal, ctx->part);
        if ((st = finish_sum(ctx, nb, 2, S_RR))) return st;   // S_RR = r.r, S_RSUM = r0.r
        hipLaunchKernelGGL(k_copy_scal_cn, dim3(1), dim3(1), 0, s, ctx->scal, S_RHO_NEW, (int)S_RSUM);
        IMPEC_HIP_CHECK(ctx, hipGetLastError());
        if ((st = fetch(ctx))) return st;
        rr = ctx->h_scal[S_RR];
        its++;
    }
    dim3 blk(64, 4), gc((ctx->N2 + 63) / 64, (ctx->N1 + 3) / 4);
    hipLaunchKernelGGL(k_cn_store, gc, blk, 0, s, x, ctx->c, ctx->N1, ctx->N2, ctx->B, ctx->lds);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    ctx->stats.last_parabolic_its = its;
    ctx->stats.last_parabolic_relres = bb > 0 ? sqrt(rr / bb) : 0.;
    if (its_out) *its_out = its;
    if (status == IMPEC_ENOCONV)
        return set_error(ctx, status, "dispersion solve: no convergence after %d iterations", its);
    if (status == IMPEC_ENAN) return set_error(ctx, status, "dispersion solve: non-finite residual");
    return IMPEC_OK;
}

int cn_operator_dump(impec_ctx *ctx, double dt, double *A9_host, double *rhs_host) {
    const CnConst k = cn_const(ctx, dt);
    const long n = (long)ctx->N1 * ctx->N2;
    int st = tensors_and_rhs(ctx, k);
    if (st) return st;
    double *dA = nullptr;
    IMPEC_HIP_CHECK(ctx, hipMalloc(&dA, sizeof(double) * 9 * n));
    dim3 blk(64, 4), gc((ctx->N2 + 63) / 64, (ctx->N1 + 3) / 4);
    hipLaunchKernelGGL(k_cn_dump, gc, blk, 0, ctx->stream, k, ctx->d11, ctx->d12, ctx->d22, ctx->d21, dA);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && A9_host) e = hipMemcpyAsync(A9_host, dA, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && rhs_host)
        e = hipMemcpyAsync(rhs_host, ctx->cn_rhs, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(dA);
    if (e != hipSuccess) return set_error(ctx, IMPEC_EHIP, "parabolic operator dump: %s", hipGetErrorString(e));
    return IMPEC_OK;
}

}  // namespace impec
