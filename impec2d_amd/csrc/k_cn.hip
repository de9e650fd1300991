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

// SetupDiffusionTensorF (MGSolver.cpp:846-882) at one x-face from u and the
// four-point average vb of v
__device__ __forceinline__ void xface_tensor(const CnConst &k, double uu, double vb, double &d11, double &d12) {
    const double u2 = uu * uu, v2 = vb * vb;
    const double s = sqrt(u2 + v2);
    double fl = k.phal / s, ft = k.phat / s;
    if (k.guard && !(s > 0)) { fl = 0.; ft = 0.; }
    const double a11 = fl * u2, a12 = fl * (uu * vb);
    const double b11 = ft * v2, b12 = ft * (-uu * vb);
    double x11 = a11 + b11;
    if (k.phdm != 0.) x11 = x11 + k.phdm;
    d11 = x11;
    d12 = a12 + b12;
}

// SetupDiffusionTensorG (MGSolver.cpp:884-919) at one y-face from the
// four-point average ub of u and v
__device__ __forceinline__ void yface_tensor(const CnConst &k, double ub, double vv, double &d22, double &d21) {
    const double u2 = ub * ub, v2 = vv * vv;
    const double s = sqrt(u2 + v2);
    double fl = k.phal / s, ft = k.phat / s;
    if (k.guard && !(s > 0)) { fl = 0.; ft = 0.; }
    const double a21 = fl * (ub * vv), a22 = fl * v2;
    const double b21 = ft * (-ub * vv), b22 = ft * u2;
    d21 = a21 + b21;
    double x22 = a22 + b22;
    if (k.phdm != 0.) x22 = x22 + k.phdm;
    d22 = x22;
}

// the eight face tensors around cell (i, j) (west / east x-faces, south /
// north y-faces) from the ghost-filled u and v, with the four-point averages
// of SetupDiffusionTensorF/G in their summation order
struct CellTensors {
    double d11m, d12m, d11p, d12p, d22m, d21m, d22p, d21p;
};
__device__ __forceinline__ CellTensors cell_tensors(const CnConst &k, const double *__restrict__ u,
                                                    const double *__restrict__ v, int i, int j) {
    const long L = k.ldv;
    const long o = (long)(i + k.B) * L + (j + k.B);
    CellTensors T;
    xface_tensor(k, u[o], 1. / 4 * (v[o - L + 1] + v[o + 1] + v[o - L] + v[o]), T.d11m, T.d12m);
    xface_tensor(k, u[o + L], 1. / 4 * (v[o + 1] + v[o + L + 1] + v[o] + v[o + L]), T.d11p, T.d12p);
    yface_tensor(k, 1. / 4 * (u[o + L - 1] + u[o + L] + u[o - 1] + u[o]), v[o], T.d22m, T.d21m);
    yface_tensor(k, 1. / 4 * (u[o + L] + u[o + L + 1] + u[o] + u[o + 1]), v[o + 1], T.d22p, T.d21p);
    return T;
}

// GetCoefficients_Parabolic (MGSolver.cpp:923-948) + folding
__device__ __forceinline__ void cn_coeffs(const CnConst &k, const CellTensors &T, int i, int j, double a[9]) {
    a[ST_NW] = k.cA * (T.d12m + T.d21p);
    a[ST_NE] = k.mcA * (T.d12p + T.d21p);
    a[ST_N] = k.mdt8 * (k.c4dy2 * T.d22p + k.c1dxdy * (-T.d12m + T.d12p));
    a[ST_SW] = k.mcA * (T.d12m + T.d21m);
    a[ST_SE] = k.cA * (T.d12p + T.d21m);
    a[ST_S] = k.mdt8 * (k.c4dy2 * T.d22m + k.c1dxdy * (-T.d12p + T.d12m));
    a[ST_W] = k.mdt8 * (k.c4dx2 * T.d11m + k.c1dxdy * (-T.d21p + T.d21m));
    a[ST_C] = 1. + k.cx2 * (T.d11p + T.d11m) + k.cy2 * (T.d22p + T.d22m);
    a[ST_E] = k.mdt8 * (k.c4dx2 * T.d11p + k.c1dxdy * (-T.d21m + T.d21p));
    cn_fold(i, j, k.N1, k.N2, a);
}

// SetRHS_Parabolic (MGSolver.cpp:1324-1370) at one cell: c at row offsets
// m/0/p (-1, 0, +1) and column offsets m/0/p, in the reference's grouping
__device__ __forceinline__ double cn_rhs_at(const CnConst &k, const CellTensors &T, double cmm, double cm0,
                                            double cmp, double c0m, double c00, double c0p, double cpm, double cp0,
                                            double cpp) {
    return c00 + k.dt8 * (
        k.c4dx2 * (T.d11p * (cp0 - c00) - T.d11m * (c00 - cm0))
        + k.c1dxdy * (T.d12p * (c0p + cpp - c0m - cpm)
                      - T.d12m * (c0p + cmp - cmm - c0m))
        + k.c1dxdy * (T.d21p * (cp0 + cpp - cmp - cm0)
                      - T.d21m * (cp0 + cpm - cm0 - cmm))
        + k.c4dy2 * (T.d22p * (c0p - c00) - T.d22m * (c00 - c0m)));
}

// The right-hand side and the Jacobi preconditioner; c, u, v ghost-filled.
// 64 x 4 threads per workgroup over (column, row), rows strided;
// per-workgroup partials of b.b (twice: ||b||^2 and rho) for the solver's
// start
__global__ void k_cn_rhs(CnConst k, const double *__restrict__ c, int lds, const double *__restrict__ u,
                         const double *__restrict__ v, double *__restrict__ rhs, float *__restrict__ diag,
                         double *__restrict__ part) {
    const HwGrid g_;
    const int j = g_.bx() * 64 + threadIdx.x;
    double sq = 0.;
    if (j < k.N2)
        for (int i = g_.by() * 4 + threadIdx.y; i < k.N1; i += g_.gy() * 4) {
            const CellTensors T = cell_tensors(k, u, v, i, j);
            const double *cc = c + (long)(i + k.B) * lds + (j + k.B);
#define C_(a, b) cc[(long)(a) * lds + (b)]
            const double r = cn_rhs_at(k, T, C_(-1, -1), C_(-1, 0), C_(-1, 1), C_(0, -1), C_(0, 0), C_(0, 1),
                                       C_(1, -1), C_(1, 0), C_(1, 1));
#undef C_
            rhs[(long)i * k.N2 + j] = r;
            double a[9];
            cn_coeffs(k, T, i, j, a);
            diag[(long)i * k.N2 + j] = (float)(1. / a[ST_C]);   // Jacobi preconditioner, rounded to FP32
            sq += r * r;
        }
    double s2[2] = {sq, sq};
    block_reduce<RED_SUM, 2>(s2);
    if (tid_linear() == 0) {
        const int blk = g_.by() * g_.gx() + g_.bx();
        part[blk] = s2[0];
        part[RED_BLOCKS + blk] = s2[1];
    }
}

// ---------------------------------------------------------------------------
// Jacobi-preconditioned BiCGStab, streamed.  Each operator application carries
// the vector updates on either side of it, so an iteration is two launches:
//   k_cn_stream<CN_S>:  s = r - alpha v,  t = A (s / D); partials t.s, t.t,
//                       r0.t, r0.s (s and t are not stored)
//   k_cn_pstep:         s and t again from r and v, then r' = s - omega t and
//                       x += alpha p/D + omega s/D (the end of the iteration),
//                       p' = r' + beta (p - omega v), v' = A (p' / D);
//                       partials r0.v', r'.r'
//   k_cn_stream<CN_P0>: the first direction p = r, v = A (p / D).
//   k_cn_stream<CN_R>:  the true residual of the iterate, partials of
//                       ||b - A x||^2 (x read from c's interior, b from r).
// rho_new = r0.r' is formed from the s-launch's partials as r0.s - omega r0.t,
// so beta is known before the p-launch; the convergence test uses the
// p-launch's r'.r', the recursive residual itself.  The p-launch forms s and t
// with the s-launch's arithmetic, bit for bit, so omega belongs to the t it
// uses; recomputing them costs an operator application and saves storing and
// re-reading two vectors (24 of 160 bytes per cell and iteration).  The last
// p-launch of a solve computes a direction nobody uses: one operator
// application per solve, against a separate vector pass per iteration.
// 1/D is stored in FP32: the preconditioner is diag(1/d~) with d~ the rounded
// inverse, applied identically to the matvec input and to the solution
// update, so the iteration solves the FP64 system to the same residual.
// A wavefront owns a strip of 64 columns (one per lane) and a chunk of rows
// and marches down the rows with the operator's input vectors of three rows in
// registers; the column neighbours come from the adjacent lanes.  The scalars
// live in device memory, so an iteration needs one host synchronisation, for
// the convergence test.
// ---------------------------------------------------------------------------
constexpr int S_RHO_NEW = 13;
#ifndef IMPEC_CN_WAVES
#define IMPEC_CN_WAVES 4
#endif
// the s-launch (k_cn_stream) in workgroups of 8 waves: 181 us against 184
// with 4 at 4096^2; the p-launch and the start lose 4-7% with 8 and keep 4
#ifndef IMPEC_CNS_WAVES
#define IMPEC_CNS_WAVES 8
#endif
// the zero-guess start (k_cn_start)
#ifndef IMPEC_CNZ_WAVES
#define IMPEC_CNZ_WAVES 4
#endif
constexpr int CN_OUT = 62, CNP_OUT = 60, CN_WAVES = IMPEC_CN_WAVES, CNS_WAVES = IMPEC_CNS_WAVES,
              CNZ_WAVES = IMPEC_CNZ_WAVES;
enum { CN_S = 1, CN_P0 = 2, CN_R = 3, CN_RX = 4 };
// build-time tuning: rows in flight per wavefront, wavefronts per launch
#ifndef IMPEC_CN_PF
#define IMPEC_CN_PF 2
#endif
// the true-residual pass with the last x update (k_cn_stream<CN_RX>, seven
// streams per row)
#ifndef IMPEC_CNRX_PF
#define IMPEC_CNRX_PF 2
#endif
// cache hints of the p-launch: 1 reads and writes x (used again only by the
// next p-launch) non-temporally, 2 also writes p'.  Both make the next
// s-launch faster (170-173 against 181 us at 4096^2) and the p-launch slower
// (316-317 against 303): off
#ifndef IMPEC_CNP_NT
#define IMPEC_CNP_NT 0
#endif
#ifndef IMPEC_CNP_PF
#define IMPEC_CNP_PF 3
#endif
#ifndef IMPEC_CN_TARGET_WAVES
#define IMPEC_CN_TARGET_WAVES 4096
#endif
// the p-launch: short chunks (about 14 rows at 4096^2) run fastest; more
// halo rows are re-read, yet the launch finishes sooner (measured with three
// rows in flight: 8192 waves 318 us, 12288 313 us, 20480 303 us, 24576
// 314 us, 32768 330 us)
#ifndef IMPEC_CNP_TARGET_WAVES
#define IMPEC_CNP_TARGET_WAVES 20480
#endif
// the one-pass start (its own chunks of rows; 8192 waves: 207 us against
// 212 at 12288 and 214 at 20480)
#ifndef IMPEC_CNS_TARGET_WAVES
#define IMPEC_CNS_TARGET_WAVES 8192
#endif
// build-time tuning: minimum waves per SIMD requested for the p-launch (1 lets
// the compiler use the registers it wants: 144, three waves per SIMD)
#ifndef IMPEC_CNP_MIN_WAVES
#define IMPEC_CNP_MIN_WAVES 1
#endif

struct CnArgs {
    CnConst k;
    const double *u, *v;                   // face velocities, ghosted (ghost width k.B, pitch ldv)
    const float *dinv;                     // rounded inverse of the folded diagonal (flat N1 x N2)
    const double *r;                       // the residual
    const double *pin, *vin;               // flat N1 x N2
    const float *r0;                       // the shadow residual, rounded to FP32 (any fixed r0 with
                                           // r0.r != 0 serves BiCGStab; 4 B per read instead of 8)
    double *x, *rout;                      // p-step: the solution (in place) and the new residual
    long ldx;                              // row pitch of x (the interior of the ghosted c)
    const double *rs;                      // CN_RX: the residual r of the last iteration's s = r - alpha v
    double *xout;                          // CN_RX: the updated solution (pitch ldx, another buffer)
    double *pout;                          // p-step, CN_P0: new p
    double *out;                           // p-step, CN_P0: new v
    const double *S;
    double *part;                          // per-wave partials
    int R, nstrip, nwave;
    int xzero;                             // p-step: x is zero (the first update of a zero-guess solve)
};

// the folded 9-point operator of row i (coefficients a) applied to y, whose
// rows i-1, i, i+1 are Y0, Y1, Y2 (column neighbours from the
// adjacent lanes; every lane of the wave executes it).  No boundary masks:
// the fold zeroes every coefficient that points outside the grid (corners and
// edges, MGSolver.cpp:1033-1296), and y is zero outside, so those terms add
// +0 -- interior results are those of the masked sum bit for bit.
__device__ __forceinline__ double cn_apply(const double a[9], double Y0, double Y1, double Y2) {
    const double yl0 = lane_prev(Y0), yl1 = lane_prev(Y1), yl2 = lane_prev(Y2);
    const double yr0 = lane_next(Y0), yr1 = lane_next(Y1), yr2 = lane_next(Y2);
    double acc = a[ST_C] * Y1;
    acc += a[ST_W] * Y0;
    acc += a[ST_SW] * yl0;
    acc += a[ST_S] * yl1;
    acc += a[ST_SE] * yl2;
    acc += a[ST_E] * Y2;
    acc += a[ST_NE] * yr2;
    acc += a[ST_N] * yr1;
    acc += a[ST_NW] * yr0;
    return acc;
}

// the face tensors streamed down a strip: push(row m) forms the x-face tensors
// of face m (from u and v of rows m-1 and m) and the y-face tensors of row m-1
// (from u of rows m-1 and m), the four-point averages of cell_tensors in its
// summation order; cell(...) then gives row m-1's eight tensors
struct FaceStream {
    double uprev, vprev, uprev_l, vprev_n;   // u, v of row m-1 and their lane shifts
    double f11a = 0., f12a = 0., f11b = 0., f12b = 0.;   // x-faces m-1 and m
    double g22 = 0., g21 = 0.;                           // y-faces of row m-1
    __device__ __forceinline__ void start(double u_, double v_) {
        uprev = u_; vprev = v_; uprev_l = lane_prev(u_); vprev_n = lane_next(v_);
    }
    __device__ __forceinline__ void push(const CnConst &k, double um, double vm) {
        const double um_l = lane_prev(um), vm_n = lane_next(vm);
        double e11, e12;
        const double vb = 1. / 4 * (vprev_n + vm_n + vprev + vm);
        xface_tensor(k, um, vb, e11, e12);
        const double ub = 1. / 4 * (um_l + um + uprev_l + uprev);
        yface_tensor(k, ub, vprev, g22, g21);
        uprev = um; vprev = vm; uprev_l = um_l; vprev_n = vm_n;
        f11a = f11b; f12a = f12b; f11b = e11; f12b = e12;
    }
    __device__ __forceinline__ CellTensors cell() const {
        return CellTensors{f11a, f12a, f11b, f12b, g22, g21, lane_next(g22), lane_next(g21)};
    }
};

// CN_S (s = r - alpha v, t = A (s / D), partials only), CN_P0 (p = r,
// v = A (p / D)), CN_R (partials of ||r - A x||^2 with r = b; nothing
// written) and CN_RX (the last iteration's update of x, as k_cn_xfinal forms
// it, into xout, and the partials of ||b - A x||^2 of the updated x: x is
// read from its own buffer and the result goes to another, so the halo rows
// a wave updates again are never ones another wave has already written);
// lanes 1..62 of a strip are written
template <int MODE>
__global__ void __launch_bounds__(64 * CNS_WAVES) k_cn_stream(CnArgs A) {
    const int lane = threadIdx.x & 63;
    const int wid = xcd_block(blockIdx.x, gridDim.x) * CNS_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wid >= A.nwave) return;
    const CnConst &k = A.k;
    const int N1 = k.N1, N2 = k.N2;
    const long L = k.ldv;
    const int strip = wid % A.nstrip, chunk = wid / A.nstrip;
    const int j = strip * CN_OUT - 1 + lane;
    const bool jv = j >= 0 && j < N2;
    const int jc = j < 0 ? 0 : (j > N2 - 1 ? N2 - 1 : j);   // cell column for loads
    const bool outl = lane >= 1 && lane <= CN_OUT && jv;
    const int i0 = chunk * A.R, i1 = i0 + A.R < N1 ? i0 + A.R : N1;
    const double alpha = (MODE == CN_S || MODE == CN_RX) ? A.S[S_ALPHA] : 0.;
    const double om = MODE == CN_RX ? A.S[S_OMEGA] : 0.;
    // raw row loads from clamped addresses (no lane-dependent branches)
    auto cell = [&](const double *p, int m) -> double {
        const int mc = m < 0 ? 0 : (m > N1 - 1 ? N1 - 1 : m);
        return p[(long)mc * N2 + jc];
    };
    auto cellf = [&](const float *p, int m) -> float {
        const int mc = m < 0 ? 0 : (m > N1 - 1 ? N1 - 1 : m);
        return p[(long)mc * N2 + jc];
    };
    // face velocities: rows -B..N1+B and columns -1..N2 of the ghosted arrays
    const int B = k.B;
    const int jg = j < -1 ? -1 : (j > N2 ? N2 : j);
    auto vel = [&](const double *p, int m) -> double {
        const int mc = m < -B ? -B : (m > N1 + B ? N1 + B : m);
        return p[(long)(mc + B) * L + (jg + B)];
    };
    // ring of prefetched rows: r and (CN_S) v of row m, 1/D, r0 and the face
    // velocities u, v of row m
    // rows in flight; a slot is refilled at the end of the step that used it
    constexpr int PF = MODE == CN_RX ? IMPEC_CNRX_PF : IMPEC_CN_PF;
    double Pa[PF] = {}, Pv[PF] = {}, P0[PF] = {}, Pu[PF] = {}, Pw[PF] = {};
    double Pr[PF] = {}, Pp[PF] = {}, Pq[PF] = {};   // CN_RX: r, p, v of the x update
    float Pd[PF] = {};
    auto cellx = [&](const double *p, int m) -> double {   // x: the interior of c, pitch ldx
        const int mc = m < 0 ? 0 : (m > N1 - 1 ? N1 - 1 : m);
        return p[(long)mc * A.ldx + jc];
    };
    auto fetch = [&](int u, int m) {
        Pa[u] = cell(A.r, m);
        if (MODE == CN_S) Pv[u] = cell(A.vin, m);
        if (MODE == CN_R || MODE == CN_RX) Pv[u] = cellx(A.x, m);
        if (MODE != CN_R) Pd[u] = cellf(A.dinv, m);
        if (MODE == CN_S || MODE == CN_P0) P0[u] = (double)cellf(A.r0, m);
        if (MODE == CN_RX) {
            Pr[u] = cell(A.rs, m);
            Pp[u] = cell(A.pin, m);
            Pq[u] = cell(A.vin, m);
        }
        Pu[u] = vel(A.u, m);
        Pw[u] = vel(A.v, m);
    };
    const int mstart = i0 - 1, mend = i1 + 1;
    const int mstop = mstart + (mend - mstart + PF - 1) / PF * PF;
#pragma unroll
    for (int u = 0; u < PF; u++) fetch(u, mstart + u);
    // windows: the operator's input y of rows m-2, m-1, m; the vector of row
    // m-1 (s or p) and r0 of row m-1
    double Y0 = 0., Y1 = 0., Y2 = 0.;
    FaceStream fs;
    fs.start(vel(A.u, mstart - 1), vel(A.v, mstart - 1));
    double wprev = 0., r0prev = 0., bprev = 0.;
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0., s4 = 0.;
    for (int mb = mstart; mb < mstop; mb += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int m = mb + u;
            const double dv = (double)Pd[u], r0v = P0[u];
            fs.push(k, Pu[u], Pw[u]);
            const bool rin = m >= 0 && m < N1;
            // the vector on row m: s, the first direction p = r, the iterate x,
            // or x after the last iteration's update (k_cn_xfinal's arithmetic)
            double wv;
            if (MODE == CN_RX) {
                const double sm = Pr[u] - alpha * Pq[u];
                const double dx_ = alpha * (Pp[u] * dv) + om * (sm * dv);
                wv = A.xzero ? dx_ : Pv[u] + dx_;
                if (rin && m >= i0 && m < i1 && outl) A.xout[(long)m * A.ldx + j] = wv;
            } else {
                wv = MODE == CN_S ? Pa[u] - alpha * Pv[u] : (MODE == CN_R ? Pv[u] : Pa[u]);
            }
            const double yv = (rin && jv) ? ((MODE == CN_R || MODE == CN_RX) ? wv : wv * dv) : 0.;
            if ((MODE == CN_S || MODE == CN_P0) && rin && m >= i0 && m < i1 && outl) {
                if (MODE == CN_S) {
                    s3 += r0v * wv;
                    s4 += wv * wv;
                } else {
                    A.pout[(long)m * N2 + j] = wv;
                    s1 += wv * wv;
                }
            }
            Y0 = Y1; Y1 = Y2; Y2 = yv;
            // row q = m - 1: GetCoefficients_Parabolic with the faces in registers
            const int q = m - 1;
            const CellTensors T = fs.cell();
            if (q >= i0 && q < i1) {
                double a[9];
                cn_coeffs(k, T, q, j, a);
                const double acc = cn_apply(a, Y0, Y1, Y2);
                if (outl) {
                    if (MODE == CN_S) {
                        s0 += acc * wprev;
                        s1 += acc * acc;
                        s2 += r0prev * acc;
                    } else if (MODE == CN_R || MODE == CN_RX) {
                        const double rv = bprev - acc;
                        s0 += rv * rv;
                    } else {
                        A.out[(long)q * N2 + j] = acc;
                        s0 += r0prev * acc;
                    }
                }
            }
            wprev = wv;
            r0prev = r0v;
            bprev = Pa[u];
            // (clamping the refill past mend - 1, as the p-launch does, fetches
            // 3% less here and runs 1-2% slower: measured)
            fetch(u, m + PF);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
        if (MODE == CN_S) {
            s2 += __shfl_xor(s2, o);
            s3 += __shfl_xor(s3, o);
            s4 += __shfl_xor(s4, o);
        }
    }
    if (lane == 0) {
        A.part[wid] = s0;
        A.part[RED_BLOCKS + wid] = s1;
        if (MODE == CN_S) {
            A.part[2 * RED_BLOCKS + wid] = s2;
            A.part[3 * RED_BLOCKS + wid] = s3;
            A.part[4 * RED_BLOCKS + wid] = s4;
        }
    }
}

// The p-launch: row m streams in r, v, 1/D and the face velocities; s and
// y_s = s/D are formed on row m, t = A y_s, r', p' and x on row m-1, and
// v' = A (p'/D) on row m-2.  Two chained operators need a halo of two rows
// and two lanes: lanes 2..61 of a strip are written, rows i0-2 .. i1+1 are
// streamed.  r' goes to a buffer other than r (neighbouring waves read r in
// their halos); x is read and written by its own lane only.
__global__ void __launch_bounds__(64 * CN_WAVES, IMPEC_CNP_MIN_WAVES) k_cn_pstep(CnArgs A) {
    const int lane = threadIdx.x & 63;
    const int wid = xcd_block(blockIdx.x, gridDim.x) * CN_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wid >= A.nwave) return;
    const CnConst &k = A.k;
    const int N1 = k.N1, N2 = k.N2;
    const long L = k.ldv;
    const int strip = wid % A.nstrip, chunk = wid / A.nstrip;
    const int j = strip * CNP_OUT - 2 + lane;
    const bool jv = j >= 0 && j < N2;
    const int jc = j < 0 ? 0 : (j > N2 - 1 ? N2 - 1 : j);
    const bool outl = lane >= 2 && lane < 2 + CNP_OUT && jv;
    const int i0 = chunk * A.R, i1 = i0 + A.R < N1 ? i0 + A.R : N1;
    const double beta = (A.S[S_RHO_NEW] / A.S[S_RHO]) * (A.S[S_ALPHA] / A.S[S_OMEGA]);
    const double om = A.S[S_OMEGA], alpha = A.S[S_ALPHA];
    auto cell = [&](const double *p, int m) -> double {
        const int mc = m < 0 ? 0 : (m > N1 - 1 ? N1 - 1 : m);
        return p[(long)mc * N2 + jc];
    };
    auto cellx = [&](const double *p, int m) -> double {   // x: the interior of c, pitch ldx
        const int mc = m < 0 ? 0 : (m > N1 - 1 ? N1 - 1 : m);
        return p[(long)mc * A.ldx + jc];
    };
    auto cellf = [&](const float *p, int m) -> float {
        const int mc = m < 0 ? 0 : (m > N1 - 1 ? N1 - 1 : m);
        return p[(long)mc * N2 + jc];
    };
    const int B = k.B;
    const int jg = j < -1 ? -1 : (j > N2 ? N2 : j);
    auto vel = [&](const double *p, int m) -> double {
        const int mc = m < -B ? -B : (m > N1 + B ? N1 + B : m);
        return p[(long)(mc + B) * L + (jg + B)];
    };
    // ring: r, v, 1/D, u, v of row m; p and x of row m-1; r0 of row m-2
    constexpr int PF = IMPEC_CNP_PF;
    double Pr[PF] = {}, Pv[PF] = {}, Pu[PF] = {}, Pw[PF] = {}, Pp[PF] = {}, Px[PF] = {}, P0[PF] = {};
    float Pd[PF] = {};
    auto fetch = [&](int u, int m) {
        Pr[u] = cell(A.r, m);
        Pv[u] = cell(A.vin, m);
        Pd[u] = cellf(A.dinv, m);
        Pu[u] = vel(A.u, m);
        Pw[u] = vel(A.v, m);
        // p is used on rows i0-1 .. i1 (the outputs and the matvec's
        // neighbours), x and r0 on the chunk's own rows: other rows re-read
        // one of those (a cache hit) instead of fetching one more
        const int mp = m - 1 < i0 - 1 ? i0 - 1 : (m - 1 > i1 ? i1 : m - 1);
        const int mx = m - 1 < i0 ? i0 : (m - 1 > i1 - 1 ? i1 - 1 : m - 1);
        const int m0 = m - 2 < i0 ? i0 : (m - 2 > i1 - 1 ? i1 - 1 : m - 2);
        Pp[u] = cell(A.pin, mp);
        if (!A.xzero) {
            if (IMPEC_CNP_NT >= 1) {
                const int mc = mx < 0 ? 0 : (mx > N1 - 1 ? N1 - 1 : mx);
                Px[u] = __builtin_nontemporal_load(A.x + (long)mc * A.ldx + jc);
            } else {
                Px[u] = cellx(A.x, mx);
            }
        }
        P0[u] = (double)cellf(A.r0, m0);
    };
    const int mstart = i0 - 2, mend = i1 + 2;
    const int mstop = mstart + (mend - mstart + PF - 1) / PF * PF;
#pragma unroll
    for (int u = 0; u < PF; u++) fetch(u, mstart + u);
    double YS0 = 0., YS1 = 0., YS2 = 0.;   // s/D of rows m-2, m-1, m
    double YP0 = 0., YP1 = 0., YP2 = 0.;   // p'/D of rows m-3, m-2, m-1
    double sprev = 0., vprev = 0., dprev = 0.;   // s, v and 1/D of row m-1
    double aprev[9] = {};                  // coefficients of row m-2
    FaceStream fs;
    fs.start(vel(A.u, mstart - 1), vel(A.v, mstart - 1));
    double s0 = 0., s1 = 0.;
    for (int mb = mstart; mb < mstop; mb += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int m = mb + u;
            fs.push(k, Pu[u], Pw[u]);
            // row m: s = r - alpha v and s/D (the s-launch's arithmetic)
            const bool rin = m >= 0 && m < N1;
            const double dv = (double)Pd[u];
            const double sm = Pr[u] - alpha * Pv[u];
            YS0 = YS1; YS1 = YS2; YS2 = (rin && jv) ? sm * dv : 0.;
            // row q = m - 1: t, r', p', x
            const int q = m - 1;
            const bool qin = q >= 0 && q < N1;
            double a[9];
            cn_coeffs(k, fs.cell(), q, j, a);
            const double t = cn_apply(a, YS0, YS1, YS2);
            const double rq = sprev - om * t;
            const double pq = rq + beta * (Pp[u] - om * vprev);
            if (qin && q >= i0 && q < i1 && outl) {
                const long o = (long)q * N2 + j;
                if (IMPEC_CNP_NT >= 2) __builtin_nontemporal_store(pq, A.pout + o);
                else A.pout[o] = pq;
                A.rout[o] = rq;
                const double dx_ = alpha * (Pp[u] * dprev) + om * (sprev * dprev);
                if (IMPEC_CNP_NT >= 1)
                    __builtin_nontemporal_store(A.xzero ? dx_ : Px[u] + dx_, A.x + (long)q * A.ldx + j);
                else A.x[(long)q * A.ldx + j] = A.xzero ? dx_ : Px[u] + dx_;
                s1 += rq * rq;
            }
            YP0 = YP1; YP1 = YP2; YP2 = (qin && jv) ? pq * dprev : 0.;
            // row q2 = m - 2: v' = A (p'/D)
            const int q2 = m - 2;
            if (q2 >= i0 && q2 < i1) {
                const double acc = cn_apply(aprev, YP0, YP1, YP2);
                if (outl) {
                    A.out[(long)q2 * N2 + j] = acc;
                    s0 += P0[u] * acc;
                }
            }
#pragma unroll
            for (int e = 0; e < 9; e++) aprev[e] = a[e];
            sprev = sm;
            vprev = Pv[u];
            dprev = dv;
            // rows past mend - 1 are never used: the refill re-reads row
            // mend - 1 (a cache hit) instead of fetching the next chunk's rows
            fetch(u, m + PF < mend ? m + PF : mend - 1);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
    }
    if (lane == 0) {
        A.part[wid] = s0;
        A.part[RED_BLOCKS + wid] = s1;
    }
}
// The start of a zero-guess solve in one pass (x0 = 0, so r = r0 = p = b):
// row m streams in c and the face velocities; the right-hand side b and the
// rounded 1/D are formed on row m-1 (k_cn_rhs's arithmetic), the first
// direction's v = A (b/D) on row m-2 (CN_P0's).  Writes b, 1/D and v;
// partials b.v and b.b.  As in the p-launch, lanes 2..61 of a strip and
// rows i0-2 .. i1+1 are streamed for the two chained stencils.
__global__ void __launch_bounds__(64 * CNZ_WAVES) k_cn_start(CnArgs A, const double *__restrict__ c, int lds,
                                                             double *__restrict__ b, float *__restrict__ dinv,
                                                             float *__restrict__ r0) {
    const int lane = threadIdx.x & 63;
    const int wid = xcd_block(blockIdx.x, gridDim.x) * CNZ_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wid >= A.nwave) return;
    const CnConst &k = A.k;
    const int N1 = k.N1, N2 = k.N2, B = k.B;
    const long L = k.ldv;
    const int strip = wid % A.nstrip, chunk = wid / A.nstrip;
    const int j = strip * CNP_OUT - 2 + lane;
    const bool jv = j >= 0 && j < N2;
    const bool outl = lane >= 2 && lane < 2 + CNP_OUT && jv;
    const int i0 = chunk * A.R, i1 = i0 + A.R < N1 ? i0 + A.R : N1;
    const int jcg = j < -B ? -B : (j > N2 + B - 1 ? N2 + B - 1 : j);   // ghosted c column for loads
    auto crow = [&](int m) -> double {
        const int mc = m < -B ? -B : (m > N1 + B - 1 ? N1 + B - 1 : m);
        return c[(long)(mc + B) * lds + (jcg + B)];
    };
    const int jg = j < -1 ? -1 : (j > N2 ? N2 : j);
    auto vel = [&](const double *p, int m) -> double {
        const int mc = m < -B ? -B : (m > N1 + B ? N1 + B : m);
        return p[(long)(mc + B) * L + (jg + B)];
    };
    constexpr int PF = IMPEC_CNP_PF;
    double Pc[PF] = {}, Pu[PF] = {}, Pw[PF] = {};
    auto fetch = [&](int u, int m) {
        Pc[u] = crow(m);
        Pu[u] = vel(A.u, m);
        Pw[u] = vel(A.v, m);
    };
    const int mstart = i0 - 2, mend = i1 + 2;
    const int mstop = mstart + (mend - mstart + PF - 1) / PF * PF;
#pragma unroll
    for (int u = 0; u < PF; u++) fetch(u, mstart + u);
    double C0 = 0., C1 = 0., C2 = 0.;      // c of rows m-2, m-1, m
    double Y0 = 0., Y1 = 0., Y2 = 0.;      // b/D of rows m-3, m-2, m-1
    double bprev = 0.;                     // r0 = b rounded to FP32, of row m-2
    double aprev[9] = {};                  // coefficients of row m-2
    FaceStream fs;
    fs.start(vel(A.u, mstart - 1), vel(A.v, mstart - 1));
    double s0 = 0., s1 = 0., s2 = 0.;
    for (int mb = mstart; mb < mstop; mb += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int m = mb + u;
            fs.push(k, Pu[u], Pw[u]);
            C0 = C1; C1 = C2; C2 = Pc[u];
            // row q = m - 1: b and 1/D
            const int q = m - 1;
            const bool qin = q >= 0 && q < N1;
            const CellTensors T = fs.cell();
            const double bq = cn_rhs_at(k, T, lane_prev(C0), C0, lane_next(C0), lane_prev(C1), C1, lane_next(C1),
                                        lane_prev(C2), C2, lane_next(C2));
            double a[9];
            cn_coeffs(k, T, q, j, a);
            const float dq = (float)(1. / a[ST_C]);
            const float r0q = (float)bq;
            if (qin && q >= i0 && q < i1 && outl) {
                const long o = (long)q * N2 + j;
                b[o] = bq;
                dinv[o] = dq;
                r0[o] = r0q;
                s1 += bq * bq;
                s2 += (double)r0q * bq;
            }
            Y0 = Y1; Y1 = Y2; Y2 = (qin && jv) ? bq * (double)dq : 0.;
            // row q2 = m - 2: v = A (b/D)
            const int q2 = m - 2;
            if (q2 >= i0 && q2 < i1) {
                const double acc = cn_apply(aprev, Y0, Y1, Y2);
                if (outl) {
                    A.out[(long)q2 * N2 + j] = acc;
                    s0 += bprev * acc;
                }
            }
#pragma unroll
            for (int e = 0; e < 9; e++) aprev[e] = a[e];
            bprev = (double)r0q;
            // rows past mend - 1 are never used: the refill re-reads row
            // mend - 1 (a cache hit) instead of fetching the next chunk's rows
            fetch(u, m + PF < mend ? m + PF : mend - 1);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o);
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
    }
    if (lane == 0) {
        A.part[wid] = s0;
        A.part[RED_BLOCKS + wid] = s1;
        A.part[2 * RED_BLOCKS + wid] = s2;
    }
}

static_assert(5 <= PART_VECS, "the s-launch writes five partial vectors");

// The last iteration's x update alone (the p-launch's arithmetic): s = r -
// alpha v, x += alpha p/D + omega s/D.  Once the s-launch's partials show the
// next residual meets the tolerance, the direction and matvec of a full
// p-launch would go unused.
__global__ void k_cn_xfinal(CnArgs A) {
    const int N1 = A.k.N1, N2 = A.k.N2;
    const double alpha = A.S[S_ALPHA], om = A.S[S_OMEGA];
    const int j = blockIdx.x * 64 + threadIdx.x;
    if (j >= N2) return;
    for (int i = blockIdx.y * 4 + threadIdx.y; i < N1; i += gridDim.y * 4) {
        const long o = (long)i * N2 + j;
        const double dv = (double)A.dinv[o];
        const double sm = A.r[o] - alpha * A.vin[o];
        const double dx_ = alpha * (A.pin[o] * dv) + om * (sm * dv);
        double *xp = A.x + (long)i * A.ldx + j;
        *xp = A.xzero ? dx_ : *xp + dx_;
    }
}

// launch shape of the element-wise BiCGStab kernels: 64 x 4 threads over
// (column, row), rows strided so that there are at most CN_VBLOCKS blocks
constexpr int CN_VBLOCKS = 2048;
static dim3 cn_vgrid(int N1, int N2, int cap = CN_VBLOCKS) {
    const int bx = (N2 + 63) / 64;
    int by = (N1 + 3) / 4;
    if ((long)bx * by > cap) by = cap / bx > 0 ? cap / bx : 1;
    return dim3(bx, by);
}
// workgroups of the right-hand side kernel (build-time tunable; at most
// RED_BLOCKS, the partial-sum capacity)
#ifndef IMPEC_CN_RHS_BLOCKS
#define IMPEC_CN_RHS_BLOCKS 2048
#endif
static_assert(IMPEC_CN_RHS_BLOCKS <= RED_BLOCKS, "partials of the right-hand side kernel");

// residual of the iterate x in c's interior: r = b - A x; partials r.r
// (twice).  STORE: also r and r0 = r (the initial guess x0 = c, or a restart
// of the recursion from the current iterate); otherwise the partials alone
// (the true residual that confirms convergence)
template <bool STORE>
__global__ void k_bicg_guess(CnConst k, const double *__restrict__ c, int lds, const double *__restrict__ u,
                             const double *__restrict__ v, const double *__restrict__ b, double *__restrict__ r,
                             float *__restrict__ r0, double *__restrict__ part) {
    const HwGrid g_;
    const int j = g_.bx() * 64 + threadIdx.x;
    double s[2] = {0., 0.};
    if (j < k.N2)
        for (int i = g_.by() * 4 + threadIdx.y; i < k.N1; i += g_.gy() * 4) {
            double a[9];
            cn_coeffs(k, cell_tensors(k, u, v, i, j), i, j, a);
            const double *cc = c + (long)(i + k.B) * lds + (j + k.B);
            const long L = lds;
            const bool we = i > 0, ea = i < k.N1 - 1, so = j > 0, no = j < k.N2 - 1;
            double acc = a[ST_C] * cc[0];
            if (we) acc += a[ST_W] * cc[-L];
            if (we && so) acc += a[ST_SW] * cc[-L - 1];
            if (so) acc += a[ST_S] * cc[-1];
            if (ea && so) acc += a[ST_SE] * cc[L - 1];
            if (ea) acc += a[ST_E] * cc[L];
            if (ea && no) acc += a[ST_NE] * cc[L + 1];
            if (no) acc += a[ST_N] * cc[1];
            if (we && no) acc += a[ST_NW] * cc[-L + 1];
            const long id = (long)i * k.N2 + j;
            const double rv = b[id] - acc;
            const float r0v = (float)rv;   // the shadow residual, rounded
            if (STORE) {
                r[id] = rv;
                r0[id] = r0v;
            }
            s[0] += rv * rv;
            s[1] += (double)r0v * rv;
        }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        const int blk = g_.by() * g_.gx() + g_.bx();
        part[blk] = s[0];
        part[RED_BLOCKS + blk] = s[1];
    }
}

// second pass of the BiCGStab reductions (NV partial vectors) and the scalar
// recurrences:
//  stage 0 (init):     rr = bb = b.b, rho = alpha = omega = 1, rho_new = b.b
//  stage 4 (guess):    rr = r.r, rho_new = r0.r of the initial residual
//  stage 5 (start):    r0.v, b.b, r0.b of a zero-guess start (r = p = b, r0 = b
//                      rounded to FP32): stages 0 and 1 at once
//  stage 1 (p-launch): r0.v, r.r; rho = rho_new; alpha = rho / r0.v; rr = r.r
//  stage 2 (s-launch): t.s, t.t, r0.t, r0.s, s.s; omega = t.s / t.t (0 when
//                      t = 0); rho_new = r0.r of r = s - omega t, as
//                      r0.s - omega r0.t; rr = r.r as s.s - 2 omega t.s + omega^2 t.t
//  stage 6 (check):    the true residual's r.r
// launched with BICG_FIN_THREADS threads: one wave per SIMD, so the loads of
// all NV vectors fit in registers (1024 threads for two vectors, four times
// the loads in flight over the p-launch's 20480 partials, ran 4% slower);
// with pub the workgroup then publishes the scalars for the host
constexpr int BICG_FIN_THREADS = 256;
template <int NV>
__device__ __forceinline__ void bicg_scalars(const double (&v)[NV], int stage, double *__restrict__ S) {
    if (stage == 0) {
        S[S_RR] = v[0];
        S[S_BB] = v[0];
        S[S_RHO] = 1.; S[S_ALPHA] = 1.; S[S_OMEGA] = 1.;
        S[S_RHO_NEW] = v[1];
    } else if (stage == 4) {
        S[S_RR] = v[0];
        S[S_RHO_NEW] = v[1];
    } else if (stage == 5 && NV >= 3) {   // a zero-guess start: v[0] = r0.v, v[1] = b.b, v[2] = r0.b
        S[S_RR] = v[1];
        S[S_BB] = v[1];
        S[S_OMEGA] = 1.;
        S[S_RHO] = v[2];
        S[S_RHO_NEW] = v[2];
        S[S_R0V] = v[0];
        S[S_ALPHA] = v[2] / v[0];
    } else if (stage == 6) {
        S[S_CNRES] = v[0];
    } else if (stage == 1) {
        S[S_R0V] = v[0];
        S[S_RHO] = S[S_RHO_NEW];
        S[S_ALPHA] = S[S_RHO] / v[0];
        S[S_RR] = v[1];
    } else if (NV >= 5) {
        S[S_TS] = v[0];
        S[S_TT] = v[1];
        const double om = v[1] > 0 ? v[0] / v[1] : 0.;
        S[S_OMEGA] = om;
        S[S_RHO_NEW] = v[3] - om * v[2];
        S[S_RR] = v[4] - 2. * om * v[0] + om * om * v[1];
    }
}
template <int NV>
__global__ void __launch_bounds__(BICG_FIN_THREADS) k_bicg_finish(const double *__restrict__ part, int nb,
                                                                        int stage, double *__restrict__ S,
                                                                        PubArgs pub) {
    double v[NV];
    strided_combine_nv<RED_SUM, NV>(part, RED_BLOCKS, nb, v);
    block_reduce<RED_SUM, NV>(v);
    if (threadIdx.x == 0) bicg_scalars<NV>(v, stage, S);
    publish_scalars(S, pub);
}
template <int NV>
static void bicg_finish(hipStream_t s, const double *part, int nb, int stage, double *S,
                        PubArgs pub = PubArgs{nullptr, 0}) {
    hipLaunchKernelGGL(k_bicg_finish<NV>, dim3(1), dim3(BICG_FIN_THREADS), 0, s, part, nb, stage, S, pub);
}

// x = 0 (a zero-guess solve that needed no iteration), pitch ldx
__global__ void k_cn_zero(int N1, int N2, double *__restrict__ x, long ldx) {
    const HwGrid g_;
    const int j = g_.bx() * 64 + threadIdx.x;
    if (j < N2)
        for (int i = g_.by() * 4 + threadIdx.y; i < N1; i += g_.gy() * 4) x[(long)i * ldx + j] = 0.;
}


// matrix dump: nine N1*N2 arrays
__global__ void k_cn_dump(CnConst k, const double *__restrict__ u, const double *__restrict__ v,
                          double *__restrict__ A) {
    const HwGrid g_;
    const int j = g_.bx() * blockDim.x + threadIdx.x;
    const int i = g_.by() * blockDim.y + threadIdx.y;
    if (i >= k.N1 || j >= k.N2) return;
    double a[9];
    cn_coeffs(k, cell_tensors(k, u, v, i, j), i, j, a);
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

// right-hand side (into cn_rhs) and rounded inverse folded diagonal (into
// cn_dg); *nblocks: the workgroups that wrote b.b partials
static int tensors_and_rhs(impec_ctx *ctx, const CnConst &k, int *nblocks = nullptr) {
    const dim3 vg = cn_vgrid(ctx->N1, ctx->N2, IMPEC_CN_RHS_BLOCKS), blk(64, 4);
    hipLaunchKernelGGL(k_cn_rhs, vg, blk, 0, ctx->stream, k, ctx->c, ctx->lds, ctx->u, ctx->v, ctx->cn_rhs, ctx->cn_dg,
                       ctx->part);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    if (nblocks) *nblocks = (int)(vg.x * vg.y);
    return IMPEC_OK;
}

// Solve_Parabolic_cc_to_c: x0 = 0 (AssignRHS_Parabolic, MGSolver.cpp:1375-1392)
// or, with cn_guess = 1, x0 = the current c; stop on
// ||rhs - A x||_2 <= para_eps ||rhs||_2 (recursive residual); solution into
// c's interior.
int cn_solve(impec_ctx *ctx, double dt, int *its_out) {
    const CnConst k = cn_const(ctx, dt);
    const int N1 = ctx->N1, N2 = ctx->N2;
    hipStream_t s = ctx->stream;
    // the solution is formed in place in the interior of c (x0 = c for cn_guess = 1)
    double *b = ctx->cn_rhs, *x = ctx->c + (long)ctx->B * ctx->lds + ctx->B, *r = ctx->cn_r;
    float *r0 = (float *)ctx->cn_r0;   // FP32 shadow residual in the FP64 buffer's first half
    // r, p, v ping-pong: a strip reads its neighbours' old values at the halo
    double *pb[2] = {ctx->cn_p, ctx->cn_y}, *vb[2] = {ctx->cn_v, ctx->cn_z}, *rb[2] = {ctx->cn_r, ctx->cn_r2};
    bool zero = ctx->P.cn_guess != 1;
    const double tol = ctx->P.para_eps;
    const int maxit = ctx->P.max_iterations > 0 ? ctx->P.max_iterations : 10000;
    CnArgs a;
    memset(&a, 0, sizeof(a));
    a.k = k;
    a.u = ctx->u; a.v = ctx->v;
    a.dinv = ctx->cn_dg;
    a.r = zero ? b : r;
    a.r0 = r0;
    a.ldx = ctx->lds;
    a.S = ctx->scal;
    a.part = ctx->part;
    a.nstrip = (N2 + CN_OUT - 1) / CN_OUT;
    {
        // at most IMPEC_CN_TARGET_WAVES wavefronts, the waves the GPU holds at
        // once at this kernel's 4 waves per SIMD; chunks of at least 8 rows
        long R = chunk_rows(N1, a.nstrip, IMPEC_CN_TARGET_WAVES, IMPEC_WAVE_CAP);
        if (R < 8) R = 8;
        if (R > N1) R = N1;
        a.R = (int)R;
    }
    a.nwave = a.nstrip * ((N1 + a.R - 1) / a.R);
    // the p-launch's strips of CNP_OUT columns and its own chunks of rows
    CnArgs a2 = a;
    a2.nstrip = (N2 + CNP_OUT - 1) / CNP_OUT;
    {
        long R = chunk_rows(N1, a2.nstrip, IMPEC_CNP_TARGET_WAVES, IMPEC_WAVE_CAP);
        if (R < 8) R = 8;
        if (R > N1) R = N1;
        a2.R = (int)R;
    }
    a2.nwave = a2.nstrip * ((N1 + a2.R - 1) / a2.R);
    if (a.nwave > RED_BLOCKS || a2.nwave > RED_BLOCKS)
        return set_error(ctx, IMPEC_EINVAL, "dispersion solve: grid too large");
    const dim3 sg((a.nwave + CNS_WAVES - 1) / CNS_WAVES), ssblk(64 * CNS_WAVES), sblk(64 * CN_WAVES);
    const dim3 pg((a2.nwave + CN_WAVES - 1) / CN_WAVES);
    int cur = 0, its = 0, status = IMPEC_OK, st;
    // the first p-launch's p: b for a zero guess, whose first direction is
    // never stored
    const double *pfirst = pb[1];
    double bb, rr;
    if (zero) {
        // x0 = 0: r = r0 = p = b, none of them formed; one pass writes b, 1/D
        // and the first direction's v, and the first p-launch writes x
        // (the start's own chunks of rows: it runs fastest with fewer, longer
        // chunks than the p-launch)
        CnArgs az = a2;
        az.out = vb[1];
        {
            long R = chunk_rows(N1, az.nstrip, IMPEC_CNS_TARGET_WAVES, IMPEC_WAVE_CAP);
            if (R < 8) R = 8;
            if (R > N1) R = N1;
            az.R = (int)R;
        }
        az.nwave = az.nstrip * ((N1 + az.R - 1) / az.R);
        prof_begin(ctx);
        hipLaunchKernelGGL(k_cn_start, dim3((az.nwave + CNZ_WAVES - 1) / CNZ_WAVES), dim3(64 * CNZ_WAVES), 0, s, az,
                           ctx->c, ctx->lds, b,
                           ctx->cn_dg, r0);
        prof_end(ctx, "cn_start");
        bicg_finish<3>(s, ctx->part, az.nwave, 5, ctx->scal, pub_next(ctx));
        IMPEC_HIP_CHECK(ctx, hipGetLastError());
        if ((st = wait_scalars(ctx))) return st;
        bb = ctx->h_scal[S_BB];
        rr = ctx->h_scal[S_RR];
        cur = 1;
        pfirst = b;
    } else {
        // x0 = the current concentration
        prof_begin(ctx);
        int nrb = 0;
        if ((st = tensors_and_rhs(ctx, k, &nrb))) return st;
        prof_end(ctx, "cn_rhs");
        const dim3 vg = cn_vgrid(N1, N2), vblk(64, 4);
        const int nvb = (int)(vg.x * vg.y);
        bicg_finish<2>(s, ctx->part, nrb, 0, ctx->scal);
        hipLaunchKernelGGL(k_bicg_guess<true>, vg, vblk, 0, s, k, ctx->c, ctx->lds, ctx->u, ctx->v, b, r, r0,
                           ctx->part);
        bicg_finish<2>(s, ctx->part, nvb, 4, ctx->scal, pub_next(ctx));
        IMPEC_HIP_CHECK(ctx, hipGetLastError());
        if ((st = wait_scalars(ctx))) return st;
        bb = ctx->h_scal[S_BB];
        rr = ctx->h_scal[S_RR];
        if (!(rr <= tol * tol * bb) && rr == rr) {   // the first direction p = r
            CnArgs ap = a;
            ap.pout = pb[1]; ap.out = vb[1];
            prof_begin(ctx);
            hipLaunchKernelGGL(k_cn_stream<CN_P0>, sg, ssblk, 0, s, ap);
            prof_end(ctx, "cn_apply_p0");
            bicg_finish<2>(s, ctx->part, a.nwave, 1, ctx->scal);
            IMPEC_HIP_CHECK(ctx, hipGetLastError());
            cur = 1;
        }
    }
    // the current residual: b before the first p-launch of a zero-guess solve
    const double *rcur = a.r;
    int rnext = 0, restarts = 0;
    double rt = 0.;   // true ||b - A x||^2 of the final iterate
    const dim3 vg = cn_vgrid(N1, N2), vblk(64, 4);
    const int nvb = (int)(vg.x * vg.y);
    // The last iteration's update of x, when its p-launch was not queued
    // ahead, is left to the true-residual pass (k_cn_stream<CN_RX>), which
    // writes the updated x into the transport temporary t1; t1 then takes
    // c's ghost cells and becomes c.  Contexts whose c and t1 addresses are
    // held elsewhere update x in place (k_cn_xfinal).
    bool x_pending = false;
    CnArgs xpend;
    memset(&xpend, 0, sizeof(xpend));
    for (;;) {
        double rr_mark = rr;   // the recursive residual when it last halved
        int stall = 0;
        bool stalled = false;
        while (!(rr <= tol * tol * bb)) {
            if (its >= maxit) { status = IMPEC_ENOCONV; break; }
            if (!(rr == rr)) { status = IMPEC_ENAN; break; }
            CnArgs as = a;
            as.r = rcur;
            as.vin = vb[cur];
            prof_begin(ctx);
            hipLaunchKernelGGL(k_cn_stream<CN_S>, sg, ssblk, 0, s, as);
            prof_end(ctx, "cn_apply_s");
            bicg_finish<5>(s, ctx->part, a.nwave, 2, ctx->scal, pub_next(ctx));
            IMPEC_HIP_CHECK(ctx, hipGetLastError());
            CnArgs ap = a2;
            ap.r = rcur; ap.pin = its == 0 ? pfirst : pb[cur]; ap.vin = vb[cur]; ap.x = x;
            ap.xzero = zero && its == 0;
            // Lookahead: while the last known residual norm^2 lies more than
            // IMPEC_CN_AHEAD times above the target, this iteration will not
            // end the solve, and its p-launch is queued before the host reads
            // the s-launch's scalars, so the device does not wait for the host.
            // Should it end the solve after all, the p-launch has updated x
            // exactly as k_cn_xfinal would (same expression, same operands);
            // its other outputs are not used.
            const bool ahead = ctx->cn_ahead >= 0. && rr > ctx->cn_ahead * (tol * tol * bb);
            int rn = rnext;
            auto p_launch = [&]() {
                if (rb[rn] == rcur) rn ^= 1;   // r' never overwrites the r it is formed from
                ap.rout = rb[rn];
                ap.pout = pb[cur ^ 1]; ap.out = vb[cur ^ 1];
                prof_begin(ctx);
                hipLaunchKernelGGL(k_cn_pstep, pg, sblk, 0, s, ap);
                prof_end(ctx, "cn_apply_p");
                bicg_finish<2>(s, ctx->part, a2.nwave, 1, ctx->scal);
            };
            if (ahead) p_launch();
            // the next residual's norm from the s-launch: the last iteration updates x only
            if ((st = wait_scalars(ctx))) return st;
            rr = ctx->h_scal[S_RR];
            // a recursion whose residual norm has not halved in
            // IMPEC_CN_STALL_ITS iterations (at the rounding floor, or stuck)
            // is checked against the true residual and restarted like a
            // converged one
            if (rr <= 0.25 * rr_mark) {
                rr_mark = rr;
                stall = 0;
            } else {
                stalled = ++stall >= IMPEC_CN_STALL_ITS;
            }
            if (rr <= tol * tol * bb || !(rr == rr) || stalled) {
                if (!ahead && rr == rr && !ctx->fixed_buffers) {
                    x_pending = true;   // formed by the true-residual pass below
                    xpend = ap;
                } else if (!ahead) {
                    prof_begin(ctx);
                    hipLaunchKernelGGL(k_cn_xfinal, cn_vgrid(N1, N2), dim3(64, 4), 0, s, ap);
                    prof_end(ctx, "cn_xfinal");
                }
                IMPEC_HIP_CHECK(ctx, hipGetLastError());
                its++;
                if (ahead) ctx->stats.cn_ahead_misses++;
                if (stalled) break;
                continue;   // the loop test ends the solve (or reports the non-finite residual)
            }
            if (!ahead) p_launch();
            IMPEC_HIP_CHECK(ctx, hipGetLastError());
            rcur = rb[rn];
            rnext = rn ^ 1;
            cur ^= 1;
            its++;
        }
        if (status != IMPEC_OK) break;
        if (zero && its == 0)   // b already met the tolerance: x = 0 was never written
            hipLaunchKernelGGL(k_cn_zero, cn_vgrid(N1, N2), dim3(64, 4), 0, s, N1, N2, x, (long)ctx->lds);
        // the recursion's residual (s.s - 2 omega t.s + omega^2 t.t from the
        // s-launch's partials) can cancel at convergence: the true residual of
        // the iterate decides, as the pressure solve's does
        const char *check_name = x_pending ? "cn_check_x" : "cn_check";
        prof_begin(ctx);
        if (x_pending) {
            CnArgs ar = a;
            ar.r = b;
            ar.x = x;
            ar.rs = xpend.r; ar.pin = xpend.pin; ar.vin = xpend.vin; ar.xzero = xpend.xzero;
            double *cn = ctx->t1;
            ar.xout = cn + (long)ctx->B * ctx->lds + ctx->B;
            hipLaunchKernelGGL(k_cn_stream<CN_RX>, sg, ssblk, 0, s, ar);
            if ((st = launch_copy_frame(ctx, ctx->c, cn))) return st;
            ctx->t1 = ctx->c;
            ctx->c = cn;
            x = cn + (long)ctx->B * ctx->lds + ctx->B;
            x_pending = false;
        } else {
            CnArgs ar = a;
            ar.r = b;
            ar.x = x;
            hipLaunchKernelGGL(k_cn_stream<CN_R>, sg, ssblk, 0, s, ar);
        }
        bicg_finish<2>(s, ctx->part, a.nwave, 6, ctx->scal, pub_next(ctx));
        prof_end(ctx, check_name);
        IMPEC_HIP_CHECK(ctx, hipGetLastError());
        if ((st = wait_scalars(ctx))) return st;
        rt = ctx->h_scal[S_CNRES];
        if (rt <= tol * tol * bb) break;
        if (!(rt == rt)) { status = IMPEC_ENAN; break; }
        if (restarts >= 3) { status = IMPEC_ENOCONV; break; }
        // restart the recursion from the current iterate: r = r0 = b - A x,
        // the first direction p = r
        restarts++;
        a.r0 = a2.r0 = r0;
        hipLaunchKernelGGL(k_bicg_guess<true>, vg, vblk, 0, s, k, ctx->c, ctx->lds, ctx->u, ctx->v, b, r, r0,
                           ctx->part);
        bicg_finish<2>(s, ctx->part, nvb, 4, ctx->scal);
        CnArgs ap = a;
        ap.r = r;
        ap.pout = pb[1]; ap.out = vb[1];
        hipLaunchKernelGGL(k_cn_stream<CN_P0>, sg, ssblk, 0, s, ap);
        bicg_finish<2>(s, ctx->part, a.nwave, 1, ctx->scal, pub_next(ctx));
        IMPEC_HIP_CHECK(ctx, hipGetLastError());
        if ((st = wait_scalars(ctx))) return st;
        rr = ctx->h_scal[S_RR];
        rcur = r;
        rnext = 1;
        cur = 1;
        zero = false;
        if (its == 0) its = 1;   // the first p-launch reads p = pb[cur], never b
    }
    ctx->stats.last_parabolic_its = its;
    ctx->stats.last_parabolic_relres = bb > 0 && rt > 0 ? sqrt(rt / bb) : 0.;
    if (its_out) *its_out = its;
    if (status == IMPEC_ENOCONV)
        return set_error(ctx, status, "dispersion solve: no convergence after %d iterations, %d restarts (relres %.3e)",
                         its, restarts, ctx->stats.last_parabolic_relres);
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
    hipLaunchKernelGGL(k_cn_dump, gc, blk, 0, ctx->stream, k, ctx->u, ctx->v, dA);
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
