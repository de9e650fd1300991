// k_pressure.hip -- the nu-weighted 9-point pressure operator, a geometric
// multigrid V-cycle on cell-centred Neumann levels, and the FP64 conjugate
// gradient loop that replaces HYPRE StructPCG + StructSMG
// (MGSolver.cpp:268-667).
//
// Formulation.  The reference pins row (1,1) to p = -q(1,1) dA / kappa
// (MGSolver.cpp:517-567).  When the folded Neumann matrix A_N is symmetric
// (nu == 1, or dx == dy with isotropic K) its column sums vanish, so the pinned
// solution equals the solution of A_N x = b~ (b~ = b with sum(b) removed from
// the pinned entry) shifted so that x(1,1) takes the pinned value.  CG then
// runs on the consistent singular Neumann system with a mean-projected
// multigrid preconditioner, which converges in a grid-independent number of
// cycles.  Otherwise the pinned row and column are eliminated and CG runs on
// the remaining rows.  Either way the returned residual is that of the
// reference's pinned system.
#include "impec_device.h"
#include <math.h>
#include <stdlib.h>
#include <type_traits>
#include <vector>

namespace impec {

// ---------------------------------------------------------------------------
// operator row
// ---------------------------------------------------------------------------
enum { ST_C = 0, ST_W, ST_SW, ST_S, ST_SE, ST_E, ST_NE, ST_N, ST_NW };

// Folding of the entries of boundary rows onto in-domain neighbours, in the
// reference's order of edge and corner overwrites (MGSolver.cpp:305-515).
__device__ __forceinline__ void fold_row(int i, int j, int n1, int n2, double a[9]) {
    const bool bi0 = i == 0, bi1 = i == n1 - 1, bj0 = j == 0, bj1 = j == n2 - 1;
    if (!(bi0 | bi1 | bj0 | bj1)) return;
    const double C = a[ST_C], W = a[ST_W], SW = a[ST_SW], S = a[ST_S], SE = a[ST_SE], E = a[ST_E], NE = a[ST_NE],
                 N = a[ST_N], NW = a[ST_NW];
    double o[9];
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

// Folded coefficients of row (i,j) from the per-cell mobility quantity
// (GetWeightedMatrixCoefficientsForCell, MGSolver.cpp:710-841).  w1/w2 point
// at cell (0,0) of arrays with pitch ldw and at least one ghost ring.
template <bool DIAG>
__device__ __forceinline__ void op_row(const OpGeom &g, const double *__restrict__ w1,
                                       const double *__restrict__ w2, int ldw, int i, int j, double a[9]) {
#define W1(a_, b_) w1[(long)(a_) * ldw + (b_)]
#define W2(a_, b_) w2[(long)(a_) * ldw + (b_)]
    const double fxm = face_mob(W1(i - 1, j), W1(i, j), g.harmonic);
    const double fxp = face_mob(W1(i, j), W1(i + 1, j), g.harmonic);
    const double fym = face_mob(W2(i, j - 1), W2(i, j), g.harmonic);
    const double fyp = face_mob(W2(i, j), W2(i, j + 1), g.harmonic);
    // GetParallelCoefficientsForCell (MGSolver.cpp:760-773)
    const double Wp = g.cx * fxm, Ep = g.cx * fxp, Sp = g.cy * fym, Np = g.cy * fyp;
    const double Cp = g.mcy * (fym + fyp) - g.cx * (fxm + fxp);
    if (!DIAG) {
        a[ST_C] = Cp; a[ST_W] = Wp; a[ST_SW] = 0.; a[ST_S] = Sp; a[ST_SE] = 0.;
        a[ST_E] = Ep; a[ST_NE] = 0.; a[ST_N] = Np; a[ST_NW] = 0.;
    } else {
        // GetTransformedMobilityTensor + GetDiagonalCoefficientsForCell (MGSolver.cpp:777-841)
        const double K11pp = vertex_mob(W1(i + 1, j), W1(i, j), W1(i, j + 1), W1(i + 1, j + 1), g.harmonic);
        const double K11mm = vertex_mob(W1(i, j - 1), W1(i - 1, j - 1), W1(i - 1, j), W1(i, j), g.harmonic);
        const double K11mp = vertex_mob(W1(i, j), W1(i - 1, j), W1(i - 1, j + 1), W1(i, j + 1), g.harmonic);
        const double K11pm = vertex_mob(W1(i + 1, j - 1), W1(i, j - 1), W1(i, j), W1(i + 1, j), g.harmonic);
        double K22pp = K11pp, K22mm = K11mm, K22mp = K11mp, K22pm = K11pm;
        if (w2 != w1) {
            K22pp = vertex_mob(W2(i + 1, j), W2(i, j), W2(i, j + 1), W2(i + 1, j + 1), g.harmonic);
            K22mm = vertex_mob(W2(i, j - 1), W2(i - 1, j - 1), W2(i - 1, j), W2(i, j), g.harmonic);
            K22mp = vertex_mob(W2(i, j), W2(i - 1, j), W2(i - 1, j + 1), W2(i, j + 1), g.harmonic);
            K22pm = vertex_mob(W2(i + 1, j - 1), W2(i, j - 1), W2(i, j), W2(i + 1, j), g.harmonic);
        }
        const double M11p = g.s2 * (K11pp * g.cosa2 + K22pp * g.sina2);
        const double M11m = g.s2 * (K11mm * g.cosa2 + K22mm * g.sina2);
        const double M22p = g.s2 * (K22mp * g.cost2 + K11mp * g.sint2);
        const double M22m = g.s2 * (K22pm * g.cost2 + K11pm * g.sint2);
        const double M12pp = g.s2 * (K22pp * g.costsina - K11pp * g.cosasint);
        const double M21mp = g.s2 * (K22mp * g.costsina - K11mp * g.cosasint);
        const double M12mm = g.s2 * (K22mm * g.costsina - K11mm * g.cosasint);
        const double M21pm = g.s2 * (K22pm * g.costsina - K11pm * g.cosasint);
        const double NEd = M11p / g.dxi2, SWd = M11m / g.dxi2, NWd = M22p / g.deta2, SEd = M22m / g.deta2;
        const double Cd = g.mcxi * (M11p + M11m) - g.ceta * (M22p + M22m);
        const double Nd = g.cnd * (M12pp + M21mp);
        const double Wd = g.mcnd * (M21mp + M12mm);
        const double Ed = g.mcnd * (M12pp + M21pm);
        const double Sd = g.cnd * (M12mm + M21pm);
        const double nu = g.nu, om = g.one_m_nu;
        a[ST_C] = nu * Cp + om * Cd;
        a[ST_W] = nu * Wp + om * Wd;
        a[ST_SW] = nu * 0. + om * SWd;
        a[ST_S] = nu * Sp + om * Sd;
        a[ST_SE] = nu * 0. + om * SEd;
        a[ST_E] = nu * Ep + om * Ed;
        a[ST_NE] = nu * 0. + om * NEd;
        a[ST_N] = nu * Np + om * Nd;
        a[ST_NW] = nu * 0. + om * NWd;
    }
    fold_row(i, j, g.n1, g.n2, a);
#undef W1
#undef W2
}

// sum_k a_k x_nb(k) in stencil order; x points at cell (0,0), pitch ld;
// out-of-domain neighbours have zero coefficients after folding and are
// skipped, so the ring contents never enter.
template <bool DIAG>
__device__ __forceinline__ double row_dot(const double a[9], const double *__restrict__ x, int ld, int i, int j,
                                          int n1, int n2) {
    const double *xc = x + (long)i * ld + j;
    double s = a[ST_C] * xc[0];
    const bool w = i > 0, e = i < n1 - 1, so = j > 0, no = j < n2 - 1;
    if (w) s += a[ST_W] * xc[-ld];
    if (DIAG && w && so) s += a[ST_SW] * xc[-ld - 1];
    if (so) s += a[ST_S] * xc[-1];
    if (DIAG && e && so) s += a[ST_SE] * xc[ld - 1];
    if (e) s += a[ST_E] * xc[ld];
    if (DIAG && e && no) s += a[ST_NE] * xc[ld + 1];
    if (no) s += a[ST_N] * xc[1];
    if (DIAG && w && no) s += a[ST_NW] * xc[-ld + 1];
    return s;
}

// sum |a_k x_k| over the row (the rounding scale of row_dot)
template <bool DIAG>
__device__ __forceinline__ double row_absdot(const double a[9], const double *__restrict__ x, int ld, int i, int j,
                                             int n1, int n2) {
    const double *xc = x + (long)i * ld + j;
    double s = fabs(a[ST_C] * xc[0]);
    const bool w = i > 0, e = i < n1 - 1, so = j > 0, no = j < n2 - 1;
    if (w) s += fabs(a[ST_W] * xc[-ld]);
    if (DIAG && w && so) s += fabs(a[ST_SW] * xc[-ld - 1]);
    if (so) s += fabs(a[ST_S] * xc[-1]);
    if (DIAG && e && so) s += fabs(a[ST_SE] * xc[ld - 1]);
    if (e) s += fabs(a[ST_E] * xc[ld]);
    if (DIAG && e && no) s += fabs(a[ST_NE] * xc[ld + 1]);
    if (no) s += fabs(a[ST_N] * xc[1]);
    if (DIAG && w && no) s += fabs(a[ST_NW] * xc[-ld + 1]);
    return s;
}

// off-diagonal part for Gauss-Seidel: b - sum_{k != C} a_k x_nb
template <bool DIAG>
__device__ __forceinline__ double gs_rhs(const double a[9], double b, const double *__restrict__ x, int ld, int i,
                                         int j, int n1, int n2) {
    const double *xc = x + (long)i * ld + j;
    double s = b;
    const bool w = i > 0, e = i < n1 - 1, so = j > 0, no = j < n2 - 1;
    if (w) s -= a[ST_W] * xc[-ld];
    if (DIAG && w && so) s -= a[ST_SW] * xc[-ld - 1];
    if (so) s -= a[ST_S] * xc[-1];
    if (DIAG && e && so) s -= a[ST_SE] * xc[ld - 1];
    if (e) s -= a[ST_E] * xc[ld];
    if (DIAG && e && no) s -= a[ST_NE] * xc[ld + 1];
    if (no) s -= a[ST_N] * xc[1];
    if (DIAG && w && no) s -= a[ST_NW] * xc[-ld + 1];
    return s;
}

// ---------------------------------------------------------------------------
// level data passed by value to the kernels
// ---------------------------------------------------------------------------
struct LevArgs {
    OpGeom g;
    const double *w1, *w2;  // at cell (0,0)
    int ldw;
    int ld;                 // pitch of x/b/r
    // level 0 of a 5-point solve: the face mobilities of op_row, formed once
    // per solve (fx(i,j) between rows i-1 and i, fy(i,j) between columns j-1
    // and j; pitch ld); null elsewhere
    const double *fx, *fy;
};

// op_row<false> from the stored face mobilities when present (the same
// values and arithmetic, without the four divisions), else op_row<false>
__device__ __forceinline__ void op_row5(const LevArgs &L, int i, int j, double a[9]) {
    if (!L.fx) {
        op_row<false>(L.g, L.w1, L.w2, L.ldw, i, j, a);
        return;
    }
    const OpGeom &g = L.g;
    const long o = (long)i * L.ld + j;
    const double fxm = L.fx[o], fxp = L.fx[o + L.ld], fym = L.fy[o], fyp = L.fy[o + 1];
    const double Wp = g.cx * fxm, Ep = g.cx * fxp, Sp = g.cy * fym, Np = g.cy * fyp;
    const double Cp = g.mcy * (fym + fyp) - g.cx * (fxm + fxp);
    a[ST_C] = Cp; a[ST_W] = Wp; a[ST_SW] = 0.; a[ST_S] = Sp; a[ST_SE] = 0.;
    a[ST_E] = Ep; a[ST_NE] = 0.; a[ST_N] = Np; a[ST_NW] = 0.;
    fold_row(i, j, g.n1, g.n2, a);
}

template <bool DIAG>
__device__ __forceinline__ void op_row_L(const LevArgs &L, int i, int j, double a[9]) {
    if (DIAG) op_row<true>(L.g, L.w1, L.w2, L.ldw, i, j, a);
    else op_row5(L, i, j, a);
}

#ifndef IMPEC_NO_FACES
#define IMPEC_NO_FACES 0
#endif
// face mobilities of the level-0 5-point operator (see LevArgs)
template <class G>
__device__ __forceinline__ void faces_body(const G &gr_, LevArgs L, double *__restrict__ fx, double *__restrict__ fy) {
    const int j = gr_.bx() * blockDim.x + threadIdx.x;
    const int i = gr_.by() * blockDim.y + threadIdx.y;
    const int n1 = L.g.n1, n2 = L.g.n2;
#define W1(a_, b_) L.w1[(long)(a_) * L.ldw + (b_)]
#define W2(a_, b_) L.w2[(long)(a_) * L.ldw + (b_)]
    if (i <= n1 && j < n2) fx[(long)i * L.ld + j] = face_mob(W1(i - 1, j), W1(i, j), L.g.harmonic);
    if (i < n1 && j <= n2) fy[(long)i * L.ld + j] = face_mob(W2(i, j - 1), W2(i, j), L.g.harmonic);
#undef W1
#undef W2
}
__global__ void k_faces(LevArgs L, double *__restrict__ fx, double *__restrict__ fy) {
    faces_body(HwGrid{}, L, fx, fy);
}

// ---------------------------------------------------------------------------
// multigrid kernels
// ---------------------------------------------------------------------------
// FINE instantiates a distinct kernel for level 0 so that profiles separate it
template <bool DIAG, bool FINE>
__global__ void k_smooth(LevArgs L, double *__restrict__ x, const double *__restrict__ b, int color) {
    const int tj = blockIdx.x * blockDim.x + threadIdx.x;
    const int ti = blockIdx.y * blockDim.y + threadIdx.y;
    int i, j;
    if (DIAG) {
        i = 2 * ti + (color >> 1);
        j = 2 * tj + (color & 1);
    } else {
        i = ti;
        j = 2 * tj + ((i + color) & 1);
    }
    if (i >= L.g.n1 || j >= L.g.n2) return;
    double a[9];
    op_row<DIAG>(L.g, L.w1, L.w2, L.ldw, i, j, a);
    const double s = gs_rhs<DIAG>(a, b[(long)i * L.ld + j], x, L.ld, i, j, L.g.n1, L.g.n2);
    x[(long)i * L.ld + j] = a[ST_C] != 0. ? s / a[ST_C] : 0.;
}

template <bool DIAG, bool FINE>
__global__ void k_residual(LevArgs L, const double *__restrict__ x, const double *__restrict__ b,
                           double *__restrict__ r) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= L.g.n1 || j >= L.g.n2) return;
    double a[9];
    op_row<DIAG>(L.g, L.w1, L.w2, L.ldw, i, j, a);
    const long o = (long)i * L.ld + j;
    r[o] = b[o] - row_dot<DIAG>(a, x, L.ld, i, j, L.g.n1, L.g.n2);
}

// 1-D weight of fine index f for coarse index I under the cell-centred
// bilinear prolongation with constant extrapolation at the ends
__device__ __forceinline__ double pw(int f, int I, int nc) {
    int I0 = f >> 1;
    int I1 = (f & 1) ? I0 + 1 : I0 - 1;
    if (I0 > nc - 1) I0 = nc - 1;
    if (I1 < 0) I1 = 0;
    if (I1 > nc - 1) I1 = nc - 1;
    return (I0 == I ? 0.75 : 0.) + (I1 == I ? 0.25 : 0.);
}

// b_c = P^T r / 4 (full weighting; the transpose of k_prolong)
__global__ void k_restrict(const double *__restrict__ r, int n1, int n2, int ldf, double *__restrict__ bc, int m1,
                           int m2, int ldc) {
    const int J = blockIdx.x * blockDim.x + threadIdx.x;
    const int I = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= m1 || J >= m2) return;
    double s = 0.;
    for (int f = 2 * I - 1; f <= 2 * I + 2; f++) {
        if (f < 0 || f >= n1) continue;
        const double wi = pw(f, I, m1);
        if (wi == 0.) continue;
        double t = 0.;
        for (int g = 2 * J - 1; g <= 2 * J + 2; g++) {
            if (g < 0 || g >= n2) continue;
            const double wj = pw(g, J, m2);
            if (wj != 0.) t += wj * r[(long)f * ldf + g];
        }
        s += wi * t;
    }
    bc[(long)I * ldc + J] = 0.25 * s;
}

// x_f += P x_c (cell-centred bilinear, constant extrapolation at the ends)
__global__ void k_prolong(double *__restrict__ x, int n1, int n2, int ldf, const double *__restrict__ xc, int m1,
                          int m2, int ldc) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= n1 || j >= n2) return;
    int I0 = i >> 1, I1 = (i & 1) ? I0 + 1 : I0 - 1;
    int J0 = j >> 1, J1 = (j & 1) ? J0 + 1 : J0 - 1;
    I0 = I0 > m1 - 1 ? m1 - 1 : I0;
    J0 = J0 > m2 - 1 ? m2 - 1 : J0;
    I1 = I1 < 0 ? 0 : (I1 > m1 - 1 ? m1 - 1 : I1);
    J1 = J1 < 0 ? 0 : (J1 > m2 - 1 ? m2 - 1 : J1);
    const double e = 0.5625 * xc[(long)I0 * ldc + J0] + 0.1875 * xc[(long)I0 * ldc + J1] +
                     0.1875 * xc[(long)I1 * ldc + J0] + 0.0625 * xc[(long)I1 * ldc + J1];
    x[(long)i * ldf + j] += e;
}

__global__ void k_fill_zero(double *__restrict__ a, long n) {
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) a[k] = 0.;
}

// coarse mobility quantity: mean over the existing fine cells of each 2x2
// block, on the interior and (clamped) ghost ring of the coarse level; wcf
// (optional, same layout) receives the interior rounded to FP32 for the
// mixed-precision cycle
template <class G>
__device__ __forceinline__ void coarsen_w_body(const G &gr_, const double *__restrict__ wf, int ldf, int n1, int n2, double *__restrict__ wc, int ldc,
                            int m1, int m2, float *__restrict__ wcf = nullptr) {
    const int J = gr_.bx() * blockDim.x + threadIdx.x - 1;
    const int I = gr_.by() * blockDim.y + threadIdx.y - 1;
    if (I > m1 || J > m2) return;
    const int Ic = I < 0 ? 0 : (I >= m1 ? m1 - 1 : I);
    const int Jc = J < 0 ? 0 : (J >= m2 ? m2 - 1 : J);
    double s = 0.;
    int cnt = 0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const int f = 2 * Ic + a, g = 2 * Jc + b;
            if (f < n1 && g < n2) {
                s += wf[(long)f * ldf + g];
                cnt++;
            }
        }
    const double v = s / cnt;
    wc[(long)I * ldc + J] = v;
    if (wcf && I >= 0 && I < m1 && J >= 0 && J < m2) wcf[(long)I * ldc + J] = (float)v;
}
__global__ void k_coarsen_w(const double *__restrict__ wf, int ldf, int n1, int n2, double *__restrict__ wc, int ldc,
                            int m1, int m2, float *__restrict__ wcf) {
    coarsen_w_body(HwGrid{}, wf, ldf, n1, n2, wc, ldc, m1, m2, wcf);
}

// Gaussian elimination with partial pivoting, then back substitution, of a
// bordered system of n = N cells: M holds the (N+1) x (N+1) matrix and, in
// column N+1, the right-hand side; sol receives the N+1 unknowns.  The sizes
// are compile-time constants and every loop is unrolled, so M stays in
// registers (with a run-time size it lives in scratch memory and each of the
// elimination's dependent accesses is a memory round trip).  The pivot row is
// exchanged by selects over the candidate rows, which keeps every index
// static.  FUSED: the updates rounded once, as fused multiply-adds; otherwise
// as a rounded product and a rounded difference.
template <int N, bool FUSED>
__device__ __forceinline__ void bordered_solve(double (&M)[N + 1][N + 2], double (&sol)[N + 1]) {
    constexpr int m = N + 1;
#pragma unroll
    for (int col = 0; col < m; col++) {
        int piv = col;
        double best = fabs(M[col][col]);
#pragma unroll
        for (int r = col + 1; r < m; r++) {
            const double a = fabs(M[r][col]);
            if (a > best) { piv = r; best = a; }
        }
#pragma unroll
        for (int r = col + 1; r < m; r++) {
            const bool sw = piv == r;
#pragma unroll
            for (int q = 0; q <= m; q++) {
                const double a = M[col][q], b = M[r][q];
                M[col][q] = sw ? b : a;
                M[r][q] = sw ? a : b;
            }
        }
        const double d = M[col][col];
        if (d == 0.) continue;
#pragma unroll
        for (int r = col + 1; r < m; r++) {
            const double f = M[r][col] / d;
#pragma unroll
            for (int q = col; q <= m; q++) M[r][q] = FUSED ? fma(-f, M[col][q], M[r][q]) : M[r][q] - f * M[col][q];
        }
    }
#pragma unroll
    for (int r = m - 1; r >= 0; r--) {
        double t = M[r][m];
#pragma unroll
        for (int q = r + 1; q < m; q++) t = FUSED ? fma(-M[r][q], sol[q], t) : t - M[r][q] * sol[q];
        sol[r] = M[r][r] != 0. ? t / M[r][r] : 0.;
    }
}

// coarsest level of N1 x N2 <= 2 x 2 cells: exact solve of the singular
// Neumann system bordered by the zero-mean condition,
// [A 1; 1^T 0][x; l] = [b; 0], on one thread (x, b stored as T; the
// elimination is FP64 in both precisions)
template <int N1, int N2, bool DIAG, typename T>
__device__ __forceinline__ void coarsest_solve(const LevArgs &L, T *__restrict__ x, const T *__restrict__ b) {
    constexpr int n = N1 * N2, m = n + 1;
    const int dI[9] = {0, -1, -1, 0, 1, 1, 1, 0, -1}, dJ[9] = {0, 0, -1, -1, -1, 0, 1, 1, 1};
    double M[m][m + 1];
#pragma unroll
    for (int r = 0; r < m; r++)
#pragma unroll
        for (int q = 0; q <= m; q++) M[r][q] = 0.;
#pragma unroll
    for (int i = 0; i < N1; i++)
#pragma unroll
        for (int j = 0; j < N2; j++) {
            const int row = i * N2 + j;
            double a[9];
            op_row<DIAG>(L.g, L.w1, L.w2, L.ldw, i, j, a);
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const int ii = i + dI[k], jj = j + dJ[k];
                if (ii < 0 || jj < 0 || ii >= N1 || jj >= N2) continue;
                M[row][ii * N2 + jj] += a[k];
            }
            M[row][n] = 1.;
            M[n][row] = 1.;
            M[row][m] = (double)b[(long)i * L.ld + j];
        }
    double sol[m];
    bordered_solve<n, false>(M, sol);
#pragma unroll
    for (int i = 0; i < N1; i++)
#pragma unroll
        for (int j = 0; j < N2; j++) x[(long)i * L.ld + j] = (T)sol[i * N2 + j];
}
template <bool DIAG, typename T = double, class G>
__device__ __forceinline__ void coarsest_body(const G &gr_, LevArgs L, T *__restrict__ x, const T *__restrict__ b) {
    if (threadIdx.x != 0 || gr_.bx() != 0) return;
    const int n1 = L.g.n1, n2 = L.g.n2;
    // the hierarchy is coarsened until both sides are at most 2
    if (n1 == 2 && n2 == 2) coarsest_solve<2, 2, DIAG>(L, x, b);
    else if (n1 == 2 && n2 == 1) coarsest_solve<2, 1, DIAG>(L, x, b);
    else if (n1 == 1 && n2 == 2) coarsest_solve<1, 2, DIAG>(L, x, b);
    else if (n1 == 1 && n2 == 1) coarsest_solve<1, 1, DIAG>(L, x, b);
}
template <bool DIAG, typename T = double>
__global__ void k_coarsest(LevArgs L, T *__restrict__ x, const T *__restrict__ b) {
    coarsest_body<DIAG, T>(HwGrid{}, L, x, b);
}

// ---------------------------------------------------------------------------
// conjugate-gradient vector kernels (level-0 ring-1 layout; x lives in p)
// ---------------------------------------------------------------------------
struct CgArgs {
    int n1, n2, ld;     // ring-1 layout pitch
    int ldp, B;         // pressure array pitch and ghost width
    long n;             // n1 * n2
    int pin_i, pin_j;
    int shift;          // consistent-Neumann formulation
    double ppin;        // pinned value
    double dpin;        // kappa / dA
    double bsum;        // sum of b = -q
};

// cell (i, j) of flat index id < n1 * n2 < 2^31 (checked at creation): 32-bit
// division, not the 64-bit sequence
__device__ __forceinline__ void cg_cell(const CgArgs &c, long id, int &i, int &j) {
    const int k = (int)id;
    i = k / c.n2;
    j = k - i * c.n2;
}

// r = b - A_N x with b = -q; pin handling; partials (r.r, sum r).  With r32
// (the FP32-vector CG's restart) the residual is stored rounded to FP32 in r32
// instead of r, the correction e is cleared, and the partials are those of the
// FP64 residual.
template <bool DIAG, class G>
__device__ __forceinline__ void cg_init_body(const G &gr_, LevArgs L, CgArgs c, double *__restrict__ p, const double *__restrict__ q,
                          double *__restrict__ r, double *__restrict__ part, float *__restrict__ r32 = nullptr,
                          float *__restrict__ e32 = nullptr) {
    double s[2] = {0., 0.};
    double *x = p + (long)c.B * c.ldp + c.B;
    for (long id = (long)gr_.bx() * blockDim.x + threadIdx.x; id < c.n; id += (long)gr_.gx() * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        double a[9];
        op_row_L<DIAG>(L, i, j, a);
        double xv[9];
        (void)xv;
        double bij = -q[id];
        const bool pin = i == c.pin_i && j == c.pin_j;
        double rv;
        if (!c.shift) {
            // x_pin = ppin for every neighbour read
            const double *xc = x + (long)i * c.ldp + j;
            double acc = 0.;
            const int dI[9] = {0, -1, -1, 0, 1, 1, 1, 0, -1}, dJ[9] = {0, 0, -1, -1, -1, 0, 1, 1, 1};
            for (int k = 0; k < 9; k++) {
                const int ii = i + dI[k], jj = j + dJ[k];
                if (ii < 0 || jj < 0 || ii >= c.n1 || jj >= c.n2) continue;
                if (!DIAG && (k == ST_SW || k == ST_SE || k == ST_NE || k == ST_NW)) continue;
                const double xn = (ii == c.pin_i && jj == c.pin_j) ? c.ppin : xc[dI[k] * c.ldp + dJ[k]];
                acc += a[k] * xn;
            }
            rv = pin ? 0. : bij - acc;
        } else {
            rv = bij - row_dot<DIAG>(a, x, c.ldp, i, j, c.n1, c.n2);
            if (pin) rv -= c.bsum;
        }
        const long o = (long)(i + 1) * c.ld + (j + 1);
        if (r32) {
            r32[o] = (float)rv;
            e32[o] = 0.f;
        } else {
            r[o] = rv;
        }
        s[0] += rv * rv;
        s[1] += rv;
    }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        part[gr_.bx()] = s[0];
        part[RED_BLOCKS + gr_.bx()] = s[1];
    }
}
template <bool DIAG>
__global__ void k_cg_init(LevArgs L, CgArgs c, double *__restrict__ p, const double *__restrict__ q,
                          double *__restrict__ r, double *__restrict__ part) {
    cg_init_body<DIAG>(HwGrid{}, L, c, p, q, r, part);
}
// the FP32-vector CG's restart from the iterate: k_cg_init into r32, e cleared
__global__ void k_cg_init32(LevArgs L, CgArgs c, double *__restrict__ p, const double *__restrict__ q,
                            float *__restrict__ r32, float *__restrict__ e32, double *__restrict__ part) {
    cg_init_body<false>(HwGrid{}, L, c, p, q, nullptr, part, r32, e32);
}

// k_cg_init of the shift formulation from the zero iterate (zero initial
// guess): p = 0 on the interior and r = b = -q (minus sum(b) at the pinned
// cell), with no operator application (A 0 = 0); the same partials in the same
// order
__global__ void k_cg_init_zero(CgArgs c, double *__restrict__ p, const double *__restrict__ q,
                               double *__restrict__ r, double *__restrict__ part) {
    double s[2] = {0., 0.};
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += (long)gridDim.x * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        p[(long)(i + c.B) * c.ldp + (j + c.B)] = 0.;
        double rv = -q[id];
        if (i == c.pin_i && j == c.pin_j) rv -= c.bsum;
        r[(long)(i + 1) * c.ld + (j + 1)] = rv;
        s[0] += rv * rv;
        s[1] += rv;
    }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        part[blockIdx.x] = s[0];
        part[RED_BLOCKS + blockIdx.x] = s[1];
    }
}

// Warm starts from the previous solutions: p = p_n + alpha d + gamma e, with
// d the network's increment (learned guess) and e = p_n - p_(n-1) the last
// step's change.  alpha and gamma minimise the 2-norm of the consistent
// residual r0 - alpha A_N d - gamma A_N e over the directions present -- the
// quantity the stopping test measures -- so the guess's residual is never
// larger than that of p_n (r0 = b~ - A_N p_n, b~ the consistent right-hand
// side; A_N e = A_N p_n - A_N p_(n-1)).
// Partials: (A_N d.r0, A_N e.r0, A_N d.A_N d, A_N d.A_N e, A_N e.A_N e).
__global__ void k_guess_dots(LevArgs L, CgArgs c, const double *__restrict__ p, const double *__restrict__ pold,
                             const double *__restrict__ q, const double *__restrict__ d, double *__restrict__ part) {
    double s[5] = {0., 0., 0., 0., 0.};
    const double *x = p + (long)c.B * c.ldp + c.B;
    const double *xo = pold ? pold + (long)c.B * c.ldp + c.B : nullptr;
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += (long)gridDim.x * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        double a[9];
        op_row5(L, i, j, a);
        const double Ap = row_dot<false>(a, x, c.ldp, i, j, c.n1, c.n2);
        double r0 = -q[id] - Ap;
        if (i == c.pin_i && j == c.pin_j) r0 -= c.bsum;
        const double Ad = d ? row_dot<false>(a, d, c.ld, i, j, c.n1, c.n2) : 0.;
        const double Ae = xo ? Ap - row_dot<false>(a, xo, c.ldp, i, j, c.n1, c.n2) : 0.;
        s[0] += Ad * r0;
        s[1] += Ae * r0;
        s[2] += Ad * Ad;
        s[3] += Ad * Ae;
        s[4] += Ae * Ae;
    }
    static_assert(5 <= PART_VECS, "five partial-sum vectors");
    block_reduce<RED_SUM, 5>(s);
    if (tid_linear() == 0)
        for (int k = 0; k < 5; k++) part[k * RED_BLOCKS + blockIdx.x] = s[k];
}

// the minimal-residual step lengths from the five products in S[S_G0..]:
// the 2 x 2 normal equations when both directions are present (the better
// single direction when they are nearly parallel), else the 1-D minimiser
__global__ void k_guess_solve(double *__restrict__ S, int use_d, int use_e) {
    const double b0 = S[S_G0], b1 = S[S_G0 + 1], m00 = S[S_G0 + 2], m01 = S[S_G0 + 3], m11 = S[S_G0 + 4];
    // a direction with non-finite products (a diverged network's increment)
    // takes no part in the guess
    if (!isfinite(b0) || !isfinite(m00) || !isfinite(m01)) use_d = 0;
    if (!isfinite(b1) || !isfinite(m11) || (use_d && !isfinite(m01))) use_e = 0;
    const double a1 = m00 > 0. ? b0 / m00 : 0., g1 = m11 > 0. ? b1 / m11 : 0.;
    double a = use_d ? a1 : 0., g = use_e ? g1 : 0.;
    if (use_d && use_e) {
        const double det = m00 * m11 - m01 * m01;
        if (det > 1e-12 * m00 * m11) {
            a = (b0 * m11 - b1 * m01) / det;
            g = (m00 * b1 - m01 * b0) / det;
        } else if (a1 * b0 >= g1 * b1) {   // residual decreases b0^2/m00 and b1^2/m11
            g = 0.;
        } else {
            a = 0.;
        }
    }
    S[S_GUESS_ALPHA] = a;
    S[S_GUESS_GAMMA] = g;
}

// p += alpha d + gamma (p - p_old) with the step lengths of k_guess_solve, or
// unit steps (project = 0: no consistent residual to minimise)
__global__ void k_guess_apply(CgArgs c, double *__restrict__ p, const double *__restrict__ pold,
                              const double *__restrict__ d, double *__restrict__ S, int project) {
    const double alpha = d ? (project ? S[S_GUESS_ALPHA] : 1.) : 0.;
    const double gamma = pold ? (project ? S[S_GUESS_GAMMA] : 1.) : 0.;
    if (!project && blockIdx.x == 0 && threadIdx.x == 0) {
        S[S_GUESS_ALPHA] = alpha;
        S[S_GUESS_GAMMA] = gamma;
    }
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += (long)gridDim.x * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        const long o = (long)(i + c.B) * c.ldp + (j + c.B);
        double v = p[o];
        if (d && alpha != 0.) v += alpha * d[(long)i * c.ld + j];   // 0 x inf would poison p
        if (pold && gamma != 0.) v += gamma * (p[o] - pold[o]);
        p[o] = v;
    }
}

// mg.b = r - mean(r) (the mean comes from S_RSUM)
__global__ void k_cg_proj_in(CgArgs c, const double *__restrict__ r, double *__restrict__ b0,
                             const double *__restrict__ S) {
    const double mean = S[S_RSUM] / (double)c.n;
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += (long)gridDim.x * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        const long o = (long)(i + 1) * c.ld + (j + 1);
        b0[o] = r[o] - mean;
    }
}

// partial sum of the V-cycle output
__global__ void k_sum(CgArgs c, const double *__restrict__ x0, double *__restrict__ part) {
    double s[1] = {0.};
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += (long)gridDim.x * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        s[0] += x0[(long)(i + 1) * c.ld + (j + 1)];
    }
    block_reduce<RED_SUM, 1>(s);
    if (tid_linear() == 0) part[blockIdx.x] = s[0];
}

// z = x0 - mean(x0) (pin entry zeroed in the eliminated formulation); r.z partials
__global__ void k_cg_z(CgArgs c, const double *__restrict__ x0, const double *__restrict__ r, double *__restrict__ z,
                       const double *__restrict__ S, double *__restrict__ part) {
    const double mean = S[S_ZSUM] / (double)c.n;
    double s[1] = {0.};
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += (long)gridDim.x * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        const long o = (long)(i + 1) * c.ld + (j + 1);
        double zv = x0[o] - mean;
        if (!c.shift && i == c.pin_i && j == c.pin_j) zv = 0.;
        z[o] = zv;
        s[0] += r[o] * zv;
    }
    block_reduce<RED_SUM, 1>(s);
    if (tid_linear() == 0) part[blockIdx.x] = s[0];
}

// alpha = rho / (d.q).  A recursion that has broken down (d.q zero or a
// non-finite scalar, which happens when it is driven below the FP64 rounding
// floor of the residual) takes no step: p and r stay as they are, and the
// solve loop restarts the recursion from the true residual.
__device__ __forceinline__ double cg_alpha(const double *__restrict__ S) {
    const double dq = S[S_DQ], a = S[S_RHO] / dq;
    return (fabs(dq) > 0. && isfinite(a)) ? a : 0.;
}
__host__ __device__ __forceinline__ bool cg_broke_down(const double *S) {
    const double dq = S[S_DQ], a = S[S_RHO] / dq;
    return !(fabs(dq) > 0.) || !isfinite(a);
}

// d = z + beta d (beta = 0 on the first iteration)
__global__ void k_cg_dir(CgArgs c, const double *__restrict__ z, double *__restrict__ d, const double *__restrict__ S,
                         int first) {
    const double beta = first ? 0. : S[S_RHO] / S[S_RHO_OLD];
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += (long)gridDim.x * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        const long o = (long)(i + 1) * c.ld + (j + 1);
        d[o] = first ? z[o] : z[o] + beta * d[o];
    }
}

// q = A_N d (pin row zeroed in the eliminated formulation); d.q partials
template <bool DIAG>
__global__ void k_cg_matvec(LevArgs L, CgArgs c, const double *__restrict__ d, double *__restrict__ qv,
                            double *__restrict__ part) {
    double s[1] = {0.};
    const double *d0 = d + c.ld + 1;
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += (long)gridDim.x * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        double a[9];
        op_row_L<DIAG>(L, i, j, a);
        double y = row_dot<DIAG>(a, d0, c.ld, i, j, c.n1, c.n2);
        if (!c.shift && i == c.pin_i && j == c.pin_j) y = 0.;
        const long o = (long)(i + 1) * c.ld + (j + 1);
        qv[o] = y;
        s[0] += d0[(long)i * c.ld + j] * y;
    }
    block_reduce<RED_SUM, 1>(s);
    if (tid_linear() == 0) part[blockIdx.x] = s[0];
}

// x += alpha d, r -= alpha q; partials (r.r, sum r)
__global__ void k_cg_update(CgArgs c, double *__restrict__ p, const double *__restrict__ d, double *__restrict__ r,
                            const double *__restrict__ qv, const double *__restrict__ S, double *__restrict__ part) {
    const double alpha = cg_alpha(S);
    double s[2] = {0., 0.};
    for (long id = (long)blockIdx.x * blockDim.x + threadIdx.x; id < c.n; id += (long)gridDim.x * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        const long o = (long)(i + 1) * c.ld + (j + 1);
        double *xp = p + (long)(i + c.B) * c.ldp + (j + c.B);
        double rv = r[o];
        if (alpha != 0.) {   // no step: a broken-down direction may hold non-finite values
            *xp = *xp + alpha * d[o];
            rv = rv - alpha * qv[o];
            r[o] = rv;
        }
        s[0] += rv * rv;
        s[1] += rv;
    }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        part[blockIdx.x] = s[0];
        part[RED_BLOCKS + blockIdx.x] = s[1];
    }
}

__global__ void k_copy_scal(double *__restrict__ S, int dst, int src) { S[dst] = S[src]; }

template <class G>
__device__ __forceinline__ void read_pin_body(const G &gr_, CgArgs c, const double *__restrict__ p, double *__restrict__ S) {
    S[S_XPIN] = p[(long)(c.pin_i + c.B) * c.ldp + (c.pin_j + c.B)];
}
__global__ void k_read_pin(CgArgs c, const double *__restrict__ p, double *__restrict__ S) {
    read_pin_body(HwGrid{}, c, p, S);
}

// x += ppin - x_pin on the interior, x_pin = ppin exactly
template <class G>
__device__ __forceinline__ void shift_body(const G &gr_, CgArgs c, double *__restrict__ p, const double *__restrict__ S) {
    const double sh = c.ppin - S[S_XPIN];
    for (long id = (long)gr_.bx() * blockDim.x + threadIdx.x; id < c.n; id += (long)gr_.gx() * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        double *xp = p + (long)(i + c.B) * c.ldp + (j + c.B);
        *xp = (i == c.pin_i && j == c.pin_j) ? c.ppin : *xp + sh;
    }
}
__global__ void k_shift(CgArgs c, double *__restrict__ p, const double *__restrict__ S) {
    shift_body(HwGrid{}, c, p, S);
}

// ||b - A x||^2 of the reference's pinned system
template <bool DIAG, class G>
__device__ __forceinline__ void true_residual_body(const G &gr_, LevArgs L, CgArgs c, const double *__restrict__ p, const double *__restrict__ q,
                                double *__restrict__ part) {
    // partials of ||b - A p||^2 and of || |b| + |A| |p| ||^2 (the rounding scale)
    double s[2] = {0., 0.};
    const double *x = p + (long)c.B * c.ldp + c.B;
    for (long id = (long)gr_.bx() * blockDim.x + threadIdx.x; id < c.n; id += (long)gr_.gx() * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        double rv, mag;
        if (i == c.pin_i && j == c.pin_j) {
            rv = -q[id] - c.dpin * x[(long)i * c.ldp + j];
            mag = fabs(q[id]) + fabs(c.dpin * x[(long)i * c.ldp + j]);
        } else {
            double a[9];
            op_row_L<DIAG>(L, i, j, a);
            rv = -q[id] - row_dot<DIAG>(a, x, c.ldp, i, j, c.n1, c.n2);
            mag = fabs(q[id]) + row_absdot<DIAG>(a, x, c.ldp, i, j, c.n1, c.n2);
        }
        s[0] += rv * rv;
        s[1] += mag * mag;
    }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        part[gr_.bx()] = s[0];
        part[RED_BLOCKS + gr_.bx()] = s[1];
    }
}
template <bool DIAG>
__global__ void k_true_residual(LevArgs L, CgArgs c, const double *__restrict__ p, const double *__restrict__ q,
                                double *__restrict__ part) {
    true_residual_body<DIAG>(HwGrid{}, L, c, p, q, part);
}

// operator dump: nine N1*N2 arrays including the pinned row (MGSolver.cpp:517-567)
template <bool DIAG>
__global__ void k_op_dump(LevArgs L, int pin_i, int pin_j, double dpin, double *__restrict__ A) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const int n1 = L.g.n1, n2 = L.g.n2;
    if (i >= n1 || j >= n2) return;
    double a[9];
    op_row<DIAG>(L.g, L.w1, L.w2, L.ldw, i, j, a);
    if (i == pin_i && j == pin_j) {
        for (int k = 0; k < 9; k++) a[k] = 0.;
        a[ST_C] = dpin;
    }
    const long M = (long)n1 * n2;
    for (int k = 0; k < 9; k++) A[k * M + (long)i * n2 + j] = a[k];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static OpGeom make_geom(const impec_ctx *ctx, int n1, int n2, double dx, double dy) {
    OpGeom g;
    g.n1 = n1; g.n2 = n2;
    g.harmonic = ctx->P.average_procedure == IMPEC_AVG_HARMONIC;
    g.nu = ctx->P.nu;
    g.one_m_nu = 1 - ctx->P.nu;
    g.diag = ctx->P.nu != 1.;
    const double dxi = dx / cos(ctx->theta), deta = dy / cos(ctx->alpha);
    g.cx = 1. / (dx * dx);
    g.cy = 1. / (dy * dy);
    g.mcy = -1. / (dy * dy);
    g.s2 = ctx->secamt2; g.cosa2 = ctx->cosa2; g.sina2 = ctx->sina2; g.cost2 = ctx->cost2; g.sint2 = ctx->sint2;
    g.costsina = ctx->costsina; g.cosasint = ctx->cosasint;
    g.dxi2 = dxi * dxi;
    g.deta2 = deta * deta;
    g.mcxi = -1. / (dxi * dxi);
    g.ceta = 1. / (deta * deta);
    g.cnd = 1. / (deta * dxi);
    g.mcnd = -1. / (deta * dxi);
    return g;
}

static LevArgs lev_args(const impec_ctx *ctx, int l) {
    LevArgs a;
    const Level &L = ctx->lev[l];
    a.g = L.g;
    a.ld = L.ld;
    a.fx = a.fy = nullptr;
    if (l == 0) {
        a.ldw = ctx->lds;
        a.w1 = ctx->w + (long)ctx->B * ctx->lds + ctx->B;
        const bool aniso = ctx->K22 && ctx->K22 != ctx->K11;
        a.w2 = aniso ? ctx->w22 + (long)ctx->B * ctx->lds + ctx->B : a.w1;
    } else {
        a.ldw = L.ld;
        a.w1 = L.w + L.ld + 1;
        a.w2 = a.w1;
    }
    return a;
}

static int round_up(int x, int m) { return (x + m - 1) / m * m; }

int pressure_setup_levels(impec_ctx *ctx) {
    int n1 = ctx->N1, n2 = ctx->N2;
    double dx = ctx->dx, dy = ctx->dy;
    int l = 0;
    for (;;) {
        Level &L = ctx->lev[l];
        memset(&L, 0, sizeof(L));
        L.n1 = n1; L.n2 = n2; L.dx = dx; L.dy = dy;
        L.ld = round_up(n2 + 2, 16);
        L.g = make_geom(ctx, n1, n2, dx, dy);
        const size_t bytes = sizeof(double) * (size_t)(n1 + 2) * L.ld;
        if (hipMalloc(&L.x, bytes) != hipSuccess || hipMalloc(&L.b, bytes) != hipSuccess ||
            hipMalloc(&L.r, bytes) != hipSuccess)
            return set_error(ctx, IMPEC_ENOMEM, "multigrid level %d allocation failed", l);
        hipMemset(L.x, 0, bytes);
        hipMemset(L.b, 0, bytes);
        hipMemset(L.r, 0, bytes);
        if (l > 0) {
            if (hipMalloc(&L.w, bytes) != hipSuccess)
                return set_error(ctx, IMPEC_ENOMEM, "multigrid level %d allocation failed", l);
            hipMemset(L.w, 0, bytes);
        }
        ctx->nlev = ++l;
        if ((n1 <= 2 && n2 <= 2) || l >= MAX_LEVELS) break;
        n1 = (n1 + 1) / 2;
        n2 = (n2 + 1) / 2;
        dx *= 2;
        dy *= 2;
    }
    // CG vectors and the face mobilities on the level-0 layout
    const size_t bytes0 = sizeof(double) * (size_t)(ctx->N1 + 2) * ctx->lev[0].ld;
    double **vecs[6] = {&ctx->pr, &ctx->pz, &ctx->pd, &ctx->pq, &ctx->fx, &ctx->fy};
    for (int k = 0; k < 6; k++) {
        if (hipMalloc(vecs[k], bytes0) != hipSuccess) return set_error(ctx, IMPEC_ENOMEM, "CG vector allocation failed");
        hipMemset(*vecs[k], 0, bytes0);
    }
    return IMPEC_OK;
}

void pressure_free_levels(impec_ctx *ctx) {
    for (int l = 0; l < ctx->nlev; l++) {
        Level &L = ctx->lev[l];
        hipFree(L.x); hipFree(L.b); hipFree(L.r);
        hipFree(L.wf); hipFree(L.wf2); hipFree(L.xf); hipFree(L.bf); hipFree(L.rf);
        if (L.w) hipFree(L.w);
    }
    ctx->nlev = 0;
    hipFree(ctx->pr); hipFree(ctx->pz); hipFree(ctx->pd); hipFree(ctx->pq);
    hipFree(ctx->fx); hipFree(ctx->fy);
    float **v32[7] = {&ctx->pr32, &ctx->pe32, &ctx->pd32, &ctx->pz32, &ctx->pq32, &ctx->fx32, &ctx->fy32};
    for (float **v : v32) {
        if (*v) hipFree(*v);
        *v = nullptr;
    }
}

static dim3 grid2(int n1, int n2, dim3 blk) { return dim3((n2 + blk.x - 1) / blk.x, (n1 + blk.y - 1) / blk.y); }

static int red_grid(long n) {
    long b = (n + 1023) / 1024;
    if (b < 1) b = 1;
    if (b > VEC_BLOCKS) b = VEC_BLOCKS;
    return (int)b;
}

// coarse-level mobility quantities for the current fine w (f32: with the FP32
// copies of the mixed-precision cycle, whose arrays exist)
static int build_hierarchy(impec_ctx *ctx, bool f32 = false) {
    dim3 blk(64, 4);
    for (int l = 1; l < ctx->nlev; l++) {
        const Level &F = ctx->lev[l - 1];
        Level &Cl = ctx->lev[l];
        const double *wf;
        int ldf;
        if (l == 1) {
            wf = ctx->w + (long)ctx->B * ctx->lds + ctx->B;
            ldf = ctx->lds;
        } else {
            wf = F.w + F.ld + 1;
            ldf = F.ld;
        }
        hipLaunchKernelGGL(k_coarsen_w, grid2(Cl.n1 + 2, Cl.n2 + 2, blk), blk, 0, ctx->stream, wf, ldf, F.n1, F.n2,
                           Cl.w + Cl.ld + 1, Cl.ld, Cl.n1, Cl.n2, f32 ? Cl.wf + Cl.ld + 1 : nullptr);
    }
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

template <bool DIAG>
static void smooth(impec_ctx *ctx, int l, int color) {
    const Level &L = ctx->lev[l];
    LevArgs a = lev_args(ctx, l);
    dim3 blk(64, 4), grd;
    if (DIAG) grd = grid2((L.n1 + 1) / 2, (L.n2 + 1) / 2, blk);
    else grd = grid2(L.n1, (L.n2 + 1) / 2, blk);
    prof_begin(ctx);
    if (l == 0) hipLaunchKernelGGL((k_smooth<DIAG, true>), grd, blk, 0, ctx->stream, a, L.x + L.ld + 1, L.b + L.ld + 1, color);
    else hipLaunchKernelGGL((k_smooth<DIAG, false>), grd, blk, 0, ctx->stream, a, L.x + L.ld + 1, L.b + L.ld + 1, color);
    prof_end(ctx, l == 0 ? "mg_smooth_l0" : "mg_coarse");
}

template <bool DIAG>
static int vcycle(impec_ctx *ctx) {
    const int nc = DIAG ? 4 : 2;
    const int npre = ctx->P.mg_pre > 0 ? ctx->P.mg_pre : 1;
    const int npost = ctx->P.mg_post > 0 ? ctx->P.mg_post : 1;
    dim3 blk(64, 4);
    const int L = ctx->nlev;
    for (int l = 0; l < L - 1; l++) {
        Level &F = ctx->lev[l];
        Level &Cl = ctx->lev[l + 1];
        const char *tag = l == 0 ? "mg_transfer_l0" : "mg_coarse";
        const long nx = (long)(F.n1 + 2) * F.ld;
        prof_begin(ctx);
        hipLaunchKernelGGL(k_fill_zero, dim3((int)((nx + 255) / 256 < 4096 ? (nx + 255) / 256 : 4096)), dim3(256), 0,
                           ctx->stream, F.x, nx);
        prof_end(ctx, tag);
        for (int s = 0; s < npre; s++)
            for (int col = 0; col < nc; col++) smooth<DIAG>(ctx, l, col);
        LevArgs a = lev_args(ctx, l);
        prof_begin(ctx);
        if (l == 0)
            hipLaunchKernelGGL((k_residual<DIAG, true>), grid2(F.n1, F.n2, blk), blk, 0, ctx->stream, a,
                               F.x + F.ld + 1, F.b + F.ld + 1, F.r + F.ld + 1);
        else
            hipLaunchKernelGGL((k_residual<DIAG, false>), grid2(F.n1, F.n2, blk), blk, 0, ctx->stream, a,
                               F.x + F.ld + 1, F.b + F.ld + 1, F.r + F.ld + 1);
        hipLaunchKernelGGL(k_restrict, grid2(Cl.n1, Cl.n2, blk), blk, 0, ctx->stream, F.r + F.ld + 1, F.n1, F.n2,
                           F.ld, Cl.b + Cl.ld + 1, Cl.n1, Cl.n2, Cl.ld);
        prof_end(ctx, tag);
    }
    {
        Level &Cl = ctx->lev[L - 1];
        LevArgs a = lev_args(ctx, L - 1);
        prof_begin(ctx);
        hipLaunchKernelGGL(k_coarsest<DIAG>, dim3(1), dim3(64), 0, ctx->stream, a, Cl.x + Cl.ld + 1, Cl.b + Cl.ld + 1);
        prof_end(ctx, "mg_coarse");
    }
    for (int l = L - 2; l >= 0; l--) {
        Level &F = ctx->lev[l];
        Level &Cl = ctx->lev[l + 1];
        prof_begin(ctx);
        hipLaunchKernelGGL(k_prolong, grid2(F.n1, F.n2, blk), blk, 0, ctx->stream, F.x + F.ld + 1, F.n1, F.n2, F.ld,
                           Cl.x + Cl.ld + 1, Cl.n1, Cl.n2, Cl.ld);
        prof_end(ctx, l == 0 ? "mg_transfer_l0" : "mg_coarse");
        for (int s = 0; s < npost; s++)
            for (int col = nc - 1; col >= 0; col--) smooth<DIAG>(ctx, l, col);
    }
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

static int fetch_scal(impec_ctx *ctx) { return sync_scalars(ctx); }

template <class G>
__device__ __forceinline__ void zero_interior_body(const G &gr_, CgArgs c, double *__restrict__ p) {
    for (long id = (long)gr_.bx() * blockDim.x + threadIdx.x; id < c.n; id += (long)gr_.gx() * blockDim.x) {
        int i, j;
        cg_cell(c, id, i, j);
        p[(long)(i + c.B) * c.ldp + (j + c.B)] = 0.;
    }
}
__global__ void k_zero_interior(CgArgs c, double *__restrict__ p) {
    zero_interior_body(HwGrid{}, c, p);
}

__global__ void k_set_pin(CgArgs c, double *__restrict__ p) {
    p[(long)(c.pin_i + c.B) * c.ldp + (c.pin_j + c.B)] = c.ppin;
}

// ===========================================================================
// Streaming V-cycle legs for the 5-point operator (nu = 1), one kernel per leg
// and level.  A wavefront owns a strip of 128 columns, two adjacent columns
// per lane, and a chunk of R rows, and marches down the rows keeping a window
// of four rows in registers; values of the neighbouring columns come from the
// adjacent lanes by DPP wave shifts.  In every row the two columns of a lane
// hold one red and one black cell, at the same offset in every lane, so the
// colour updates of a step have no lane divergence.  Each strip recomputes 4
// columns on either side and writes the middle 120; each chunk recomputes a
// few rows above and below.  Colour 0 (red) is (i + j) even.
//   down, step k: red from a zero guess on row k, black on row k-1, residual
//                 on row k-2, full weighting into the coarse rows completed
//                 by that residual row;
//   up,   step k: x + P e on row k, black on row k-1, red on row k-2.
// The coefficients are formed from the mobility quantity w as in op_row<false>
// with the boundary faces dropped (the Neumann folding), and the diagonal is
// minus the sum of the four faces.
// ===========================================================================
#ifndef IMPEC_SW_WAVES
#define IMPEC_SW_WAVES 4
#endif
constexpr int SW_WAVES = IMPEC_SW_WAVES;
// A strip is 128 columns, two per lane.  Each Gauss-Seidel sweep of a leg
// spoils one more column from the strip's edges (the neighbours beyond them
// are not in the wave), so a leg with NS red-black pairs keeps NS + 1 lanes
// of halo on each side: NS = 1 writes 120 columns, NS = 2 writes 116.
template <int NS>
struct Strip {
    static constexpr int HALO_LANES = NS + 1;
    static constexpr int HALO = 2 * HALO_LANES;   // columns
    static constexpr int OUT = 128 - 2 * HALO;    // written columns
};
constexpr int SW_OUT = Strip<1>::OUT, SW_HALO = Strip<1>::HALO;

// T is the storage and arithmetic type of the cycle: double (FP64) or float
// (the FP32 V-cycles of the mixed-precision solve, inside the FP64 CG)
template <typename T>
struct FusedArgsT {
    OpGeom g;
    const T *w, *w2;        // mobility quantity (x faces, y faces) at cell (0,0), pitch ldw
    int ldw;
    int ld;                 // pitch of this level's x/b (ring-1 arrays at cell (0,0))
    const T *b;             // level rhs (levels >= 1)
    const double *r;        // level 0: CG residual (FP64), b = r - S[S_RSUM]/n
    const float *r32;       // level 0: the FP32 residual of the FP32-vector CG (R32 legs)
    const double *S;
    double inv_n;
    T *x;                   // this level's x (down: written; up: read)
    T *xo;                  // up: corrected, post-smoothed x (a different array:
                            // neighbouring strips and chunks read x)
    const T *xc;            // up: coarse-level solution at (0,0)
    T *bc;                  // down: coarse-level b at (0,0)
    int m1, m2, ldc;        // coarse level
    double *part;           // up at level 0: per-wave partials (sum x, sum r x), FP64
    int R, nstrip, nwave;   // rows per chunk, strips, wavefronts
};
using FusedArgs = FusedArgsT<double>;

// 1/d from the hardware reciprocal estimate refined by one Newton-Raphson
// step, r' = r + r (1 - d r).  Only the multigrid preconditioner uses it: its
// rounding differs from IEEE division but is the same every time for the same
// d, so the preconditioner stays a fixed linear operator.
__device__ __forceinline__ double recip(double d) {
    const double r = __builtin_amdgcn_rcp(d);
    const double e = fma(-d, r, 1.);
    return fma(r, e, r);
}
__device__ __forceinline__ float recip(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    const float e = fmaf(-d, r, 1.f);
    return fmaf(r, e, r);
}

// face mobility as face_mob, with recip in place of the division
template <typename T>
__device__ __forceinline__ T face_mob_pc(T wl, T wr, int harmonic) {
    return harmonic ? T(2) * recip(wl + wr) : T(0.5) * (wl + wr);
}

// scaled face coefficients of one row for a lane's two columns g0, g0+1:
// E = face to the next row, S0 = face left of g0, M = between g0 and g0+1,
// N1 = face right of g0+1 (zero where the face is a domain boundary)
template <typename T>
struct FRowT {
    T E0, E1, S0, M, N1;
};

struct Lane {
    int g0;
    int c0, c1;        // g0, g0+1 clamped into the level (every load is issued)
    bool v0, v1, vR;   // columns g0, g0+1 and the pairs (g0,g0+1), (g0+1,g0+2) inside
};

template <typename T>
__device__ __forceinline__ FRowT<T> row_faces(const OpGeom &g, const Lane &L, int m, T a0, T a1, T an0, T an1, T y0,
                                              T y1) {
#pragma clang fp contract(fast)
    const bool rin = m >= 0 && m < g.n1, rnx = m >= 0 && m + 1 < g.n1;
    const T cx = T(g.cx), cy = T(g.cy);
    FRowT<T> f;
    const T e0 = cx * face_mob_pc(a0, an0, g.harmonic), e1 = cx * face_mob_pc(a1, an1, g.harmonic);
    const T yR = lane_next(y0);
    const T fm = cy * face_mob_pc(y0, y1, g.harmonic), fn = cy * face_mob_pc(y1, yR, g.harmonic);
    f.E0 = (rnx && L.v0) ? e0 : T(0);
    f.E1 = (rnx && L.v1) ? e1 : T(0);
    f.M = (rin && L.v1) ? fm : T(0);
    f.N1 = (rin && L.vR) ? fn : T(0);
    f.S0 = lane_prev(f.N1);
    return f;
}

// one Gauss-Seidel update of the cell at column offset c of a row with faces
// F, face above Wf (E of the previous row), neighbours above/below xa/xb and
// the row's two values xr[2]
template <int C, typename T>
__device__ __forceinline__ T gs_cell(const FRowT<T> &F, const T Wf[2], T b, T xa, T xb, const T xr[2]) {
#pragma clang fp contract(fast)
    T Wc, Ec, Sc, Nc, xS, xN;
    if (C == 0) {
        Wc = Wf[0]; Ec = F.E0; Sc = F.S0; Nc = F.M;
        xS = lane_prev(xr[1]); xN = xr[1];
    } else {
        Wc = Wf[1]; Ec = F.E1; Sc = F.M; Nc = F.N1;
        xS = xr[0]; xN = lane_next(xr[0]);
    }
    const T Cc = -((Wc + Ec) + (Sc + Nc));
    T s = b;
    s -= Wc * xa;
    s -= Sc * xS;
    s -= Ec * xb;
    s -= Nc * xN;
    return Cc != T(0) ? s * recip(Cc) : T(0);
}

struct Chunk {
    int i0, i1;
    bool out;   // the lane writes (lanes 2..61)
};

// wave index of the calling wave in a launch of SW_WAVES-wave workgroups;
// wave-uniform by construction, and readfirstlane lets the compiler know, so
// the march loop and the row masks are scalar
__device__ __forceinline__ int hw_wave() {
    return blockIdx.x * SW_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}
// the same with the workgroups renumbered by xcd_block, for the launches over
// one grid: the strips of a chunk, which share their edge columns, then run on
// one XCD
__device__ __forceinline__ int xcd_wave() {
    return xcd_block(blockIdx.x, gridDim.x) * SW_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// strip and chunk of wave `wid` of a leg's launch over one level, for legs of
// NS red-black pairs
template <int NS, typename Args>
__device__ __forceinline__ bool wave_setup(const Args &A, int wid, Lane &L, Chunk &K) {
    const int lane = threadIdx.x & 63;
    if (wid >= A.nwave) return false;
    const int strip = wid % A.nstrip, chunk = wid / A.nstrip;
    const int n2 = A.g.n2;
    L.g0 = strip * Strip<NS>::OUT - Strip<NS>::HALO + 2 * lane;
    L.v0 = L.g0 >= 0 && L.g0 < n2;
    L.v1 = L.g0 >= 0 && L.g0 + 1 < n2;
    L.vR = L.g0 >= 0 && L.g0 + 2 < n2;
    L.c0 = L.g0 < 0 ? 0 : (L.g0 > n2 - 1 ? n2 - 1 : L.g0);
    L.c1 = L.g0 + 1 < 0 ? 0 : (L.g0 + 1 > n2 - 1 ? n2 - 1 : L.g0 + 1);
    K.i0 = chunk * A.R;
    K.i1 = K.i0 + A.R < A.g.n1 ? K.i0 + A.R : A.g.n1;
    K.out = lane >= Strip<NS>::HALO_LANES && lane < 64 - Strip<NS>::HALO_LANES;
    return true;
}

// the lane's two values of row m of p, loaded unconditionally from a clamped
// address, so that no lane-dependent branch separates a prefetch from its
// use; mask_pair replaces the values outside the level when they are used
// (a select at load time would wait for the load there).
template <typename V>
__device__ __forceinline__ void load_pair(const V *__restrict__ p, int ldp, const Lane &L, int m, int n1, V &a,
                                          V &b) {
    const int mc = m < 0 ? 0 : (m > n1 - 1 ? n1 - 1 : m);
    const V *q = p + (long)mc * ldp;
    a = q[L.c0];
    b = q[L.c1];
}
template <typename V>
__device__ __forceinline__ void mask_pair(const Lane &L, int m, int n1, V fill, V &a, V &b) {
    const bool rin = m >= 0 && m < n1;
    a = (rin && L.v0) ? a : fill;
    b = (rin && L.v1) ? b : fill;
}

// prefetch distance in rows (one slot per unrolled step), per arithmetic
// type; build-time tunable.  The FP64 legs run at 3 waves per SIMD with 2 rows
// in flight (4 rows take them to 203 VGPRs and 2 waves, 10% slower); the FP32
// legs are fastest with 4.
#ifndef IMPEC_SW_PF
#define IMPEC_SW_PF 2
#endif
#ifndef IMPEC_SW_PF_F32
#define IMPEC_SW_PF_F32 4
#endif
// the FP64 legs of the coarser levels (4 rows in flight measured slower: level
// 1's down leg 51 us against 44 at 4096^2)
#ifndef IMPEC_SW_PF_COARSE
#define IMPEC_SW_PF_COARSE 2
#endif
static_assert(IMPEC_SW_PF % 2 == 0 && IMPEC_SW_PF_F32 % 2 == 0 && IMPEC_SW_PF_COARSE % 2 == 0,
              "the unrolled steps take the row parity (colour, prolongation neighbour) from the slot index, so the "
              "prefetch depth must be even");
template <typename T, bool FINE0 = true>
struct SwPf {
    static constexpr int value =
        std::is_same<T, float>::value ? IMPEC_SW_PF_F32 : (FINE0 ? IMPEC_SW_PF : IMPEC_SW_PF_COARSE);
};

template <bool FINE0, bool ANISO, typename T, int NS = 1, bool R32 = false>
__device__ __forceinline__ void mg_down_body(const FusedArgsT<T> &A, const int wid) {
#pragma clang fp contract(fast)
    // NS red-black pairs: red from zero on row k, then the other 2 NS - 1
    // colour sweeps on rows k-1 .. k-2NS+1, the residual on row k-2NS
    constexpr int NX = 2 + 2 * NS;   // x and W-face window: rows k-NX+1 .. k
    constexpr int NF = 1 + 2 * NS;   // face and rhs window: rows k-NF+1 .. k
    constexpr int LAG = 2 * NS;      // residual and output row k - LAG
    constexpr int SW_PF = SwPf<T, FINE0>::value;
    Lane L;
    Chunk K;
    if (!wave_setup<NS>(A, wid, L, K)) return;
    const int n1 = A.g.n1, n2 = A.g.n2;
    constexpr bool aniso = ANISO;
    const double mean = FINE0 ? A.S[S_RSUM] * A.inv_n : 0.;
    // level 0 reads the CG residual: FP64, or FP32 for the FP32-vector CG (R32)
    using Src = typename std::conditional<FINE0, typename std::conditional<R32, float, double>::type, T>::type;
    const Src *src;
    if constexpr (FINE0 && R32) src = A.r32;
    else if constexpr (FINE0) src = A.r;
    else src = A.b;
    // full weighting: this lane's coarse column and the weights of its four fine columns
    const int J = L.g0 >> 1;
    const bool wcol = K.out && L.g0 >= 0 && J < A.m2;
    const T wjL = T((L.g0 >= 1 && L.g0 - 1 < n2) ? pw(L.g0 - 1, J, A.m2) : 0.);
    const T wj0 = T(L.v0 ? pw(L.g0, J, A.m2) : 0.);
    const T wj1 = T(L.v1 ? pw(L.g0 + 1, J, A.m2) : 0.);
    const T wjR = T(L.vR ? pw(L.g0 + 2, J, A.m2) : 0.);
    const int Ilo = K.i0 >> 1, Ihi = (K.i1 + 1) >> 1;
    const int kstart = K.i0 - 2 - LAG, kend = 2 * Ihi + LAG;

    T PW[SW_PF][2], PY[SW_PF][2];
    Src PB[SW_PF][2];
    auto fetch = [&](int u, int k) {   // inputs of step k: w(k+1), w2(k), b(k)
        load_pair(A.w, A.ldw, L, k + 1, n1, PW[u][0], PW[u][1]);
        if (aniso) load_pair(A.w2, A.ldw, L, k, n1, PY[u][0], PY[u][1]);
        load_pair(src, A.ld, L, k, n1, PB[u][0], PB[u][1]);
    };
#pragma unroll
    for (int u = 0; u < SW_PF; u++) fetch(u, kstart + u);
    T wc0, wc1;
    load_pair(A.w, A.ldw, L, kstart, n1, wc0, wc1);
    mask_pair(L, kstart, n1, T(1), wc0, wc1);

    T X[NX][2] = {};   // rows k-NX+1 .. k
    T B[NF][2] = {};   // rows k-NF+1 .. k
    FRowT<T> F[NF];    // rows k-NF+1 .. k
    T E[NX][2] = {};   // E of rows k-NX+1 .. k
    for (int q = 0; q < NF; q++) F[q] = FRowT<T>{T(0), T(0), T(0), T(0), T(0)};
    T acc_c = T(0), acc_n = T(0);

    // whole groups of SW_PF steps (no exit inside the unrolled group, so the
    // prefetch ring stays in place); steps past kend write nothing
    const int kstop = kstart + (kend - kstart + SW_PF) / SW_PF * SW_PF;
    for (int kb = kstart; kb < kstop; kb += SW_PF) {
#pragma unroll
        for (int u = 0; u < SW_PF; u++) {
            const int k = kb + u;
            T wn0 = PW[u][0], wn1 = PW[u][1];
            mask_pair(L, k + 1, n1, T(1), wn0, wn1);
            T y0 = wc0, y1 = wc1;
            if (aniso) {
                y0 = PY[u][0];
                y1 = PY[u][1];
                mask_pair(L, k, n1, T(1), y0, y1);
            }
            T bk[2];
            bk[0] = L.v0 && k >= 0 && k < n1 ? T(PB[u][0] - mean) : T(0);
            bk[1] = L.v1 && k >= 0 && k < n1 ? T(PB[u][1] - mean) : T(0);
            const FRowT<T> fk = row_faces(A.g, L, k, wc0, wc1, wn0, wn1, y0, y1);
            wc0 = wn0;
            wc1 = wn1;
#pragma unroll
            for (int q = 0; q < NX - 1; q++) {
                X[q][0] = X[q + 1][0]; X[q][1] = X[q + 1][1];
                E[q][0] = E[q + 1][0]; E[q][1] = E[q + 1][1];
            }
#pragma unroll
            for (int q = 0; q < NF - 1; q++) {
                F[q] = F[q + 1];
                B[q][0] = B[q + 1][0]; B[q][1] = B[q + 1][1];
            }
            F[NF - 1] = fk;
            B[NF - 1][0] = bk[0]; B[NF - 1][1] = bk[1];
            E[NX - 1][0] = fk.E0; E[NX - 1][1] = fk.E1;
            // red from zero on row k (its black cell stays zero)
            const int c = u & 1;   // k & 1 (kstart is even)
            {
                const T C0 = -((E[NX - 2][0] + fk.E0) + (fk.S0 + fk.M));
                const T C1 = -((E[NX - 2][1] + fk.E1) + (fk.M + fk.N1));
                if (c == 0) {
                    X[NX - 1][0] = C0 != T(0) ? bk[0] * recip(C0) : T(0);
                    X[NX - 1][1] = T(0);
                } else {
                    X[NX - 1][0] = T(0);
                    X[NX - 1][1] = C1 != T(0) ? bk[1] * recip(C1) : T(0);
                }
            }
            // black on row k-1, then red on row k-2, black on row k-3 (NS = 2);
            // on each of these rows the cell of colour the sweep updates is
            // the lane's cell c
#pragma unroll
            for (int l = 1; l < 2 * NS; l++) {
                const int xi = NX - 1 - l, fi = NF - 1 - l;
                if (c == 0) X[xi][0] = gs_cell<0>(F[fi], E[xi - 1], B[fi][0], X[xi - 1][0], X[xi + 1][0], X[xi]);
                else X[xi][1] = gs_cell<1>(F[fi], E[xi - 1], B[fi][1], X[xi - 1][1], X[xi + 1][1], X[xi]);
            }
            // residual on row k-LAG (window rows 0..2 of X, 0 of F, B and E)
            T r0, r1;
            {
                const FRowT<T> &fr = F[0];
                const T xl = lane_prev(X[1][1]), xr = lane_next(X[1][0]);
                T W = E[0][0], Ee = fr.E0, S = fr.S0, N = fr.M;
                T Cc = -((W + Ee) + (S + N));
                T s = Cc * X[1][0];
                s += W * X[0][0];
                s += S * xl;
                s += Ee * X[2][0];
                s += N * X[1][1];
                r0 = B[0][0] - s;
                W = E[0][1]; Ee = fr.E1; S = fr.M; N = fr.N1;
                Cc = -((W + Ee) + (S + N));
                s = Cc * X[1][1];
                s += W * X[0][1];
                s += S * X[1][0];
                s += Ee * X[2][1];
                s += N * xr;
                r1 = B[0][1] - s;
            }
            const int f = k - LAG;
            if (K.out && f >= K.i0 && f < K.i1) {
                T *xp = A.x + (long)f * A.ld + L.g0;
                if (L.v0) xp[0] = X[1][0];
                if (L.v1) xp[1] = X[1][1];
            }
            // full weighting (P^T / 4, the transpose of the prolongation)
            {
                const T rL = lane_prev(r1), rR = lane_next(r0);
                const T h = ((wjL * rL + wj0 * r0) + wj1 * r1) + wjR * rR;
                const bool fin = f >= 0 && f < n1;
                const int cur = (f - 1) >> 1;
                const T wc = T(fin ? pw(f, cur, A.m1) : 0.), wn = T(fin ? pw(f, cur + 1, A.m1) : 0.);
                if (wc != T(0)) acc_c += wc * h;
                if (wn != T(0)) acc_n += wn * h;
                if (c == 0) {   // f even: coarse row cur is complete
                    if (wcol && cur >= Ilo && cur < Ihi) A.bc[(long)cur * A.ldc + J] = T(0.25) * acc_c;
                    acc_c = acc_n;
                    acc_n = T(0);
                }
            }
            // refill the slot once its registers are dead (an earlier refill
            // would copy them out and wait for the load that filled them);
            // steps past kend re-read the inputs of step kend (cache hits)
            fetch(u, k + SW_PF <= kend ? k + SW_PF : kend);
        }
    }
}

template <bool FINE0, bool ANISO, typename T, int NS, bool R32 = false>
__global__ void __launch_bounds__(64 * SW_WAVES) k_mg_down(FusedArgsT<T> A) {
    static_assert(FINE0 || !R32, "only level 0 reads the CG residual");
    mg_down_body<FINE0, ANISO, T, NS, R32>(A, xcd_wave());
}

template <bool FINE0, bool ANISO, typename T, int NS = 1, bool R32 = false>
__device__ __forceinline__ void mg_up_body(const FusedArgsT<T> &A, const int wid) {
#pragma clang fp contract(fast)
    // NS black-red pairs after the correction of row k: black on row k-1, red
    // on row k-2, ..., red on row k-2NS, the output row
    constexpr int NX = 2 + 2 * NS;
    constexpr int NF = 1 + 2 * NS;
    constexpr int LAG = 2 * NS;
    constexpr int SW_PF = SwPf<T, FINE0>::value;
    Lane L;
    Chunk K;
    if (!wave_setup<NS>(A, wid, L, K)) return;
    const int n1 = A.g.n1, m1 = A.m1, m2 = A.m2;
    constexpr bool aniso = ANISO;
    const double mean = FINE0 ? A.S[S_RSUM] * A.inv_n : 0.;
    // level 0 reads the CG residual: FP64, or FP32 for the FP32-vector CG (R32)
    using Src = typename std::conditional<FINE0, typename std::conditional<R32, float, double>::type, T>::type;
    const Src *src;
    if constexpr (FINE0 && R32) src = A.r32;
    else if constexpr (FINE0) src = A.r;
    else src = A.b;
    // prolongation: this lane's coarse column J and its clamped neighbours
    const int J = L.g0 >> 1;
    const bool jok = L.g0 >= 0 && J < m2;
    const bool jfirst = J == 0, jlast = J == m2 - 1;
    const int Jc = J < 0 ? 0 : (J > m2 - 1 ? m2 - 1 : J);
    auto coarse = [&](int I) -> T {   // raw; select with jok at use
        I = I < 0 ? 0 : (I > m1 - 1 ? m1 - 1 : I);
        return A.xc[(long)I * A.ldc + Jc];
    };
    const int kstart = K.i0 - LAG, kend = K.i1 + LAG - 1;

    T PX[SW_PF][2], PW[SW_PF][2], PY[SW_PF][2], PC[SW_PF][2];
    Src PB[SW_PF][2];
    // inputs of step k: x(k), w(k+1), w2(k), b(k) and the coarse rows I0 = k>>1,
    // I1 = I0 -+ 1 of the prolongation (re-read for every fine row; cached)
    auto fetch = [&](int u, int k) {
        load_pair(A.x, A.ld, L, k, n1, PX[u][0], PX[u][1]);
        load_pair(A.w, A.ldw, L, k + 1, n1, PW[u][0], PW[u][1]);
        if (aniso) load_pair(A.w2, A.ldw, L, k, n1, PY[u][0], PY[u][1]);
        load_pair(src, A.ld, L, k, n1, PB[u][0], PB[u][1]);
        PC[u][0] = coarse(k >> 1);
        PC[u][1] = coarse((u & 1) ? (k >> 1) + 1 : (k >> 1) - 1);
    };
#pragma unroll
    for (int u = 0; u < SW_PF; u++) fetch(u, kstart + u);
    T wc0, wc1;
    load_pair(A.w, A.ldw, L, kstart, n1, wc0, wc1);
    mask_pair(L, kstart, n1, T(1), wc0, wc1);

    T X[NX][2] = {};     // rows k-NX+1 .. k
    Src B[NF][2] = {};   // raw rhs rows k-NF+1 .. k
    FRowT<T> F[NF];
    T E[NX][2] = {};
    for (int q = 0; q < NF; q++) F[q] = FRowT<T>{T(0), T(0), T(0), T(0), T(0)};
    double acc0 = 0., acc1 = 0.;

    // whole groups of SW_PF steps (no exit inside the unrolled group, so the
    // prefetch ring stays in place); steps past kend write nothing
    const int kstop = kstart + (kend - kstart + SW_PF) / SW_PF * SW_PF;
    for (int kb = kstart; kb < kstop; kb += SW_PF) {
#pragma unroll
        for (int u = 0; u < SW_PF; u++) {
            const int k = kb + u;
            const int c = u & 1;   // k & 1
            const T xk0 = PX[u][0], xk1 = PX[u][1];
            T wn0 = PW[u][0], wn1 = PW[u][1];
            mask_pair(L, k + 1, n1, T(1), wn0, wn1);
            T y0 = wc0, y1 = wc1;
            if (aniso) {
                y0 = PY[u][0];
                y1 = PY[u][1];
                mask_pair(L, k, n1, T(1), y0, y1);
            }
            Src bk0 = PB[u][0], bk1 = PB[u][1];
            mask_pair(L, k, n1, Src(0), bk0, bk1);
            const T o0 = jok ? PC[u][0] : T(0), o1 = jok ? PC[u][1] : T(0);
            const FRowT<T> fk = row_faces(A.g, L, k, wc0, wc1, wn0, wn1, y0, y1);
            wc0 = wn0;
            wc1 = wn1;
#pragma unroll
            for (int q = 0; q < NX - 1; q++) {
                X[q][0] = X[q + 1][0]; X[q][1] = X[q + 1][1];
                E[q][0] = E[q + 1][0]; E[q][1] = E[q + 1][1];
            }
#pragma unroll
            for (int q = 0; q < NF - 1; q++) {
                F[q] = F[q + 1];
                B[q][0] = B[q + 1][0]; B[q][1] = B[q + 1][1];
            }
            F[NF - 1] = fk;
            B[NF - 1][0] = bk0; B[NF - 1][1] = bk1;
            E[NX - 1][0] = fk.E0; E[NX - 1][1] = fk.E1;
            // x + P e on row k from coarse rows I0 (o0) and I1 (o1)
            {
                const T pl0 = lane_prev(o0), pl1 = lane_prev(o1), pr0 = lane_next(o0), pr1 = lane_next(o1);
                const T l0 = jfirst ? o0 : pl0, l1 = jfirst ? o1 : pl1;
                const T r0 = jlast ? o0 : pr0, r1 = jlast ? o1 : pr1;
                const T e0 = T(0.5625) * o0 + T(0.1875) * l0 + T(0.1875) * o1 + T(0.0625) * l1;
                const T e1 = T(0.5625) * o0 + T(0.1875) * r0 + T(0.1875) * o1 + T(0.0625) * r1;
                const bool rin = k >= 0 && k < n1;
                X[NX - 1][0] = (rin && L.v0) ? xk0 + e0 : T(0);
                X[NX - 1][1] = (rin && L.v1) ? xk1 + e1 : T(0);
            }
            // black on row k-1, red on row k-2 (then black on k-3, red on k-4
            // for NS = 2); on each row the updated cell is the lane's cell c
#pragma unroll
            for (int l = 1; l <= 2 * NS; l++) {
                const int xi = NX - 1 - l, fi = NF - 1 - l;
                if (c == 0)
                    X[xi][0] = gs_cell<0>(F[fi], E[xi - 1], T(B[fi][0] - mean), X[xi - 1][0], X[xi + 1][0], X[xi]);
                else
                    X[xi][1] = gs_cell<1>(F[fi], E[xi - 1], T(B[fi][1] - mean), X[xi - 1][1], X[xi + 1][1], X[xi]);
            }
            const int f = k - LAG;
            if (K.out && f >= K.i0 && f < K.i1) {
                T *xp = A.xo + (long)f * A.ld + L.g0;
                if (L.v0) xp[0] = X[1][0];
                if (L.v1) xp[1] = X[1][1];
                if (FINE0) {   // FP64 partials of the cycle output
                    if (L.v0) { acc0 += double(X[1][0]); acc1 += double(B[0][0]) * double(X[1][0]); }
                    if (L.v1) { acc0 += double(X[1][1]); acc1 += double(B[0][1]) * double(X[1][1]); }
                }
            }
            fetch(u, k + SW_PF <= kend ? k + SW_PF : kend);   // the slot's registers are dead here; past kend: re-reads
        }
    }
    if (FINE0) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            acc0 += __shfl_xor(acc0, o);
            acc1 += __shfl_xor(acc1, o);
        }
        if ((threadIdx.x & 63) == 0) {
            A.part[wid] = acc0;
            A.part[RED_BLOCKS + wid] = acc1;
        }
    }
}

template <bool FINE0, bool ANISO, typename T, int NS, bool R32 = false>
__global__ void __launch_bounds__(64 * SW_WAVES) k_mg_up(FusedArgsT<T> A) {
    static_assert(FINE0 || !R32, "only level 0 reads the CG residual");
    mg_up_body<FINE0, ANISO, T, NS, R32>(A, xcd_wave());
}

// ===========================================================================
// The bottom of the streaming V-cycle: from the first level of at most
// BOTTOM_N x BOTTOM_N cells down to the coarsest and back up in one
// workgroup, with every level's x, b, r, w and faces in LDS.  Per cell it does
// what the streaming legs do (faces from w by the same expressions, red then black
// from a zero guess, residual, full weighting; bilinear prolongation, black,
// red), and on the coarsest level k_coarsest's bordered solve of the legs'
// operator (faces from w, boundary faces dropped) assembled from LDS;
// barriers separate the colour and transfer phases.  It replaces two launches
// per small level, each a latency chain, and the one-thread coarsest launch.
// ===========================================================================
constexpr int BOTTOM_N = 32, BOTTOM_LEVELS = 8, BOTTOM_THREADS = 1024;
constexpr int BOTTOM_CELLS = 1024 + 256 + 64 + 16 + 4;   // levels of at most 32, 16, 8, 4, 2 cells a side

template <typename T>
struct BottomLev {
    const T *w;      // mobility quantity at cell (0,0), pitch ldw
    int ldw, n1, n2, harmonic;
    double cx, cy;
};
template <typename T>
struct BottomArgs {
    int nl;                        // levels, the last one the coarsest
    BottomLev<T> lev[BOTTOM_LEVELS];
    const T *b;                    // right-hand side of the first level at (0,0), pitch ldb
    int ldb;
    T *out;                        // its corrected, smoothed x at (0,0), pitch ldo
    int ldo;
    int ns;                        // red-black pairs per leg, as the streaming legs
    const double *S;               // loop state (ensembles)
};

// the four faces of cell (i,j) of an n1 x n2 level (zero on the boundary)
template <typename T>
__device__ __forceinline__ void bottom_faces(const T *w, int n1, int n2, int i, int j, T cx, T cy, int harmonic,
                                             T &W, T &E, T &Sf, T &N) {
#pragma clang fp contract(fast)
    const T wc = w[i * n2 + j];
    W = i > 0 ? cx * face_mob_pc(w[(i - 1) * n2 + j], wc, harmonic) : T(0);
    E = i < n1 - 1 ? cx * face_mob_pc(wc, w[(i + 1) * n2 + j], harmonic) : T(0);
    Sf = j > 0 ? cy * face_mob_pc(w[i * n2 + j - 1], wc, harmonic) : T(0);
    N = j < n2 - 1 ? cy * face_mob_pc(wc, w[i * n2 + j + 1], harmonic) : T(0);
}

// the bottom's coarsest level of N1 x N2 cells: the bordered solve of the
// legs' operator (faces from w, boundary faces dropped), w, b and x at the
// level's place in LDS
template <int N1, int N2, typename T>
__device__ __forceinline__ void bottom_coarsest(const T *w, const T *b, T *x, T cx, T cy, int hm) {
    constexpr int n = N1 * N2, m = n + 1;
    double M[m][m + 1];
#pragma unroll
    for (int r = 0; r < m; r++)
#pragma unroll
        for (int q = 0; q <= m; q++) M[r][q] = 0.;
#pragma unroll
    for (int i = 0; i < N1; i++)
#pragma unroll
        for (int j = 0; j < N2; j++) {
            const int row = i * N2 + j;
            T W, E, Sf, N;
            bottom_faces(w, N1, N2, i, j, cx, cy, hm, W, E, Sf, N);
            M[row][row] += (double)(-((W + E) + (Sf + N)));
            if (i > 0) M[row][row - N2] += (double)W;
            if (i < N1 - 1) M[row][row + N2] += (double)E;
            if (j > 0) M[row][row - 1] += (double)Sf;
            if (j < N2 - 1) M[row][row + 1] += (double)N;
            M[row][n] = 1.;
            M[n][row] = 1.;
            M[row][m] = (double)b[row];
        }
    double sol[m];
    bordered_solve<n, true>(M, sol);
#pragma unroll
    for (int k = 0; k < n; k++) x[k] = (T)sol[k];
}

// the faces of cell id = (i, j) of an n1 x n2 level from the per-level face
// arrays in LDS: fw the face to row i-1, fs the face to column j-1 (each zero
// on the boundary), so the east and north faces are the neighbours' fw and fs;
// C the diagonal
template <typename T>
struct BFaces {
    T W, E, S, N, C;
};
template <typename T>
__device__ __forceinline__ BFaces<T> bottom_cell_faces(const T *fw, const T *fs, int n1, int n2, int i, int j,
                                                       int id) {
#pragma clang fp contract(fast)
    BFaces<T> f;
    f.W = fw[id];
    f.E = i < n1 - 1 ? fw[id + n2] : T(0);
    f.S = fs[id];
    f.N = j < n2 - 1 ? fs[id + 1] : T(0);
    f.C = -((f.W + f.E) + (f.S + f.N));
    return f;
}

// Gauss-Seidel update of cell id = (i, j) from the stored faces and the
// stored reciprocal rc of the diagonal, in gs_cell's operation order
template <typename T>
__device__ __forceinline__ T bottom_gs_f(const T *x, const T *fw, const T *fs, const T *rc, T bv, int n1, int n2,
                                         int i, int j, int id) {
#pragma clang fp contract(fast)
    const BFaces<T> f = bottom_cell_faces(fw, fs, n1, n2, i, j, id);
    const T xa = i > 0 ? x[id - n2] : T(0), xb = i < n1 - 1 ? x[id + n2] : T(0);
    const T xS = j > 0 ? x[id - 1] : T(0), xN = j < n2 - 1 ? x[id + 1] : T(0);
    T s = bv;
    s -= f.W * xa;
    s -= f.S * xS;
    s -= f.E * xb;
    s -= f.N * xN;
    return f.C != T(0) ? s * rc[id] : T(0);
}

// Every level has at most BOTTOM_THREADS cells, so a thread holds at most one
// cell of each and computes its (i, j) once per level.  The faces of all
// levels and the reciprocals of their diagonals are formed once, after w is
// loaded, and read from LDS by every sweep.  Level offsets in LDS are running
// sums (an array of them indexed by the level would live in scratch memory).
static_assert(BOTTOM_N * BOTTOM_N <= BOTTOM_THREADS, "one bottom cell per thread and level");
template <typename T>
__device__ void mg_bottom_body(const BottomArgs<T> &A) {
#pragma clang fp contract(fast)
    __shared__ T sx[BOTTOM_CELLS], sb[BOTTOM_CELLS], sr[BOTTOM_CELLS], sw[BOTTOM_CELLS];
    __shared__ T sfw[BOTTOM_CELLS], sfs[BOTTOM_CELLS], src[BOTTOM_CELLS];
    const int t = threadIdx.x;
    const int nl = A.nl;
    for (int l = 0, o = 0; l < nl; l++) {
        const int n1 = A.lev[l].n1, n2 = A.lev[l].n2, ldw = A.lev[l].ldw;
        const T *w = A.lev[l].w;
        if (t < n1 * n2) {
            const int i = t / n2, j = t - i * n2;
            sw[o + t] = w[(long)i * ldw + j];
            if (l == 0) sb[t] = A.b[(long)i * A.ldb + j];
        }
        o += n1 * n2;
    }
    __syncthreads();
    for (int l = 0, o = 0; l < nl; l++) {
        const int n1 = A.lev[l].n1, n2 = A.lev[l].n2, hm = A.lev[l].harmonic;
        const T cx = T(A.lev[l].cx), cy = T(A.lev[l].cy);
        if (t < n1 * n2) {
            const int i = t / n2, j = t - i * n2;
            T W, E, Sf, N;
            bottom_faces(sw + o, n1, n2, i, j, cx, cy, hm, W, E, Sf, N);
            const T Cc = -((W + E) + (Sf + N));
            sfw[o + t] = W;
            sfs[o + t] = Sf;
            src[o + t] = Cc != T(0) ? recip(Cc) : T(0);
        }
        o += n1 * n2;
    }
    __syncthreads();
    // down: red from zero, black, residual, full weighting
    int o = 0;   // offset of level l
    for (int l = 0; l + 1 < nl; l++) {
        const int n1 = A.lev[l].n1, n2 = A.lev[l].n2, n = n1 * n2;
        T *x = sx + o, *r = sr + o;
        const T *b = sb + o, *fw = sfw + o, *fs = sfs + o, *rc = src + o;
        const bool act = t < n;
        const int i = t / n2, j = t - i * n2;
        if (act) {
            T v = T(0);
            if (((i + j) & 1) == 0) {
                const BFaces<T> f = bottom_cell_faces(fw, fs, n1, n2, i, j, t);
                v = f.C != T(0) ? b[t] * rc[t] : T(0);
            }
            x[t] = v;
        }
        __syncthreads();
        if (act && ((i + j) & 1)) x[t] = bottom_gs_f(x, fw, fs, rc, b[t], n1, n2, i, j, t);
        __syncthreads();
        for (int pr = 1; pr < A.ns; pr++)   // further red-black pairs
            for (int c = 0; c <= 1; c++) {
                if (act && ((i + j) & 1) == c) x[t] = bottom_gs_f(x, fw, fs, rc, b[t], n1, n2, i, j, t);
                __syncthreads();
            }
        if (act) {
            const BFaces<T> f = bottom_cell_faces(fw, fs, n1, n2, i, j, t);
            T s = f.C * x[t];
            s += f.W * (i > 0 ? x[t - n2] : T(0));
            s += f.S * (j > 0 ? x[t - 1] : T(0));
            s += f.E * (i < n1 - 1 ? x[t + n2] : T(0));
            s += f.N * (j < n2 - 1 ? x[t + 1] : T(0));
            r[t] = b[t] - s;
        }
        __syncthreads();
        const int m1 = A.lev[l + 1].n1, m2 = A.lev[l + 1].n2;
        T *bc = sb + o + n;
        if (t < m1 * m2) {
            const int I = t / m2, J = t - I * m2;
            const T wL = T((2 * J - 1 >= 0 && 2 * J - 1 < n2) ? pw(2 * J - 1, J, m2) : 0.);
            const T w0 = T(2 * J < n2 ? pw(2 * J, J, m2) : 0.);
            const T w1 = T(2 * J + 1 < n2 ? pw(2 * J + 1, J, m2) : 0.);
            const T wR = T(2 * J + 2 < n2 ? pw(2 * J + 2, J, m2) : 0.);
            T acc = T(0);
            for (int f = 2 * I - 1; f <= 2 * I + 2; f++) {
                if (f < 0 || f >= n1) continue;
                const T wi = T(pw(f, I, m1));
                if (wi == T(0)) continue;
                const T *rf = r + f * n2;
                const T rL = wL != T(0) ? rf[2 * J - 1] : T(0), r0 = w0 != T(0) ? rf[2 * J] : T(0);
                const T r1 = w1 != T(0) ? rf[2 * J + 1] : T(0), rR = wR != T(0) ? rf[2 * J + 2] : T(0);
                const T h = ((wL * rL + w0 * r0) + w1 * r1) + wR * rR;
                acc += wi * h;
            }
            bc[t] = T(0.25) * acc;
        }
        __syncthreads();
        o += n;
    }
    // coarsest (o is its offset): k_coarsest's bordered solve (FP64
    // elimination) of the operator the legs use on that level, assembled from
    // the faces in LDS
    if (t == 0) {
        const int l = nl - 1, n1 = A.lev[l].n1, n2 = A.lev[l].n2, hm = A.lev[l].harmonic;
        const T cx = T(A.lev[l].cx), cy = T(A.lev[l].cy);
        const T *b = sb + o, *w = sw + o;
        T *x = sx + o;
        // the hierarchy is coarsened until both sides are at most 2
        if (n1 == 2 && n2 == 2) bottom_coarsest<2, 2>(w, b, x, cx, cy, hm);
        else if (n1 == 2 && n2 == 1) bottom_coarsest<2, 1>(w, b, x, cx, cy, hm);
        else if (n1 == 1 && n2 == 2) bottom_coarsest<1, 2>(w, b, x, cx, cy, hm);
        else if (n1 == 1 && n2 == 1) bottom_coarsest<1, 1>(w, b, x, cx, cy, hm);
    }
    __syncthreads();
    // up: x + P e, black, red
    for (int l = nl - 2; l >= 0; l--) {
        const int n1 = A.lev[l].n1, n2 = A.lev[l].n2, n = n1 * n2;
        const int m1 = A.lev[l + 1].n1, m2 = A.lev[l + 1].n2;
        const int oc = o;   // offset of level l + 1
        o -= n;
        T *x = sx + o;
        const T *xc = sx + oc, *b = sb + o, *fw = sfw + o, *fs = sfs + o, *rc = src + o;
        const bool act = t < n;
        const int i = t / n2, j = t - i * n2;
        if (act) {
            int I0 = i >> 1, I1 = (i & 1) ? I0 + 1 : I0 - 1;
            I0 = I0 > m1 - 1 ? m1 - 1 : I0;
            I1 = I1 < 0 ? 0 : (I1 > m1 - 1 ? m1 - 1 : I1);
            const int J = j >> 1;
            const int Jn = (j & 1) ? (J == m2 - 1 ? J : J + 1) : (J == 0 ? J : J - 1);
            const T o0 = xc[I0 * m2 + J], o1 = xc[I1 * m2 + J], l0 = xc[I0 * m2 + Jn], l1 = xc[I1 * m2 + Jn];
            const T e = T(0.5625) * o0 + T(0.1875) * l0 + T(0.1875) * o1 + T(0.0625) * l1;
            x[t] = x[t] + e;
        }
        __syncthreads();
        for (int pr = 0; pr < A.ns; pr++)
            for (int c = 1; c >= 0; c--) {
                if (act && ((i + j) & 1) == c) x[t] = bottom_gs_f(x, fw, fs, rc, b[t], n1, n2, i, j, t);
                __syncthreads();
            }
    }
    {
        const int n1 = A.lev[0].n1, n2 = A.lev[0].n2;
        if (t < n1 * n2) {
            const int i = t / n2, j = t - i * n2;
            A.out[(long)i * A.ldo + j] = sx[t];
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(BOTTOM_THREADS) k_mg_bottom(BottomArgs<T> A) {
    mg_bottom_body<T>(A);
}

// first level handled by k_mg_bottom, or -1 when the hierarchy has no level of
// at most BOTTOM_N x BOTTOM_N cells above its coarsest
static int bottom_level(const impec_ctx *ctx) {
    if (ctx->P.mg_variant == 2) return -1;
    for (int l = 1; l + 1 < ctx->nlev; l++)
        if (ctx->lev[l].n1 <= BOTTOM_N && ctx->lev[l].n2 <= BOTTOM_N) {
            if (ctx->nlev - l > BOTTOM_LEVELS) return -1;
            int cells = 0;
            for (int q = l; q < ctx->nlev; q++) cells += ctx->lev[q].n1 * ctx->lev[q].n2;
            return cells <= BOTTOM_CELLS ? l : -1;
        }
    return -1;
}

// levels l < top_level run the streaming legs
static int top_level(const impec_ctx *ctx) {
    const int lb = bottom_level(ctx);
    return lb > 0 ? lb : ctx->nlev - 1;
}

// sums of the V-cycle output and rho = r.z with z = x0 - mean(x0)
template <class G>
__device__ __forceinline__ void finish_rho_body(const G &gr_, const double *__restrict__ part, int nblocks, double inv_n, double *__restrict__ S) {
    double v[2];
    strided_combine_nv<RED_SUM, 2>(part, RED_BLOCKS, nblocks, v);
    block_reduce<RED_SUM, 2>(v);
    if (threadIdx.x == 0) {
        S[S_ZSUM] = v[0];
        S[S_RX] = v[1];
        S[S_RHO_OLD] = S[S_RHO];
        S[S_RHO] = v[1] - v[0] * inv_n * S[S_RSUM];
    }
}
__global__ void k_finish_rho(const double *__restrict__ part, int nblocks, double inv_n, double *__restrict__ S) {
    finish_rho_body(HwGrid{}, part, nblocks, inv_n, S);
}

// tiled 2-D launch shape for the level-0 vector kernels: 64 x 4 threads,
// rows strided so that the grid has at most RED_BLOCKS workgroups
struct Grid2 {
    dim3 grid, block;
    int nblocks;
};
static Grid2 vec_grid(int n1, int n2) {
    Grid2 g;
    g.block = dim3(64, 4);
    const int bx = (n2 + 63) / 64;
    int by = (n1 + 3) / 4;
    if (by * bx > VEC_BLOCKS) by = VEC_BLOCKS / bx > 0 ? VEC_BLOCKS / bx : 1;
    g.grid = dim3(bx, by);
    g.nblocks = bx * by;
    return g;
}

// Direction update and matvec of one CG iteration: d_new = (x0 - mean) +
// beta d_old (d_new = x0 - mean on a (re)start), q = A_N d_new (pinned row
// zeroed in the eliminated formulation) and the d_new.q partials.
// A wavefront owns 64 columns, one per lane, writes the middle 62 (lanes
// 1..62) and marches down a chunk of R rows, keeping the new direction of
// rows i-1, i, i+1 in registers; column neighbours come from the adjacent
// lanes by DPP wave shifts and the inputs of the next DM_PF rows are in
// flight (branch-free clamped loads).  Each cell's x0 and d_old are read once
// instead of five times.  Per-wave partials of d.q.
#ifndef IMPEC_DM_PF
#define IMPEC_DM_PF 4
#endif
constexpr int DM_OUT = 62, DM_PF = IMPEC_DM_PF;
#ifndef IMPEC_DM_WAVES
#define IMPEC_DM_WAVES 4096
#endif
// the FP32-vector CG's pass (half the bytes per cell): shorter chunks, 6144
// waves (2.8% of a step of that mode faster than 4096; the FP64 pass loses 1%
// with them)
#ifndef IMPEC_DM_WAVES_F32
#define IMPEC_DM_WAVES_F32 6144
#endif

// DT: storage of d and q, double for the FP64 CG, float for the FP32-vector CG
// (whose matvec reads FP32 copies of the face mobilities, fx32 / fy32)
template <typename T, typename DT = double>
struct DirMvArgs {
    LevArgs L;
    CgArgs c;
    const T *x0;          // V-cycle output, level-0 ring-1 layout
    DT *d0, *d1;          // the two direction arrays (ensembles pick by iteration parity)
    DT *q;
    const float *fx32, *fy32;
    const double *S;
    double *part;
    int R, nstrip, nwave;
};

// op_row5 with FP32 face mobilities (the FP32-vector CG's operator)
__device__ __forceinline__ void op_row5_f32(const LevArgs &L, const float *fx, const float *fy, int i, int j,
                                            double a[9]) {
    const OpGeom &g = L.g;
    const long o = (long)i * L.ld + j;
    const double fxm = fx[o], fxp = fx[o + L.ld], fym = fy[o], fyp = fy[o + 1];
    const double Wp = g.cx * fxm, Ep = g.cx * fxp, Sp = g.cy * fym, Np = g.cy * fyp;
    const double Cp = g.mcy * (fym + fyp) - g.cx * (fxm + fxp);
    a[ST_C] = Cp; a[ST_W] = Wp; a[ST_SW] = 0.; a[ST_S] = Sp; a[ST_SE] = 0.;
    a[ST_E] = Ep; a[ST_NE] = 0.; a[ST_N] = Np; a[ST_NW] = 0.;
    fold_row(i, j, g.n1, g.n2, a);
}

template <typename T, typename DT = double>
__device__ __forceinline__ void dir_mv_body(const DirMvArgs<T, DT> &A, const DT *__restrict__ dold,
                                            DT *__restrict__ dnew, const int first, const int wid) {
    if (wid >= A.nwave) return;
    const CgArgs &c = A.c;
    const int n1 = c.n1, n2 = c.n2;
    const int lane = threadIdx.x & 63;
    const int strip = wid % A.nstrip, chunk = wid / A.nstrip;
    const int j = strip * DM_OUT - 1 + lane;
    const bool jv = j >= 0 && j < n2;
    const int jc = j < 0 ? 0 : (j > n2 - 1 ? n2 - 1 : j);
    const bool outl = lane >= 1 && lane <= DM_OUT && jv;
    const int i0 = chunk * A.R, i1 = i0 + A.R < n1 ? i0 + A.R : n1;
    const double mean = A.S[S_ZSUM] / (double)c.n;
    const double beta = first ? 0. : A.S[S_RHO] / A.S[S_RHO_OLD];
    const long ld = c.ld;
    const T *__restrict__ x0 = A.x0;
    auto row_of = [&](int i) -> long { const int ic = i < 0 ? 0 : (i > n1 - 1 ? n1 - 1 : i); return (long)(ic + 1) * ld + (jc + 1); };
    // the new direction at (i, j) from its loaded inputs (0 outside the level)
    auto dn = [&](int i, double xv, double dv) -> double {
        if (i < 0 || i >= n1 || !jv) return 0.;
        if (!c.shift && i == c.pin_i && j == c.pin_j) return 0.;
        // rounded to the storage type: the matvec applies A to the stored d
        return (double)(DT)(first ? xv - mean : (xv - mean) + beta * dv);
    };
    double RX[DM_PF], RD[DM_PF];
#pragma unroll
    for (int u = 0; u < DM_PF; u++) {
        const long o = row_of(i0 + 1 + u);
        RX[u] = (double)x0[o];
        RD[u] = first ? 0. : dold[o];
    }
    double dm, dcur;
    {
        const long om = row_of(i0 - 1), oc = row_of(i0);
        dm = dn(i0 - 1, (double)x0[om], first ? 0. : dold[om]);
        dcur = dn(i0, (double)x0[oc], first ? 0. : dold[oc]);
    }
    double acc = 0.;
    for (int kb = i0; kb < i1; kb += DM_PF) {
#pragma unroll
        for (int u = 0; u < DM_PF; u++) {
            const int i = kb + u;
            const double dnext = dn(i + 1, RX[u], RD[u]);
            {   // refill the slot with row i + 1 + DM_PF (rows past i1 are never used: re-read row i1)
                const long o = row_of(i + 1 + DM_PF <= i1 ? i + 1 + DM_PF : i1);
                RX[u] = (double)x0[o];
                RD[u] = first ? 0. : dold[o];
            }
            if (i < i1) {
                double a[9];
                if constexpr (std::is_same<DT, float>::value) op_row5_f32(A.L, A.fx32, A.fy32, i, jc, a);
                else op_row5(A.L, i, jc, a);
                const double dl = lane_prev(dcur), dr = lane_next(dcur);
                double y = a[ST_C] * dcur;
                if (i > 0) y += a[ST_W] * dm;
                if (j > 0) y += a[ST_S] * dl;
                if (i < n1 - 1) y += a[ST_E] * dnext;
                if (j < n2 - 1) y += a[ST_N] * dr;
                if (!c.shift && i == c.pin_i && j == c.pin_j) y = 0.;
                if (outl) {
                    const long o = (long)(i + 1) * ld + (j + 1);
                    const double yq = (double)(DT)y;   // q as stored
                    dnew[o] = (DT)dcur;
                    A.q[o] = (DT)yq;
                    acc += dcur * yq;
                }
            }
            dm = dcur;
            dcur = dnext;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) A.part[wid] = acc;
}

template <typename T, typename DT = double>
__global__ void __launch_bounds__(64 * SW_WAVES) k_dir_mv(DirMvArgs<T, DT> A, const DT *__restrict__ dold,
                                                          DT *__restrict__ dnew, int first) {
    dir_mv_body<T, DT>(A, dold, dnew, first, xcd_wave());
}

// The pinned system's true residual, streamed like k_dir_mv: a wave owns
// DM_OUT columns (one per lane, a halo lane on each side) and a chunk of R
// rows; it reads x and q of each row once, keeps rows i-1, i, i+1 of x in
// registers and takes the column neighbours by DPP.  Per cell the arithmetic
// of true_residual_body (op_row5, the terms of row_dot and of row_absdot in
// their order); per-wave partials of ||b - A x||^2 and || |b| + |A| |x| ||^2.
// (A grid-stride loop over flat indices, reading x five times, ran at
// 2.8 TB/s.)  gx: workgroups of SW_WAVES waves (the ensemble's batch units).
struct TresArgs {
    LevArgs L;
    CgArgs c;
    const double *p, *q;
    double *part;
    int R, nstrip, nwave;
    int gx, gy;
};
static TresArgs tres_args(const LevArgs &a0, const CgArgs &c, const double *p, const double *q, double *part) {
    TresArgs A;
    memset(&A, 0, sizeof(A));
    A.L = a0;
    A.c = c;
    A.p = p;
    A.q = q;
    A.part = part;
    A.nstrip = (c.n2 + DM_OUT - 1) / DM_OUT;
    long R = ((long)c.n1 * A.nstrip + IMPEC_DM_WAVES - 1) / IMPEC_DM_WAVES;
    if (R < 8) R = 8;
    A.R = (int)R;
    A.nwave = A.nstrip * ((c.n1 + A.R - 1) / A.R);
    A.gx = (A.nwave + SW_WAVES - 1) / SW_WAVES;
    A.gy = 1;
    return A;
}
constexpr int TRES_PF = 4;
__device__ __forceinline__ void tres_rows_body(const TresArgs &A, const int wid) {
    if (wid >= A.nwave) return;
    const CgArgs &c = A.c;
    const int n1 = c.n1, n2 = c.n2;
    const int lane = threadIdx.x & 63;
    const int strip = wid % A.nstrip, chunk = wid / A.nstrip;
    const int j = strip * DM_OUT - 1 + lane;
    const bool jv = j >= 0 && j < n2;
    const int jc = j < 0 ? 0 : (j > n2 - 1 ? n2 - 1 : j);
    const bool outl = lane >= 1 && lane <= DM_OUT && jv;
    const int i0 = chunk * A.R, i1 = i0 + A.R < n1 ? i0 + A.R : n1;
    const double *__restrict__ x = A.p + (long)c.B * c.ldp + c.B;
    auto xrow = [&](int i) -> double {
        const int ic = i < 0 ? 0 : (i > n1 - 1 ? n1 - 1 : i);
        return x[(long)ic * c.ldp + jc];
    };
    double RX[TRES_PF], RQ[TRES_PF];
#pragma unroll
    for (int u = 0; u < TRES_PF; u++) {
        RX[u] = xrow(i0 + 1 + u);
        const int ic = i0 + u < n1 ? i0 + u : n1 - 1;
        RQ[u] = A.q[(long)ic * n2 + jc];
    }
    double xm = xrow(i0 - 1), xc = xrow(i0);
    double s0 = 0., s1 = 0.;
    for (int kb = i0; kb < i1; kb += TRES_PF) {
#pragma unroll
        for (int u = 0; u < TRES_PF; u++) {
            const int i = kb + u;
            const double xn = RX[u], qv = RQ[u];
            {   // refill the slot (rows past the chunk re-read its last row)
                const int in = i + 1 + TRES_PF <= i1 ? i + 1 + TRES_PF : i1;
                RX[u] = xrow(in);
                const int iq = i + TRES_PF < i1 ? i + TRES_PF : i1 - 1;
                RQ[u] = A.q[(long)iq * n2 + jc];
            }
            const double xl = lane_prev(xc), xr = lane_next(xc);
            if (i < i1 && outl) {
                double rv, mag;
                if (i == c.pin_i && j == c.pin_j) {
                    rv = -qv - c.dpin * xc;
                    mag = fabs(qv) + fabs(c.dpin * xc);
                } else {
                    double a[9];
                    op_row5(A.L, i, j, a);
                    const bool w = i > 0, e = i < n1 - 1, so = j > 0, no = j < n2 - 1;
                    double sd = a[ST_C] * xc;
                    if (w) sd += a[ST_W] * xm;
                    if (so) sd += a[ST_S] * xl;
                    if (e) sd += a[ST_E] * xn;
                    if (no) sd += a[ST_N] * xr;
                    double sa = fabs(a[ST_C] * xc);
                    if (w) sa += fabs(a[ST_W] * xm);
                    if (so) sa += fabs(a[ST_S] * xl);
                    if (e) sa += fabs(a[ST_E] * xn);
                    if (no) sa += fabs(a[ST_N] * xr);
                    rv = -qv - sd;
                    mag = fabs(qv) + sa;
                }
                s0 += rv * rv;
                s1 += mag * mag;
            }
            xm = xc;
            xc = xn;
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
__global__ void __launch_bounds__(64 * SW_WAVES) k_true_rows(TresArgs A) { tres_rows_body(A, xcd_wave()); }

// --- the FP32-vector CG (precision IMPEC_PREC_MIXED32) ----------------------
// e += alpha d, r -= alpha q in FP32 storage (alpha and the arithmetic in
// FP64, each result rounded once); partials (r.r, sum r) of the stored r
__global__ void k_cg_update32(CgArgs c, float *__restrict__ e, const float *__restrict__ d, float *__restrict__ r,
                              const float *__restrict__ qv, const double *__restrict__ S, double *__restrict__ part) {
    const double alpha = cg_alpha(S);
    const int j = blockIdx.x * 64 + threadIdx.x;
    double s[2] = {0., 0.};
    if (j < c.n2)
        for (int i = blockIdx.y * 4 + threadIdx.y; i < c.n1; i += gridDim.y * 4) {
            const long o = (long)(i + 1) * c.ld + (j + 1);
            float rv = r[o];
            if (alpha != 0.) {   // no step: a broken-down direction may hold non-finite values
                e[o] = (float)((double)e[o] + alpha * (double)d[o]);
                rv = (float)((double)r[o] - alpha * (double)qv[o]);
                r[o] = rv;
            }
            s[0] += (double)rv * (double)rv;
            s[1] += (double)rv;
        }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        const int blk = blockIdx.y * gridDim.x + blockIdx.x;
        part[blk] = s[0];
        part[RED_BLOCKS + blk] = s[1];
    }
}

// a reliable update's restart of the FP32 recursion: r32 = the FP64 residual
// rounded, e = 0; partials (r32.r32, sum r32)
__global__ void k_cg_restart32(CgArgs c, const double *__restrict__ r64, float *__restrict__ r32,
                               float *__restrict__ e, double *__restrict__ part) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    double s[2] = {0., 0.};
    if (j < c.n2)
        for (int i = blockIdx.y * 4 + threadIdx.y; i < c.n1; i += gridDim.y * 4) {
            const long o = (long)(i + 1) * c.ld + (j + 1);
            const float rv = (float)r64[o];
            r32[o] = rv;
            e[o] = 0.f;
            s[0] += (double)rv * (double)rv;
            s[1] += (double)rv;
        }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        const int blk = blockIdx.y * gridDim.x + blockIdx.x;
        part[blk] = s[0];
        part[RED_BLOCKS + blk] = s[1];
    }
}

// a reliable update's FP64 accumulation: p += e (pressure pitch ldp, ghost B)
__global__ void k_cg_fold32(CgArgs c, double *__restrict__ p, const float *__restrict__ e) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    if (j < c.n2)
        for (int i = blockIdx.y * 4 + threadIdx.y; i < c.n1; i += gridDim.y * 4) {
            double *xp = p + (long)(i + c.B) * c.ldp + (j + c.B);
            *xp = *xp + (double)e[(long)(i + 1) * c.ld + (j + 1)];
        }
}


// x += alpha d, r -= alpha q; partials (r.r, sum r).  IMPEC_CGU_NT: cache
// hints; 2 reads d, q and p and writes p non-temporally (none of them is read
// again before the next direction pass), which leaves the Infinity Cache to r,
// the next V-cycle's input: its first leg 116 against 124 us at 4096^2, the
// step 0.2% faster; 1 leaves d cached; 0 none
#ifndef IMPEC_CGU_NT
#define IMPEC_CGU_NT 2
#endif
template <class G>
__device__ __forceinline__ void cg_update2_body(const G &gr_, CgArgs c, double *__restrict__ p, const double *__restrict__ d, double *__restrict__ r,
                             const double *__restrict__ qv, const double *__restrict__ S, double *__restrict__ part) {
    const double alpha = cg_alpha(S);
    const int j = gr_.bx() * 64 + threadIdx.x;
    double s[2] = {0., 0.};
    for (int i = gr_.by() * 4 + threadIdx.y; i < c.n1; i += gr_.gy() * 4) {
        if (j >= c.n2) break;
        const long o = (long)(i + 1) * c.ld + (j + 1);
        double *xp = p + (long)(i + c.B) * c.ldp + (j + c.B);
        double rv = r[o];
        if (alpha != 0.) {   // no step: a broken-down direction may hold non-finite values
#if IMPEC_CGU_NT
            const double dv = IMPEC_CGU_NT > 1 ? __builtin_nontemporal_load(d + o) : d[o];
            __builtin_nontemporal_store(__builtin_nontemporal_load(xp) + alpha * dv, xp);
            rv = rv - alpha * __builtin_nontemporal_load(qv + o);
#else
            *xp = *xp + alpha * d[o];
            rv = rv - alpha * qv[o];
#endif
            r[o] = rv;
        }
        s[0] += rv * rv;
        s[1] += rv;
    }
    block_reduce<RED_SUM, 2>(s);
    if (tid_linear() == 0) {
        const int blk = gr_.by() * gr_.gx() + gr_.bx();
        part[blk] = s[0];
        part[RED_BLOCKS + blk] = s[1];
    }
}
__global__ void k_cg_update2(CgArgs c, double *__restrict__ p, const double *__restrict__ d, double *__restrict__ r,
                             const double *__restrict__ qv, const double *__restrict__ S, double *__restrict__ part) {
    cg_update2_body(HwGrid{}, c, p, d, r, qv, S, part);
}

// the finest level's target in the FP64 cycle (3072 waves, 3 per SIMD: 118 and
// 149 us per leg at 4096^2 against 131 and 160 at 2048)
#ifndef IMPEC_MG_WAVES
#define IMPEC_MG_WAVES 3072
#endif
// the finest level's target in the FP32 cycle (its legs are compute-bound and
// gain from more waves; the FP64 legs and the small levels lose from shorter
// chunks)
#ifndef IMPEC_MG_WAVES0_F32
#define IMPEC_MG_WAVES0_F32 6144
#endif
// the target of the FP64 cycle's coarser levels (level 1 and below): with two
// rows in flight and 3 waves per SIMD, 2048 waves (1.4% of a step faster
// than 1024; with four rows in flight 1024 had been the faster)
#ifndef IMPEC_MG_WAVES_COARSE
#define IMPEC_MG_WAVES_COARSE 2048
#endif
// the FP64 cycle's level 1 (0: IMPEC_MG_WAVES_COARSE; 1024, 4096 and 6144
// measured 0.4%, 0.2% and 0.7% slower per step than 2048 at 4096^2)
#ifndef IMPEC_MG_WAVES_L1
#define IMPEC_MG_WAVES_L1 0
#endif
// the FP32 cycle's coarser levels
#ifndef IMPEC_MG_WAVES_COARSE_F32
#define IMPEC_MG_WAVES_COARSE_F32 2048
#endif
// rows per chunk of the streaming legs: about IMPEC_MG_WAVES wavefronts on large
// levels, short chunks on small ones; even, since the march starts on an
// even row.  The small levels' legs are a latency chain whose length is the
// chunk plus the halo rows: on levels of at most IMPEC_MG_SMALL_ROWS rows
// chunks of 2 rows make the step 0.4% faster than 4; on larger levels the
// halo work of 2-row chunks costs more than it saves (512^2: 30 us per
// V-cycle against 26), and they keep 4
#ifndef IMPEC_MG_MINROWS
#define IMPEC_MG_MINROWS 2
#endif
#ifndef IMPEC_MG_SMALL_ROWS
#define IMPEC_MG_SMALL_ROWS 256
#endif
static_assert(IMPEC_MG_MINROWS % 2 == 0 && IMPEC_MG_MINROWS >= 2, "chunks of the streaming legs start on even rows");
static int stream_rows(const impec_ctx *ctx, int l, int n1, int nstrip, bool f32) {
    if (l == 0 && ctx->P.mg_rows > 0) return (ctx->P.mg_rows + 1) & ~1;
    const long target = f32 ? (l == 0 ? IMPEC_MG_WAVES0_F32 : IMPEC_MG_WAVES_COARSE_F32)
                            : (l == 0 ? IMPEC_MG_WAVES
                                      : (l == 1 && IMPEC_MG_WAVES_L1 > 0 ? IMPEC_MG_WAVES_L1 : IMPEC_MG_WAVES_COARSE));
    // the FP64 finest level's target is the waves the GPU holds at once (3 per
    // SIMD): no more chunks than that, or the excess runs as a round of its own
    long r = chunk_rows(n1, nstrip, target, IMPEC_WAVE_CAP && l == 0 && !f32);
    r = (r + 1) & ~1L;
    const long minr = n1 <= IMPEC_MG_SMALL_ROWS ? IMPEC_MG_MINROWS : (IMPEC_MG_MINROWS > 4 ? IMPEC_MG_MINROWS : 4);
    if (r < minr) r = minr;
    if (r > 64) r = 64;
    return (int)r;
}

// the level arrays a V-cycle of storage type T uses
template <typename T> struct LevelArrays;
template <> struct LevelArrays<double> {
    static double *x(const Level &L) { return L.x; }
    static double *b(const Level &L) { return L.b; }
    static double *r(const Level &L) { return L.r; }
};
template <> struct LevelArrays<float> {
    static float *x(const Level &L) { return L.xf; }
    static float *b(const Level &L) { return L.bf; }
    static float *r(const Level &L) { return L.rf; }
};

// red-black pairs per leg of the streaming legs: 2 for mg_pre = mg_post = 2,
// else 1 (the single-workgroup bottom of the cycle always runs one pair)
static int mg_ns(const impec_ctx *ctx) { return ctx->P.mg_pre == 2 && ctx->P.mg_post == 2 ? 2 : 1; }

template <typename T>
static FusedArgsT<T> fused_args(impec_ctx *ctx, int l) {
    using LA = LevelArrays<T>;
    FusedArgsT<T> A;
    memset(&A, 0, sizeof(A));
    const Level &L = ctx->lev[l];
    LevArgs la = lev_args(ctx, l);
    A.g = L.g;
    if constexpr (std::is_same<T, double>::value) {
        A.w = la.w1;
        A.w2 = la.w2;
        A.ldw = la.ldw;
    } else {   // FP32 copies, ring-1 layout of this level
        A.w = L.wf + L.ld + 1;
        A.w2 = (l == 0 && L.wf2) ? L.wf2 + L.ld + 1 : A.w;
        A.ldw = L.ld;
    }
    A.ld = L.ld;
    A.b = LA::b(L) + L.ld + 1;
    A.r = ctx->pr + L.ld + 1;
    A.r32 = ctx->pr32 ? ctx->pr32 + L.ld + 1 : nullptr;
    A.S = ctx->scal;
    A.inv_n = 1.0 / ((double)ctx->N1 * ctx->N2);
    A.x = LA::x(L) + L.ld + 1;
    A.xo = LA::r(L) + L.ld + 1;   // the fused cycle leaves each level's result in r
    if (l + 1 < ctx->nlev) {
        const Level &Cl = ctx->lev[l + 1];
        // the coarsest level is solved in place into x, the others end in r
        A.xc = (l + 2 == ctx->nlev ? LA::x(Cl) : LA::r(Cl)) + Cl.ld + 1;
        A.bc = LA::b(Cl) + Cl.ld + 1;
        A.m1 = Cl.n1;
        A.m2 = Cl.n2;
        A.ldc = Cl.ld;
    }
    A.part = ctx->part;
    const int out = mg_ns(ctx) == 2 ? Strip<2>::OUT : Strip<1>::OUT;
    A.nstrip = (L.n2 + out - 1) / out;
    A.R = stream_rows(ctx, l, L.n1, A.nstrip, std::is_same<T, float>::value);
    A.nwave = A.nstrip * ((L.n1 + A.R - 1) / A.R);
    return A;
}

template <typename T>
static dim3 stream_grid(const FusedArgsT<T> &A) { return dim3((A.nwave + SW_WAVES - 1) / SW_WAVES); }

// one V(1,1) cycle on the level-0 rhs r - mean(r) in storage type T; the result
// x0 is left in lev[0]'s r array of that type, with the per-wave FP64 partials
// of sum(x0) and r.x0 for k_finish_rho
// arguments of k_mg_bottom for levels lb .. nlev-1 in storage type T
template <typename T>
static BottomArgs<T> bottom_args(impec_ctx *ctx, int lb) {
    using LA = LevelArrays<T>;
    BottomArgs<T> A;
    memset(&A, 0, sizeof(A));
    A.nl = ctx->nlev - lb;
    for (int l = lb; l < ctx->nlev; l++) {
        const Level &L = ctx->lev[l];
        const FusedArgsT<T> f = fused_args<T>(ctx, l);
        BottomLev<T> &b = A.lev[l - lb];
        b.w = f.w;
        b.ldw = f.ldw;
        b.n1 = L.n1;
        b.n2 = L.n2;
        b.harmonic = L.g.harmonic;
        b.cx = L.g.cx;
        b.cy = L.g.cy;
    }
    const Level &B = ctx->lev[lb];
    A.b = LA::b(B) + B.ld + 1;
    A.ldb = B.ld;
    A.out = LA::r(B) + B.ld + 1;
    A.ldo = B.ld;
    A.ns = mg_ns(ctx);
    A.S = ctx->scal;
    return A;
}

// r32: level 0 reads the FP32 residual of the FP32-vector CG (T = float only)
// skip_first: the finest level's down leg is already queued (vcycle_first_leg)
template <typename T, int NS>
static int vcycle_fused_ns(impec_ctx *ctx, bool r32, bool skip_first = false) {
    constexpr bool F = std::is_same<T, float>::value;
    using LA = LevelArrays<T>;
    const int L = ctx->nlev;
    const dim3 blk(64 * SW_WAVES);
    const int lb = bottom_level(ctx);
    const int ltop = lb > 0 ? lb : L - 1;
    for (int l = skip_first ? 1 : 0; l < ltop; l++) {
        FusedArgsT<T> A = fused_args<T>(ctx, l);
        prof_begin(ctx);
        if (F && r32 && l == 0 && A.w2 != A.w)
            hipLaunchKernelGGL((k_mg_down<true, true, T, NS, F>), stream_grid(A), blk, 0, ctx->stream, A);
        else if (F && r32 && l == 0)
            hipLaunchKernelGGL((k_mg_down<true, false, T, NS, F>), stream_grid(A), blk, 0, ctx->stream, A);
        else if (l == 0 && A.w2 != A.w)
            hipLaunchKernelGGL((k_mg_down<true, true, T, NS>), stream_grid(A), blk, 0, ctx->stream, A);
        else if (l == 0) hipLaunchKernelGGL((k_mg_down<true, false, T, NS>), stream_grid(A), blk, 0, ctx->stream, A);
        else hipLaunchKernelGGL((k_mg_down<false, false, T, NS>), stream_grid(A), blk, 0, ctx->stream, A);
        prof_end(ctx, l == 0 ? "mg_down_l0" : "mg_coarse");
    }
    if (lb > 0) {
        prof_begin(ctx);
        hipLaunchKernelGGL((k_mg_bottom<T>), dim3(1), dim3(BOTTOM_THREADS), 0, ctx->stream, bottom_args<T>(ctx, lb));
        prof_end(ctx, "mg_coarse");
    } else {
        Level &Cl = ctx->lev[L - 1];
        LevArgs a = lev_args(ctx, L - 1);
        prof_begin(ctx);
        hipLaunchKernelGGL((k_coarsest<false, T>), dim3(1), dim3(64), 0, ctx->stream, a, LA::x(Cl) + Cl.ld + 1,
                           LA::b(Cl) + Cl.ld + 1);
        prof_end(ctx, "mg_coarse");
    }
    for (int l = ltop - 1; l >= 0; l--) {
        FusedArgsT<T> A = fused_args<T>(ctx, l);
        prof_begin(ctx);
        if (F && r32 && l == 0 && A.w2 != A.w)
            hipLaunchKernelGGL((k_mg_up<true, true, T, NS, F>), stream_grid(A), blk, 0, ctx->stream, A);
        else if (F && r32 && l == 0)
            hipLaunchKernelGGL((k_mg_up<true, false, T, NS, F>), stream_grid(A), blk, 0, ctx->stream, A);
        else if (l == 0 && A.w2 != A.w)
            hipLaunchKernelGGL((k_mg_up<true, true, T, NS>), stream_grid(A), blk, 0, ctx->stream, A);
        else if (l == 0) hipLaunchKernelGGL((k_mg_up<true, false, T, NS>), stream_grid(A), blk, 0, ctx->stream, A);
        else hipLaunchKernelGGL((k_mg_up<false, false, T, NS>), stream_grid(A), blk, 0, ctx->stream, A);
        prof_end(ctx, l == 0 ? "mg_up_l0" : "mg_coarse");
    }
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

template <typename T>
static int vcycle_fused(impec_ctx *ctx, bool r32 = false, bool skip_first = false) {
    return mg_ns(ctx) == 2 ? vcycle_fused_ns<T, 2>(ctx, r32, skip_first) : vcycle_fused_ns<T, 1>(ctx, r32, skip_first);
}

// the finest level's down leg of vcycle_fused (FP64 CG residual), queued on
// its own: it reads only the CG residual and writes the finest level's x and
// the next level's right-hand side, scratch of the preconditioner
template <typename T, int NS>
static int vcycle_first_leg_ns(impec_ctx *ctx) {
    const dim3 blk(64 * SW_WAVES);
    FusedArgsT<T> A = fused_args<T>(ctx, 0);
    prof_begin(ctx);
    if (A.w2 != A.w) hipLaunchKernelGGL((k_mg_down<true, true, T, NS>), stream_grid(A), blk, 0, ctx->stream, A);
    else hipLaunchKernelGGL((k_mg_down<true, false, T, NS>), stream_grid(A), blk, 0, ctx->stream, A);
    prof_end(ctx, "mg_down_l0");
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}
template <typename T>
static int vcycle_first_leg(impec_ctx *ctx) {
    return mg_ns(ctx) == 2 ? vcycle_first_leg_ns<T, 2>(ctx) : vcycle_first_leg_ns<T, 1>(ctx);
}

// FP32 copy of a mobility array (rows x cols cells from (0,0) of each)
template <class G>
__device__ __forceinline__ void to_f32_body(const G &gr_, const double *__restrict__ src, int lds, float *__restrict__ dst, int ldd, int rows,
                         int cols) {
    const int j = gr_.bx() * blockDim.x + threadIdx.x;
    const int i = gr_.by() * blockDim.y + threadIdx.y;
    if (i < rows && j < cols) dst[(long)i * ldd + j] = (float)src[(long)i * lds + j];
}
__global__ void k_to_f32(const double *__restrict__ src, int lds, float *__restrict__ dst, int ldd, int rows,
                         int cols) {
    to_f32_body(HwGrid{}, src, lds, dst, ldd, rows, cols);
}

// FP32 level arrays of the mixed-precision cycle (first use)
static int mixed_alloc(impec_ctx *ctx) {
    const bool aniso = ctx->K22 && ctx->K22 != ctx->K11;
    for (int l = 0; l < ctx->nlev; l++) {
        Level &L = ctx->lev[l];
        const size_t bytes = sizeof(float) * (size_t)(L.n1 + 2) * L.ld;
        float **arrs[5] = {&L.wf, &L.xf, &L.bf, &L.rf, &L.wf2};
        const int na = (l == 0 && aniso) ? 5 : 4;
        for (int k = 0; k < na; k++) {
            if (*arrs[k]) continue;
            if (hipMalloc(arrs[k], bytes) != hipSuccess)
                return set_error(ctx, IMPEC_ENOMEM, "mixed-precision level %d allocation failed", l);
            hipMemset(*arrs[k], 0, bytes);
        }
    }
    return IMPEC_OK;
}

// FP32 copies of levels [0, lend) of this solve's mobility hierarchy
static int mixed_convert(impec_ctx *ctx, int lend) {
    const bool aniso = ctx->K22 && ctx->K22 != ctx->K11;
    dim3 blk(64, 4);
    for (int l = 0; l < lend; l++) {
        Level &L = ctx->lev[l];
        LevArgs la = lev_args(ctx, l);
        hipLaunchKernelGGL(k_to_f32, grid2(L.n1, L.n2, blk), blk, 0, ctx->stream, la.w1, la.ldw, L.wf + L.ld + 1,
                           L.ld, L.n1, L.n2);
        if (l == 0 && L.wf2 && aniso)
            hipLaunchKernelGGL(k_to_f32, grid2(L.n1, L.n2, blk), blk, 0, ctx->stream, la.w2, la.ldw,
                               L.wf2 + L.ld + 1, L.ld, L.n1, L.n2);
    }
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

// FP32 level arrays and the FP32 copies of the whole hierarchy
static int mixed_setup(impec_ctx *ctx) {
    const int st = mixed_alloc(ctx);
    return st ? st : mixed_convert(ctx, ctx->nlev);
}

// the FP32-vector CG's arrays (first use) and this solve's FP32 face copies
static int mixed32_setup(impec_ctx *ctx) {
    const size_t bytes = sizeof(float) * (size_t)(ctx->N1 + 2) * ctx->lev[0].ld;
    float **v32[7] = {&ctx->pr32, &ctx->pe32, &ctx->pd32, &ctx->pz32, &ctx->pq32, &ctx->fx32, &ctx->fy32};
    for (float **v : v32) {
        if (*v) continue;
        if (hipMalloc(v, bytes) != hipSuccess) {
            *v = nullptr;
            return set_error(ctx, IMPEC_ENOMEM, "FP32 CG vector allocation failed");
        }
        hipMemset(*v, 0, bytes);
    }
    // fx: faces i = 0..N1 of columns 0..N2-1; fy: rows 0..N1-1, faces j = 0..N2
    const int ld = ctx->lev[0].ld;
    const dim3 blk(64, 4);
    hipLaunchKernelGGL(k_to_f32, grid2(ctx->N1 + 1, ctx->N2 + 1, blk), blk, 0, ctx->stream, ctx->fx, ld, ctx->fx32, ld,
                       ctx->N1 + 1, ctx->N2 + 1);
    hipLaunchKernelGGL(k_to_f32, grid2(ctx->N1 + 1, ctx->N2 + 1, blk), blk, 0, ctx->stream, ctx->fy, ld, ctx->fy32, ld,
                       ctx->N1 + 1, ctx->N2 + 1);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

// restart of the FP32 recursion from the FP64 residual in pr: r32 = r, e = 0,
// S_RR and S_RSUM of r32
static int cg_restart32(impec_ctx *ctx, const CgArgs &c) {
    const Grid2 vg = vec_grid(ctx->N1, ctx->N2);
    hipLaunchKernelGGL(k_cg_restart32, vg.grid, vg.block, 0, ctx->stream, c, ctx->pr, ctx->pr32, ctx->pe32, ctx->part);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return finish_sum(ctx, vg.nblocks, 2, S_RR);
}

// r = b~ - A_N x recomputed from the current iterate (the CG start-up
// kernel), with ||r||^2 returned; also (re)starts the recursion
template <bool DIAG>
static int cg_residual(impec_ctx *ctx, const LevArgs &a0, const CgArgs &c, int nb, double *rr) {
    prof_begin(ctx);
    hipLaunchKernelGGL(k_cg_init<DIAG>, dim3(nb), dim3(256), 0, ctx->stream, a0, c, ctx->p, ctx->q, ctx->pr, ctx->part);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    int st = finish_sum(ctx, nb, 2, S_RR, pub_next(ctx));
    prof_end(ctx, "cg_init");
    if (st) return st;
    if ((st = wait_scalars(ctx))) return st;
    *rr = ctx->h_scal[S_RR];
    return IMPEC_OK;
}

// one CG iteration with the one-kernel-per-pass V-cycle
template <bool DIAG>
static int cg_iteration_unfused(impec_ctx *ctx, const LevArgs &a0, const CgArgs &c, int nb, bool first) {
    const Level &L0 = ctx->lev[0];
    hipStream_t s = ctx->stream;
    int st;
    prof_begin(ctx);
    hipLaunchKernelGGL(k_cg_proj_in, dim3(nb), dim3(256), 0, s, c, ctx->pr, L0.b, ctx->scal);
    prof_end(ctx, "cg_vector");
    if ((st = vcycle<DIAG>(ctx))) return st;
    prof_begin(ctx);
    hipLaunchKernelGGL(k_sum, dim3(nb), dim3(256), 0, s, c, L0.x, ctx->part);
    if ((st = finish_sum(ctx, nb, 1, S_ZSUM))) return st;
    hipLaunchKernelGGL(k_cg_z, dim3(nb), dim3(256), 0, s, c, L0.x, ctx->pr, ctx->pz, ctx->scal, ctx->part);
    if ((st = finish_sum(ctx, nb, 1, S_RHO))) return st;
    hipLaunchKernelGGL(k_cg_dir, dim3(nb), dim3(256), 0, s, c, ctx->pz, ctx->pd, ctx->scal, first);
    prof_end(ctx, "cg_vector");
    prof_begin(ctx);
    hipLaunchKernelGGL(k_cg_matvec<DIAG>, dim3(nb), dim3(256), 0, s, a0, c, ctx->pd, ctx->pq, ctx->part);
    prof_end(ctx, "cg_matvec");
    if ((st = finish_sum(ctx, nb, 1, S_DQ))) return st;
    prof_begin(ctx);
    hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(256), 0, s, c, ctx->p, ctx->pd, ctx->pr, ctx->pq, ctx->scal,
                       ctx->part);
    if ((st = finish_sum(ctx, nb, 2, S_RR))) return st;
    hipLaunchKernelGGL(k_copy_scal, dim3(1), dim3(1), 0, s, ctx->scal, (int)S_RHO_OLD, (int)S_RHO);
    prof_end(ctx, "cg_vector");
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

template <typename T>
static DirMvArgs<T> dir_mv_args(impec_ctx *ctx, const LevArgs &a0, const CgArgs &c, long waves = IMPEC_DM_WAVES) {
    DirMvArgs<T> A;
    memset(&A, 0, sizeof(A));
    A.L = a0;
    A.c = c;
    A.x0 = LevelArrays<T>::r(ctx->lev[0]);
    A.d0 = ctx->pd;
    A.d1 = ctx->pz;
    A.q = ctx->pq;
    A.S = ctx->scal;
    A.part = ctx->part;
    A.nstrip = (c.n2 + DM_OUT - 1) / DM_OUT;
    // no more chunks than the target: 4096 waves are 1024 workgroups, four
    // per CU; a few more put a fifth on some CUs, which then finish last
    long R = chunk_rows(c.n1, A.nstrip, waves, IMPEC_WAVE_CAP);
    if (R < 8) R = 8;
    A.R = (int)R;
    A.nwave = A.nstrip * ((c.n1 + A.R - 1) / A.R);
    return A;
}

// one CG iteration with the streaming V-cycle (nu = 1) in storage type T
// (double: plain FP64; float: the FP32 cycle of the mixed-precision solve);
// the CG vectors, the residual and every reduction stay FP64; d ping-pongs.
// first_leg_queued: the V-cycle's first leg was queued ahead (vcycle_first_leg)
template <typename T>
static int cg_iteration_fused(impec_ctx *ctx, const LevArgs &a0, const CgArgs &c, bool first, double **dold,
                              double **dnew, bool first_leg_queued = false) {
    const Level &L0 = ctx->lev[0];
    hipStream_t s = ctx->stream;
    const Grid2 vg = vec_grid(ctx->N1, ctx->N2);
    const int nwave0 = fused_args<T>(ctx, 0).nwave;
    const double inv_n = 1.0 / (double)c.n;
    int st;
    if ((st = vcycle_fused<T>(ctx, false, first_leg_queued))) return st;
    hipLaunchKernelGGL(k_finish_rho, dim3(1), dim3(1024), 0, s, ctx->part, nwave0, inv_n, ctx->scal);
    const DirMvArgs<T> dm = dir_mv_args<T>(ctx, a0, c);
    prof_begin(ctx);
    hipLaunchKernelGGL((k_dir_mv<T>), dim3((dm.nwave + SW_WAVES - 1) / SW_WAVES), dim3(64 * SW_WAVES), 0, s, dm, *dold,
                       *dnew, first ? 1 : 0);
    prof_end(ctx, "cg_dir_matvec");
    if ((st = finish_sum(ctx, dm.nwave, 1, S_DQ))) return st;
    prof_begin(ctx);
    hipLaunchKernelGGL(k_cg_update2, vg.grid, vg.block, 0, s, c, ctx->p, *dnew, ctx->pr, ctx->pq, ctx->scal,
                       ctx->part);
    prof_end(ctx, "cg_update");
    if ((st = finish_sum(ctx, vg.nblocks, 2, S_RR, pub_next(ctx)))) return st;
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    double *t = *dold;
    *dold = *dnew;
    *dnew = t;
    return IMPEC_OK;
}

// one iteration of the FP32-vector CG (IMPEC_PREC_MIXED32): the FP32 V-cycle
// on the FP32 residual, the FP32 direction and matvec (FP32 faces), the FP32
// updates of the residual and of the correction accumulated since the last
// reliable update; every dot product and scalar in FP64
static int cg_iteration_mixed32(impec_ctx *ctx, const LevArgs &a0, const CgArgs &c, bool first, float **dold,
                                float **dnew) {
    hipStream_t s = ctx->stream;
    const Grid2 vg = vec_grid(ctx->N1, ctx->N2);
    const int nwave0 = fused_args<float>(ctx, 0).nwave;
    const double inv_n = 1.0 / (double)c.n;
    int st;
    if ((st = vcycle_fused<float>(ctx, true))) return st;
    hipLaunchKernelGGL(k_finish_rho, dim3(1), dim3(1024), 0, s, ctx->part, nwave0, inv_n, ctx->scal);
    const DirMvArgs<float> d64 = dir_mv_args<float>(ctx, a0, c, IMPEC_DM_WAVES_F32);
    DirMvArgs<float, float> dm;
    memset(&dm, 0, sizeof(dm));
    dm.L = d64.L; dm.c = d64.c; dm.x0 = d64.x0;
    dm.d0 = ctx->pd32; dm.d1 = ctx->pz32; dm.q = ctx->pq32;
    dm.fx32 = ctx->fx32; dm.fy32 = ctx->fy32;
    dm.S = d64.S; dm.part = d64.part;
    dm.R = d64.R; dm.nstrip = d64.nstrip; dm.nwave = d64.nwave;
    prof_begin(ctx);
    hipLaunchKernelGGL((k_dir_mv<float, float>), dim3((dm.nwave + SW_WAVES - 1) / SW_WAVES), dim3(64 * SW_WAVES), 0,
                       s, dm, *dold, *dnew, first ? 1 : 0);
    prof_end(ctx, "cg_dir_matvec");
    if ((st = finish_sum(ctx, dm.nwave, 1, S_DQ))) return st;
    prof_begin(ctx);
    hipLaunchKernelGGL(k_cg_update32, vg.grid, vg.block, 0, s, c, ctx->pe32, *dnew, ctx->pr32, ctx->pq32, ctx->scal,
                       ctx->part);
    prof_end(ctx, "cg_update");
    if ((st = finish_sum(ctx, vg.nblocks, 2, S_RR, pub_next(ctx)))) return st;
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    float *t = *dold;
    *dold = *dnew;
    *dnew = t;
    return IMPEC_OK;
}

// the FP32-vector CG's (re)start from the iterate in p: the FP64 residual of
// the shift formulation (||r||^2 into *rr) stored rounded in r32, e cleared
static int cg_residual32(impec_ctx *ctx, const LevArgs &a0, const CgArgs &c, int nb, double *rr) {
    prof_begin(ctx);
    hipLaunchKernelGGL(k_cg_init32, dim3(nb), dim3(256), 0, ctx->stream, a0, c, ctx->p, ctx->q, ctx->pr32, ctx->pe32,
                       ctx->part);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    int st = finish_sum(ctx, nb, 2, S_RR, pub_next(ctx));
    prof_end(ctx, "cg_init");
    if (st || (st = wait_scalars(ctx))) return st;
    *rr = ctx->h_scal[S_RR];
    return IMPEC_OK;
}

// a reliable update of the FP32-vector CG: p += e in FP64, then the restart
// from the FP64 residual; the direction is kept
static int cg_reliable32(impec_ctx *ctx, const LevArgs &a0, const CgArgs &c, int nb, double *rr) {
    const Grid2 vg = vec_grid(ctx->N1, ctx->N2);
    prof_begin(ctx);
    hipLaunchKernelGGL(k_cg_fold32, vg.grid, vg.block, 0, ctx->stream, c, ctx->p, ctx->pe32);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    prof_end(ctx, "cg_reliable");
    return cg_residual32(ctx, a0, c, nb, rr);
}

// a reliable update falls due when the FP32 residual has dropped by this factor
// (in norm) since the last one, or meets the tolerance
#ifndef IMPEC_RU_DELTA
#define IMPEC_RU_DELTA 1e-4
#endif

// Multigrid-preconditioned CG for the reference's pinned system.
//  phase 0: consistent Neumann formulation (when A_N is symmetric), then the
//           shift to the pinned value;
//  phase 1: the pinned row and column eliminated, from the current iterate;
//           runs only when phase 0 was not applicable or the shifted iterate's
//           pinned residual exceeds the tolerance (the assembled rows sum to
//           zero only to rounding, so a large shift leaves a small residual).
// Each phase confirms convergence on a recomputed residual and restarts from
// it when the recursion has drifted.
template <bool DIAG, bool FUSED, typename T = double>
static int solve_cg(impec_ctx *ctx, int *its_out, double *relres_out) {
    constexpr bool F32 = std::is_same<T, float>::value;
    int st = launch_mobility(ctx);
    if (st) return st;
    // mixed precision: the coarsening writes the FP32 copies of levels >= 1
    if (F32 && (st = mixed_alloc(ctx))) return st;
    st = build_hierarchy(ctx, F32);
    if (st) return st;
    if (F32 && (st = mixed_convert(ctx, 1))) return st;
    // the FP32-vector CG with reliable FP64 updates (streaming 5-point path)
    const bool m32 = std::is_same<T, float>::value && FUSED && !DIAG && ctx->P.precision == IMPEC_PREC_MIXED32;
    const Level &L0 = ctx->lev[0];
    LevArgs a0 = lev_args(ctx, 0);
    if (!DIAG && (!IMPEC_NO_FACES || m32)) {
        dim3 blk(64, 4);
        hipLaunchKernelGGL(k_faces, grid2(ctx->N1 + 1, ctx->N2 + 1, blk), blk, 0, ctx->stream, a0, ctx->fx, ctx->fy);
        IMPEC_HIP_CHECK(ctx, hipGetLastError());
        a0.fx = ctx->fx;
        a0.fy = ctx->fy;
    }
    if (m32 && (st = mixed32_setup(ctx))) return st;
    CgArgs c;
    c.n1 = ctx->N1; c.n2 = ctx->N2; c.ld = L0.ld; c.ldp = ctx->lds; c.B = ctx->B;
    c.n = (long)ctx->N1 * ctx->N2;
    c.pin_i = ctx->pin_i; c.pin_j = ctx->pin_j;
    c.dpin = ctx->kappa / ctx->dA;           // pinned diagonal kappa/dA (MGSolver.cpp:555)
    c.ppin = -ctx->q_pin / c.dpin;           // its solution for b = -q
    c.bsum = ctx->bsum;
    const int nb = red_grid(c.n);
    hipStream_t s = ctx->stream;
    const double tol = ctx->P.eps;
    const double bb = ctx->bnorm * ctx->bnorm;
    const int maxit = ctx->P.max_iterations > 0 ? ctx->P.max_iterations : 10000;

    const int guess = ctx->P.initial_guess;
    const bool warm = guess != IMPEC_GUESS_ZERO && ctx->have_prev_p;
    // the zero iterate: in the shift formulation its residual is b itself, formed
    // together with p = 0 (k_cg_init_zero) instead of zeroing p and applying A
    const bool zero_start = !warm && ctx->symmetric;
    if (!warm && !zero_start) hipLaunchKernelGGL(k_zero_interior, dim3(nb), dim3(256), 0, s, c, ctx->p);
    else if (guess == IMPEC_GUESS_LEARNED || guess == IMPEC_GUESS_EXTRAPOLATE) {
        prof_begin(ctx);
        double *d = nullptr;   // the network's increment, in pd (free until the CG starts)
        if (guess == IMPEC_GUESS_LEARNED && ctx->net) {
            d = ctx->pd + L0.ld + 1;
            st = cnn_increment(ctx, d, L0.ld);
            if (st) return st;
        }
        const double *pold = (ctx->p_old && ctx->n_solves >= 2) ? ctx->p_old : nullptr;
        const bool project = !DIAG && ctx->symmetric;
        if (d || pold) {
            if (project) {
                hipLaunchKernelGGL(k_guess_dots, dim3(nb), dim3(256), 0, s, a0, c, ctx->p, pold, ctx->q, d, ctx->part);
                if ((st = finish_sum(ctx, nb, 5, S_G0))) return st;
                hipLaunchKernelGGL(k_guess_solve, dim3(1), dim3(1), 0, s, ctx->scal, d ? 1 : 0, pold ? 1 : 0);
            }
            hipLaunchKernelGGL(k_guess_apply, dim3(nb), dim3(256), 0, s, c, ctx->p, pold, d, ctx->scal, project ? 1 : 0);
            IMPEC_HIP_CHECK(ctx, hipGetLastError());
        }
        prof_end(ctx, "cnn_guess");
    }
    int its = 0, status = IMPEC_OK;
    double relres = 0., rr = 0.;
    double floor_rel = 0., tol_eff = tol;   // the true residual's rounding floor, max(tol, floor)
    double *dold = ctx->pd, *dnew = ctx->pz;
    float *dold32 = ctx->pd32, *dnew32 = ctx->pz32;
    double rr_ref = 0.;   // FP64 residual norm^2 at the FP32 recursion's last restart
    int restarts_all = 0, elim_its = 0;
    for (int phase = ctx->symmetric ? 0 : 1; phase < 2; phase++) {
        c.shift = phase == 0;
        const int its_at_start = its;
        // the FP32-vector CG runs the consistent phase; the elimination phase, a
        // short continuation at the FP64 rounding floor of a worse-conditioned
        // system, keeps FP64 vectors (its FP32 recursion stalls there)
        const bool ru = m32 && c.shift;
        if (phase == 0 && zero_start) {
            prof_begin(ctx);
            hipLaunchKernelGGL(k_cg_init_zero, dim3(nb), dim3(256), 0, s, c, ctx->p, ctx->q, ctx->pr, ctx->part);
            IMPEC_HIP_CHECK(ctx, hipGetLastError());
            st = finish_sum(ctx, nb, 2, S_RR, pub_next(ctx));
            prof_end(ctx, "cg_init");
            if (st || (st = wait_scalars(ctx))) return st;
            rr = ctx->h_scal[S_RR];
            if (ru && (st = cg_restart32(ctx, c))) return st;
        } else if (ru) {
            if ((st = cg_residual32(ctx, a0, c, nb, &rr))) return st;
        } else if ((st = cg_residual<DIAG>(ctx, a0, c, nb, &rr))) {
            return st;
        }
        rr_ref = rr;
        // p to the reference's pinned system (the shift, or the pinned value) and
        // that system's true relative residual
        auto finish_phase = [&]() -> int {
            if (c.shift) {
                hipLaunchKernelGGL(k_read_pin, dim3(1), dim3(1), 0, s, c, ctx->p, ctx->scal);
                hipLaunchKernelGGL(k_shift, dim3(nb), dim3(256), 0, s, c, ctx->p, ctx->scal);
            } else {
                hipLaunchKernelGGL(k_set_pin, dim3(1), dim3(1), 0, s, c, ctx->p);
            }
            int nparts = nb;
            const TresArgs ta = tres_args(a0, c, ctx->p, ctx->q, ctx->part);
            if (!DIAG && ta.nwave <= RED_BLOCKS) {   // the 5-point operator: the streamed pass
                hipLaunchKernelGGL(k_true_rows, dim3(ta.gx), dim3(64 * SW_WAVES), 0, s, ta);
                nparts = ta.nwave;
            } else {
                hipLaunchKernelGGL(k_true_residual<DIAG>, dim3(nb), dim3(256), 0, s, a0, c, ctx->p, ctx->q,
                                   ctx->part);
            }
            IMPEC_HIP_CHECK(ctx, hipGetLastError());
            int e = finish_sum(ctx, nparts, 2, S_TT, pub_next(ctx));
            if (!e) e = wait_scalars(ctx);
            if (!e) {
                relres = ctx->bnorm > 0 ? sqrt(ctx->h_scal[S_TT]) / ctx->bnorm : 0.;
                floor_rel = ctx->bnorm > 0 ? IMPEC_FLOOR_C * 0x1p-53 * sqrt(ctx->h_scal[S_FLOOR]) / ctx->bnorm : 0.;
                tol_eff = tol > floor_rel ? tol : floor_rel;
            }
            return e;
        };
        int since = 0, restarts = 0;
        double rr_check = -1.;   // true residual norm^2 at the previous restart (-1: none yet)
        bool finished = false, at_floor = false;
        for (;;) {
            bool broke = false;
            double rr_mark = rr;   // the recursive residual when it last halved
            int stall = 0;
            // Lookahead (streaming iterations with FP64 vectors): while the last
            // known residual norm^2 lies more than cg_ahead times above the
            // target, the next V-cycle's first leg is queued before the host
            // reads the iteration's scalars, so the device does not wait for
            // the host.  The leg reads only r; every path below that changes r
            // leaves the loop, and the next entry queues the leg afresh.
            bool leg_queued = false;
            const bool lookahead = !ru && FUSED && !DIAG && ctx->cg_ahead >= 0. && top_level(ctx) >= 1;
            while (!(rr <= tol * tol * bb)) {
                if (its >= maxit) { status = IMPEC_ENOCONV; break; }
                if (!(rr == rr)) { status = IMPEC_ENAN; break; }
                // the streaming iterations publish the scalars from their
                // last reduction; the unfused one is published separately
                if (ru) st = cg_iteration_mixed32(ctx, a0, c, since == 0, &dold32, &dnew32);
                else if (FUSED && !DIAG) st = cg_iteration_fused<T>(ctx, a0, c, since == 0, &dold, &dnew, leg_queued);
                else st = cg_iteration_unfused<DIAG>(ctx, a0, c, nb, since == 0);
                if (st) return st;
                const bool ahead = lookahead && rr > ctx->cg_ahead * (tol * tol * bb);
                if (ahead && (st = vcycle_first_leg<T>(ctx))) return st;
                leg_queued = ahead;
                if ((st = (ru || (FUSED && !DIAG)) ? wait_scalars(ctx) : fetch_scal(ctx))) return st;
                its++;
                since++;
                // the recursion broke down (it took no step): restart it from
                // the true residual as on convergence
                if (cg_broke_down(ctx->h_scal)) {
                    broke = true;
                    break;
                }
                rr = ctx->h_scal[S_RR];
                // FP32-vector CG: the correction goes into p and the FP64 residual
                // replaces the FP32 one when the latter has fallen by IMPEC_RU_DELTA,
                // meets the tolerance or is not finite; only an FP64 residual ends
                // the loop
                if (ru && (rr <= tol * tol * bb || rr <= IMPEC_RU_DELTA * IMPEC_RU_DELTA * rr_ref || !(rr == rr))) {
                    if ((st = cg_reliable32(ctx, a0, c, nb, &rr))) return st;
                    rr_ref = rr;
                }
                if (rr <= 0.25 * rr_mark) {
                    rr_mark = rr;
                    stall = 0;
                } else if (++stall >= IMPEC_STALL_ITS && rr == rr && !(rr <= tol * tol * bb)) {
                    // the pinned system's true residual and floor (p moves by a
                    // constant in the consistent phase, its pinned value is kept
                    // in the elimination phase: the recursion is unchanged)
                    if (ru) {   // the FP32 corrections into p first
                        if ((st = cg_reliable32(ctx, a0, c, nb, &rr))) return st;
                        rr_ref = rr;
                    }
                    if ((st = finish_phase())) return st;
                    if (relres <= tol_eff) {
                        at_floor = true;
                        break;
                    }
                    const double gate = IMPEC_STALL_GATE * floor_rel;
                    if (rr <= gate * gate * bb) break;   // restart from the true residual
                    rr_mark = rr;
                    stall = 0;
                }
            }
            if (leg_queued) ctx->stats.cg_ahead_misses++;   // queued for an iteration that did not run
            if (at_floor) {
                finished = true;
                break;
            }
            if (ru && (status != IMPEC_OK || broke)) {   // keep the corrections made so far
                const Grid2 vg = vec_grid(ctx->N1, ctx->N2);
                hipLaunchKernelGGL(k_cg_fold32, vg.grid, vg.block, 0, s, c, ctx->p, ctx->pe32);
                hipLaunchKernelGGL(k_cg_restart32, vg.grid, vg.block, 0, s, c, ctx->pr, ctx->pr32, ctx->pe32,
                                   ctx->part);
            }
            if (status != IMPEC_OK) break;
            if (c.shift) {
                // the consistent formulation confirms convergence on the pinned
                // system itself: shift, then its true residual.  When that fails
                // (drift of the recursion, or a shift too large for the rows'
                // rounding) the elimination phase continues from this iterate
                // and restarts on drift itself.
                if ((st = finish_phase())) return st;
                finished = true;
                break;
            }
            const double rr_before = rr_check;
            if ((st = cg_residual<DIAG>(ctx, a0, c, nb, &rr))) return st;
            if (ru) {
                if ((st = cg_restart32(ctx, c))) return st;
                rr_ref = rr;
            }
            rr_check = rr;
            // restart while each restart at least halves the true residual
            // (up to IMPEC_MAX_RESTARTS), and at least 3 times
            const bool progress = rr_before < 0. || rr <= 0.25 * rr_before;
            if (rr <= tol_eff * tol_eff * bb || restarts >= IMPEC_MAX_RESTARTS || (restarts >= 3 && !progress)) break;
            restarts++;
            since = 0;
        }
        if (!finished && (st = finish_phase())) return st;
        restarts_all += restarts;
        if (phase == 1) elim_its = its - its_at_start;
        if (status != IMPEC_OK || relres <= tol_eff) break;
    }
    // the elimination phase is the last: its restarts are spent and the
    // pinned system's true residual is still above the tolerance and above
    // what FP64 can resolve
    bool missed = false;
    if (status == IMPEC_OK && !(relres <= tol_eff)) {
        status = IMPEC_ENOCONV;
        missed = true;
    }
    const bool used_d = warm && guess == IMPEC_GUESS_LEARNED && ctx->net;
    const bool used_e = warm && (guess == IMPEC_GUESS_LEARNED || guess == IMPEC_GUESS_EXTRAPOLATE) && ctx->p_old &&
                        ctx->n_solves >= 2;
    ctx->stats.last_guess_alpha = used_d ? ctx->h_scal[S_GUESS_ALPHA] : 0.;
    ctx->stats.last_guess_gamma = used_e ? ctx->h_scal[S_GUESS_GAMMA] : 0.;
    ctx->have_prev_p = 1;
    ctx->n_solves += 1;
    ctx->stats.last_its = its;
    ctx->stats.last_vcycles = its;
    ctx->stats.last_relres = relres;
    ctx->stats.last_relres_floor = floor_rel;
    ctx->stats.last_restarts = restarts_all;
    ctx->stats.last_elim_its = elim_its;
    ctx->stats.total_its += its;
    ctx->stats.calls += 1;
    if (its_out) *its_out = its;
    if (relres_out) *relres_out = relres;
    if (missed)
        return set_error(ctx, status,
                         "pressure solve: true residual %.3e above eps %.3e after %d iterations and its restarts",
                         relres, tol, its);
    if (status == IMPEC_ENOCONV)
        return set_error(ctx, status, "pressure solve: no convergence after %d iterations (relres %.3e)", its, relres);
    if (status == IMPEC_ENAN) return set_error(ctx, status, "pressure solve: non-finite residual");
    return IMPEC_OK;
}

int pressure_solve(impec_ctx *ctx, int *its, double *relres) {
    // the pressure before this solve (training pairs and the learned guess)
    if (ctx->p_old && ctx->p_prev) {   // p_old <- the pressure before the previous solve
        double *t = ctx->p_old;
        ctx->p_old = ctx->p_prev;
        ctx->p_prev = t;
    }
    if (ctx->p_prev && ctx->have_prev_p)
        IMPEC_HIP_CHECK(ctx, hipMemcpyAsync(ctx->p_prev, ctx->p, sizeof(double) * (size_t)(ctx->N1 + 2 * ctx->B) *
                                                                     ctx->lds,
                                            hipMemcpyDeviceToDevice, ctx->stream));
    if (ctx->P.nu != 1.) return solve_cg<true, false>(ctx, its, relres);
    const bool fused_ok = ((ctx->P.mg_pre <= 1 && ctx->P.mg_post <= 1) || mg_ns(ctx) == 2) &&
                          fused_args<double>(ctx, 0).nwave <= RED_BLOCKS && !ctx->force_unfused;
    // mixed precision: FP32 V-cycles as the preconditioner of the FP64 CG (a
    // Krylov-accelerated defect correction: residual, reductions, update and
    // the coarsest solve in FP64); on the streaming 5-point path only
    if (fused_ok && (ctx->P.precision == IMPEC_PREC_MIXED || ctx->P.precision == IMPEC_PREC_MIXED32))
        return solve_cg<false, true, float>(ctx, its, relres);
    if (fused_ok) return solve_cg<false, true>(ctx, its, relres);
    return solve_cg<false, false>(ctx, its, relres);
}

int pressure_operator_dump(impec_ctx *ctx, double *A9_host) {
    int st = launch_mobility(ctx);
    if (st) return st;
    LevArgs a0 = lev_args(ctx, 0);
    const size_t bytes = sizeof(double) * 9 * (size_t)ctx->N1 * ctx->N2;
    double *dA = nullptr;
    IMPEC_HIP_CHECK(ctx, hipMalloc(&dA, bytes));
    dim3 blk(64, 4);
    const double dpin = ctx->kappa / ctx->dA;
    if (ctx->P.nu != 1.)
        hipLaunchKernelGGL(k_op_dump<true>, grid2(ctx->N1, ctx->N2, blk), blk, 0, ctx->stream, a0, ctx->pin_i,
                           ctx->pin_j, dpin, dA);
    else
        hipLaunchKernelGGL(k_op_dump<false>, grid2(ctx->N1, ctx->N2, blk), blk, 0, ctx->stream, a0, ctx->pin_i,
                           ctx->pin_j, dpin, dA);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(A9_host, dA, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(dA);
    if (e != hipSuccess) return set_error(ctx, IMPEC_EHIP, "operator dump: %s", hipGetErrorString(e));
    return IMPEC_OK;
}


// ===========================================================================
// ensembles: the 5-point streaming solve of every member in batched launches.
// Each member runs the CG of solve_cg (consistent Neumann formulation) with
// its loop state on the device: kb_ctl_* make, per member, the decisions the
// host loop of solve_cg makes (restart on drift, iteration limit, non-finite
// residual), so a member's iterates, residuals and iteration count are those
// of its own solve_cg.
// ===========================================================================
#define ENS_CHECK(e, call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ens_set_error((e), IMPEC_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,   \
                                 hipGetErrorString(e_));                                       \
    } while (0)

struct FacesRec {
    LevArgs L;
    double *fx, *fy;
    int gx, gy;
};
struct CoarsenRec {
    const double *wf;
    int ldf, n1, n2;
    double *wc;
    int ldc, m1, m2, gx, gy;
};
struct ToF32Rec {
    const double *src;
    int lds;
    float *dst;
    int ldd, rows, cols, gx, gy;
};
struct ZeroRec {
    CgArgs c;
    double *p;
    const double *S;
    int gx, gy;
};
struct CgInitRec {
    LevArgs L;
    CgArgs c;
    double *p;
    const double *q;
    double *r, *part;
    const double *S;
    int gx, gy;
};
template <typename T>
struct CoarsestRec {
    LevArgs L;
    T *x;
    const T *b;
    const double *S;
};
struct RhoRec {
    const double *part;
    int nblocks;
    double inv_n;
    double *S;
};
struct Upd2Rec {
    CgArgs c;
    double *p, *d0, *d1, *r;
    const double *q, *S;
    double *part;
    int gx, gy;
};
struct CtlRec {
    double *S;
    int always_zero;
};
struct PinRec {
    CgArgs c;
    double *p, *S;
};
struct ShiftRec {
    CgArgs c;
    double *p;
    const double *S;
    int gx, gy;
};

template <typename T>
struct LegTabs {
    Batch<FusedArgsT<T>> down[MAX_LEVELS], up[MAX_LEVELS];
    Batch<CoarsestRec<T>> coarsest;    // members without a bottom kernel
    Batch<BottomArgs<T>> bottom;       // members with one
    Batch<DirMvArgs<T>> dir;
    void free_all() {
        for (int l = 0; l < MAX_LEVELS; l++) {
            batch_free(down[l]);
            batch_free(up[l]);
        }
        batch_free(coarsest);
        batch_free(bottom);
        batch_free(dir);
    }
};

// The members of a batch that are in a given loop state (ActList), compacted
// on the device after every state change, so that the iteration kernels visit
// only the workgroups (waves) of members still iterating.
struct ListDesc {
    double *const *S;    // scalar arrays of the batch's members
    const int *bstart;   // the batch's unit offsets
    int nmem;
    ActList act;
};

struct EnsPressTabs {
    int mixed;
    int lmax;                           // most levels of any member
    // compacted lists: iter[0] the level-0 vector update, iter[1 + l] the legs
    // of level l, iter[lmax] the direction update and matvec; check: the
    // residual recomputation
    ListDesc *d_iter = nullptr, *d_check = nullptr;
    int n_iter = 0;
    std::vector<ListDesc> h_iter, h_check;
    std::vector<void *> list_mem;
    Batch<FacesRec> faces;
    Batch<CoarsenRec> coarsen[MAX_LEVELS];
    Batch<ToF32Rec> tof32[MAX_LEVELS];
    Batch<ZeroRec> zero;
    Batch<CgInitRec> init;
    Batch<FinishRec> fin_init, fin_dq, fin_rr, fin_tt;
    LegTabs<double> l64;
    LegTabs<float> l32;
    Batch<RhoRec> rho;
    Batch<Upd2Rec> upd;
    Batch<CtlRec> ctl;
    Batch<PinRec> pin;
    Batch<ShiftRec> shift;
    Batch<TresArgs> tres;
};

__device__ __forceinline__ bool in_state(const double *S, int need) { return need < 0 || S[S_EST] == (double)need; }

__global__ void kb_faces(const FacesRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        faces_body(G, R.L, R.fx, R.fy);
    });
}
__global__ void kb_coarsen_w(const CoarsenRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        coarsen_w_body(G, R.wf, R.ldf, R.n1, R.n2, R.wc, R.ldc, R.m1, R.m2);
    });
}
__global__ void kb_to_f32(const ToF32Rec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        to_f32_body(G, R.src, R.lds, R.dst, R.ldd, R.rows, R.cols);
    });
}
__global__ void kb_zero_p(const ZeroRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        if (R.S[S_ZEROP] == 0.) return;
        zero_interior_body(G, R.c, R.p);
    });
}
__global__ void kb_cg_init(const CgInitRec *__restrict__ tab, const int *__restrict__ start, int nmem, int need) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        if (!in_state(R.S, need)) return;
        cg_init_body<false>(G, R.L, R.c, R.p, R.q, R.r, R.part);
    });
}
template <bool FINE0, typename T, int NS>
__global__ void __launch_bounds__(64 * SW_WAVES)
    kb_mg_down(const FusedArgsT<T> *__restrict__ tab, const int *__restrict__ start, int nmem, int total, int need) {
    const int gw = hw_wave();
    if (gw >= total) return;
    const int m = member_of(start, nmem, gw);
    const FusedArgsT<T> A = tab[m];
    if (!in_state(A.S, need)) return;
    mg_down_body<FINE0, false, T, NS>(A, gw - start[m]);
}
template <bool FINE0, typename T, int NS>
__global__ void __launch_bounds__(64 * SW_WAVES)
    kb_mg_up(const FusedArgsT<T> *__restrict__ tab, const int *__restrict__ start, int nmem, int total, int need) {
    const int gw = hw_wave();
    if (gw >= total) return;
    const int m = member_of(start, nmem, gw);
    const FusedArgsT<T> A = tab[m];
    if (!in_state(A.S, need)) return;
    mg_up_body<FINE0, false, T, NS>(A, gw - start[m]);
}
template <typename T>
__global__ void __launch_bounds__(BOTTOM_THREADS) kb_mg_bottom(const BottomArgs<T> *__restrict__ tab, int need) {
    const BottomArgs<T> &A = tab[blockIdx.x];
    if (!in_state(A.S, need)) return;
    mg_bottom_body<T>(A);
}
template <typename T>
__global__ void kb_coarsest(const CoarsestRec<T> *__restrict__ tab, int need) {
    const CoarsestRec<T> R = tab[blockIdx.x];
    if (!in_state(R.S, need)) return;
    coarsest_body<false, T>(VGrid{0, 0, 1, 1}, R.L, R.x, R.b);
}
__global__ void __launch_bounds__(1024) kb_finish_rho(const RhoRec *__restrict__ tab, int need) {
    const RhoRec R = tab[blockIdx.x];
    if (!in_state(R.S, need)) return;
    finish_rho_body(VGrid{0, 0, 1, 1}, R.part, R.nblocks, R.inv_n, R.S);
}
__global__ void kb_update2(const Upd2Rec *__restrict__ tab, const int *__restrict__ start, int nmem, int need) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        if (!in_state(R.S, need)) return;
        const bool odd = ((long long)R.S[S_EITS]) & 1;
        cg_update2_body(G, R.c, R.p, odd ? R.d0 : R.d1, R.r, R.q, R.S, R.part);
    });
}
// after the first residual: iterate unless converged (solve_cg's loop entry)
__global__ void kb_ctl_first(const CtlRec *__restrict__ tab, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double *S = tab[m].S;
    const double rr = S[S_RR];
    S[S_EITS] = 0.;
    S[S_ESINCE] = 0.;
    S[S_ERST] = 0.;
    S[S_ERRPREV] = -1.;
    S[S_EMARK] = rr;
    S[S_ESTALL] = 0.;
    S[S_ESTATUS] = 0.;
    double est = EST_ITER;
    if (rr <= S[S_ETOL]) est = EST_DONE;
    else if (S[S_EMAXIT] <= 0.) { est = EST_DONE; S[S_ESTATUS] = 1.; }
    else if (!(rr == rr)) { est = EST_DONE; S[S_ESTATUS] = 2.; }
    S[S_EST] = est;
    if (!tab[m].always_zero) S[S_ZEROP] = 0.;
}
// after an iteration: count it; converged, broken-down and stalled recursions
// (residual norm not halved in IMPEC_STALL_ITS iterations) get their residual
// recomputed
__global__ void kb_ctl_iter(const CtlRec *__restrict__ tab, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double *S = tab[m].S;
    if (S[S_EST] != (double)EST_ITER) return;
    const double its = S[S_EITS] + 1.;
    S[S_EITS] = its;
    S[S_ESINCE] = S[S_ESINCE] + 1.;
    const double rr = S[S_RR];
    bool stalled = false;
    if (rr <= 0.25 * S[S_EMARK]) {
        S[S_EMARK] = rr;
        S[S_ESTALL] = 0.;
    } else {
        S[S_ESTALL] = S[S_ESTALL] + 1.;
        stalled = S[S_ESTALL] >= (double)IMPEC_STALL_ITS;
    }
    if (rr <= S[S_ETOL] || cg_broke_down(S)) S[S_EST] = EST_CHECK;
    else if (its >= S[S_EMAXIT]) { S[S_EST] = EST_DONE; S[S_ESTATUS] = 1.; }
    else if (!(rr == rr)) { S[S_EST] = EST_DONE; S[S_ESTATUS] = 2.; }
    else if (stalled) S[S_EST] = EST_CHECK;
}
// after a recomputed residual: done, or restart the recursion
__global__ void kb_ctl_check(const CtlRec *__restrict__ tab, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double *S = tab[m].S;
    if (S[S_EST] != (double)EST_CHECK) return;
    const double rr = S[S_RR];
    // restart while each restart at least halves the true residual (up to
    // IMPEC_MAX_RESTARTS), and at least 3 times, as the single-context loop
    const double before = S[S_ERRPREV];
    S[S_ERRPREV] = rr;
    const bool progress = before < 0. || rr <= 0.25 * before;
    if (rr <= S[S_ETOL] || S[S_ERST] >= (double)IMPEC_MAX_RESTARTS || (S[S_ERST] >= 3. && !progress)) {
        S[S_EST] = EST_DONE;
        return;
    }
    S[S_ERST] = S[S_ERST] + 1.;
    S[S_ESINCE] = 0.;
    S[S_EMARK] = rr;
    S[S_ESTALL] = 0.;
    if (S[S_EITS] >= S[S_EMAXIT]) { S[S_EST] = EST_DONE; S[S_ESTATUS] = 1.; }
    else if (!(rr == rr)) { S[S_EST] = EST_DONE; S[S_ESTATUS] = 2.; }
    else S[S_EST] = EST_ITER;
}
__global__ void kb_read_pin(const PinRec *__restrict__ tab, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const PinRec R = tab[m];
    read_pin_body(VGrid{0, 0, 1, 1}, R.c, R.p, R.S);
}
__global__ void kb_shift(const ShiftRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        shift_body(G, R.c, R.p, R.S);
    });
}
__global__ void __launch_bounds__(64 * SW_WAVES) kb_true_rows(const TresArgs *__restrict__ tab,
                                                               const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const TresArgs &R, const VGrid &G) {
        tres_rows_body(R, G.bx() * SW_WAVES + (int)(threadIdx.x >> 6));
    });
}

// one list per workgroup: flags of the members in state `need`, their units,
// block-wide inclusive scans, compacted output
__global__ void __launch_bounds__(256) kb_compact(const ListDesc *__restrict__ lists, int need) {
    const ListDesc d = lists[blockIdx.x];
    if (d.nmem == 0) return;
    const int t = threadIdx.x;
    __shared__ int sf[256], su[256];
    __shared__ int base_n, base_u;
    if (t == 0) {
        base_n = 0;
        base_u = 0;
    }
    __syncthreads();
    for (int c0 = 0; c0 < d.nmem; c0 += 256) {
        const int m = c0 + t;
        int f = 0, u = 0;
        if (m < d.nmem && d.S[m][S_EST] == (double)need) {
            f = 1;
            u = d.bstart[m + 1] - d.bstart[m];
        }
        sf[t] = f;
        su[t] = u;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int a = t >= off ? sf[t - off] : 0, b = t >= off ? su[t - off] : 0;
            __syncthreads();
            sf[t] += a;
            su[t] += b;
            __syncthreads();
        }
        const int bn = base_n, bu = base_u;
        if (f) {
            const int pos = bn + sf[t] - 1;
            d.act.member[pos] = m;
            d.act.start[pos] = bu + su[t] - u;
        }
        __syncthreads();
        if (t == 255) {
            base_n = bn + sf[255];
            base_u = bu + su[255];
        }
        __syncthreads();
    }
    if (t == 0) {
        d.act.start[base_n] = base_u;
        d.act.n[0] = base_n;
    }
}


// per-wave version of act_for for the streaming legs: wave w of the launch
// takes a contiguous range of the listed members' waves
template <class F>
__device__ __forceinline__ void act_wave_for(const ActList &A, F &&f) {
    const int nact = A.n[0];
    if (nact == 0) return;
    const int total = A.start[nact], parts = (int)gridDim.x * SW_WAVES;
    const int per = (total + parts - 1) / parts;
    int vw = hw_wave() * per;
    const int end = vw + per < total ? vw + per : total;
    if (vw >= end) return;
    int k = member_of(A.start, nact, vw);
    int kend = A.start[k + 1];
    for (; vw < end; vw++) {
        if (vw >= kend) {
            k++;
            kend = A.start[k + 1];
        }
        f(A.member[k], vw - A.start[k]);
    }
}

template <bool FINE0, typename T, int NS>
__global__ void __launch_bounds__(64 * SW_WAVES) kp_mg_down(const FusedArgsT<T> *__restrict__ tab, ActList act) {
    act_wave_for(act, [&](int m, int loc) {
        const FusedArgsT<T> A = tab[m];
        mg_down_body<FINE0, false, T, NS>(A, loc);
    });
}
template <bool FINE0, typename T, int NS>
__global__ void __launch_bounds__(64 * SW_WAVES) kp_mg_up(const FusedArgsT<T> *__restrict__ tab, ActList act) {
    act_wave_for(act, [&](int m, int loc) {
        const FusedArgsT<T> A = tab[m];
        mg_up_body<FINE0, false, T, NS>(A, loc);
    });
}
// the direction arrays alternate every iteration of a solve: (d_old, d_new) =
// (d0, d1) on even iterations and (d1, d0) on odd ones, as in solve_cg
template <typename T>
__global__ void __launch_bounds__(64 * SW_WAVES) kp_dir_mv(const DirMvArgs<T> *__restrict__ tab, ActList act) {
    act_wave_for(act, [&](int m, int loc) {
        const DirMvArgs<T> &A = tab[m];
        const bool odd = ((long long)A.S[S_EITS]) & 1;
        const int first = A.S[S_ESINCE] == 0.;
        dir_mv_body<T>(A, odd ? A.d1 : A.d0, odd ? A.d0 : A.d1, first, loc);
    });
}
__global__ void kp_update2(const Upd2Rec *__restrict__ tab, ActList act) {
    act_for(tab, act, [&](const Upd2Rec &R, const VGrid &G) {
        const bool odd = ((long long)R.S[S_EITS]) & 1;
        cg_update2_body(G, R.c, R.p, odd ? R.d0 : R.d1, R.r, R.q, R.S, R.part);
    });
}
__global__ void kp_cg_init(const CgInitRec *__restrict__ tab, ActList act) {
    act_for(tab, act, [&](const CgInitRec &R, const VGrid &G) { cg_init_body<false>(G, R.L, R.c, R.p, R.q, R.r, R.part); });
}

// persistent grid sizes of the compacted launches (workgroups)
constexpr int PERSIST_LEG_WG = 2048, PERSIST_VEC_WG = 4096;

template <typename T>
static void free_legs(LegTabs<T> &t) { t.free_all(); }

void ens_press_free(impec_ensemble *e) {
    if (!e->pres) return;
    EnsPressTabs *t = e->pres;
    batch_free(t->faces);
    for (int l = 0; l < MAX_LEVELS; l++) {
        batch_free(t->coarsen[l]);
        batch_free(t->tof32[l]);
    }
    batch_free(t->zero); batch_free(t->init);
    batch_free(t->fin_init); batch_free(t->fin_dq); batch_free(t->fin_rr); batch_free(t->fin_tt);
    free_legs(t->l64);
    free_legs(t->l32);
    batch_free(t->rho); batch_free(t->upd); batch_free(t->ctl); batch_free(t->pin); batch_free(t->shift);
    batch_free(t->tres);
    for (void *p : t->list_mem) (void)hipFree(p);
    delete t;
    e->pres = nullptr;
}

// CgArgs of solve_cg's consistent-Neumann phase
static CgArgs cg_args(const impec_ctx *ctx) {
    const Level &L0 = ctx->lev[0];
    CgArgs c;
    c.n1 = ctx->N1; c.n2 = ctx->N2; c.ld = L0.ld; c.ldp = ctx->lds; c.B = ctx->B;
    c.n = (long)ctx->N1 * ctx->N2;
    c.pin_i = ctx->pin_i; c.pin_j = ctx->pin_j;
    c.dpin = ctx->kappa / ctx->dA;
    c.ppin = -ctx->q_pin / c.dpin;
    c.bsum = ctx->bsum;
    c.shift = 1;
    return c;
}

template <typename T>
static int build_legs(impec_ensemble *e, LegTabs<T> &t) {
    using LA = LevelArrays<T>;
    const int n = e->n;
    t.free_all();
    const int lmax = e->pres->lmax;
    for (int l = 0; l + 1 < lmax; l++) {
        std::vector<FusedArgsT<T>> recs(n);
        std::vector<int> units(n, 0);
        for (int k = 0; k < n; k++) {
            impec_ctx *ctx = e->m[k];
            if (l >= top_level(ctx)) continue;
            recs[k] = fused_args<T>(ctx, l);
            units[k] = recs[k].nwave;
        }
        ENS_CHECK(e, batch_upload(t.down[l], recs, units));
        ENS_CHECK(e, batch_upload(t.up[l], recs, units));
    }
    std::vector<CoarsestRec<T>> cr(n);
    std::vector<BottomArgs<T>> br(n);
    std::vector<DirMvArgs<T>> dr(n);
    std::vector<int> u_co(n), u_bo(n), u_vec(n);
    for (int k = 0; k < n; k++) {
        impec_ctx *ctx = e->m[k];
        Level &Cl = ctx->lev[ctx->nlev - 1];
        cr[k] = CoarsestRec<T>{lev_args(ctx, ctx->nlev - 1), LA::x(Cl) + Cl.ld + 1, LA::b(Cl) + Cl.ld + 1, ctx->scal};
        const int lb = bottom_level(ctx);
        if (lb > 0) br[k] = bottom_args<T>(ctx, lb);
        u_bo[k] = lb > 0;
        u_co[k] = lb <= 0;
        LevArgs a0 = lev_args(ctx, 0);
        if (!IMPEC_NO_FACES) {
            a0.fx = ctx->fx;
            a0.fy = ctx->fy;
        }
        dr[k] = dir_mv_args<T>(ctx, a0, cg_args(ctx));
        u_vec[k] = dr[k].nwave;
    }
    ENS_CHECK(e, batch_upload(t.coarsest, cr, u_co));
    ENS_CHECK(e, batch_upload(t.bottom, br, u_bo));
    ENS_CHECK(e, batch_upload(t.dir, dr, u_vec));
    return IMPEC_OK;
}

// a compacted list over the members of a batch (those with units > 0, in
// member order, as batch_upload keeps them)
static int make_list(impec_ensemble *e, const int *bstart, const std::vector<int> &units, ListDesc &d) {
    EnsPressTabs *t = e->pres;
    std::vector<double *> S;
    for (int k = 0; k < e->n; k++)
        if (units[k] > 0) S.push_back(e->m[k]->scal);
    d = ListDesc();
    d.nmem = (int)S.size();
    d.bstart = bstart;
    if (d.nmem == 0) return IMPEC_OK;
    void *pS, *pm, *ps, *pn;
    ENS_CHECK(e, hipMalloc(&pS, sizeof(double *) * S.size()));
    t->list_mem.push_back(pS);
    ENS_CHECK(e, hipMalloc(&pm, sizeof(int) * S.size()));
    t->list_mem.push_back(pm);
    ENS_CHECK(e, hipMalloc(&ps, sizeof(int) * (S.size() + 1)));
    t->list_mem.push_back(ps);
    ENS_CHECK(e, hipMalloc(&pn, sizeof(int)));
    t->list_mem.push_back(pn);
    ENS_CHECK(e, hipMemcpy(pS, S.data(), sizeof(double *) * S.size(), hipMemcpyHostToDevice));
    ENS_CHECK(e, hipMemset(pn, 0, sizeof(int)));
    ENS_CHECK(e, hipMemset(ps, 0, sizeof(int)));
    d.S = (double *const *)pS;
    d.act.member = (int *)pm;
    d.act.start = (int *)ps;
    d.act.n = (int *)pn;
    return IMPEC_OK;
}

static int build_lists(impec_ensemble *e) {
    EnsPressTabs *t = e->pres;
    const int n = e->n;
    t->h_iter.assign(t->lmax + 1, ListDesc());
    t->h_check.assign(1, ListDesc());
    std::vector<int> all(n, 1);
    int st;
    if ((st = make_list(e, t->upd.start, all, t->h_iter[0]))) return st;
    const int *dstart = t->mixed ? t->l32.dir.start : t->l64.dir.start;
    if ((st = make_list(e, dstart, all, t->h_iter[t->lmax]))) return st;
    for (int l = 0; l + 1 < t->lmax; l++) {
        std::vector<int> u(n, 0);
        for (int k = 0; k < n; k++) u[k] = l < top_level(e->m[k]);
        const int *ls = t->mixed ? t->l32.down[l].start : t->l64.down[l].start;
        if ((st = make_list(e, ls, u, t->h_iter[1 + l]))) return st;
    }
    if ((st = make_list(e, t->init.start, all, t->h_check[0]))) return st;
    t->n_iter = t->lmax + 1;
    void *pi, *pc;
    ENS_CHECK(e, hipMalloc(&pi, sizeof(ListDesc) * t->h_iter.size()));
    t->list_mem.push_back(pi);
    ENS_CHECK(e, hipMalloc(&pc, sizeof(ListDesc)));
    t->list_mem.push_back(pc);
    ENS_CHECK(e, hipMemcpy(pi, t->h_iter.data(), sizeof(ListDesc) * t->h_iter.size(), hipMemcpyHostToDevice));
    ENS_CHECK(e, hipMemcpy(pc, t->h_check.data(), sizeof(ListDesc), hipMemcpyHostToDevice));
    t->d_iter = (ListDesc *)pi;
    t->d_check = (ListDesc *)pc;
    return IMPEC_OK;
}

int ens_press_build(impec_ensemble *e) {
    ens_press_free(e);
    EnsPressTabs *t = new EnsPressTabs();
    e->pres = t;
    const int n = e->n;
    t->mixed = e->m[0]->P.precision == IMPEC_PREC_MIXED;
    t->lmax = 0;
    for (int k = 0; k < n; k++) {
        impec_ctx *ctx = e->m[k];
        if (ctx->nlev > t->lmax) t->lmax = ctx->nlev;
        if (t->mixed) {
            int st = mixed_setup(ctx);   // FP32 level arrays
            if (st) return ens_set_error(e, st, "member %d: %s", k, ctx->err);
        }
    }
    dim3 blk(64, 4);
    // hierarchy and FP32 copies, level by level
    for (int l = 0; l < t->lmax; l++) {
        std::vector<CoarsenRec> cr(n);
        std::vector<ToF32Rec> fr(n);
        std::vector<int> uc(n, 0), uf(n, 0);
        for (int k = 0; k < n; k++) {
            impec_ctx *ctx = e->m[k];
            if (l >= ctx->nlev) continue;
            Level &L = ctx->lev[l];
            if (l >= 1) {   // build_hierarchy
                const Level &F = ctx->lev[l - 1];
                const double *wf = l == 1 ? ctx->w + (long)ctx->B * ctx->lds + ctx->B : F.w + F.ld + 1;
                const int ldf = l == 1 ? ctx->lds : F.ld;
                const dim3 g = grid2(L.n1 + 2, L.n2 + 2, blk);
                cr[k] = CoarsenRec{wf, ldf, F.n1, F.n2, L.w + L.ld + 1, L.ld, L.n1, L.n2, (int)g.x, (int)g.y};
                uc[k] = g.x * g.y;
            }
            if (t->mixed) {   // mixed_setup's copies
                LevArgs la = lev_args(ctx, l);
                const dim3 g = grid2(L.n1, L.n2, blk);
                fr[k] = ToF32Rec{la.w1, la.ldw, L.wf + L.ld + 1, L.ld, L.n1, L.n2, (int)g.x, (int)g.y};
                uf[k] = g.x * g.y;
            }
        }
        ENS_CHECK(e, batch_upload(t->coarsen[l], cr, uc));
        ENS_CHECK(e, batch_upload(t->tof32[l], fr, uf));
    }
    std::vector<FacesRec> fa(n);
    std::vector<ZeroRec> zr(n);
    std::vector<CgInitRec> ir(n);
    std::vector<FinishRec> fi(n), fd(n), frr(n), ft(n);
    std::vector<RhoRec> rh(n);
    std::vector<Upd2Rec> up(n);
    std::vector<CtlRec> ct(n);
    std::vector<PinRec> pr(n);
    std::vector<ShiftRec> sh(n);
    std::vector<TresArgs> tr(n);
    std::vector<int> u_fa(n), u_red(n), u_vec(n), u_tres(n), one(n, 1);
    std::vector<double> init(S_COUNT);
    for (int k = 0; k < n; k++) {
        impec_ctx *ctx = e->m[k];
        const Level &L0 = ctx->lev[0];
        LevArgs a0 = lev_args(ctx, 0);
        const dim3 gf = grid2(ctx->N1 + 1, ctx->N2 + 1, blk);
        fa[k] = FacesRec{a0, ctx->fx, ctx->fy, (int)gf.x, (int)gf.y};
        u_fa[k] = IMPEC_NO_FACES ? 0 : gf.x * gf.y;
        if (!IMPEC_NO_FACES) {
            a0.fx = ctx->fx;
            a0.fy = ctx->fy;
        }
        const CgArgs c = cg_args(ctx);
        const int nb = red_grid(c.n);
        const Grid2 vg = vec_grid(ctx->N1, ctx->N2);
        zr[k] = ZeroRec{c, ctx->p, ctx->scal, nb, 1};
        ir[k] = CgInitRec{a0, c, ctx->p, ctx->q, ctx->pr, ctx->part, ctx->scal, nb, 1};
        u_red[k] = nb;
        u_vec[k] = vg.nblocks;
        fi[k] = FinishRec{ctx->part, ctx->scal, nb, 2, (int)S_RR, (int)RED_SUM};
        fd[k] = FinishRec{ctx->part, ctx->scal, dir_mv_args<double>(ctx, a0, c).nwave, 1, (int)S_DQ, (int)RED_SUM};
        frr[k] = FinishRec{ctx->part, ctx->scal, vg.nblocks, 2, (int)S_RR, (int)RED_SUM};
        const TresArgs ta = tres_args(a0, c, ctx->p, ctx->q, ctx->part);
        if (ta.nwave > RED_BLOCKS) return ens_set_error(e, IMPEC_EINVAL, "member %d: grid too wide", k);
        ft[k] = FinishRec{ctx->part, ctx->scal, ta.nwave, 2, (int)S_TT, (int)RED_SUM};
        const int nwave0 = t->mixed ? fused_args<float>(ctx, 0).nwave : fused_args<double>(ctx, 0).nwave;
        rh[k] = RhoRec{ctx->part, nwave0, 1.0 / (double)c.n, ctx->scal};
        up[k] = Upd2Rec{c, ctx->p, ctx->pd, ctx->pz, ctx->pr, ctx->pq, ctx->scal, ctx->part, (int)vg.grid.x,
                        (int)vg.grid.y};
        ct[k] = CtlRec{ctx->scal, ctx->P.initial_guess == IMPEC_GUESS_ZERO};
        pr[k] = PinRec{c, ctx->p, ctx->scal};
        sh[k] = ShiftRec{c, ctx->p, ctx->scal, nb, 1};
        tr[k] = ta;
        u_tres[k] = ta.gx;
        // loop constants of solve_cg
        const double tol = ctx->P.eps, bb = ctx->bnorm * ctx->bnorm;
        double v[4];
        v[0] = tol * tol * bb;   // S_ETOL
        v[1] = ctx->P.max_iterations > 0 ? ctx->P.max_iterations : 10000;
        ENS_CHECK(e, hipMemcpy(ctx->scal + S_ETOL, v, 2 * sizeof(double), hipMemcpyHostToDevice));
        const double zp = (ctx->P.initial_guess == IMPEC_GUESS_PREVIOUS && ctx->have_prev_p) ? 0. : 1.;
        ENS_CHECK(e, hipMemcpy(ctx->scal + S_ZEROP, &zp, sizeof(double), hipMemcpyHostToDevice));
    }
    ENS_CHECK(e, batch_upload(t->faces, fa, u_fa));
    ENS_CHECK(e, batch_upload(t->zero, zr, u_red));
    ENS_CHECK(e, batch_upload(t->init, ir, u_red));
    ENS_CHECK(e, batch_upload(t->fin_init, fi, one));
    ENS_CHECK(e, batch_upload(t->fin_dq, fd, one));
    ENS_CHECK(e, batch_upload(t->fin_rr, frr, one));
    ENS_CHECK(e, batch_upload(t->fin_tt, ft, one));
    ENS_CHECK(e, batch_upload(t->rho, rh, one));
    ENS_CHECK(e, batch_upload(t->upd, up, u_vec));
    ENS_CHECK(e, batch_upload(t->ctl, ct, one));
    ENS_CHECK(e, batch_upload(t->pin, pr, one));
    ENS_CHECK(e, batch_upload(t->shift, sh, u_red));
    ENS_CHECK(e, batch_upload(t->tres, tr, u_tres));
    int st = t->mixed ? build_legs<float>(e, t->l32) : build_legs<double>(e, t->l64);
    if (st) return st;
    if ((st = build_lists(e))) return st;
    ENS_CHECK(e, hipDeviceSynchronize());
    return IMPEC_OK;
}

template <class E, class... A>
static hipError_t launch_b(void (*k)(const E *, const int *, int, A...), const Batch<E> &b, dim3 blk,
                           hipStream_t s, A... a) {
    if (b.nmem == 0) return hipSuccess;
    hipLaunchKernelGGL(k, dim3(batch_grid(b.total)), blk, 0, s, b.tab, b.start, b.nmem, a...);
    return hipGetLastError();
}
template <class E>
static hipError_t launch_m(void (*k)(const E *, int), const Batch<E> &b, dim3 blk, hipStream_t s, int a) {
    if (b.nmem == 0) return hipSuccess;
    hipLaunchKernelGGL(k, dim3(b.nmem), blk, 0, s, b.tab, a);
    return hipGetLastError();
}
template <class E>
static hipError_t launch_t(void (*k)(const E *, int), const Batch<E> &b, hipStream_t s) {
    if (b.nmem == 0) return hipSuccess;
    hipLaunchKernelGGL(k, dim3((b.nmem + 63) / 64), dim3(64), 0, s, b.tab, b.nmem);
    return hipGetLastError();
}

// the legs of one V-cycle of every member in state `need` (-1: all); the
// members share the number of red-black pairs per leg
template <typename T, int NS>
static int ens_vcycle_ns(impec_ensemble *e, const LegTabs<T> &t, int need) {
    const dim3 blk(64 * SW_WAVES);
    const int lmax = e->pres->lmax;
    hipStream_t s = e->stream;
    for (int l = 0; l + 1 < lmax; l++) {
        const Batch<FusedArgsT<T>> &b = t.down[l];
        if (b.nmem == 0) continue;
        const dim3 g((b.total + SW_WAVES - 1) / SW_WAVES);
        if (l == 0) hipLaunchKernelGGL((kb_mg_down<true, T, NS>), g, blk, 0, s, b.tab, b.start, b.nmem, b.total, need);
        else hipLaunchKernelGGL((kb_mg_down<false, T, NS>), g, blk, 0, s, b.tab, b.start, b.nmem, b.total, need);
    }
    ENS_CHECK(e, launch_m(kb_coarsest<T>, t.coarsest, dim3(64), s, need));
    ENS_CHECK(e, launch_m(kb_mg_bottom<T>, t.bottom, dim3(BOTTOM_THREADS), s, need));
    for (int l = lmax - 2; l >= 0; l--) {
        const Batch<FusedArgsT<T>> &b = t.up[l];
        if (b.nmem == 0) continue;
        const dim3 g((b.total + SW_WAVES - 1) / SW_WAVES);
        if (l == 0) hipLaunchKernelGGL((kb_mg_up<true, T, NS>), g, blk, 0, s, b.tab, b.start, b.nmem, b.total, need);
        else hipLaunchKernelGGL((kb_mg_up<false, T, NS>), g, blk, 0, s, b.tab, b.start, b.nmem, b.total, need);
    }
    ENS_CHECK(e, hipGetLastError());
    return IMPEC_OK;
}
template <typename T>
static int ens_vcycle(impec_ensemble *e, const LegTabs<T> &t, int need) {
    return mg_ns(e->m[0]) == 2 ? ens_vcycle_ns<T, 2>(e, t, need) : ens_vcycle_ns<T, 1>(e, t, need);
}

int ens_solve_begin(impec_ensemble *e) {
    EnsPressTabs *t = e->pres;
    hipStream_t s = e->stream;
    const dim3 blk(64, 4);
    for (int l = 1; l < t->lmax; l++) ENS_CHECK(e, launch_b(kb_coarsen_w, t->coarsen[l], blk, s));
    if (t->mixed)
        for (int l = 0; l < t->lmax; l++) ENS_CHECK(e, launch_b(kb_to_f32, t->tof32[l], blk, s));
    ENS_CHECK(e, launch_b(kb_faces, t->faces, blk, s));
    ENS_CHECK(e, launch_b(kb_zero_p, t->zero, dim3(256), s));
    ENS_CHECK(e, launch_b(kb_cg_init, t->init, dim3(256), s, -1));
    ENS_CHECK(e, ens_finish(s, t->fin_init, -1));
    ENS_CHECK(e, launch_t(kb_ctl_first, t->ctl, s));
    hipLaunchKernelGGL(kb_compact, dim3(t->n_iter), dim3(256), 0, s, t->d_iter, (int)EST_ITER);
    ENS_CHECK(e, hipGetLastError());
    return IMPEC_OK;
}

static int persist_grid(int units, int cap) { return units < cap ? units : cap; }

// one CG iteration of the members still iterating (compacted launches)
template <typename T>
static int ens_cg_slot_t(impec_ensemble *e, const LegTabs<T> &lt) {
    EnsPressTabs *t = e->pres;
    hipStream_t s = e->stream;
    const dim3 lblk(64 * SW_WAVES), vblk(64, 4);
    const int lmax = t->lmax;
    const int ns = mg_ns(e->m[0]);
    for (int l = 0; l + 1 < lmax; l++) {
        const Batch<FusedArgsT<T>> &b = lt.down[l];
        if (b.nmem == 0) continue;
        const dim3 g(persist_grid((b.total + SW_WAVES - 1) / SW_WAVES, PERSIST_LEG_WG));
        const ActList &a = t->h_iter[1 + l].act;
        if (ns == 2) {
            if (l == 0) hipLaunchKernelGGL((kp_mg_down<true, T, 2>), g, lblk, 0, s, b.tab, a);
            else hipLaunchKernelGGL((kp_mg_down<false, T, 2>), g, lblk, 0, s, b.tab, a);
        } else {
            if (l == 0) hipLaunchKernelGGL((kp_mg_down<true, T, 1>), g, lblk, 0, s, b.tab, a);
            else hipLaunchKernelGGL((kp_mg_down<false, T, 1>), g, lblk, 0, s, b.tab, a);
        }
    }
    ENS_CHECK(e, launch_m(kb_coarsest<T>, lt.coarsest, dim3(64), s, (int)EST_ITER));
    ENS_CHECK(e, launch_m(kb_mg_bottom<T>, lt.bottom, dim3(BOTTOM_THREADS), s, (int)EST_ITER));
    for (int l = lmax - 2; l >= 0; l--) {
        const Batch<FusedArgsT<T>> &b = lt.up[l];
        if (b.nmem == 0) continue;
        const dim3 g(persist_grid((b.total + SW_WAVES - 1) / SW_WAVES, PERSIST_LEG_WG));
        const ActList &a = t->h_iter[1 + l].act;
        if (ns == 2) {
            if (l == 0) hipLaunchKernelGGL((kp_mg_up<true, T, 2>), g, lblk, 0, s, b.tab, a);
            else hipLaunchKernelGGL((kp_mg_up<false, T, 2>), g, lblk, 0, s, b.tab, a);
        } else {
            if (l == 0) hipLaunchKernelGGL((kp_mg_up<true, T, 1>), g, lblk, 0, s, b.tab, a);
            else hipLaunchKernelGGL((kp_mg_up<false, T, 1>), g, lblk, 0, s, b.tab, a);
        }
    }
    ENS_CHECK(e, hipGetLastError());
    ENS_CHECK(e, launch_m(kb_finish_rho, t->rho, dim3(1024), s, (int)EST_ITER));
    const ActList &av = t->h_iter[0].act;
    hipLaunchKernelGGL(kp_dir_mv<T>, dim3(persist_grid((lt.dir.total + SW_WAVES - 1) / SW_WAVES, PERSIST_LEG_WG)), lblk,
                       0, s, lt.dir.tab, t->h_iter[lmax].act);
    ENS_CHECK(e, hipGetLastError());
    ENS_CHECK(e, ens_finish(s, t->fin_dq, EST_ITER));
    hipLaunchKernelGGL(kp_update2, dim3(persist_grid(t->upd.total, PERSIST_VEC_WG)), vblk, 0, s, t->upd.tab, av);
    ENS_CHECK(e, hipGetLastError());
    ENS_CHECK(e, ens_finish(s, t->fin_rr, EST_ITER));
    ENS_CHECK(e, launch_t(kb_ctl_iter, t->ctl, s));
    hipLaunchKernelGGL(kb_compact, dim3(1), dim3(256), 0, s, t->d_check, (int)EST_CHECK);
    hipLaunchKernelGGL(kp_cg_init, dim3(persist_grid(t->init.total, PERSIST_VEC_WG)), dim3(256), 0, s, t->init.tab,
                       t->h_check[0].act);
    ENS_CHECK(e, hipGetLastError());
    ENS_CHECK(e, ens_finish(s, t->fin_init, EST_CHECK));
    ENS_CHECK(e, launch_t(kb_ctl_check, t->ctl, s));
    hipLaunchKernelGGL(kb_compact, dim3(t->n_iter), dim3(256), 0, s, t->d_iter, (int)EST_ITER);
    ENS_CHECK(e, hipGetLastError());
    return IMPEC_OK;
}

int ens_cg_slot(impec_ensemble *e) {
    return e->pres->mixed ? ens_cg_slot_t<float>(e, e->pres->l32) : ens_cg_slot_t<double>(e, e->pres->l64);
}

int ens_solve_end(impec_ensemble *e) {
    EnsPressTabs *t = e->pres;
    hipStream_t s = e->stream;
    ENS_CHECK(e, launch_t(kb_read_pin, t->pin, s));
    ENS_CHECK(e, launch_b(kb_shift, t->shift, dim3(256), s));
    ENS_CHECK(e, launch_b(kb_true_rows, t->tres, dim3(64 * SW_WAVES), s));
    ENS_CHECK(e, ens_finish(s, t->fin_tt, -1));
    return IMPEC_OK;
}

int ens_vcycle_time(impec_ensemble *e, int reps, float *ms) {
    hipEvent_t a, b;
    ENS_CHECK(e, hipEventCreate(&a));
    ENS_CHECK(e, hipEventCreate(&b));
    int st = IMPEC_OK;
    for (int r = 0; r < reps + 1 && !st; r++) {
        if (r == 1) (void)hipEventRecord(a, e->stream);
        st = e->pres->mixed ? ens_vcycle<float>(e, e->pres->l32, -1) : ens_vcycle<double>(e, e->pres->l64, -1);
    }
    (void)hipEventRecord(b, e->stream);
    hipError_t he = hipEventSynchronize(b);
    if (!st && he == hipSuccess) he = hipEventElapsedTime(ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st) return st;
    ENS_CHECK(e, he);
    return IMPEC_OK;
}

}  // namespace impec
