// k_field.hip -- ghost fills, mobility, Darcy face velocities, CFL reduction,
// trapezoidal-rule well sources and diagnostics reductions (FP64).
#include "impec_device.h"
#include <chrono>
#include <stdio.h>
#include <stdarg.h>
#include <vector>

namespace impec {

// ---------------------------------------------------------------------------
// ghost fills
// ---------------------------------------------------------------------------

// FillGhostCellsNeumannScalar / FillGhostCellsPeriodicScalar (Util.cpp:215-324):
// every ghost cell of the ghosted scalar array copies the interior cell its
// even mirror (or periodic image) maps to; corners compose both maps, which
// is what the reference's corner loops produce.  Cell (i, j), interior
// indices, -B <= i < N1 + B, -B <= j < N2 + B; interior cells are left alone.
__device__ __forceinline__ void fill_scalar_cell(double *__restrict__ a, int N1, int N2, int B, int ld, int periodic,
                                                 int i, int j) {
    const bool gi = i < 0 || i >= N1, gj = j < 0 || j >= N2;
    if (!gi && !gj) return;
    int si, sj;
    if (periodic) {
        si = i < 0 ? i + N1 : (i >= N1 ? i - N1 : i);
        sj = j < 0 ? j + N2 : (j >= N2 ? j - N2 : j);
    } else {
        si = i < 0 ? -1 - i : (i >= N1 ? 2 * N1 - 1 - i : i);
        sj = j < 0 ? -1 - j : (j >= N2 ? 2 * N2 - 1 - j : j);
    }
    a[(size_t)(i + B) * ld + (j + B)] = a[(size_t)(si + B) * ld + (sj + B)];
}
// over the whole ghosted array (the ensemble's batched launch)
template <class G>
__device__ __forceinline__ void fill_scalar_body(const G &gr_, double *__restrict__ a, int N1, int N2, int B, int ld, int periodic) {
    const int j = gr_.bx() * blockDim.x + threadIdx.x - B;
    const int i = gr_.by() * blockDim.y + threadIdx.y - B;
    if (i >= N1 + B || j >= N2 + B) return;
    fill_scalar_cell(a, N1, N2, B, ld, periodic, i, j);
}

// FillGhostCellsNeumannVelocity (Util.cpp:331-438).  u (x-faces) is odd in x
// about the faces i = 0 and i = N1 and even in y; v (y-faces) is odd in y about
// j = 0 and j = N2 and even in x.  Entries the reference never writes
// (u at j = N2+B, v at i = N1+B) are left untouched.  Cell (i, j),
// -B <= i <= N1 + B, -B <= j <= N2 + B.
__device__ __forceinline__ void fill_velocity_cell(double *__restrict__ u, double *__restrict__ v, int N1, int N2,
                                                   int B, int ld, int i, int j) {
    const size_t o = (size_t)(i + B) * ld + (j + B);
    // u component
    if (!((i >= 0 && i <= N1 && j >= 0 && j < N2) || j == N2 + B)) {
        const int jm = j < 0 ? -1 - j : (j >= N2 ? 2 * N2 - 1 - j : j);
        double val;
        if (i < 0) val = -u[(size_t)(-i + B) * ld + (jm + B)];
        else if (i > N1) val = -u[(size_t)(2 * N1 - i + B) * ld + (jm + B)];
        else val = u[(size_t)(i + B) * ld + (jm + B)];
        u[o] = val;
    }
    // v component
    if (!((i >= 0 && i < N1 && j >= 0 && j <= N2) || i == N1 + B)) {
        const int im = i < 0 ? -1 - i : (i >= N1 ? 2 * N1 - 1 - i : i);
        double val;
        if (j < 0) val = -v[(size_t)(im + B) * ld + (-j + B)];
        else if (j > N2) val = -v[(size_t)(im + B) * ld + (2 * N2 - j + B)];
        else val = v[(size_t)(im + B) * ld + (j + B)];
        v[o] = val;
    }
}
template <class G>
__device__ __forceinline__ void fill_velocity_body(const G &gr_, double *__restrict__ u, double *__restrict__ v, int N1, int N2, int B,
                                int ld) {
    const int j = gr_.bx() * blockDim.x + threadIdx.x - B;
    const int i = gr_.by() * blockDim.y + threadIdx.y - B;
    if (i > N1 + B || j > N2 + B) return;
    fill_velocity_cell(u, v, N1, N2, B, ld, i, j);
}

// The cells of the frame between an outer rectangle [r0, r1) x [c0, c1) and
// an inner one [ir0, ir1) x [ic0, ic1): the rows above and below the inner
// rectangle at full width, then the columns left and right of it, row by
// row.  A ghost fill launches one thread per frame cell instead of one per
// cell of the array (at 4096^2 about 33 thousand threads instead of 17
// million: the launch of the whole grid, not the copies, was its cost).
struct Frame {
    int r0, r1, c0, c1, ir0, ir1, ic0, ic1;
    long cells() const {
        return (long)(ir0 - r0 + r1 - ir1) * (c1 - c0) + (long)(ir1 - ir0) * (ic0 - c0 + c1 - ic1);
    }
};
__device__ __forceinline__ bool frame_cell(const Frame &f, long t, int &i, int &j) {
    const long W = f.c1 - f.c0;
    const long top = (long)(f.ir0 - f.r0) * W, bottom = (long)(f.r1 - f.ir1) * W;
    if (t < top) {
        i = f.r0 + (int)(t / W);
        j = f.c0 + (int)(t % W);
        return true;
    }
    t -= top;
    if (t < bottom) {
        i = f.ir1 + (int)(t / W);
        j = f.c0 + (int)(t % W);
        return true;
    }
    t -= bottom;
    const int left = f.ic0 - f.c0, side = left + (f.c1 - f.ic1);
    if (side <= 0 || t >= (long)(f.ir1 - f.ir0) * side) return false;
    i = f.ir0 + (int)(t / side);
    const int k = (int)(t % side);
    j = k < left ? f.c0 + k : f.ic1 + (k - left);
    return true;
}
constexpr int FRAME_THREADS = 256;

__global__ void __launch_bounds__(FRAME_THREADS) k_fill_scalar(Frame f, double *__restrict__ a, int N1, int N2,
                                                               int B, int ld, int periodic) {
    int i, j;
    if (frame_cell(f, (long)blockIdx.x * FRAME_THREADS + threadIdx.x, i, j))
        fill_scalar_cell(a, N1, N2, B, ld, periodic, i, j);
}
__global__ void __launch_bounds__(FRAME_THREADS) k_fill_velocity(Frame f, double *__restrict__ u,
                                                                 double *__restrict__ v, int N1, int N2, int B,
                                                                 int ld) {
    int i, j;
    if (frame_cell(f, (long)blockIdx.x * FRAME_THREADS + threadIdx.x, i, j))
        fill_velocity_cell(u, v, N1, N2, B, ld, i, j);
}

// the ghost frame of a ghosted scalar array copied into another of the same
// layout (the buffer that takes over as c keeps the ghost cells c had)
__global__ void __launch_bounds__(FRAME_THREADS) k_copy_frame(Frame f, const double *__restrict__ src,
                                                              double *__restrict__ dst, int B, int ld) {
    int i, j;
    if (frame_cell(f, (long)blockIdx.x * FRAME_THREADS + threadIdx.x, i, j)) {
        const size_t o = (size_t)(i + B) * ld + (j + B);
        dst[o] = src[o];
    }
}
int launch_copy_frame(impec_ctx *ctx, const double *src, double *dst) {
    const int N1 = ctx->N1, N2 = ctx->N2, B = ctx->B;
    const Frame f{-B, N1 + B, -B, N2 + B, 0, N1, 0, N2};
    const long n = f.cells();
    if (n == 0) return IMPEC_OK;
    hipLaunchKernelGGL(k_copy_frame, dim3((unsigned)((n + FRAME_THREADS - 1) / FRAME_THREADS)), dim3(FRAME_THREADS),
                       0, ctx->stream, f, src, dst, B, ctx->lds);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

int launch_fill_scalar(impec_ctx *ctx, double *a, int periodic) {
    const int N1 = ctx->N1, N2 = ctx->N2, B = ctx->B;
    const Frame f{-B, N1 + B, -B, N2 + B, 0, N1, 0, N2};
    const long n = f.cells();
    if (n == 0) return IMPEC_OK;
    prof_begin(ctx);
    hipLaunchKernelGGL(k_fill_scalar, dim3((unsigned)((n + FRAME_THREADS - 1) / FRAME_THREADS)), dim3(FRAME_THREADS),
                       0, ctx->stream, f, a, N1, N2, B, ctx->lds, periodic);
    prof_end(ctx, "fill_scalar");
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

int launch_fill_velocity(impec_ctx *ctx) {
    // every cell outside the common interior [0, N1) x [0, N2) of u and v;
    // fill_velocity_cell leaves the interior ones of each component alone
    const int N1 = ctx->N1, N2 = ctx->N2, B = ctx->B;
    const Frame f{-B, N1 + B + 1, -B, N2 + B + 1, 0, N1, 0, N2};
    const long n = f.cells();
    prof_begin(ctx);
    hipLaunchKernelGGL(k_fill_velocity, dim3((unsigned)((n + FRAME_THREADS - 1) / FRAME_THREADS)),
                       dim3(FRAME_THREADS), 0, ctx->stream, f, ctx->u, ctx->v, N1, N2, B, ctx->ldv);
    prof_end(ctx, "fill_velocity");
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

// ---------------------------------------------------------------------------
// mobility: w = mu(c)/K (harmonic) or K/mu(c) (mean concentration) on every
// cell of the ghosted array.  Face and vertex mobilities are formed from w
// with the reference's expressions wherever they are needed.
// ---------------------------------------------------------------------------
template <class G>
__device__ __forceinline__ void mobility_body(const G &gr_, const double *__restrict__ c, const double *__restrict__ K11,
                           const double *__restrict__ K22, double *__restrict__ w, double *__restrict__ w22,
                           int rows, int cols, int ld, double mob1over4, double mu_res, int harmonic) {
    const int j = gr_.bx() * blockDim.x + threadIdx.x;
    const int i = gr_.by() * blockDim.y + threadIdx.y;
    if (i >= rows || j >= cols) return;
    const size_t o = (size_t)i * ld + j;
    const double mu = mu_of(c[o], mob1over4, mu_res);
    const double k11 = K11 ? K11[o] : 1.;
    w[o] = harmonic ? mu / k11 : k11 / mu;
    if (w22) {
        const double k22 = K22[o];
        w22[o] = harmonic ? mu / k22 : k22 / mu;
    }
}
__global__ void k_mobility(const double *__restrict__ c, const double *__restrict__ K11,
                           const double *__restrict__ K22, double *__restrict__ w, double *__restrict__ w22,
                           int rows, int cols, int ld, double mob1over4, double mu_res, int harmonic) {
    mobility_body(HwGrid{}, c, K11, K22, w, w22, rows, cols, ld, mob1over4, mu_res, harmonic);
}

int launch_mobility(impec_ctx *ctx) {
    const int rows = ctx->N1 + 2 * ctx->B, cols = ctx->N2 + 2 * ctx->B;
    dim3 blk(64, 4), grd((cols + 63) / 64, (rows + 3) / 4);
    const int aniso = ctx->K22 != nullptr && ctx->K22 != ctx->K11;
    prof_begin(ctx);
    hipLaunchKernelGGL(k_mobility, grd, blk, 0, ctx->stream, ctx->c, ctx->K11, ctx->K22, ctx->w,
                       aniso ? ctx->w22 : nullptr, rows, cols, ctx->lds, ctx->mob1over4, ctx->P.mu_res,
                       ctx->P.average_procedure == IMPEC_AVG_HARMONIC);
    prof_end(ctx, "mobility");
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

// ---------------------------------------------------------------------------
// Darcy face velocities (EllipticSolver.cpp:541-683)
// ---------------------------------------------------------------------------
struct VelConst {
    int N1, N2, B, lds, ldv, harmonic;
    double nu, one_m_nu, dx, dy, idxi, ideta;
    double s2, cosa2, sina2, cost2, sint2, costsina, cosasint, cosa, sint, sina, cost;
};

// rotated-frame velocity w at the vertex whose lower-left cell is (a, b)
// (M_im12j + GetDiagonalW_im12j, EllipticSolver.cpp:554-574)
__device__ __forceinline__ void vertex_w(const VelConst &k, const double *__restrict__ p,
                                         const double *__restrict__ w11, const double *__restrict__ w22, int a,
                                         int b, double &w1, double &w2) {
    const int ld = k.lds, B = k.B;
#define P_(i, j) p[(size_t)((i) + B) * ld + ((j) + B)]
#define W1(i, j) w11[(size_t)((i) + B) * ld + ((j) + B)]
#define W2(i, j) w22[(size_t)((i) + B) * ld + ((j) + B)]
    // cells: (a,b) = (i-1,j-1), (a+1,b+1) = (i,j)
    const double d1dp = k.idxi * (P_(a + 1, b + 1) - P_(a, b));
    const double d2dp = k.ideta * (P_(a, b + 1) - P_(a + 1, b));
    const double K11 = vertex_mob(W1(a + 1, b), W1(a, b), W1(a, b + 1), W1(a + 1, b + 1), k.harmonic);
    const double K22 = vertex_mob(W2(a + 1, b), W2(a, b), W2(a, b + 1), W2(a + 1, b + 1), k.harmonic);
    const double M11 = k.s2 * (K11 * k.cosa2 + K22 * k.sina2);
    const double M12 = k.s2 * (K22 * k.costsina - K11 * k.cosasint);
    const double M22 = k.s2 * (K11 * k.sint2 + K22 * k.cost2);
    w1 = -d1dp * M11 - d2dp * M12;
    w2 = -d1dp * M12 - d2dp * M22;
#undef P_
#undef W1
#undef W2
}

// PAR: nu = 1, the diagonal-grid term (weight 1 - nu = 0) is not formed; the
// stored value nu * par equals nu * par + 0 * dg for every finite dg
template <class G, bool PAR = false>
__device__ __forceinline__ void velocity_body(const G &gr_, VelConst k, const double *__restrict__ p, const double *__restrict__ w11,
                           const double *__restrict__ w22, double *__restrict__ u, double *__restrict__ v) {
    const int j = gr_.bx() * blockDim.x + threadIdx.x;
    const int i = gr_.by() * blockDim.y + threadIdx.y;
    if (i > k.N1 || j > k.N2) return;
    const int ld = k.lds, B = k.B;
#define P_(i, j) p[(size_t)((i) + B) * ld + ((j) + B)]
#define W1(i, j) w11[(size_t)((i) + B) * ld + ((j) + B)]
#define W2(i, j) w22[(size_t)((i) + B) * ld + ((j) + B)]
    double w1m = 0., w2m = 0.;  // vertex (i-1/2, j-1/2), shared by u and v
    if (!PAR) vertex_w(k, p, w11, w22, i - 1, j - 1, w1m, w2m);
    if (j < k.N2) {
        // GetParallelU1_im12j + GetDiagonalU1_im12j
        const double lam = face_mob(W1(i - 1, j), W1(i, j), k.harmonic);
        const double par = -lam * (P_(i, j) - P_(i - 1, j)) / k.dx;
        if (PAR) {
            u[(size_t)(i + B) * k.ldv + (j + B)] = k.nu * par;
        } else {
        double w1p, w2p;
        vertex_w(k, p, w11, w22, i - 1, j, w1p, w2p);
        const double u1m = k.cost * w1m - k.sina * w2m;
        const double u1p = k.cost * w1p - k.sina * w2p;
        const double dg = 0.5 * (u1p + u1m);
        u[(size_t)(i + B) * k.ldv + (j + B)] = k.nu * par + k.one_m_nu * dg;
        }
    }
    if (i < k.N1) {
        // GetParallelU2_ijm12 + GetDiagonalU2_ijm12
        const double lam = face_mob(W2(i, j - 1), W2(i, j), k.harmonic);
        const double par = -lam * (P_(i, j) - P_(i, j - 1)) / k.dy;
        if (PAR) {
            v[(size_t)(i + B) * k.ldv + (j + B)] = k.nu * par;
        } else {
        double w1p, w2p;
        vertex_w(k, p, w11, w22, i, j - 1, w1p, w2p);
        const double u2m = k.sint * w1m + k.cosa * w2m;
        const double u2p = k.sint * w1p + k.cosa * w2p;
        const double dg = 0.5 * (u2m + u2p);
        v[(size_t)(i + B) * k.ldv + (j + B)] = k.nu * par + k.one_m_nu * dg;
        }
    }
#undef P_
#undef W1
#undef W2
}
template <bool PAR>
__global__ void k_velocity(VelConst k, const double *__restrict__ p, const double *__restrict__ w11,
                           const double *__restrict__ w22, double *__restrict__ u, double *__restrict__ v) {
    velocity_body<HwGrid, PAR>(HwGrid{}, k, p, w11, w22, u, v);
}

int launch_velocity(impec_ctx *ctx) {
    // CalcVelocities fills p's ghosts first (EllipticSolver.cpp:663)
    int st = launch_fill_scalar(ctx, ctx->p, 0);
    if (st) return st;
    VelConst k;
    k.N1 = ctx->N1; k.N2 = ctx->N2; k.B = ctx->B; k.lds = ctx->lds; k.ldv = ctx->ldv;
    k.harmonic = ctx->P.average_procedure == IMPEC_AVG_HARMONIC;
    k.nu = ctx->P.nu; k.one_m_nu = 1. - ctx->P.nu;
    k.dx = ctx->dx; k.dy = ctx->dy;
    k.idxi = 1. / ctx->dxi; k.ideta = 1. / ctx->deta;
    k.s2 = ctx->secamt2; k.cosa2 = ctx->cosa2; k.sina2 = ctx->sina2; k.cost2 = ctx->cost2; k.sint2 = ctx->sint2;
    k.costsina = ctx->costsina; k.cosasint = ctx->cosasint;
    k.cosa = ctx->cosa; k.sint = ctx->sint; k.sina = ctx->sina; k.cost = ctx->cost;
    const double *w22 = ctx->K22 && ctx->K22 != ctx->K11 ? ctx->w22 : ctx->w;
    dim3 blk(64, 4), grd((ctx->N2 + 1 + 63) / 64, (ctx->N1 + 1 + 3) / 4);
    prof_begin(ctx);
    if (ctx->P.nu == 1.)
        hipLaunchKernelGGL(k_velocity<true>, grd, blk, 0, ctx->stream, k, ctx->p, ctx->w, w22, ctx->u, ctx->v);
    else hipLaunchKernelGGL(k_velocity<false>, grd, blk, 0, ctx->stream, k, ctx->p, ctx->w, w22, ctx->u, ctx->v);
    prof_end(ctx, "velocity");
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return launch_fill_velocity(ctx);
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------

// second pass: scal[dst+k] = combine over nblocks partials of value k
template <int OP, class G>
__device__ __forceinline__ void finish_body(const G &gr_, const double *__restrict__ part, int nblocks, int nv, double *__restrict__ scal,
                         int dst) {
    for (int k = 0; k < nv; k++) {
        double v[1] = {strided_combine<OP>(part + k * RED_BLOCKS, nblocks)};
        block_reduce<OP, 1>(v);
        if (threadIdx.x == 0) scal[dst + k] = v[0];
        __syncthreads();
    }
}
template <int OP>
__global__ void k_finish(const double *__restrict__ part, int nblocks, int nv, double *__restrict__ scal,
                         int dst, PubArgs pub) {
    finish_body<OP>(HwGrid{}, part, nblocks, nv, scal, dst);
    publish_scalars(scal, pub);
}

// The same combine in two launches, for many partials: one workgroup would
// read them all through one compute unit (about 24 us for 2 x 65536).  The
// first launch computes the FIN_THREADS per-thread shares of k_finish
// (strided_combine_at with the same thread index and count, 64 loads in
// flight) on 16 workgroups; the second runs k_finish's block reduction over
// them, so the result is k_finish's bit for bit.
constexpr int FIN_STAGE_WG = 64;
template <int OP>
__global__ void __launch_bounds__(FIN_STAGE_WG) k_finish_shares(const double *__restrict__ part, int nblocks, int nv,
                                                               double *__restrict__ mid) {
    const int t = blockIdx.x * FIN_STAGE_WG + threadIdx.x;
    for (int k = 0; k < nv; k++)
        mid[k * FIN_THREADS + t] = strided_combine_at<OP, 8>(part + (long)k * RED_BLOCKS, nblocks, t, FIN_THREADS);
}
template <int OP>
__global__ void __launch_bounds__(FIN_THREADS) k_finish_reduce(const double *__restrict__ mid, int nv,
                                                              double *__restrict__ scal, int dst, PubArgs pub) {
    for (int k = 0; k < nv; k++) {
        double v[1] = {mid[k * FIN_THREADS + threadIdx.x]};
        block_reduce<OP, 1>(v);
        if (threadIdx.x == 0) scal[dst + k] = v[0];
        __syncthreads();
    }
    publish_scalars(scal, pub);
}

template <int OP>
static int finish_op(impec_ctx *ctx, int nblocks, int nv, int dst, PubArgs pub = PubArgs{nullptr, 0}) {
    if (nblocks > 8 * FIN_THREADS && nv <= PART_VECS) {
        hipLaunchKernelGGL(k_finish_shares<OP>, dim3(FIN_THREADS / FIN_STAGE_WG), dim3(FIN_STAGE_WG), 0, ctx->stream,
                           ctx->part, nblocks, nv, ctx->part_mid);
        hipLaunchKernelGGL(k_finish_reduce<OP>, dim3(1), dim3(FIN_THREADS), 0, ctx->stream, ctx->part_mid, nv,
                           ctx->scal, dst, pub);
    } else {
        hipLaunchKernelGGL(k_finish<OP>, dim3(1), dim3(FIN_THREADS), 0, ctx->stream, ctx->part, nblocks, nv, ctx->scal,
                           dst, pub);
    }
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

int finish_sum(impec_ctx *ctx, int nblocks, int nv, int dst, PubArgs pub) {
    return finish_op<RED_SUM>(ctx, nblocks, nv, dst, pub);
}

static int red_blocks(long long n) {
    long long b = (n + 1023) / 1024;
    if (b < 1) b = 1;
    if (b > VEC_BLOCKS) b = VEC_BLOCKS;
    return (int)b;
}

// CFL (HyperbolicSolver.cpp:100-198): value 0 = max centred speed (combined),
// value 1 = min advective dt over cells (test mode), value 2 = min diffusive dt
struct CflConst {
    int N1, N2, B, ld, mode, diff;
    double dx, dy, al, at;
};

template <class G>
__device__ __forceinline__ void cfl_body(const G &gr_, CflConst k, const double *__restrict__ u, const double *__restrict__ v,
                      double *__restrict__ part) {
    const long long n = (long long)k.N1 * k.N2;
    double mx[1] = {0.0}, mn[2] = {DBL_MAX, 1.e10};
    for (long long id = (long long)gr_.bx() * blockDim.x + threadIdx.x; id < n;
         id += (long long)gr_.gx() * blockDim.x) {
        // fewer than 2^31 cells (checked at creation): 32-bit index arithmetic
        // instead of a 64-bit division per cell
        const int id32 = (int)id, i = id32 / k.N2, j = id32 - i * k.N2;
        const size_t o = (size_t)(i + k.B) * k.ld + (j + k.B);
        const double u0 = u[o], u1 = u[o + k.ld], v0 = v[o], v1 = v[o + 1];
        double uc = fabs(0.5 * (u0 + u1));
        double vc = fabs(0.5 * (v0 + v1));
        if (uc < DBL_MIN) uc = DBL_MIN;
        if (vc < DBL_MIN) vc = DBL_MIN;
        if (mx[0] < uc) mx[0] = uc;
        if (mx[0] < vc) mx[0] = vc;
        if (k.mode) {
            const double a = k.dx / uc, b = k.dy / vc;
            const double dt = b < a ? b : a;
            if (dt < mn[0]) mn[0] = dt;
        }
        if (k.diff) {
            const double vij = 0.5 * (v0 + v1), uij = 0.5 * (u0 + u1);
            const double vij2 = vij * vij, uij2 = uij * uij;
            const double absu = sqrt(vij2 + uij2);
            const double d11 = k.dx / absu * (k.al * uij2 + k.at * vij2);
            const double d22 = k.dx / absu * (k.al * vij2 + k.at * uij2);
            const double dt = 1. / 2 * k.dx * k.dx / (d11 + d22);
            if (dt < mn[1]) mn[1] = dt;
        }
    }
    block_reduce<RED_MAX, 1>(mx);
    block_reduce<RED_MIN, 2>(mn);
    if (tid_linear() == 0) {
        part[0 * RED_BLOCKS + gr_.bx()] = mx[0];
        part[1 * RED_BLOCKS + gr_.bx()] = mn[0];
        part[2 * RED_BLOCKS + gr_.bx()] = mn[1];
    }
}
__global__ void k_cfl(CflConst k, const double *__restrict__ u, const double *__restrict__ v,
                      double *__restrict__ part) {
    cfl_body(HwGrid{}, k, u, v, part);
}

// The scalars into the mapped host mirror, then the sequence number: every
// thread makes its stores visible system-wide before the barrier, and one
// thread then releases the number the host waits for.
__global__ void k_publish(const double *__restrict__ S, PubArgs pub) { publish_scalars(S, pub); }

PubArgs pub_next(impec_ctx *ctx) { return PubArgs{ctx->d_hscal, ++ctx->pub_seq}; }

int sync_scalars(impec_ctx *ctx) {
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, ctx->stream, ctx->scal, pub_next(ctx));
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return wait_scalars(ctx);
}

int wait_scalars(impec_ctx *ctx) {
    const unsigned long long seq = ctx->pub_seq;
    // the host polls the sequence number (a few microseconds sooner than a
    // copy and a stream synchronisation); past 50 ms it leaves the wait to
    // the runtime, which also reports an earlier failure of the stream
    unsigned long long *flag = (unsigned long long *)(ctx->h_scal + S_COUNT);
    const auto t0 = std::chrono::steady_clock::now();
    while (__atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq) {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
            IMPEC_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
            if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq)
                return set_error(ctx, IMPEC_EHIP, "scalars of the stream were not published");
            break;
        }
    }
    return IMPEC_OK;
}

int launch_cfl(impec_ctx *ctx, double *dt_out) {
    CflConst k;
    k.N1 = ctx->N1; k.N2 = ctx->N2; k.B = ctx->B; k.ld = ctx->ldv;
    k.mode = ctx->P.is_test;
    k.diff = ctx->P.diffusion && ctx->P.para_method != IMPEC_PARA_CN && !ctx->P.is_test;
    k.dx = ctx->dx; k.dy = ctx->dy; k.al = ctx->P.alpha_l; k.at = ctx->P.alpha_t;
    const int nb = red_blocks((long long)ctx->N1 * ctx->N2);
    prof_begin(ctx);
    hipLaunchKernelGGL(k_cfl, dim3(nb), dim3(256), 0, ctx->stream, k, ctx->u, ctx->v, ctx->part);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    int st = finish_op<RED_MAX>(ctx, nb, 1, S_MAXVEL);
    if (!st) {
        // min over the two remaining partial rows: shift them down with an offset
        hipLaunchKernelGGL(k_finish<RED_MIN>, dim3(1), dim3(256), 0, ctx->stream, ctx->part + RED_BLOCKS, nb, 2,
                           ctx->scal, (int)S_MINDT, pub_next(ctx));
        IMPEC_HIP_CHECK(ctx, hipGetLastError());
    }
    prof_end(ctx, "cfl");
    if (st) return st;
    if ((st = wait_scalars(ctx))) return st;
    const double *S = ctx->h_scal;
    double dt;
    if (ctx->P.is_test) {
        // MaxStableAdvectionTimestep (HyperbolicSolver.cpp:100-121)
        dt = S[S_MINDT] * ctx->P.cfl;
    } else {
        // MaxStableTimestepCombined (HyperbolicSolver.cpp:123-198)
        const double maxv = S[S_MAXVEL];
        const double a = ctx->dx / maxv, b = ctx->dy / maxv;
        double dt_adv = b < a ? b : a;
        dt_adv *= ctx->P.cfl;
        double dt_diff = DBL_MAX;
        if (ctx->P.diffusion) dt_diff = S[S_MINDT_DIFF] * ctx->P.dt_diff_limiter;
        if (ctx->P.para_method == IMPEC_PARA_CN) dt = dt_adv;
        else dt = dt_adv < dt_diff ? dt_adv : dt_diff;
    }
    *dt_out = dt;
    return IMPEC_OK;
}

// ---------------------------------------------------------------------------
// TRIntegrator::Integrate over the source cells only (TRIntegrator.cpp:49-71);
// cells with q == 0 are untouched by the reference loop.  The produced-mass
// sum runs over the sinks in row-major order on one thread when there are few
// of them (the reference's order), as a block tree otherwise.
// ---------------------------------------------------------------------------
template <class G>
__device__ __forceinline__ void tr_body(const G &gr_, double *__restrict__ c, const int *__restrict__ idx, const double *__restrict__ qv, int nsrc,
                     int N2, int B, int ld, double dt, double hdt, double inj, double *__restrict__ scal,
                     double *__restrict__ md) {
    const int t = threadIdx.x;
    for (int k = t; k < nsrc; k += blockDim.x) {
        const int cell = idx[k];
        const int i = cell / N2, j = cell - i * N2;
        double *cij = &c[(size_t)(i + B) * ld + (j + B)];
        const double qij = qv[k];
        double d = 0.;
        if (qij < 0) {
            const double old = *cij;
            *cij = *cij * (1 + hdt * qij) / (1 - hdt * qij);
            d = *cij - old;
        } else if (qij > 0) {
            *cij += dt * inj * qij;
        }
        md[k] = d;
    }
    __syncthreads();
    double tmo[1] = {0.};
    if (nsrc <= 4096) {
        if (t == 0)
            for (int k = 0; k < nsrc; k++)
                if (qv[k] < 0) tmo[0] += md[k];
        __syncthreads();
    } else {
        for (int k = t; k < nsrc; k += blockDim.x) tmo[0] += md[k];
        block_reduce<RED_SUM, 1>(tmo);
    }
    if (t == 0) {
        scal[S_TMO] = tmo[0];
        scal[S_OUTFLUX] += fabs(tmo[0]);
    }
}
__global__ void k_tr(double *__restrict__ c, const int *__restrict__ idx, const double *__restrict__ qv, int nsrc,
                     int N2, int B, int ld, double dt, double hdt, double inj, double *__restrict__ scal,
                     double *__restrict__ md) {
    tr_body(HwGrid{}, c, idx, qv, nsrc, N2, B, ld, dt, hdt, inj, scal, md);
}

int launch_tr(impec_ctx *ctx, double dt) {
    if (ctx->nsrc == 0) return IMPEC_OK;
    const double hdt = 0.5 * dt;
    prof_begin(ctx);
    hipLaunchKernelGGL(k_tr, dim3(1), dim3(1024), 0, ctx->stream, ctx->c, ctx->src_idx, ctx->src_q, ctx->nsrc,
                       ctx->N2, ctx->B, ctx->lds, dt, hdt, ctx->P.inj_fluid_conc, ctx->scal, ctx->tr_md);
    prof_end(ctx, "tr_source");
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

// ---------------------------------------------------------------------------
// diagnostics (Impec2DSolver.cpp:530-620)
// ---------------------------------------------------------------------------
__global__ void k_diag(const double *__restrict__ c, const double *__restrict__ u, const double *__restrict__ v,
                       const double *__restrict__ q, int N1, int N2, int B, int lds, int ldv, double dx, double dy,
                       double *__restrict__ part) {
    const long long n = (long long)N1 * N2;
    double s[3] = {0., 0., 0.};
    for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n;
         id += (long long)gridDim.x * blockDim.x) {
        const int id32 = (int)id, i = id32 / N2, j = id32 - i * N2;   // fewer than 2^31 cells
        s[0] += c[(size_t)(i + B) * lds + (j + B)];
        const size_t o = (size_t)(i + B) * ldv + (j + B);
        const double div = (v[o + 1] - v[o]) / dy + (u[o + ldv] - u[o]) / dx;
        s[1] += fabs(div - q[id]);
        double nf = 0.;
        if (j == 0) nf += fabs(v[o]);
        if (j == N2 - 1) nf += fabs(v[o + 1]);
        if (i == 0) nf += fabs(u[o]);
        if (i == N1 - 1) nf += fabs(u[o + ldv]);
        s[2] += nf;
    }
    block_reduce<RED_SUM, 3>(s);
    if (tid_linear() == 0)
        for (int k = 0; k < 3; k++) part[k * RED_BLOCKS + blockIdx.x] = s[k];   // 3 <= PART_VECS
}

int launch_diagnostics(impec_ctx *ctx, int which) {
    (void)which;
    const int nb = red_blocks((long long)ctx->N1 * ctx->N2);
    prof_begin(ctx);
    hipLaunchKernelGGL(k_diag, dim3(nb), dim3(256), 0, ctx->stream, ctx->c, ctx->u, ctx->v, ctx->q, ctx->N1,
                       ctx->N2, ctx->B, ctx->lds, ctx->ldv, ctx->dx, ctx->dy, ctx->part);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    int st = finish_sum(ctx, nb, 3, S_MASS, pub_next(ctx));
    prof_end(ctx, "diagnostics");
    if (st) return st;
    // S_MASS, S_MASS+1 (div sum), S_MASS+2 (noflow sum)
    return wait_scalars(ctx);
}

__global__ void k_minmax(const double *__restrict__ c, long long n, double *__restrict__ part) {
    double mx[1] = {-DBL_MAX}, mn[1] = {DBL_MAX};
    for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n;
         id += (long long)gridDim.x * blockDim.x) {
        const double x = c[id];
        mx[0] = x > mx[0] ? x : mx[0];
        mn[0] = x < mn[0] ? x : mn[0];
    }
    block_reduce<RED_MAX, 1>(mx);
    block_reduce<RED_MIN, 1>(mn);
    if (tid_linear() == 0) {
        part[blockIdx.x] = mx[0];
        part[RED_BLOCKS + blockIdx.x] = mn[0];
    }
}

// max/min over the whole ghosted array, like blitz max(c)/min(c); the row
// padding beyond N2+2B columns is excluded by walking rows of the pitch.
__global__ void k_minmax_rows(const double *__restrict__ c, int rows, int cols, int ld, double *__restrict__ part) {
    const long long n = (long long)rows * cols;
    double mx[1] = {-DBL_MAX}, mn[1] = {DBL_MAX};
    for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < n;
         id += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(id / cols), j = (int)(id - (long long)i * cols);
        const double x = c[(size_t)i * ld + j];
        mx[0] = x > mx[0] ? x : mx[0];
        mn[0] = x < mn[0] ? x : mn[0];
    }
    block_reduce<RED_MAX, 1>(mx);
    block_reduce<RED_MIN, 1>(mn);
    if (tid_linear() == 0) {
        part[blockIdx.x] = mx[0];
        part[RED_BLOCKS + blockIdx.x] = mn[0];
    }
}

int launch_minmax(impec_ctx *ctx, double *cmax, double *cmin) {
    const int rows = ctx->N1 + 2 * ctx->B, cols = ctx->N2 + 2 * ctx->B;
    const int nb = red_blocks((long long)rows * cols);
    hipLaunchKernelGGL(k_minmax_rows, dim3(nb), dim3(256), 0, ctx->stream, ctx->c, rows, cols, ctx->lds, ctx->part);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    int st = finish_op<RED_MAX>(ctx, nb, 1, S_CMAX);
    if (st) return st;
    hipLaunchKernelGGL(k_finish<RED_MIN>, dim3(1), dim3(256), 0, ctx->stream, ctx->part + RED_BLOCKS, nb, 1,
                       ctx->scal, (int)S_CMIN, pub_next(ctx));
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    if (int e = wait_scalars(ctx)) return e;
    *cmax = ctx->h_scal[S_CMAX];
    *cmin = ctx->h_scal[S_CMIN];
    return IMPEC_OK;
}


// ===========================================================================
// ensembles: the kernels above over every member in one launch each
// ===========================================================================
int red_blocks_of(long long n) { return red_blocks(n); }

int ens_set_error(impec_ensemble *e, int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(e->err, sizeof(e->err), fmt, ap);
    va_end(ap);
    return code;
}

#define ENS_CHECK(e, call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ens_set_error((e), IMPEC_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,   \
                                 hipGetErrorString(e_));                                       \
    } while (0)

__global__ void __launch_bounds__(1024) kb_finish(const FinishRec *__restrict__ tab, int need) {
    const FinishRec r = tab[blockIdx.x];
    if (need >= 0 && r.S[S_EST] != (double)need) return;
    const VGrid g{0, 0, 1, 1};
    if (r.op == RED_MAX) finish_body<RED_MAX>(g, r.part, r.nblocks, r.nv, r.S, r.dst);
    else finish_body<RED_SUM>(g, r.part, r.nblocks, r.nv, r.S, r.dst);
}

hipError_t ens_finish(hipStream_t s, const Batch<FinishRec> &b, int need) {
    if (b.nmem == 0) return hipSuccess;
    hipLaunchKernelGGL(kb_finish, dim3(b.nmem), dim3(1024), 0, s, b.tab, need);
    return hipGetLastError();
}

struct FillRec {
    double *a;
    int N1, N2, B, ld, gx, gy;
};
struct MobRec {
    const double *c, *K11, *K22;
    double *w, *w22;
    int rows, cols, ld, harmonic;
    double mob1over4, mu_res;
    int gx, gy;
};
struct VelRec {
    VelConst k;
    const double *p, *w11, *w22;
    double *u, *v;
    int gx, gy;
};
struct FillVRec {
    double *u, *v;
    int N1, N2, B, ld, gx, gy;
};
struct CflRec {
    CflConst k;
    const double *u, *v;
    double *part;
    int gx, gy;
};
struct DtRec {
    double *S;
    double dx, dy, cfl;
};
struct TrRec {
    double *c;
    const int *idx;
    const double *qv;
    int nsrc, N2, B, ld;
    double inj;
    double *S, *md;
    int gx, gy;
};

struct EnsFieldTabs {
    Batch<FillRec> fill_c, fill_p;
    Batch<MobRec> mob;
    Batch<VelRec> vel;
    Batch<FillVRec> fillv;
    Batch<CflRec> cfl;
    Batch<FinishRec> cfl_fin;
    Batch<DtRec> dt;
    Batch<TrRec> tr;
};

__global__ void kb_fill_scalar(const FillRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        fill_scalar_body(G, R.a, R.N1, R.N2, R.B, R.ld, 0);
    });
}
__global__ void kb_mobility(const MobRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        mobility_body(G, R.c, R.K11, R.K22, R.w, R.w22, R.rows, R.cols, R.ld, R.mob1over4, R.mu_res, R.harmonic);
    });
}
// ensemble members are nu = 1 (checked at creation): the single-context
// kernel's nu = 1 path
__global__ void kb_velocity(const VelRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        velocity_body<VGrid, true>(G, R.k, R.p, R.w11, R.w22, R.u, R.v);
    });
}
__global__ void kb_fill_velocity(const FillVRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        fill_velocity_body(G, R.u, R.v, R.N1, R.N2, R.B, R.ld);
    });
}
__global__ void kb_cfl(const CflRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        cfl_body(G, R.k, R.u, R.v, R.part);
    });
}
// MaxStableTimestepCombined without dispersion limit (HyperbolicSolver.cpp:123-198),
// the expressions of launch_cfl
__global__ void kb_dt(const DtRec *__restrict__ tab, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const DtRec r = tab[m];
    const double maxv = r.S[S_MAXVEL];
    const double a = r.dx / maxv, b = r.dy / maxv;
    double dt = b < a ? b : a;
    dt *= r.cfl;
    r.S[S_DT] = dt;
}
// TRIntegrator over half the member's step (HyperbolicSplitSolver.cpp:162-198)
__global__ void kb_tr(const TrRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        const double dt = 0.5 * R.S[S_DT];
        const double hdt = 0.5 * dt;
        tr_body(G, R.c, R.idx, R.qv, R.nsrc, R.N2, R.B, R.ld, dt, hdt, R.inj, R.S, R.md);
    });
}

template <class E>
static hipError_t launch_batch(void (*k)(const E *, const int *, int), const Batch<E> &b, dim3 blk, hipStream_t s) {
    if (b.nmem == 0) return hipSuccess;
    hipLaunchKernelGGL(k, dim3(batch_grid(b.total)), blk, 0, s, b.tab, b.start, b.nmem);
    return hipGetLastError();
}

void ens_field_free(impec_ensemble *e) {
    if (!e->field) return;
    EnsFieldTabs *t = e->field;
    batch_free(t->fill_c); batch_free(t->fill_p); batch_free(t->mob); batch_free(t->vel); batch_free(t->fillv);
    batch_free(t->cfl); batch_free(t->cfl_fin); batch_free(t->dt); batch_free(t->tr);
    delete t;
    e->field = nullptr;
}

int ens_field_build(impec_ensemble *e) {
    ens_field_free(e);
    EnsFieldTabs *t = new EnsFieldTabs();
    e->field = t;
    const int n = e->n;
    std::vector<FillRec> fc(n), fp(n);
    std::vector<MobRec> mb(n);
    std::vector<VelRec> vl(n);
    std::vector<FillVRec> fv(n);
    std::vector<CflRec> cf(n);
    std::vector<FinishRec> ff(n);
    std::vector<DtRec> dt(n);
    std::vector<TrRec> tr(n);
    std::vector<int> u_fill(n), u_mob(n), u_vel(n), u_fv(n), u_cfl(n), u_one(n, 1), u_tr(n);
    for (int k = 0; k < n; k++) {
        impec_ctx *ctx = e->m[k];
        const int N1 = ctx->N1, N2 = ctx->N2, B = ctx->B;
        // launch_fill_scalar
        fc[k] = FillRec{ctx->c, N1, N2, B, ctx->lds, (N2 + 2 * B + 63) / 64, (N1 + 2 * B + 3) / 4};
        fp[k] = fc[k];
        fp[k].a = ctx->p;
        u_fill[k] = fc[k].gx * fc[k].gy;
        // launch_mobility
        const int rows = N1 + 2 * B, cols = N2 + 2 * B;
        const int aniso = ctx->K22 != nullptr && ctx->K22 != ctx->K11;
        mb[k] = MobRec{ctx->c, ctx->K11, ctx->K22, ctx->w, aniso ? ctx->w22 : nullptr, rows, cols, ctx->lds,
                       ctx->P.average_procedure == IMPEC_AVG_HARMONIC, ctx->mob1over4, ctx->P.mu_res,
                       (cols + 63) / 64, (rows + 3) / 4};
        u_mob[k] = mb[k].gx * mb[k].gy;
        // launch_velocity
        VelConst &vk = vl[k].k;
        vk.N1 = N1; vk.N2 = N2; vk.B = B; vk.lds = ctx->lds; vk.ldv = ctx->ldv;
        vk.harmonic = ctx->P.average_procedure == IMPEC_AVG_HARMONIC;
        vk.nu = ctx->P.nu; vk.one_m_nu = 1. - ctx->P.nu;
        vk.dx = ctx->dx; vk.dy = ctx->dy;
        vk.idxi = 1. / ctx->dxi; vk.ideta = 1. / ctx->deta;
        vk.s2 = ctx->secamt2; vk.cosa2 = ctx->cosa2; vk.sina2 = ctx->sina2; vk.cost2 = ctx->cost2;
        vk.sint2 = ctx->sint2; vk.costsina = ctx->costsina; vk.cosasint = ctx->cosasint;
        vk.cosa = ctx->cosa; vk.sint = ctx->sint; vk.sina = ctx->sina; vk.cost = ctx->cost;
        vl[k].p = ctx->p;
        vl[k].w11 = ctx->w;
        vl[k].w22 = ctx->K22 && ctx->K22 != ctx->K11 ? ctx->w22 : ctx->w;
        vl[k].u = ctx->u;
        vl[k].v = ctx->v;
        vl[k].gx = (N2 + 1 + 63) / 64;
        vl[k].gy = (N1 + 1 + 3) / 4;
        u_vel[k] = vl[k].gx * vl[k].gy;
        // launch_fill_velocity
        fv[k] = FillVRec{ctx->u, ctx->v, N1, N2, B, ctx->ldv, (N2 + 1 + 2 * B + 63) / 64, (N1 + 1 + 2 * B + 3) / 4};
        u_fv[k] = fv[k].gx * fv[k].gy;
        // launch_cfl (combined mode, no dispersion limit)
        CflConst &ck = cf[k].k;
        ck.N1 = N1; ck.N2 = N2; ck.B = B; ck.ld = ctx->ldv; ck.mode = 0; ck.diff = 0;
        ck.dx = ctx->dx; ck.dy = ctx->dy; ck.al = ctx->P.alpha_l; ck.at = ctx->P.alpha_t;
        cf[k].u = ctx->u;
        cf[k].v = ctx->v;
        cf[k].part = ctx->part;
        cf[k].gx = red_blocks((long long)N1 * N2);
        cf[k].gy = 1;
        u_cfl[k] = cf[k].gx;
        ff[k] = FinishRec{ctx->part, ctx->scal, cf[k].gx, 1, (int)S_MAXVEL, (int)RED_MAX};
        dt[k] = DtRec{ctx->scal, ctx->dx, ctx->dy, ctx->P.cfl};
        // launch_tr
        tr[k] = TrRec{ctx->c, ctx->src_idx, ctx->src_q, ctx->nsrc, N2, B, ctx->lds, ctx->P.inj_fluid_conc, ctx->scal,
                      ctx->tr_md, 1, 1};
        u_tr[k] = ctx->nsrc > 0 ? 1 : 0;
    }
    ENS_CHECK(e, batch_upload(t->fill_c, fc, u_fill));
    ENS_CHECK(e, batch_upload(t->fill_p, fp, u_fill));
    ENS_CHECK(e, batch_upload(t->mob, mb, u_mob));
    ENS_CHECK(e, batch_upload(t->vel, vl, u_vel));
    ENS_CHECK(e, batch_upload(t->fillv, fv, u_fv));
    ENS_CHECK(e, batch_upload(t->cfl, cf, u_cfl));
    ENS_CHECK(e, batch_upload(t->cfl_fin, ff, u_one));
    ENS_CHECK(e, batch_upload(t->dt, dt, u_one));
    ENS_CHECK(e, batch_upload(t->tr, tr, u_tr));
    return IMPEC_OK;
}

int ens_fill_c(impec_ensemble *e) {
    ENS_CHECK(e, launch_batch(kb_fill_scalar, e->field->fill_c, dim3(64, 4), e->stream));
    return IMPEC_OK;
}

int ens_mobility(impec_ensemble *e) {
    ENS_CHECK(e, launch_batch(kb_mobility, e->field->mob, dim3(64, 4), e->stream));
    return IMPEC_OK;
}

int ens_velocity_dt(impec_ensemble *e) {
    EnsFieldTabs *t = e->field;
    ENS_CHECK(e, launch_batch(kb_fill_scalar, t->fill_p, dim3(64, 4), e->stream));
    ENS_CHECK(e, launch_batch(kb_velocity, t->vel, dim3(64, 4), e->stream));
    ENS_CHECK(e, launch_batch(kb_fill_velocity, t->fillv, dim3(64, 4), e->stream));
    ENS_CHECK(e, launch_batch(kb_cfl, t->cfl, dim3(256), e->stream));
    ENS_CHECK(e, ens_finish(e->stream, t->cfl_fin, -1));
    hipLaunchKernelGGL(kb_dt, dim3((e->n + 63) / 64), dim3(64), 0, e->stream, t->dt.tab, t->dt.nmem);
    ENS_CHECK(e, hipGetLastError());
    return IMPEC_OK;
}

int ens_tr(impec_ensemble *e) {
    ENS_CHECK(e, launch_batch(kb_tr, e->field->tr, dim3(1024), e->stream));
    return IMPEC_OK;
}

}  // namespace impec
