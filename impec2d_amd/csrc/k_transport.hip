// k_transport.hip -- dimension-split transport sweeps: MUSCL-Hancock with the
// ratio-form limiter and upwind Riemann flux (MUSCLSplitSolver.cpp:54-189),
// first-order upwind (UpwindSplitSolver.cpp:43-65), conservative updates
// (HyperbolicSplitSolver.cpp:113-159) and the splitting modes
// (HyperbolicSplitSolver.cpp:51-106).  FP64; the one-pass parallel
// splitting also in FP32 arithmetic (reconstruction, limiter, half step and
// fluxes) with FP64 states and updates, the mixed-precision configuration.
//
// Sweeps are out of place: cells a sweep updates are written from the input
// state, every other cell of the ghosted array is copied, which equals the
// reference's in-place update because its fluxes are all formed before the
// update loop.
#include "impec_device.h"
#include <vector>

namespace impec {

// the sweep's constants in the arithmetic type T of its fluxes
template <typename T>
struct SweepConstT {
    int N1, N2, B, lds, ldv;
    T hdtdx;   // 0.5 * dt / dx   (MUSCLSplitSolver.cpp:118)
    T dtdx;    // dt / dx         (HyperbolicSplitSolver.cpp:131)
    T phi;     // porosity
};
using SweepConst = SweepConstT<double>;

template <bool PHI1, typename T>
__device__ __forceinline__ T flux(T vel, T conc, T phi) {
    // HyperbolicSolver::Flux (HyperbolicSolver.cpp:208-210); x / 1.0 == x
    return PHI1 ? vel * conc : vel * conc / phi;
}

template <bool PHI1, typename T>
__device__ __forceinline__ T upwind(T speed, T UL, T UR, T phi) {
    return (speed > 0) ? flux<PHI1>(speed, UL, phi) : flux<PHI1>(speed, UR, phi);
}

template <typename T> __device__ __forceinline__ T max_finite();
template <> __device__ __forceinline__ double max_finite<double>() { return DBL_MAX; }
template <> __device__ __forceinline__ float max_finite<float>() { return FLT_MAX; }

// Reconstruct + GetSlope + SlopeLimiter (MUSCLSplitSolver.cpp:54-93).  The
// limiter min(2r/(1+r), 2/(1+r)) is evaluated with one division: both
// candidates share the rounded denominator 1+r and correctly rounded division
// is monotone in the numerator, so the smaller quotient is the one whose
// numerator min(2r, 2) is smaller -- bit-identical to evaluating both.
template <typename T>
__device__ __forceinline__ void reconstruct(T cm, T c0, T cp, T &UL, T &UR) {
    const T dprev = c0 - cm, dnext = cp - c0;
    T lim = 0;
    if (dprev * dnext > 0) {
        T r = dprev / dnext;
        if (r > max_finite<T>()) r = max_finite<T>();
        const T num = r < T(1) ? T(2) * r : T(2);
        lim = num / (T(1) + r);
    }
    const T d = lim * T(0.5) * (dprev + dnext);
    UR = c0 + T(0.5) * d;
    UL = c0 - T(0.5) * d;
}

// half-step evolved interface states of cell k: left state of face k+1 and
// right state of face k
template <bool PHI1, typename T>
__device__ __forceinline__ void hancock(T cm, T c0, T cp, T vl, T vr, const SweepConstT<T> &k, T &left_of_next,
                                        T &right_of_this) {
    T UL, UR;
    reconstruct<T>(cm, c0, cp, UL, UR);
    const T df = k.hdtdx * (flux<PHI1>(vl, UL, k.phi) - flux<PHI1>(vr, UR, k.phi));
    left_of_next = UR + df;
    right_of_this = UL + df;
}

// X sweep (Advect_X): cells i in [0, N1), j in [-B, N2+B) are updated.
template <bool MUSCL, bool PHI1, class G>
__device__ __forceinline__ void sweep_x_body(const G &gr_, SweepConst k, const double *__restrict__ cin, double *__restrict__ cout,
                          const double *__restrict__ u) {
    const int j = gr_.bx() * blockDim.x + threadIdx.x - k.B;
    const int i = gr_.by() * blockDim.y + threadIdx.y - k.B;
    if (i >= k.N1 + k.B || j >= k.N2 + k.B) return;
    const long o = (long)(i + k.B) * k.lds + (j + k.B);
    if (i < 0 || i >= k.N1) {
        cout[o] = cin[o];
        return;
    }
    const long ov = (long)(i + k.B) * k.ldv + (j + k.B);
    const long L = k.lds, LV = k.ldv;
    double Fi, Fip1;
    if (MUSCL) {
        const double cm2 = cin[o - 2 * L], cm1 = cin[o - L], c0 = cin[o], cp1 = cin[o + L], cp2 = cin[o + 2 * L];
        const double um1 = u[ov - LV], u0 = u[ov], up1 = u[ov + LV], up2 = u[ov + 2 * LV];
        double lA, rA, lB, rB, lC, rC;
        hancock<PHI1>(cm2, cm1, c0, um1, u0, k, lA, rA);   // cell i-1
        hancock<PHI1>(cm1, c0, cp1, u0, up1, k, lB, rB);   // cell i
        hancock<PHI1>(c0, cp1, cp2, up1, up2, k, lC, rC);  // cell i+1
        (void)rA; (void)lC;
        Fi = upwind<PHI1>(u0, lA, rB, k.phi);
        Fip1 = upwind<PHI1>(up1, lB, rC, k.phi);
    } else {
        Fi = upwind<PHI1>(u[ov], cin[o - L], cin[o], k.phi);
        Fip1 = upwind<PHI1>(u[ov + LV], cin[o], cin[o + L], k.phi);
    }
    cout[o] = cin[o] + k.dtdx * (Fi - Fip1);
}
template <bool MUSCL, bool PHI1>
__global__ void k_sweep_x(SweepConst k, const double *__restrict__ cin, double *__restrict__ cout,
                          const double *__restrict__ u) {
    sweep_x_body<MUSCL, PHI1>(HwGrid{}, k, cin, cout, u);
}

// Y sweep (Advect_Y): cells i in [-B, N1+B), j in [0, N2) are updated.
template <bool MUSCL, bool PHI1, class G>
__device__ __forceinline__ void sweep_y_body(const G &gr_, SweepConst k, const double *__restrict__ cin, double *__restrict__ cout,
                          const double *__restrict__ v) {
    const int j = gr_.bx() * blockDim.x + threadIdx.x - k.B;
    const int i = gr_.by() * blockDim.y + threadIdx.y - k.B;
    if (i >= k.N1 + k.B || j >= k.N2 + k.B) return;
    const long o = (long)(i + k.B) * k.lds + (j + k.B);
    if (j < 0 || j >= k.N2) {
        cout[o] = cin[o];
        return;
    }
    const long ov = (long)(i + k.B) * k.ldv + (j + k.B);
    double Gj, Gjp1;
    if (MUSCL) {
        const double cm2 = cin[o - 2], cm1 = cin[o - 1], c0 = cin[o], cp1 = cin[o + 1], cp2 = cin[o + 2];
        const double vm1 = v[ov - 1], v0 = v[ov], vp1 = v[ov + 1], vp2 = v[ov + 2];
        double lA, rA, lB, rB, lC, rC;
        hancock<PHI1>(cm2, cm1, c0, vm1, v0, k, lA, rA);
        hancock<PHI1>(cm1, c0, cp1, v0, vp1, k, lB, rB);
        hancock<PHI1>(c0, cp1, cp2, vp1, vp2, k, lC, rC);
        (void)rA; (void)lC;
        Gj = upwind<PHI1>(v0, lA, rB, k.phi);
        Gjp1 = upwind<PHI1>(vp1, lB, rC, k.phi);
    } else {
        Gj = upwind<PHI1>(v[ov], cin[o - 1], cin[o], k.phi);
        Gjp1 = upwind<PHI1>(v[ov + 1], cin[o], cin[o + 1], k.phi);
    }
    cout[o] = cin[o] + k.dtdx * (Gj - Gjp1);
}
template <bool MUSCL, bool PHI1>
__global__ void k_sweep_y(SweepConst k, const double *__restrict__ cin, double *__restrict__ cout,
                          const double *__restrict__ v) {
    sweep_y_body<MUSCL, PHI1>(HwGrid{}, k, cin, cout, v);
}

// The intercell fluxes alone, with the sweeps' arithmetic
// (CalculateIntercellFlux_F, MUSCLSplitSolver.cpp:102-145 / UpwindSplitSolver.cpp:43-51):
// F(i,j) for faces i in [0, N1], j in [-B, N2+B) into F[i * (N2+2B) + j + B].
template <bool MUSCL, bool PHI1>
__global__ void k_flux_x(SweepConst k, const double *__restrict__ c, const double *__restrict__ u,
                         double *__restrict__ F) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x - k.B;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > k.N1 || j >= k.N2 + k.B) return;
    const long o = (long)(i + k.B) * k.lds + (j + k.B);
    const long ov = (long)(i + k.B) * k.ldv + (j + k.B);
    const long L = k.lds, LV = k.ldv;
    double Fi;
    if (MUSCL) {
        double lA, rA, lB, rB;
        hancock<PHI1>(c[o - 2 * L], c[o - L], c[o], u[ov - LV], u[ov], k, lA, rA);   // cell i-1
        hancock<PHI1>(c[o - L], c[o], c[o + L], u[ov], u[ov + LV], k, lB, rB);       // cell i
        (void)rA; (void)lB;
        Fi = upwind<PHI1>(u[ov], lA, rB, k.phi);
    } else {
        Fi = upwind<PHI1>(u[ov], c[o - L], c[o], k.phi);
    }
    F[(long)i * (k.N2 + 2 * k.B) + (j + k.B)] = Fi;
}

// G(i,j) for i in [-B, N1+B), faces j in [0, N2] into G[(i+B) * (N2+1) + j]
// (CalculateIntercellFlux_G, MUSCLSplitSolver.cpp:147-189 / UpwindSplitSolver.cpp:53-61)
template <bool MUSCL, bool PHI1>
__global__ void k_flux_y(SweepConst k, const double *__restrict__ c, const double *__restrict__ v,
                         double *__restrict__ G) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y - k.B;
    if (i >= k.N1 + k.B || j > k.N2) return;
    const long o = (long)(i + k.B) * k.lds + (j + k.B);
    const long ov = (long)(i + k.B) * k.ldv + (j + k.B);
    double Gj;
    if (MUSCL) {
        double lA, rA, lB, rB;
        hancock<PHI1>(c[o - 2], c[o - 1], c[o], v[ov - 1], v[ov], k, lA, rA);
        hancock<PHI1>(c[o - 1], c[o], c[o + 1], v[ov], v[ov + 1], k, lB, rB);
        (void)rA; (void)lB;
        Gj = upwind<PHI1>(v[ov], lA, rB, k.phi);
    } else {
        Gj = upwind<PHI1>(v[ov], c[o - 1], c[o], k.phi);
    }
    G[(long)(i + k.B) * (k.N2 + 1) + j] = Gj;
}

// c = 0.5 * (a + b) over the whole ghosted array (HyperbolicSplitSolver.cpp:97-101);
// the ensemble's average (single contexts average inside k_advect_parallel)
template <class G>
__device__ __forceinline__ void average_body(const G &gr_, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ c, int rows,
                          int cols, int ld) {
    const int j = gr_.bx() * blockDim.x + threadIdx.x;
    const int i = gr_.by() * blockDim.y + threadIdx.y;
    if (i >= rows || j >= cols) return;
    const long o = (long)i * ld + j;
    c[o] = 0.5 * (a[o] + b[o]);
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, int rows, int cols, int ld) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= rows || j >= cols) return;
    const long o = (long)i * ld + j;
    b[o] = a[o];
}

// ===========================================================================
// The symmetric parallel splitting in one streaming pass
// (HyperbolicSplitSolver.cpp:51-106, mode SYMMETRIC_PARALLEL): X then Y, and
// Y then X, from the same state, averaged over the whole ghosted array.  A
// wavefront owns a strip of 64 columns (one per lane, 60 written) and marches
// down a chunk of rows.  Per row k it Y-sweeps the loaded row (YX path, column
// neighbours from lanes +-1, +-2 by DPP shifts); on row k-2 it X-sweeps the
// loaded state (XY path) and the Y-swept rows (YX path) from register windows
// of five rows, Y-sweeps the XY row, averages and writes.  Every update is the
// per-cell arithmetic of k_sweep_x / k_sweep_y with the same operands (ghost
// rows Y-swept, ghost columns X-swept, as there), so the result is bitwise that
// of the four sweeps and the average.  It reads c, u and v once and writes
// the averaged state once: 32 B per cell instead of five passes of 24.
// ===========================================================================
constexpr int PS_OUT = 60, PS_WAVES = 4;
#ifndef IMPEC_PS_PF
#define IMPEC_PS_PF 2
#endif
#ifndef IMPEC_PS_TARGET_WAVES
#define IMPEC_PS_TARGET_WAVES 4096
#endif

// T: the arithmetic of the reconstructions, half steps and fluxes; states,
// updates and the average stay FP64 whatever T is
template <typename T>
struct ParArgsT {
    SweepConstT<T> kx, ky;      // x sweeps (dt / dx), y sweeps (dt / dy)
    const double *c, *u, *v;    // ghosted: c pitch lds, u and v pitch ldv
    double *out;                // the averaged state, pitch lds
    int R, nstrip, nwave;       // output rows per chunk, strips, wavefronts
};
using ParArgs = ParArgsT<double>;

// Y sweep of one row held across the lanes (column j in lane j), with the
// arithmetic of k_sweep_y: each cell's half-step states and each face's flux
// are formed once, in the lane of the cell (of the face's right cell), and
// reach the neighbours by lane shifts.  Cell j's update needs lanes j-2..j+2.
// the conservative update c + dtdx (F_left - F_right).  FP64: the reference's
// expression.  FP32 fluxes: each face's dtdx F is rounded once and both cells
// of the face take that same value, so the increments telescope and the FP64
// state keeps its mass to FP64 rounding
__device__ __forceinline__ double flux_update(double c, double dtdx, double Fl, double Fr) {
    return c + dtdx * (Fl - Fr);
}
__device__ __forceinline__ double flux_update(double c, float dtdx, float Fl, float Fr) {
    return c + ((double)(dtdx * Fl) - (double)(dtdx * Fr));
}

template <bool MUSCL, bool PHI1, typename T>
__device__ __forceinline__ double ysweep_lanes(const SweepConstT<T> &k, double c, double v) {
    const T ct = (T)c, vt = (T)v;   // no-ops in FP64
    const T cl = lane_prev(ct);
    T G;   // flux through the face left of the lane's cell
    if (MUSCL) {
        T lo, ro;
        hancock<PHI1>(cl, ct, lane_next(ct), vt, lane_next(vt), k, lo, ro);
        G = upwind<PHI1>(vt, lane_prev(lo), ro, k.phi);
    } else {
        G = upwind<PHI1>(vt, cl, ct, k.phi);
    }
    return flux_update(c, k.dtdx, G, lane_next(G));
}

// X sweep of a lane's column as rows stream past, with the arithmetic of
// k_sweep_x: push(row k, faces k-1 and k) forms the half-step states of row
// k-1 and the flux through face k-1, and returns the update of row k-2.
// After push(row k) the fields hold:
template <typename T>
struct XStream {
    double cm2 = 0., cm1 = 0., c0 = 0.;   // rows k-2, k-1, k (the states, FP64)
    T lo = 0;                             // left state of face k (from row k-1)
    T F = 0;                              // flux through face k-1
    template <bool MUSCL, bool PHI1>
    __device__ __forceinline__ double push(const SweepConstT<T> &k, double ck, double ukm1, double uk) {
        cm2 = cm1; cm1 = c0; c0 = ck;
        T Fn;   // flux through face k-1
        if (MUSCL) {
            T lo_n, ro_n;
            hancock<PHI1>((T)cm2, (T)cm1, (T)c0, (T)ukm1, (T)uk, k, lo_n, ro_n);
            Fn = upwind<PHI1>((T)ukm1, lo, ro_n, k.phi);
            lo = lo_n;
        } else {
            Fn = upwind<PHI1>((T)ukm1, (T)cm2, (T)cm1, k.phi);
        }
        const double out = flux_update(cm2, k.dtdx, F, Fn);
        F = Fn;
        return out;
    }
};

template <bool MUSCL, bool PHI1, typename T>
__device__ __forceinline__ void advect_parallel_body(const ParArgsT<T> &A, const int wid) {
    const int lane = threadIdx.x & 63;
    if (wid >= A.nwave) return;
    const SweepConstT<T> &kx = A.kx, &ky = A.ky;
    const int N1 = kx.N1, N2 = kx.N2, B = kx.B;
    const long LS = kx.lds, LV = kx.ldv;
    const int strip = wid % A.nstrip, chunk = wid / A.nstrip;
    const int j = strip * PS_OUT - 2 - B + lane;                 // ghosted column of the lane
    const bool outl = lane >= 2 && lane < 2 + PS_OUT && j < N2 + B;
    const int jc = j < -B ? -B : (j > N2 + B - 1 ? N2 + B - 1 : j);   // cell column for loads
    const int jf = j < -B ? -B : (j > N2 + B ? N2 + B : j);           // face column for loads
    const bool ycol = j >= 0 && j < N2;                          // Y sweeps update this column
    const int i0 = -B + chunk * A.R, i1 = i0 + A.R < N1 + B ? i0 + A.R : N1 + B;
    auto cload = [&](int m) -> double {
        const int mc = m < -B ? -B : (m > N1 + B - 1 ? N1 + B - 1 : m);
        return A.c[(long)(mc + B) * LS + (jc + B)];
    };
    auto fload = [&](const double *p, int m) -> double {
        const int mc = m < -B ? -B : (m > N1 + B ? N1 + B : m);
        return p[(long)(mc + B) * LV + (jf + B)];
    };
    constexpr int PF = IMPEC_PS_PF;
    double Pc[PF] = {}, Pu[PF] = {}, Pv[PF] = {};
    auto fetch = [&](int s_, int m) {
        Pc[s_] = cload(m);
        Pu[s_] = fload(A.u, m);
        Pv[s_] = fload(A.v, m);
    };
    // output row i = k - 2 needs rows i-2 .. i+2 of the state and faces i-1 .. i+2
    const int kstart = i0 - 2, kend = i1 + 2;
    const int kstop = kstart + (kend - kstart + PF - 1) / PF * PF;
#pragma unroll
    for (int s_ = 0; s_ < PF; s_++) fetch(s_, kstart + s_);
    XStream<T> xy, yx;                    // X sweeps of the state (XY) and of its Y sweep (YX)
    double ukm1 = 0., vkm1 = 0., vkm2 = 0.;
    for (int kb = kstart; kb < kstop; kb += PF) {
#pragma unroll
        for (int s_ = 0; s_ < PF; s_++) {
            const int k = kb + s_;
            const double ck = Pc[s_], uk = Pu[s_], vk = Pv[s_];
            const int i = k - 2;
            const bool xrow = i >= 0 && i < N1;                  // X sweeps update this row
            // XY: X sweep of the state, then Y sweep of row i
            const double x1 = xy.template push<MUSCL, PHI1>(kx, ck, ukm1, uk);
            const double c1 = xrow ? x1 : xy.cm2;
            const double y2 = ysweep_lanes<MUSCL, PHI1>(ky, c1, vkm2);
            const double c2 = ycol ? y2 : c1;
            // YX: Y sweep of row k, then X sweep of row i
            const double y3 = ysweep_lanes<MUSCL, PHI1>(ky, ck, vk);
            const double x4 = yx.template push<MUSCL, PHI1>(kx, ycol ? y3 : ck, ukm1, uk);
            const double c4 = xrow ? x4 : yx.cm2;
            if (outl && i >= i0 && i < i1) A.out[(long)(i + B) * LS + (j + B)] = 0.5 * (c2 + c4);
            ukm1 = uk;
            vkm2 = vkm1;
            vkm1 = vk;
            fetch(s_, k + PF < kend ? k + PF : kend - 1);   // past kend - 1: a re-read (cache hit)
        }
    }
}
template <bool MUSCL, bool PHI1, typename T>
__global__ void __launch_bounds__(64 * PS_WAVES) k_advect_parallel(ParArgsT<T> A) {
    advect_parallel_body<MUSCL, PHI1>(A, blockIdx.x * PS_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6));
}

// The ensemble's one-pass parallel splitting: one launch over every member's
// wavefronts (member m owns waves [start[m], start[m+1])), the step's factors
// formed from the member's device dt as sweep_const forms them
struct ParRec {
    ParArgs a;       // dt factors zero; out is the member's t1
    double dx, dy;
    const double *S;
};
template <bool MUSCL, bool PHI1>
__global__ void __launch_bounds__(64 * PS_WAVES)
    kp_advect_parallel(const ParRec *__restrict__ tab, const int *__restrict__ start, int nmem, int total) {
    const int gw = blockIdx.x * PS_WAVES + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (gw >= total) return;
    const int m = member_of(start, nmem, gw);
    ParRec R = tab[m];
    const double dt = R.S[S_DT];
    R.a.kx.hdtdx = 0.5 * dt / R.dx;
    R.a.kx.dtdx = dt / R.dx;
    R.a.ky.hdtdx = 0.5 * dt / R.dy;
    R.a.ky.dtdx = dt / R.dy;
    advect_parallel_body<MUSCL, PHI1>(R.a, gw - start[m]);
}

static SweepConst sweep_const(const impec_ctx *ctx, double dt, double h) {
    SweepConst k;
    k.N1 = ctx->N1; k.N2 = ctx->N2; k.B = ctx->B; k.lds = ctx->lds; k.ldv = ctx->ldv;
    k.hdtdx = 0.5 * dt / h;
    k.dtdx = dt / h;
    k.phi = ctx->P.porosity;
    return k;
}

// the sweep constants rounded to the arithmetic type T
template <typename T>
static SweepConstT<T> sweep_const_t(const impec_ctx *ctx, double dt, double h) {
    const SweepConst d = sweep_const(ctx, dt, h);
    SweepConstT<T> k;
    k.N1 = d.N1; k.N2 = d.N2; k.B = d.B; k.lds = d.lds; k.ldv = d.ldv;
    k.hdtdx = (T)d.hdtdx;
    k.dtdx = (T)d.dtdx;
    k.phi = (T)d.phi;
    return k;
}

// arguments of the one-pass parallel splitting of c into out: strips of
// PS_OUT columns, chunks of rows for about IMPEC_PS_TARGET_WAVES wavefronts
template <typename T = double>
static ParArgsT<T> par_args(const impec_ctx *ctx, double dt, const double *c, double *out) {
    ParArgsT<T> a;
    a.kx = sweep_const_t<T>(ctx, dt, ctx->dx);
    a.ky = sweep_const_t<T>(ctx, dt, ctx->dy);
    a.c = c; a.u = ctx->u; a.v = ctx->v; a.out = out;
    const int rows = ctx->N1 + 2 * ctx->B, cols = ctx->N2 + 2 * ctx->B;
    a.nstrip = (cols + PS_OUT - 1) / PS_OUT;
    long R = ((long)rows * a.nstrip + IMPEC_PS_TARGET_WAVES - 1) / IMPEC_PS_TARGET_WAVES;
    if (R < 8) R = 8;
    if (R > rows) R = rows;
    a.R = (int)R;
    a.nwave = a.nstrip * ((rows + a.R - 1) / a.R);
    return a;
}

template <bool MUSCL, bool PHI1>
static void launch_x(impec_ctx *ctx, const SweepConst &k, const double *cin, double *cout) {
    dim3 blk(64, 4), grd((k.N2 + 2 * k.B + 63) / 64, (k.N1 + 2 * k.B + 3) / 4);
    hipLaunchKernelGGL((k_sweep_x<MUSCL, PHI1>), grd, blk, 0, ctx->stream, k, cin, cout, ctx->u);
}
template <bool MUSCL, bool PHI1>
static void launch_y(impec_ctx *ctx, const SweepConst &k, const double *cin, double *cout) {
    dim3 blk(64, 4), grd((k.N2 + 2 * k.B + 63) / 64, (k.N1 + 2 * k.B + 3) / 4);
    hipLaunchKernelGGL((k_sweep_y<MUSCL, PHI1>), grd, blk, 0, ctx->stream, k, cin, cout, ctx->v);
}

int advect_x(impec_ctx *ctx, const double *cin, double *cout, double dt) {
    SweepConst k = sweep_const(ctx, dt, ctx->dx);
    const bool muscl = ctx->P.hyper_method == IMPEC_HYPER_MUSCL, phi1 = ctx->P.porosity == 1.;
    prof_begin(ctx);
    if (muscl && phi1) launch_x<true, true>(ctx, k, cin, cout);
    else if (muscl) launch_x<true, false>(ctx, k, cin, cout);
    else if (phi1) launch_x<false, true>(ctx, k, cin, cout);
    else launch_x<false, false>(ctx, k, cin, cout);
    prof_end(ctx, "sweep_x");
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

int advect_y(impec_ctx *ctx, const double *cin, double *cout, double dt) {
    SweepConst k = sweep_const(ctx, dt, ctx->dy);
    const bool muscl = ctx->P.hyper_method == IMPEC_HYPER_MUSCL, phi1 = ctx->P.porosity == 1.;
    prof_begin(ctx);
    if (muscl && phi1) launch_y<true, true>(ctx, k, cin, cout);
    else if (muscl) launch_y<true, false>(ctx, k, cin, cout);
    else if (phi1) launch_y<false, true>(ctx, k, cin, cout);
    else launch_y<false, false>(ctx, k, cin, cout);
    prof_end(ctx, "sweep_y");
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

// fluxes of the current c and face velocities into F (device, layouts of
// k_flux_x / k_flux_y)
int intercell_flux(impec_ctx *ctx, int dir, double dt, double *F) {
    SweepConst k = sweep_const(ctx, dt, dir == 0 ? ctx->dx : ctx->dy);
    const bool muscl = ctx->P.hyper_method == IMPEC_HYPER_MUSCL, phi1 = ctx->P.porosity == 1.;
    dim3 blk(64, 4);
    if (dir == 0) {
        dim3 grd((k.N2 + 2 * k.B + 63) / 64, (k.N1 + 1 + 3) / 4);
        if (muscl && phi1) hipLaunchKernelGGL((k_flux_x<true, true>), grd, blk, 0, ctx->stream, k, ctx->c, ctx->u, F);
        else if (muscl) hipLaunchKernelGGL((k_flux_x<true, false>), grd, blk, 0, ctx->stream, k, ctx->c, ctx->u, F);
        else if (phi1) hipLaunchKernelGGL((k_flux_x<false, true>), grd, blk, 0, ctx->stream, k, ctx->c, ctx->u, F);
        else hipLaunchKernelGGL((k_flux_x<false, false>), grd, blk, 0, ctx->stream, k, ctx->c, ctx->u, F);
    } else {
        dim3 grd((k.N2 + 1 + 63) / 64, (k.N1 + 2 * k.B + 3) / 4);
        if (muscl && phi1) hipLaunchKernelGGL((k_flux_y<true, true>), grd, blk, 0, ctx->stream, k, ctx->c, ctx->v, F);
        else if (muscl) hipLaunchKernelGGL((k_flux_y<true, false>), grd, blk, 0, ctx->stream, k, ctx->c, ctx->v, F);
        else if (phi1) hipLaunchKernelGGL((k_flux_y<false, true>), grd, blk, 0, ctx->stream, k, ctx->c, ctx->v, F);
        else hipLaunchKernelGGL((k_flux_y<false, false>), grd, blk, 0, ctx->stream, k, ctx->c, ctx->v, F);
    }
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

int copy_scalar(impec_ctx *ctx, const double *src, double *dst) {
    const int rows = ctx->N1 + 2 * ctx->B, cols = ctx->N2 + 2 * ctx->B;
    dim3 blk(64, 4), grd((cols + 63) / 64, (rows + 3) / 4);
    hipLaunchKernelGGL(k_copy, grd, blk, 0, ctx->stream, src, dst, rows, cols, ctx->lds);
    IMPEC_HIP_CHECK(ctx, hipGetLastError());
    return IMPEC_OK;
}

template <typename T>
static void launch_parallel(impec_ctx *ctx, const ParArgsT<T> &a, bool muscl, bool phi1) {
    const dim3 g((a.nwave + PS_WAVES - 1) / PS_WAVES), blk(64 * PS_WAVES);
    if (muscl && phi1) hipLaunchKernelGGL((k_advect_parallel<true, true, T>), g, blk, 0, ctx->stream, a);
    else if (muscl) hipLaunchKernelGGL((k_advect_parallel<true, false, T>), g, blk, 0, ctx->stream, a);
    else if (phi1) hipLaunchKernelGGL((k_advect_parallel<false, true, T>), g, blk, 0, ctx->stream, a);
    else hipLaunchKernelGGL((k_advect_parallel<false, false, T>), g, blk, 0, ctx->stream, a);
}

// AdvectSplit (HyperbolicSplitSolver.cpp:51-106) after its FillBoundaries
int advect_split(impec_ctx *ctx, double dt) {
    int st = launch_fill_scalar(ctx, ctx->c, ctx->P.is_test);
    if (st) return st;
    double *c = ctx->c, *t1 = ctx->t1, *t2 = ctx->t2, *t3 = ctx->t3;
    switch (ctx->P.dim_splitting) {
    case IMPEC_SPLIT_ALTERNATING:
        if (ctx->prev_first_sweep == 1) {
            if ((st = advect_x(ctx, c, t1, dt)) || (st = advect_y(ctx, t1, c, dt))) return st;
            ctx->prev_first_sweep = 0;
        } else {
            if ((st = advect_y(ctx, c, t1, dt)) || (st = advect_x(ctx, t1, c, dt))) return st;
            ctx->prev_first_sweep = 1;
        }
        break;
    case IMPEC_SPLIT_X_FIRST:
        if ((st = advect_x(ctx, c, t1, dt)) || (st = advect_y(ctx, t1, c, dt))) return st;
        ctx->prev_first_sweep = 0;
        break;
    case IMPEC_SPLIT_Y_FIRST:
        if ((st = advect_y(ctx, c, t1, dt)) || (st = advect_x(ctx, t1, c, dt))) return st;
        ctx->prev_first_sweep = 1;
        break;
    case IMPEC_SPLIT_STRANG:
        if ((st = advect_x(ctx, c, t1, 0.5 * dt)) || (st = advect_y(ctx, t1, t2, dt)) ||
            (st = advect_x(ctx, t2, c, 0.5 * dt)))
            return st;
        ctx->prev_first_sweep = 0;
        break;
    default: {
        // X then Y and Y then X from c, averaged: one streaming pass into t1,
        // which becomes c (the two buffers exchange roles; every user of c
        // reads ctx->c when it launches), or is copied back into c where the
        // addresses are held elsewhere
        (void)t2; (void)t3;
        const int rows = ctx->N1 + 2 * ctx->B;
        const bool muscl = ctx->P.hyper_method == IMPEC_HYPER_MUSCL, phi1 = ctx->P.porosity == 1.;
        prof_begin(ctx);
        if (ctx->P.transport32) launch_parallel<float>(ctx, par_args<float>(ctx, dt, c, t1), muscl, phi1);
        else launch_parallel<double>(ctx, par_args<double>(ctx, dt, c, t1), muscl, phi1);
        prof_end(ctx, "advect_parallel");
        IMPEC_HIP_CHECK(ctx, hipGetLastError());
        if (ctx->fixed_buffers) {
            IMPEC_HIP_CHECK(ctx, hipMemcpyAsync(c, t1, sizeof(double) * (size_t)rows * ctx->lds,
                                                hipMemcpyDeviceToDevice, ctx->stream));
        } else {
            ctx->c = t1;
            ctx->t1 = c;
        }
        ctx->prev_first_sweep = 0;
    }
    }
    return IMPEC_OK;
}


// ===========================================================================
// ensembles: the symmetric parallel splitting of AdvectSplit
// (HyperbolicSplitSolver.cpp:51-106) for every member in one launch per
// sweep, with each member's time step read from its S_DT
// ===========================================================================
#define ENS_CHECK(e, call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ens_set_error((e), IMPEC_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,   \
                                 hipGetErrorString(e_));                                       \
    } while (0)

struct AvgRec {
    const double *a, *b;
    double *c;
    int rows, cols, ld, gx, gy;
};

struct EnsTransTabs {
    int muscl, phi1;
    Batch<ParRec> par;      // c -> t1, both orders averaged (units: wavefronts)
    Batch<AvgRec> back;     // t1 -> c as 0.5 (t1 + t1), which is t1 exactly
};

__global__ void kb_average(const AvgRec *__restrict__ tab, const int *__restrict__ start, int nmem) {
    batch_for(tab, start, nmem, [&](const auto &R, const VGrid &G) {
        average_body(G, R.a, R.b, R.c, R.rows, R.cols, R.ld);
    });
}

void ens_trans_free(impec_ensemble *e) {
    if (!e->trans) return;
    EnsTransTabs *t = e->trans;
    batch_free(t->par); batch_free(t->back);
    delete t;
    e->trans = nullptr;
}

int ens_trans_build(impec_ensemble *e) {
    ens_trans_free(e);
    EnsTransTabs *t = new EnsTransTabs();
    e->trans = t;
    const int n = e->n;
    t->muscl = e->m[0]->P.hyper_method == IMPEC_HYPER_MUSCL;
    t->phi1 = e->m[0]->P.porosity == 1.;
    std::vector<ParRec> pr(n);
    std::vector<AvgRec> bk(n);
    std::vector<int> wv(n), u(n);
    for (int k = 0; k < n; k++) {
        impec_ctx *ctx = e->m[k];
        pr[k] = ParRec{par_args(ctx, 0., ctx->c, ctx->t1), ctx->dx, ctx->dy, ctx->scal};
        wv[k] = pr[k].a.nwave;
        const int gx = (ctx->N2 + 2 * ctx->B + 63) / 64, gy = (ctx->N1 + 2 * ctx->B + 3) / 4;
        bk[k] = AvgRec{ctx->t1, ctx->t1, ctx->c, ctx->N1 + 2 * ctx->B, ctx->N2 + 2 * ctx->B, ctx->lds, gx, gy};
        u[k] = gx * gy;
    }
    ENS_CHECK(e, batch_upload(t->par, pr, wv));
    ENS_CHECK(e, batch_upload(t->back, bk, u));
    return IMPEC_OK;
}

template <bool MUSCL, bool PHI1>
static int ens_parallel(impec_ensemble *e) {
    EnsTransTabs *t = e->trans;
    hipStream_t s = e->stream;
    if (t->par.nmem)
        hipLaunchKernelGGL((kp_advect_parallel<MUSCL, PHI1>), dim3((t->par.total + PS_WAVES - 1) / PS_WAVES),
                           dim3(64 * PS_WAVES), 0, s, t->par.tab, t->par.start, t->par.nmem, t->par.total);
    if (t->back.nmem)
        hipLaunchKernelGGL(kb_average, dim3(batch_grid(t->back.total)), dim3(64, 4), 0, s, t->back.tab,
                           t->back.start, t->back.nmem);
    ENS_CHECK(e, hipGetLastError());
    return IMPEC_OK;
}

int ens_advect_parallel(impec_ensemble *e) {
    const EnsTransTabs *t = e->trans;
    if (t->muscl && t->phi1) return ens_parallel<true, true>(e);
    if (t->muscl) return ens_parallel<true, false>(e);
    if (t->phi1) return ens_parallel<false, true>(e);
    return ens_parallel<false, false>(e);
}

}  // namespace impec
