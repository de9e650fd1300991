// impec_internal.h -- device state of one context and the kernel launchers
// shared by the .hip translation units of libimpec_hip.so (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include "../../include/impec_hip.h"
#include "impec_batch.h"

namespace impec {

constexpr int GHOST = 2;          // num_ghostcells of every shipped config
constexpr int MAX_LEVELS = 24;
constexpr int RED_BLOCKS = 65536; // max partial sums per reduction pass
constexpr int PART_VECS = 6;      // partial-sum vectors (stride RED_BLOCKS) in a context's part buffer
constexpr int FIN_THREADS = 1024; // threads of the one-workgroup combine of partials
#ifndef IMPEC_VEC_BLOCKS
#define IMPEC_VEC_BLOCKS 16384
#endif
// workgroups of the grid-stride reduction kernels (their partials are summed
// by one workgroup); build-time tunable.  At 4096^2 the CG update with 16384
// workgroups (4 rows per thread) takes 160 us against 177 with 65536 (one)
constexpr int VEC_BLOCKS = IMPEC_VEC_BLOCKS;

// Scalars kept in device memory so that solver loops need no host round trip
// except the convergence test.
enum Scal {
    S_RHO = 0, S_RHO_OLD = 1, S_DQ = 2, S_RR = 3, S_RSUM = 4, S_ZSUM = 5, S_ALPHA = 6, S_BETA = 7,
    S_OMEGA = 8, S_TS = 9, S_TT = 10,
    S_FLOOR = 11,     // with S_TT: || |b| + |A| |p| ||^2, the rounding scale of the true residual
    S_XPIN = 12, S_RX = 14, S_BB = 15, S_R0V = 19,
    S_MASS = 16, S_DIV = 17, S_NOFLOW = 18,          // written together by the diagnostics pass
    S_MAXVEL = 20, S_MINDT = 21, S_MINDT_DIFF = 22,  // CFL pass
    S_OUTFLUX = 24, S_TMO = 25, S_CMAX = 26, S_CMIN = 27,
    S_GUESS_ALPHA = 28,                              // step length of the learned increment
    S_GUESS_GAMMA = 29,                              // step length of the last step's change
    S_CNRES = 30,                                    // true ||b - A x||^2 of the dispersion solve
    // per-member control of an ensemble's pressure solve (device-side loop state)
    S_EST = 32,        // 0 done, 1 iterating, 2 residual recheck due
    S_EITS = 33,       // iterations of this solve
    S_ESINCE = 34,     // iterations since the last (re)start of the recursion
    S_ERST = 35,       // restarts
    S_EPAR = 36,       // which of the two direction arrays is the current one
    S_ETOL = 37,       // tol^2 ||b||^2
    S_EMAXIT = 38,     // iteration limit
    S_ESTATUS = 39,    // 0 ok, 1 limit reached, 2 non-finite residual
    S_DT = 40,         // time step of the ensemble step
    S_ZEROP = 41,      // start the next solve from zero
    S_G0 = 42,         // 42..46: partial products of the warm-start projection
    S_ERRPREV = 47,    // true residual norm^2 at the previous restart check (-1: none)
    S_EMARK = 48,      // recursive residual norm^2 when it last halved (in norm)
    S_ESTALL = 49,     // iterations since then
    S_COUNT = 50
};
enum { EST_DONE = 0, EST_ITER = 1, EST_CHECK = 2 };
// The FP64 rounding floor of the pressure solve's true residual: evaluating
// b - A p rounds every row at about u (|b| + |A| |p|), so a residual norm below
// IMPEC_FLOOR_C u || |b| + |A| |p| || cannot be told from zero and counts as
// converged whatever the tolerance (u = 2^-53).
#ifndef IMPEC_FLOOR_C
#define IMPEC_FLOOR_C 8.0
#endif
// restarts of the pressure solve's recursion from the true residual: at least
// 3, more while each one at least halves the residual norm
#ifndef IMPEC_MAX_RESTARTS
#define IMPEC_MAX_RESTARTS 12
#endif
// A recursion whose residual norm has not halved in IMPEC_STALL_ITS iterations
// has its true residual and rounding floor evaluated: converged at the floor,
// it stops; stalled within IMPEC_STALL_GATE times the floor, it is restarted
// from the true residual (a recursion continued at the floor drifts away from
// the true residual); stalled above that (a slow stretch of a poorly
// preconditioned system), it goes on.
#ifndef IMPEC_STALL_ITS
#define IMPEC_STALL_ITS 20
#endif
#ifndef IMPEC_STALL_GATE
#define IMPEC_STALL_GATE 4.0
#endif
// dispersion solve: the p-launch of an iteration is queued before the host
// reads the s-launch's residual while the last known residual norm^2 lies more
// than this factor above the target (one BiCGStab iteration reduces it by
// about 10^3 on the benchmark workload)
#ifndef IMPEC_CN_AHEAD
#define IMPEC_CN_AHEAD 1e6
#endif
// pressure solve: the next V-cycle's first leg is queued before the host reads
// the CG iteration's residual while the last known residual norm^2 lies more
// than this factor above the target (an iteration reduces it by about 10^2.5)
#ifndef IMPEC_CG_AHEAD
#define IMPEC_CG_AHEAD 1e6
#endif
// the dispersion solve's BiCGStab (about 7 iterations per solve at 4096^2): a recursion
// whose residual norm has not halved in this many iterations gets the true
// residual check and restart of a converged one
#ifndef IMPEC_CN_STALL_ITS
#define IMPEC_CN_STALL_ITS 50
#endif

// Geometry and constants of the nu-weighted pressure operator on one level.
// Constant factors are formed on the host with the reference's expressions
// (MGSolver.cpp:760-804), so device code multiplies by exactly the same
// doubles the reference computes.
struct OpGeom {
    int n1, n2;          // cells of the level
    int harmonic;        // average_procedure
    int diag;            // nu < 1: rotated stencil present
    double nu, one_m_nu; // nu, (1 - nu)
    double cx, cy;       // 1./(dx*dx), 1./(dy*dy)
    double mcy;          // -1./(dy*dy)
    double s2, cosa2, sina2, cost2, sint2, costsina, cosasint;
    double dxi2, deta2;  // dxi*dxi, deta*deta (divisors of NE/SW/NW/SE)
    double mcxi, ceta;   // -1./(dxi*dxi), 1./(deta*deta)
    double cnd, mcnd;    // 1./(deta*dxi), -1./(deta*dxi)
};

struct Level {
    int n1, n2, ld;      // cells, row pitch of ring-1 arrays
    double dx, dy;
    OpGeom g;
    double *w;           // inverse-mobility (harmonic) or mobility, ring 1 (level >= 1)
    double *x, *b, *r;   // ring-1 arrays (n1 + 2) x ld
    // FP32 copies for the mixed-precision V-cycle (allocated on first use):
    // w (level 0: of ctx->w; wf2 of ctx->w22 when K is anisotropic), x, b, r
    float *wf, *wf2, *xf, *bf, *rf;
};

struct ProfEntry {
    const char *name;
    long long launches;
    double ms;
};

}  // namespace impec

struct impec_ctx {
    impec_params P;
    int N1, N2, B;
    int lds, ldv;              // row pitch of scalar / velocity device arrays
    double dx, dy, dA;
    double alpha, theta, dxi, deta;
    double secamt2, cosa2, sina2, costsina, cosasint, cost2, sint2, cosa, sint, sina, cost;
    double mob1over4;
    double kappa;
    int device;
    hipStream_t stream;

    // fields (scalar layout: (N1+2B) rows; velocity: (N1+1+2B) rows)
    double *c, *p, *u, *v, *q;
    double *K11, *K22;         // nullptr when K == 1
    double *w, *w22;           // mobility quantity of the fine level (scalar layout)
    double *t1, *t2, *t3;      // transport temporaries (scalar layout)
    // the addresses of c and t1 are held elsewhere (an ensemble's launch
    // records): the transport copies its result into c instead of exchanging
    // the two buffers
    int fixed_buffers;
    double cn_ahead;           // dispersion lookahead factor (negative: off)
    double cg_ahead;           // pressure lookahead factor (negative: off)

    // pressure solve
    int nlev;
    impec::Level lev[impec::MAX_LEVELS];
    double *pr, *pz, *pd, *pq; // PCG vectors on the level-0 ring-1 layout
    double *fx, *fy;           // level-0 face mobilities of the 5-point operator (per solve)
    // the FP32-vector CG (IMPEC_PREC_MIXED32), level-0 ring-1 layout: residual,
    // correction since the last reliable update, direction pair, matvec, and
    // FP32 copies of fx / fy
    float *pr32, *pe32, *pd32, *pz32, *pq32, *fx32, *fy32;
    double *scal;              // device scalars (Scal)
    double *part;              // reduction partials
    double *part_mid;          // per-thread shares of a two-stage finish (FIN_THREADS x PART_VECS)
    double *h_scal;            // pinned host mirror of scal (mapped: kernels publish into it)
    double *d_hscal;           // its device address
    unsigned long long pub_seq;   // sequence number of the last publication (slot S_COUNT of h_scal)
    int symmetric;             // shift formulation usable
    int force_unfused;         // use the unfused multigrid kernels (testing)
    int pin_i, pin_j;
    double bnorm, bsum;        // ||-q||_2 and sum(-q), exact host sums
    double q_pin;              // q at the pinned cell
    int have_prev_p;

    // Crank-Nicolson
    double *cn_rhs, *cn_r, *cn_r0, *cn_p, *cn_v, *cn_r2, *cn_y, *cn_z;   // the solution forms in c; r ping-pongs with cn_r2
    float *cn_dg;                 // rounded inverse of the folded CN diagonal

    // sources
    double *tr_md;             // per-source mass change scratch
    int nsrc;
    int *src_idx;              // cell index i*N2+j, row-major order
    double *src_q;

    int prev_first_sweep;      // 0 X_SWEEP, 1 Y_SWEEP (HyperbolicSolver.cpp:84)
    impec_stats stats;
    char err[512];

    // learned initial guess of the pressure solve
    impec_cnn *net;            // not owned
    double *p_prev;            // pressure before the latest solve (scalar layout)
    double *p_old;             // pressure before the solve before that
    int n_solves;              // pressure solves so far

    int prof_on;
    impec::ProfEntry prof[32];
    int nprof;
    hipEvent_t ev0, ev1;
    int own_stream;            // 0 when the stream belongs to an ensemble
};

namespace impec {
struct EnsFieldTabs;           // batched-launch tables of k_field.hip
struct EnsPressTabs;           // of k_pressure.hip
struct EnsTransTabs;           // of k_transport.hip
struct EnsSnap;                // member inputs the tables were built for
}

// An ensemble: member contexts on one stream, advanced together by batched
// launches (one launch per pass over all members) with the pressure solve's
// loop state kept per member on the device.
struct impec_ensemble {
    int n;
    impec_ctx **m;             // member contexts (owned), sharing `stream`
    hipStream_t stream;
    int ready;                 // tables built for the current member state
    int use_graphs;            // replay captured graphs (else launch eagerly)
    int slots;                 // CG iterations per replay of the loop graph
    int have_graphs;
    long long graph_nodes;     // kernel nodes of the captured stages (checked at capture)
    hipGraphExec_t g_pre, g_loop, g_end, g_post;
    impec::EnsFieldTabs *field;
    impec::EnsPressTabs *pres;
    impec::EnsTransTabs *trans;
    double **d_S;              // device array of the members' scalar arrays
    double *d_out, *h_out;     // per-member gathered scalars (device / pinned host)
    int *d_active, *h_active;  // members still iterating
    impec::EnsSnap *snap;      // per member
    long long max_loop;        // replays of the loop graph before giving up
    char err[512];
};

namespace impec {

int set_error(impec_ctx *ctx, int code, const char *fmt, ...);
// the last error of a context-free entry point on this thread (noalloc.hip)
const char *noalloc_last_error();
int ctx_create(const impec_params *P, hipStream_t ext, impec_ctx **out);

#define IMPEC_HIP_CHECK(ctx, call)                                                            \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return ::impec::set_error((ctx), IMPEC_EHIP, "%s:%d %s: %s", __FILE__, __LINE__,  \
                                      #call, hipGetErrorString(e_));                          \
    } while (0)

// profiling brackets around launches of one kernel class
void prof_begin(impec_ctx *ctx);
void prof_end(impec_ctx *ctx, const char *name);

// Rows per chunk of a streaming launch over n rows in nstrip strips (one wave
// per strip and chunk), aimed at `target` waves.  R = ceil(n nstrip / target)
// leaves ceil(n / R) chunks, which can overshoot the target by up to a row of
// chunks; when the target is exactly the waves the GPU holds at once, the
// overshoot waits for free slots and runs as a short round of its own after
// the rest (the dispersion s-launch at 4096^2: 520 workgroups for 512 slots);
// when the target is a whole number of workgroups per CU, the overshoot gives
// some CUs one more workgroup than the others, and they finish last.
// `capped` chooses R so that nstrip ceil(n / R) <= target instead.
#ifndef IMPEC_WAVE_CAP
#define IMPEC_WAVE_CAP 1
#endif
inline long chunk_rows(long n, long nstrip, long target, bool capped) {
    if (capped && nstrip <= target) {
        const long per = target / nstrip;   // chunks per strip
        return (n + per - 1) / per;
    }
    return (n * nstrip + target - 1) / target;
}

inline size_t sidx(const impec_ctx *ctx, int i, int j) {
    return (size_t)(i + ctx->B) * ctx->lds + (j + ctx->B);
}

// ---- field kernels (k_field.hip) ----
int launch_fill_scalar(impec_ctx *ctx, double *a, int periodic);
int launch_fill_velocity(impec_ctx *ctx);
int launch_copy_frame(impec_ctx *ctx, const double *src, double *dst);   // ghost cells of a scalar array
int launch_mobility(impec_ctx *ctx);
int cnn_increment(impec_ctx *ctx, double *d, int ldd);
int launch_velocity(impec_ctx *ctx);
int launch_cfl(impec_ctx *ctx, double *dt);
int launch_tr(impec_ctx *ctx, double dt);
int launch_diagnostics(impec_ctx *ctx, int which);  // results in h_scal after sync
// h_scal <- scal once the stream's work so far has finished (a publishing
// kernel and a host poll, with a stream synchronisation as the fallback)
int sync_scalars(impec_ctx *ctx);
// A finishing kernel (one workgroup) publishes the scalars itself when given
// the mapped mirror and the next sequence number, in place of the separate
// publishing launch: after its own stores the workgroup copies scal into the
// mirror and releases the number.  pub_next reserves the number; wait_scalars
// then waits for it as sync_scalars does.
struct PubArgs {
    double *h;                 // the mapped mirror, or null: no publication
    unsigned long long seq;
};
PubArgs pub_next(impec_ctx *ctx);
int wait_scalars(impec_ctx *ctx);
__device__ __forceinline__ void publish_scalars(const double *S, const PubArgs &p) {
    if (!p.h) return;   // a kernel argument: the whole workgroup returns
    __syncthreads();    // the workgroup's stores to S are complete and visible
    for (int k = threadIdx.x; k < S_COUNT; k += blockDim.x) p.h[k] = S[k];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store((unsigned long long *)(p.h + S_COUNT), p.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
int launch_minmax(impec_ctx *ctx, double *cmax, double *cmin);

// ---- pressure (k_pressure.hip) ----
int pressure_setup_levels(impec_ctx *ctx);
void pressure_free_levels(impec_ctx *ctx);
int pressure_solve(impec_ctx *ctx, int *its, double *relres);
int pressure_operator_dump(impec_ctx *ctx, double *A9_host);

// ---- transport (k_transport.hip) ----
int advect_x(impec_ctx *ctx, const double *cin, double *cout, double dt);
int advect_y(impec_ctx *ctx, const double *cin, double *cout, double dt);
int advect_split(impec_ctx *ctx, double dt);
int intercell_flux(impec_ctx *ctx, int dir, double dt, double *F);
int copy_scalar(impec_ctx *ctx, const double *src, double *dst);

// ---- dispersion (k_cn.hip) ----
int cn_solve(impec_ctx *ctx, double dt, int *its);
int cn_operator_dump(impec_ctx *ctx, double dt, double *A9_host, double *rhs_host);

// ---- reductions (k_field.hip) ----
// sum of NV per-block partial vectors into scal[dst .. dst+NV)
int finish_sum(impec_ctx *ctx, int nblocks, int nv, int dst, PubArgs pub = PubArgs{nullptr, 0});

// ---- ensembles: batched launches (k_field.hip, k_pressure.hip, k_transport.hip) ----
int ens_set_error(impec_ensemble *e, int code, const char *fmt, ...);
int red_blocks_of(long long n);   // workgroups of a grid-stride reduction over n cells
// one reduction finish per member (workgroup m sums member m's partials)
struct FinishRec {
    const double *part;
    double *S;
    int nblocks, nv, dst, op;
};
hipError_t ens_finish(hipStream_t s, const Batch<FinishRec> &b, int need);
int ens_field_build(impec_ensemble *e);
void ens_field_free(impec_ensemble *e);
int ens_fill_c(impec_ensemble *e);
int ens_mobility(impec_ensemble *e);
int ens_velocity_dt(impec_ensemble *e);   // fill p, velocities, fill v, CFL time step into S_DT
int ens_tr(impec_ensemble *e);            // trapezoidal source step over S_DT / 2
int ens_press_build(impec_ensemble *e);
void ens_press_free(impec_ensemble *e);
int ens_solve_begin(impec_ensemble *e);   // hierarchy, faces, guess, first residual
int ens_cg_slot(impec_ensemble *e);       // one CG iteration (members still iterating)
int ens_solve_end(impec_ensemble *e);     // shift to the pinned value, true residual
int ens_vcycle_time(impec_ensemble *e, int reps, float *ms);   // level-0..coarsest legs, all members
int ens_trans_build(impec_ensemble *e);
void ens_trans_free(impec_ensemble *e);
int ens_advect_parallel(impec_ensemble *e);   // the four sweeps and the average

}  // namespace impec
