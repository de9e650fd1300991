/*
 * impec_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64, single-threaded restatement of the Impec2D IMPEC step
 * (pw359/Impec2D, the src directory).  It is the parity oracle for the HIP library in
 * impec2d_amd/ and the CPU baseline timed by bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links, imports or calls it.
 *
 * Every function cites the reference lines it restates.  Floating-point
 * expressions keep the reference's operand order (C/C++ left-to-right
 * evaluation), and the file is compiled with -ffp-contract=off, so that the
 * per-cell arithmetic is the reference's arithmetic bit for bit.
 *
 * Deviation (documented): the reference solves its linear systems with HYPRE
 * 2.8.0b StructPCG preconditioned by one StructSMG V(1,1) cycle
 * (MGSolver.cpp:594-629).  HYPRE is not available; this file restates HYPRE's
 * PCG with its two-norm stopping rule (||r||_2 <= tol ||b||_2) and replaces
 * the SMG preconditioner with a geometric multigrid V-cycle written here.
 * The linear systems themselves are assembled exactly as the reference does.
 * The non-symmetric Crank-Nicolson system is solved with BiCGStab to the same
 * two-norm criterion.  Solutions therefore agree with the reference's only to
 * the solver tolerance ("parity unpinned" below it).
 *
 * Extensions (off by default, flagged in or_params): max_steps, per-cell K
 * arrays, molecular dispersion d_m, absolute dispersivities.
 *
 * Threads: the per-cell loops that write disjoint outputs run under OpenMP
 * (or_set_threads), so every result is the same bit for bit whatever the
 * thread count.  Sums over vectors of 65536 or more entries add fixed chunks
 * of 16384 in order; shorter ones, and every other reduction, stay serial.
 *
 * Array layouts are the reference's (GenSimDefs.h:22-31, Impec2DSolver.cpp:151-182):
 *   scalar field  (N1+2b) x (N2+2b) doubles, row-major, j contiguous, base -b
 *   velocity      (N1+1+2b) x (N2+1+2b) x 2 doubles, component last, base -b
 *   q             N1*N2 doubles, index i*N2+j
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Planted mistakes (tests only).  or_set_mutation(m) makes one restated      */
/* formula deliberately wrong, so that the tests that pin this oracle can     */
/* show they fail on it; 0 (the default) is the faithful restatement.         */
/* ------------------------------------------------------------------------- */
enum {
    MUT_NONE = 0,
    MUT_CN_SWAP_ALPHA = 1,      /* alpha_l and alpha_t exchanged in the tensors   */
    MUT_CN_DROP_PHI = 2,        /* porosity left out of the tensors               */
    MUT_CN_AVG_F = 3,           /* x-face four-point average: v(i+1,j+1) for v(i-1,j+1) */
    MUT_CN_AVG_G = 4,           /* y-face four-point average: u(i-1,j-1) for u(i+1,j-1) */
    MUT_CN_DM_CROSS = 5,        /* d_m added to d12 / d21 as well                  */
    MUT_CN_RHS_CROSS = 6,       /* right-hand side: sign of the d12 cross term     */
    MUT_SPLIT_XFIRST = 7,       /* x_sweep_first runs the y sweep first            */
    MUT_SPLIT_YFIRST = 8,       /* y_sweep_first runs the x sweep first            */
    MUT_SPLIT_STRANG = 9,       /* strang: full x steps around the y step          */
    MUT_SPLIT_ALT = 10,         /* alternating: the first sweep never changes      */
    MUT_SPLIT_PARALLEL = 11,    /* parallel: the Y-then-X copy left out (no average) */
    MUT_COUNT
};
static int or_mutation = MUT_NONE;
int or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int or_set_mutation(int m)
{
    if (m < 0 || m >= MUT_COUNT) return -1;
    or_mutation = m;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* parameters (SimParams.h:29-164 field meanings)                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    int N1, N2, b;            /* x_cells, y_cells, num_ghostcells            */
    int stype;                /* 0 parallel, 1 diagonal (S_Type)             */
    int avgproc;              /* 0 harmonic, 1 mean_concentration            */
    int dim_split;            /* 0 alternating 1 x_first 2 y_first 3 strang 4 parallel */
    int hyper;                /* 0 muscl, 1 upwind (first order)            */
    int diffusion;            /* hyperbolic.diffusion                        */
    int para_cn;              /* 1 crank-nicolson, 0 explicit                */
    int is_test;              /* periodic BC + advection-only CFL            */
    int testno;               /* elliptic test number (pin kappa)            */
    int alpha_abs;            /* extension: alpha given in absolute units    */
    int precond;              /* 0 multigrid, 1 jacobi (oracle solver only)  */
    int max_iter;
    double x_lb, x_ub, y_lb, y_ub;
    double nu, mu_res, mu_inj;
    double porosity, cfl, inj_fluid_conc;
    double alpha_l, alpha_t, dt_diff_limit, d_m;
    double eps, para_eps;
    double I, well_rc;
    double tstart, tfinal;
} or_params;

typedef struct {
    or_params P;
    int N1, N2, b, ld, vld;
    double dx, dy, dA;
    double alpha, theta, dxi, deta;
    double secamt, secamt2, cosa2, sina2, costsina, cosasint, cost2, sint2;
    double cosa, sint, sina, cost;
    double mob1over4;
    const double *K11, *K22;  /* ghosted scalar layout, or NULL for K == 1  */
    /* state */
    double total_mass_outflux;
    int prev_first_sweep;     /* 0 X_SWEEP, 1 Y_SWEEP                        */
    int last_its, last_para_its;
    double last_relres, last_para_relres;
    int last_vcycles;
    double kappa;             /* pin coefficient (MGSolver.cpp:528-548)      */
} or_ctx;

#define S_(ctx, i, j) (((i) + (ctx)->b) * (ctx)->ld + ((j) + (ctx)->b))
#define V_(ctx, i, j, k) ((((i) + (ctx)->b) * (ctx)->vld + ((j) + (ctx)->b)) * 2 + (k))

size_t or_ctx_size(void) { return sizeof(or_ctx); }
size_t or_params_size(void) { return sizeof(or_params); }

/* EllipticSolver ctor trig setup (EllipticSolver.cpp:59-100) and
 * SimParams derived quantities (SimParams.cpp:445-452). */
int or_init(or_ctx *ctx, const or_params *P, const double *K11, const double *K22)
{
    memset(ctx, 0, sizeof(*ctx));
    ctx->P = *P;
    ctx->N1 = P->N1; ctx->N2 = P->N2; ctx->b = P->b;
    if (P->N1 < 2 || P->N2 < 2 || P->b < 2) return -1;
    ctx->ld = P->N2 + 2 * P->b;
    ctx->vld = P->N2 + 1 + 2 * P->b;
    ctx->dx = (P->x_ub - P->x_lb) / P->N1;
    ctx->dy = (P->y_ub - P->y_lb) / P->N2;
    ctx->dA = ctx->dx * ctx->dy;
    ctx->alpha = atan(ctx->dx / ctx->dy);
    ctx->theta = atan(ctx->dy / ctx->dx);
    ctx->dxi = ctx->dx / cos(ctx->theta);
    ctx->deta = ctx->dy / cos(ctx->alpha);
    ctx->secamt = 1. / cos(ctx->alpha - ctx->theta);
    ctx->secamt2 = ctx->secamt * ctx->secamt;
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
    ctx->mob1over4 = pow(P->mu_res / P->mu_inj, 0.25);   /* EllipticSolver.cpp:126 */
    ctx->K11 = K11;
    ctx->K22 = K22 ? K22 : K11;
    ctx->prev_first_sweep = 1;                            /* HyperbolicSolver.cpp:84 */
    ctx->kappa = 1.;
    return 0;
}

void or_set_kappa(or_ctx *ctx, double kappa) { ctx->kappa = kappa; }
double or_total_mass_outflux(const or_ctx *ctx) { return ctx->total_mass_outflux; }
int or_last_its(const or_ctx *ctx) { return ctx->last_its; }
int or_last_para_its(const or_ctx *ctx) { return ctx->last_para_its; }
double or_last_relres(const or_ctx *ctx) { return ctx->last_relres; }
int or_last_vcycles(const or_ctx *ctx) { return ctx->last_vcycles; }

/* ------------------------------------------------------------------------- */
/* ghost fills (Util.cpp)                                                     */
/* ------------------------------------------------------------------------- */

/* FillGhostCellsNeumannScalar (Util.cpp:283-324); arr has extents
 * (N1+2n) x (N2+2n), base -n. */
void or_fill_neumann_scalar(const or_ctx *ctx, double *arr)
{
    const int N1 = ctx->N1, N2 = ctx->N2, n = ctx->b;
    const int xN = N1 - 1, yN = N2 - 1;
    for (int i = 0; i <= xN; i++)
        for (int j = 0; j < n; j++) {
            arr[S_(ctx, i, yN + j + 1)] = arr[S_(ctx, i, yN - j)];
            arr[S_(ctx, i, -(j + 1))] = arr[S_(ctx, i, j)];
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= yN; j++) {
            arr[S_(ctx, -(i + 1), j)] = arr[S_(ctx, i, j)];
            arr[S_(ctx, xN + i + 1, j)] = arr[S_(ctx, xN - i, j)];
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            arr[S_(ctx, -(i + 1), -(j + 1))] = arr[S_(ctx, i, j)];
            arr[S_(ctx, xN + i + 1, -(j + 1))] = arr[S_(ctx, xN - i, j)];
            arr[S_(ctx, -(i + 1), yN + j + 1)] = arr[S_(ctx, i, yN - j)];
            arr[S_(ctx, xN + i + 1, yN + j + 1)] = arr[S_(ctx, xN - i, yN - j)];
        }
}

/* FillGhostCellsPeriodicScalar (Util.cpp:215-256). */
void or_fill_periodic_scalar(const or_ctx *ctx, double *arr)
{
    const int N1 = ctx->N1, N2 = ctx->N2, n = ctx->b;
    const int xN = N1 - 1, yN = N2 - 1;
    for (int i = 0; i <= xN; i++)
        for (int j = 0; j < n; j++) {
            arr[S_(ctx, i, yN + j + 1)] = arr[S_(ctx, i, j)];
            arr[S_(ctx, i, -(j + 1))] = arr[S_(ctx, i, yN - j)];
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= yN; j++) {
            arr[S_(ctx, -(i + 1), j)] = arr[S_(ctx, xN - i, j)];
            arr[S_(ctx, xN + i + 1, j)] = arr[S_(ctx, i, j)];
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            arr[S_(ctx, -(i + 1), -(j + 1))] = arr[S_(ctx, xN - i, yN - j)];
            arr[S_(ctx, xN + i + 1, -(j + 1))] = arr[S_(ctx, i, yN - j)];
            arr[S_(ctx, -(i + 1), yN + j + 1)] = arr[S_(ctx, xN - i, j)];
            arr[S_(ctx, xN + i + 1, yN + j + 1)] = arr[S_(ctx, i, j)];
        }
}

/* HyperbolicSolver::FillBoundaries (HyperbolicSolver.cpp:56-61). */
void or_fill_boundaries(const or_ctx *ctx, double *c)
{
    if (ctx->P.is_test) or_fill_periodic_scalar(ctx, c);
    else or_fill_neumann_scalar(ctx, c);
}

/* FillGhostCellsNeumannVelocity (Util.cpp:331-438); v has extents
 * (N1+1+2n) x (N2+1+2n) x 2, base -n, so xN = N1 and yN = N2. */
void or_fill_neumann_velocity(const or_ctx *ctx, double *v)
{
    const int n = ctx->b, xN = ctx->N1, yN = ctx->N2;
#define VV(i, j, k) v[V_(ctx, i, j, k)]
    for (int i = 0; i <= xN; i++)
        for (int j = 0; j < n; j++) {
            VV(i, -(j + 1), 0) = VV(i, j, 0);
            VV(i, yN + j, 0) = VV(i, yN - 1 - j, 0);
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            VV(-(i + 1), yN + j, 0) = -VV(i + 1, yN + j, 0);
            VV(-(i + 1), -(j + 1), 0) = -VV(i + 1, -(j + 1), 0);
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            VV(xN + i + 1, yN + j, 0) = -VV(xN - 1 - i, yN + j, 0);
            VV(xN + i + 1, -(j + 1), 0) = -VV(xN - i - 1, -(j + 1), 0);
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < yN; j++) {
            VV(-(i + 1), j, 0) = -VV(i + 1, j, 0);
            VV(xN + i + 1, j, 0) = -VV(xN - i - 1, j, 0);
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j <= yN; j++) {
            VV(-(i + 1), j, 1) = VV(i, j, 1);
            VV(xN + i, j, 1) = VV(xN - i - 1, j, 1);
        }
    for (int i = 0; i < xN; i++)
        for (int j = 0; j < n; j++) {
            VV(i, -(j + 1), 1) = -VV(i, j + 1, 1);
            VV(i, yN + j + 1, 1) = -VV(i, yN - 1 - j, 1);
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            VV(-(i + 1), yN + 1 + j, 1) = VV(i, yN + 1 + j, 1);
            VV(-(i + 1), -(j + 1), 1) = VV(i, -(j + 1), 1);
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            VV(xN + i, yN + 1 + j, 1) = VV(xN - 1 - i, yN + j + 1, 1);
            VV(xN + i, -(j + 1), 1) = VV(xN - 1 - i, -(j + 1), 1);
        }
#undef VV
}

/* ------------------------------------------------------------------------- */
/* mobility (EllipticSolver.cpp:478-524, 692-754)                            */
/* ------------------------------------------------------------------------- */

/* EllipticSolver::mu (EllipticSolver.cpp:692-695). */
double or_mu(const or_ctx *ctx, double c)
{
    double x = (1. - c + ctx->mob1over4 * c);
    return 1. / (x * x * x * x) * ctx->P.mu_res;
}

static inline double K11_at(const or_ctx *ctx, int i, int j)
{
    return ctx->K11 ? ctx->K11[S_(ctx, i, j)] : 1.;
}
static inline double K22_at(const or_ctx *ctx, int i, int j)
{
    return ctx->K22 ? ctx->K22[S_(ctx, i, j)] : 1.;
}

/* KoverMu_11_im12j (EllipticSolver.cpp:478-500): face between cells
 * (i-1,j) and (i,j). */
static double kom11_face(const or_ctx *ctx, const double *c, int i, int j)
{
    double K11_im1j = K11_at(ctx, i - 1, j);
    double K11_ij = K11_at(ctx, i, j);
    double mu_ij = or_mu(ctx, c[S_(ctx, i, j)]);
    double mu_im1j = or_mu(ctx, c[S_(ctx, i - 1, j)]);
    if (ctx->P.avgproc == 0)
        return 2. / (mu_im1j / K11_im1j + mu_ij / K11_ij);
    return 1. / 2 * (K11_im1j / mu_im1j + K11_ij / mu_ij);
}

/* KoverMu_22_ijm12 (EllipticSolver.cpp:502-524). */
static double kom22_face(const or_ctx *ctx, const double *c, int i, int j)
{
    double K22_ijm1 = K22_at(ctx, i, j - 1);
    double K22_ij = K22_at(ctx, i, j);
    double mu_ij = or_mu(ctx, c[S_(ctx, i, j)]);
    double mu_ijm1 = or_mu(ctx, c[S_(ctx, i, j - 1)]);
    if (ctx->P.avgproc == 0)
        return 2. / (mu_ijm1 / K22_ijm1 + mu_ij / K22_ij);
    return 1. / 2 * (K22_ijm1 / mu_ijm1 + K22_ij / mu_ij);
}

/* KoverMu_11_ip12jp12 / KoverMu_22_ip12jp12 (EllipticSolver.cpp:697-754):
 * vertex whose lower-left cell is (i,j); every caller in the reference passes
 * the four cells in the order (i,j), (i+1,j), (i,j+1), (i+1,j+1). */
static double kom_vertex(const or_ctx *ctx, const double *c, int i, int j, int comp)
{
    double Kip1j = comp == 0 ? K11_at(ctx, i + 1, j) : K22_at(ctx, i + 1, j);
    double Kij = comp == 0 ? K11_at(ctx, i, j) : K22_at(ctx, i, j);
    double Kijp1 = comp == 0 ? K11_at(ctx, i, j + 1) : K22_at(ctx, i, j + 1);
    double Kip1jp1 = comp == 0 ? K11_at(ctx, i + 1, j + 1) : K22_at(ctx, i + 1, j + 1);
    double mu_ip1j = or_mu(ctx, c[S_(ctx, i + 1, j)]);
    double mu_ij = or_mu(ctx, c[S_(ctx, i, j)]);
    double mu_ijp1 = or_mu(ctx, c[S_(ctx, i, j + 1)]);
    double mu_ip1jp1 = or_mu(ctx, c[S_(ctx, i + 1, j + 1)]);
    if (ctx->P.avgproc == 0)
        return 4. / (mu_ip1j / Kip1j + mu_ij / Kij + mu_ijp1 / Kijp1 + mu_ip1jp1 / Kip1jp1);
    return 1. / 4 * (Kip1j / mu_ip1j + Kij / mu_ij + Kijp1 / mu_ijp1 + Kip1jp1 / mu_ip1jp1);
}

/* ------------------------------------------------------------------------- */
/* 9-point pressure operator (MGSolver.cpp:268-570, 710-841)                  */
/* stencil order {C, W, SW, S, SE, E, NE, N, NW} (MGSolver.cpp:164-166)       */
/* ------------------------------------------------------------------------- */
enum { ST_C, ST_W, ST_SW, ST_S, ST_SE, ST_E, ST_NE, ST_N, ST_NW };
static const int ST_OFF[9][2] = {{0, 0}, {-1, 0}, {-1, -1}, {0, -1}, {1, -1},
                                 {1, 0}, {1, 1}, {0, 1}, {-1, 1}};

/* GetParallelCoefficientsForCell (MGSolver.cpp:760-773). */
static void par_coeffs(const or_ctx *ctx, const double *c, int i, int j,
                       double *W, double *S, double *C, double *N, double *E)
{
    const double dx = ctx->dx, dy = ctx->dy;
    double K11_im12j = kom11_face(ctx, c, i, j);
    double K11_ip12j = kom11_face(ctx, c, i + 1, j);
    double K22_ijm12 = kom22_face(ctx, c, i, j);
    double K22_ijp12 = kom22_face(ctx, c, i, j + 1);
    *W = 1. / (dx * dx) * K11_im12j;
    *E = 1. / (dx * dx) * K11_ip12j;
    *S = 1. / (dy * dy) * K22_ijm12;
    *N = 1. / (dy * dy) * K22_ijp12;
    *C = -1. / (dy * dy) * (K22_ijm12 + K22_ijp12) - 1. / (dx * dx) * (K11_im12j + K11_ip12j);
}

/* GetDiagonalCoefficientsForCell + GetTransformedMobilityTensor
 * (MGSolver.cpp:777-841). */
static void diag_coeffs(const or_ctx *ctx, const double *c, int i, int j, double d[9])
{
    const double s2 = ctx->secamt2;
    double K11_pp = kom_vertex(ctx, c, i, j, 0);
    double K11_mm = kom_vertex(ctx, c, i - 1, j - 1, 0);
    double K11_mp = kom_vertex(ctx, c, i - 1, j, 0);
    double K11_pm = kom_vertex(ctx, c, i, j - 1, 0);
    double K22_pp = kom_vertex(ctx, c, i, j, 1);
    double K22_mm = kom_vertex(ctx, c, i - 1, j - 1, 1);
    double K22_mp = kom_vertex(ctx, c, i - 1, j, 1);
    double K22_pm = kom_vertex(ctx, c, i, j - 1, 1);

    double M11_ip12j = s2 * (K11_pp * ctx->cosa2 + K22_pp * ctx->sina2);
    double M11_im12j = s2 * (K11_mm * ctx->cosa2 + K22_mm * ctx->sina2);
    double M22_ijp12 = s2 * (K22_mp * ctx->cost2 + K11_mp * ctx->sint2);
    double M22_ijm12 = s2 * (K22_pm * ctx->cost2 + K11_pm * ctx->sint2);
    double M12_ip12j = s2 * (K22_pp * ctx->costsina - K11_pp * ctx->cosasint);
    double M21_ijp12 = s2 * (K22_mp * ctx->costsina - K11_mp * ctx->cosasint);
    double M12_im12j = s2 * (K22_mm * ctx->costsina - K11_mm * ctx->cosasint);
    double M21_ijm12 = s2 * (K22_pm * ctx->costsina - K11_pm * ctx->cosasint);

    const double dxi = ctx->dxi, deta = ctx->deta;
    d[ST_NE] = M11_ip12j / (dxi * dxi);
    d[ST_SW] = M11_im12j / (dxi * dxi);
    d[ST_NW] = M22_ijp12 / (deta * deta);
    d[ST_SE] = M22_ijm12 / (deta * deta);
    d[ST_C] = -1. / (dxi * dxi) * (M11_ip12j + M11_im12j)
              - 1. / (deta * deta) * (M22_ijp12 + M22_ijm12);
    d[ST_N] = 1. / (deta * dxi) * (M12_ip12j + M21_ijp12);
    d[ST_W] = -1. / (deta * dxi) * (M21_ijp12 + M12_im12j);
    d[ST_E] = -1. / (deta * dxi) * (M12_ip12j + M21_ijm12);
    d[ST_S] = 1. / (deta * dxi) * (M12_im12j + M21_ijm12);
}

/* GetWeightedMatrixCoefficientsForCell (MGSolver.cpp:710-757). */
static void weighted_coeffs(const or_ctx *ctx, const double *c, int i, int j, double a[9])
{
    double p[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double d[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    par_coeffs(ctx, c, i, j, &p[ST_W], &p[ST_S], &p[ST_C], &p[ST_N], &p[ST_E]);
    diag_coeffs(ctx, c, i, j, d);
    const double nu = ctx->P.nu;
    for (int k = 0; k < 9; k++) a[k] = nu * p[k] + (1 - nu) * d[k];
}

/* Neumann folding of the nine entries of boundary cells, applied in the
 * reference's order: y=lb, y=ub, x=lb, x=ub edges, then the four corners
 * (MGSolver.cpp:305-515; the same pattern for the parabolic matrix at
 * MGSolver.cpp:1033-1296).  `a` holds the unfolded coefficients. */
static void fold_boundary(int i, int j, int N1, int N2, const double a[9], double out[9])
{
    const double C = a[ST_C], W = a[ST_W], SW = a[ST_SW], S = a[ST_S], SE = a[ST_SE],
                 E = a[ST_E], NE = a[ST_NE], N = a[ST_N], NW = a[ST_NW];
    for (int k = 0; k < 9; k++) out[k] = a[k];
    /* edges: later edges overwrite earlier ones in corner cells, exactly as
     * the reference's successive SetValues calls do */
    if (j == 0) {
        out[0] = C + S; out[1] = W + SW; out[2] = 0; out[3] = 0; out[4] = 0;
        out[5] = E + SE; out[6] = NE; out[7] = N; out[8] = NW;
    }
    if (j == N2 - 1) {
        out[0] = C + N; out[1] = W + NW; out[2] = SW; out[3] = S; out[4] = SE;
        out[5] = E + NE; out[6] = 0; out[7] = 0; out[8] = 0;
    }
    if (i == 0) {
        out[0] = C + W; out[1] = 0; out[2] = 0; out[3] = S + SW; out[4] = SE;
        out[5] = E; out[6] = NE; out[7] = N + NW; out[8] = 0;
    }
    if (i == N1 - 1) {
        out[0] = C + E; out[1] = W; out[2] = SW; out[3] = S + SE; out[4] = 0;
        out[5] = 0; out[6] = 0; out[7] = N + NE; out[8] = NW;
    }
    if (i == 0 && j == 0) {
        out[0] = C + W + S + SW; out[1] = 0; out[2] = 0; out[3] = 0; out[4] = 0;
        out[5] = E + SE; out[6] = NE; out[7] = N + NW; out[8] = 0;
    }
    if (i == 0 && j == N2 - 1) {
        out[0] = C + W + N + NW; out[1] = 0; out[2] = 0; out[3] = S + SW; out[4] = SE;
        out[5] = E + NE; out[6] = 0; out[7] = 0; out[8] = 0;
    }
    if (i == N1 - 1 && j == N2 - 1) {
        out[0] = C + N + NE + E; out[1] = W + NW; out[2] = SW; out[3] = S + SE; out[4] = 0;
        out[5] = 0; out[6] = 0; out[7] = 0; out[8] = 0;
    }
    if (i == N1 - 1 && j == 0) {
        out[0] = C + S + E + SE; out[1] = W + SW; out[2] = 0; out[3] = 0; out[4] = 0;
        out[5] = 0; out[6] = 0; out[7] = N + NE; out[8] = NW;
    }
}

/* SetMatrixCoefficients_Elliptic (MGSolver.cpp:268-570).  A holds 9 arrays
 * of N1*N2 doubles (stencil order above), c is ghost-filled.  The pinned row
 * (1,1) gets kappa/dA on the diagonal (MGSolver.cpp:517-567). */
void or_assemble_elliptic(const or_ctx *ctx, const double *c, double *A)
{
    const int N1 = ctx->N1, N2 = ctx->N2, M = N1 * N2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) {
            double a[9], f[9];
            weighted_coeffs(ctx, c, i, j, a);
            fold_boundary(i, j, N1, N2, a, f);
            for (int k = 0; k < 9; k++) A[k * M + i * N2 + j] = f[k];
        }
    const int pi_ = 1, pj_ = 1;
    for (int k = 0; k < 9; k++) A[k * M + pi_ * N2 + pj_] = 0;
    A[ST_C * M + pi_ * N2 + pj_] = ctx->kappa / ctx->dA;
}

/* y = A x for a 9-array stencil matrix; x, y are flat N1*N2 arrays.  Entries
 * pointing outside the grid are zero after folding and are skipped. */
void or_matvec9(int N1, int N2, const double *A, const double *x, double *y)
{
    const int M = N1 * N2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) {
            double s = 0;
            for (int k = 0; k < 9; k++) {
                int ii = i + ST_OFF[k][0], jj = j + ST_OFF[k][1];
                if (ii < 0 || jj < 0 || ii >= N1 || jj >= N2) continue;
                s += A[k * M + i * N2 + j] * x[ii * N2 + jj];
            }
            y[i * N2 + j] = s;
        }
}

/* ------------------------------------------------------------------------- */
/* geometric multigrid preconditioner (replaces HYPRE StructSMG)              */
/* Neumann operators on every level, rediscretised from a coarsened          */
/* per-cell mobility; V(1,1) with symmetric colour-ordered Gauss-Seidel,     */
/* full-weighting restriction (transpose of bilinear prolongation / 4).      */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n1, n2;
    double dx, dy;
    double *A;      /* 9 * n1 * n2 */
    double *x, *b, *r;
} or_level;

typedef struct {
    int nlev;
    or_level L[32];
    int ncolor;
} or_mg;

/* per-cell stored mobility quantity: mu/K (harmonic) or K/mu (mean conc.) */
static double wcell(const or_ctx *ctx, double c, double K)
{
    double mu = or_mu(ctx, c);
    return ctx->P.avgproc == 0 ? mu / K : K / mu;
}

/* coefficients of a rediscretised level from a ghosted per-cell w array
 * (extents (n1+2) x (n2+2), base -1); same formulas as the fine operator */
static void level_assemble(const or_ctx *ctx, or_level *L, const double *w, int ldw)
{
#define W_(i, j) w[((i) + 1) * ldw + ((j) + 1)]
    const int n1 = L->n1, n2 = L->n2, M = n1 * n2;
    const double dx = L->dx, dy = L->dy;
    const double dxi = dx / cos(ctx->theta), deta = dy / cos(ctx->alpha);
    const int harm = ctx->P.avgproc == 0;
    const double s2 = ctx->secamt2, nu = ctx->P.nu;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) {
            double fx[2], fy[2], vt[4];
            for (int s = 0; s < 2; s++) {
                double wl = W_(i - 1 + s, j), wr = W_(i + s, j);
                fx[s] = harm ? 2. / (wl + wr) : 1. / 2 * (wl + wr);
                wl = W_(i, j - 1 + s); wr = W_(i, j + s);
                fy[s] = harm ? 2. / (wl + wr) : 1. / 2 * (wl + wr);
            }
            /* vertices: 0 = (i,j) pp, 1 = (i-1,j-1) mm, 2 = (i-1,j) mp, 3 = (i,j-1) pm */
            const int vo[4][2] = {{0, 0}, {-1, -1}, {-1, 0}, {0, -1}};
            for (int s = 0; s < 4; s++) {
                int a = i + vo[s][0], bb = j + vo[s][1];
                double sum = W_(a + 1, bb) + W_(a, bb) + W_(a, bb + 1) + W_(a + 1, bb + 1);
                vt[s] = harm ? 4. / sum : 1. / 4 * sum;
            }
            double p[9] = {0}, d[9];
            p[ST_W] = 1. / (dx * dx) * fx[0];
            p[ST_E] = 1. / (dx * dx) * fx[1];
            p[ST_S] = 1. / (dy * dy) * fy[0];
            p[ST_N] = 1. / (dy * dy) * fy[1];
            p[ST_C] = -1. / (dy * dy) * (fy[0] + fy[1]) - 1. / (dx * dx) * (fx[0] + fx[1]);
            double M11p = s2 * (vt[0] * ctx->cosa2 + vt[0] * ctx->sina2);
            double M11m = s2 * (vt[1] * ctx->cosa2 + vt[1] * ctx->sina2);
            double M22p = s2 * (vt[2] * ctx->cost2 + vt[2] * ctx->sint2);
            double M22m = s2 * (vt[3] * ctx->cost2 + vt[3] * ctx->sint2);
            double M12pp = s2 * (vt[0] * ctx->costsina - vt[0] * ctx->cosasint);
            double M21mp = s2 * (vt[2] * ctx->costsina - vt[2] * ctx->cosasint);
            double M12mm = s2 * (vt[1] * ctx->costsina - vt[1] * ctx->cosasint);
            double M21pm = s2 * (vt[3] * ctx->costsina - vt[3] * ctx->cosasint);
            d[ST_NE] = M11p / (dxi * dxi);
            d[ST_SW] = M11m / (dxi * dxi);
            d[ST_NW] = M22p / (deta * deta);
            d[ST_SE] = M22m / (deta * deta);
            d[ST_C] = -1. / (dxi * dxi) * (M11p + M11m) - 1. / (deta * deta) * (M22p + M22m);
            d[ST_N] = 1. / (deta * dxi) * (M12pp + M21mp);
            d[ST_W] = -1. / (deta * dxi) * (M21mp + M12mm);
            d[ST_E] = -1. / (deta * dxi) * (M12pp + M21pm);
            d[ST_S] = 1. / (deta * dxi) * (M12mm + M21pm);
            double a[9], f[9];
            for (int k = 0; k < 9; k++) a[k] = nu * p[k] + (1 - nu) * d[k];
            fold_boundary(i, j, n1, n2, a, f);
            for (int k = 0; k < 9; k++) L->A[k * M + i * n2 + j] = f[k];
        }
#undef W_
}

static void mg_free(or_mg *mg)
{
    for (int l = 0; l < mg->nlev; l++) {
        free(mg->L[l].A); free(mg->L[l].x); free(mg->L[l].b); free(mg->L[l].r);
    }
    mg->nlev = 0;
}

/* Builds the hierarchy.  Level 0 uses the reference's assembled (unpinned)
 * matrix; coarser levels are rediscretised from w averaged over 2x2 blocks. */
static int mg_setup(const or_ctx *ctx, or_mg *mg, const double *A0, const double *c)
{
    memset(mg, 0, sizeof(*mg));
    mg->ncolor = ctx->P.nu == 1. ? 2 : 4;
    int n1 = ctx->N1, n2 = ctx->N2, l = 0;
    /* level-0 w with a ghost ring of 1, from the ghost-filled c */
    int ldw = n2 + 2;
    double *w = malloc(sizeof(double) * (n1 + 2) * ldw);
    for (int i = -1; i <= n1; i++)
        for (int j = -1; j <= n2; j++)
            w[(i + 1) * ldw + (j + 1)] = wcell(ctx, c[S_(ctx, i, j)], K11_at(ctx, i, j));
    double dx = ctx->dx, dy = ctx->dy;
    for (;;) {
        or_level *L = &mg->L[l];
        L->n1 = n1; L->n2 = n2; L->dx = dx; L->dy = dy;
        size_t M = (size_t)n1 * n2;
        L->A = malloc(sizeof(double) * 9 * M);
        L->x = calloc(M, sizeof(double));
        L->b = calloc(M, sizeof(double));
        L->r = calloc(M, sizeof(double));
        if (l == 0) memcpy(L->A, A0, sizeof(double) * 9 * M);
        else level_assemble(ctx, L, w, ldw);
        mg->nlev = ++l;
        if ((n1 <= 2 && n2 <= 2) || l >= 30) break;
        int m1 = (n1 + 1) / 2, m2 = (n2 + 1) / 2;
        int ldc = m2 + 2;
        double *wc = malloc(sizeof(double) * (m1 + 2) * ldc);
        for (int I = 0; I < m1; I++)
            for (int J = 0; J < m2; J++) {
                double s = 0; int cnt = 0;
                for (int a = 0; a < 2; a++)
                    for (int bb = 0; bb < 2; bb++) {
                        int fi = 2 * I + a, fj = 2 * J + bb;
                        if (fi < n1 && fj < n2) { s += w[(fi + 1) * ldw + (fj + 1)]; cnt++; }
                    }
                wc[(I + 1) * ldc + (J + 1)] = s / cnt;
            }
        for (int I = -1; I <= m1; I++)
            for (int J = -1; J <= m2; J++) {
                int ii = I < 0 ? 0 : (I >= m1 ? m1 - 1 : I);
                int jj = J < 0 ? 0 : (J >= m2 ? m2 - 1 : J);
                wc[(I + 1) * ldc + (J + 1)] = wc[(ii + 1) * ldc + (jj + 1)];
            }
        free(w);
        w = wc; ldw = ldc; n1 = m1; n2 = m2; dx *= 2; dy *= 2;
    }
    free(w);
    return 0;
}

static int color_of(int i, int j, int ncolor)
{
    return ncolor == 2 ? ((i + j) & 1) : ((i & 1) * 2 + (j & 1));
}

static void gs_color(or_level *L, int color, int ncolor)
{
    const int n1 = L->n1, n2 = L->n2, M = n1 * n2;
    /* the cells of one colour do not read each other (or read them with a
     * zero coefficient), so the colour sweep runs in any order */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) {
            if (color_of(i, j, ncolor) != color) continue;
            const int id = i * n2 + j;
            double s = L->b[id];
            for (int k = 1; k < 9; k++) {
                int ii = i + ST_OFF[k][0], jj = j + ST_OFF[k][1];
                if (ii < 0 || jj < 0 || ii >= n1 || jj >= n2) continue;
                s -= L->A[k * M + id] * L->x[ii * n2 + jj];
            }
            double diag = L->A[id];
            L->x[id] = diag != 0 ? s / diag : 0;
        }
}

static void project_mean(double *x, int n)
{
    double s = 0;
    for (int k = 0; k < n; k++) s += x[k];
    s /= n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) x[k] -= s;
}

/* bilinear cell-centred prolongation weight of coarse index I for fine f */
static void prolong1d(int f, int nc, int *I0, int *I1)
{
    int I = f / 2;
    int Io = (f & 1) ? I + 1 : I - 1;
    if (Io < 0) Io = 0;
    if (Io > nc - 1) Io = nc - 1;
    if (I > nc - 1) I = nc - 1;
    *I0 = I; *I1 = Io;   /* weights 3/4 and 1/4 */
}

static void vcycle(or_mg *mg, int l)
{
    or_level *L = &mg->L[l];
    const int n1 = L->n1, n2 = L->n2, M = n1 * n2, nc = mg->ncolor;
    memset(L->x, 0, sizeof(double) * M);
    if (l == mg->nlev - 1) {
        /* coarsest Neumann problem: symmetric sweeps + mean projection */
        for (int it = 0; it < 50; it++) {
            for (int c = 0; c < nc; c++) gs_color(L, c, nc);
            for (int c = nc - 1; c >= 0; c--) gs_color(L, c, nc);
        }
        project_mean(L->x, M);
        return;
    }
    for (int c = 0; c < nc; c++) gs_color(L, c, nc);
    or_matvec9(n1, n2, L->A, L->x, L->r);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < M; k++) L->r[k] = L->b[k] - L->r[k];
    or_level *Lc = &mg->L[l + 1];
    const int m1 = Lc->n1, m2 = Lc->n2;
    /* full weighting, the transpose of the prolongation: each fine residual
     * r/4 goes to its four coarse cells with weights 9/16, 3/16, 3/16, 1/16.
     * Every coarse cell gathers its contributions in the order a row-major
     * scatter over the fine cells would add them (fine rows, then columns,
     * then the four targets in turn), so the sums do not depend on threads. */
#pragma omp parallel for schedule(static)
    for (int I = 0; I < m1; I++)
        for (int J = 0; J < m2; J++) {
            double acc = 0;
            const int i0 = 2 * I - 3 > 0 ? 2 * I - 3 : 0, i1 = 2 * I + 4 < n1 - 1 ? 2 * I + 4 : n1 - 1;
            const int j0 = 2 * J - 3 > 0 ? 2 * J - 3 : 0, j1 = 2 * J + 4 < n2 - 1 ? 2 * J + 4 : n2 - 1;
            for (int i = i0; i <= i1; i++) {
                int Ia, Ib; prolong1d(i, m1, &Ia, &Ib);
                if (Ia != I && Ib != I) continue;
                for (int j = j0; j <= j1; j++) {
                    int Ja, Jb; prolong1d(j, m2, &Ja, &Jb);
                    if (Ja != J && Jb != J) continue;
                    double r = L->r[i * n2 + j] * 0.25;
                    if (Ia == I && Ja == J) acc += 0.5625 * r;
                    if (Ia == I && Jb == J) acc += 0.1875 * r;
                    if (Ib == I && Ja == J) acc += 0.1875 * r;
                    if (Ib == I && Jb == J) acc += 0.0625 * r;
                }
            }
            Lc->b[I * m2 + J] = acc;
        }
    vcycle(mg, l + 1);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n1; i++) {
        int Ia, Ib; prolong1d(i, m1, &Ia, &Ib);
        for (int j = 0; j < n2; j++) {
            int Ja, Jb; prolong1d(j, m2, &Ja, &Jb);
            L->x[i * n2 + j] += 0.5625 * Lc->x[Ia * m2 + Ja] + 0.1875 * Lc->x[Ia * m2 + Jb]
                                + 0.1875 * Lc->x[Ib * m2 + Ja] + 0.0625 * Lc->x[Ib * m2 + Jb];
        }
    }
    for (int c = nc - 1; c >= 0; c--) gs_color(L, c, nc);
}

/* ------------------------------------------------------------------------- */
/* linear solves                                                              */
/* ------------------------------------------------------------------------- */
#define DOT_CHUNK 16384
static double dot(const double *a, const double *b, int n)
{
    if (n < 4 * DOT_CHUNK) {
        double s = 0;
        for (int k = 0; k < n; k++) s += a[k] * b[k];
        return s;
    }
    const int nch = (n + DOT_CHUNK - 1) / DOT_CHUNK;
    double *part = malloc(sizeof(double) * nch);
#pragma omp parallel for schedule(static)
    for (int q = 0; q < nch; q++) {
        const int k1 = (q + 1) * DOT_CHUNK < n ? (q + 1) * DOT_CHUNK : n;
        double s = 0;
        for (int k = q * DOT_CHUNK; k < k1; k++) s += a[k] * b[k];
        part[q] = s;
    }
    double s = 0;
    for (int q = 0; q < nch; q++) s += part[q];
    free(part);
    return s;
}

/* exact symmetry test of a folded 9-array stencil matrix */
static int is_symmetric9(int N1, int N2, const double *A)
{
    static const int opp[9] = {ST_C, ST_E, ST_NE, ST_N, ST_NW, ST_W, ST_SW, ST_S, ST_SE};
    const int M = N1 * N2;
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++)
            for (int k = 1; k < 9; k++) {
                int ii = i + ST_OFF[k][0], jj = j + ST_OFF[k][1];
                if (ii < 0 || jj < 0 || ii >= N1 || jj >= N2) continue;
                if (A[k * M + i * N2 + j] != A[opp[k] * M + ii * N2 + jj]) return 0;
            }
    return 1;
}

/* Pressure solve: HYPRE StructPCG semantics (MGSolver.cpp:596-621) with the
 * pinned row eliminated symmetrically.  A is the assembled matrix with the
 * pin row (MGSolver.cpp:517-567); A_N (its unpinned Neumann counterpart) is
 * obtained by reassembling.  x is initialised by the caller (zero in the
 * reference, MGSolver.cpp:573-590).  Returns 0 on convergence, 1 if max_iter
 * was hit (the reference does not detect that case). */
static int pcg_pinned(or_ctx *ctx, const double *A, const double *A_N, const double *c,
                      const double *b, double *x, double tol)
{
    const int N1 = ctx->N1, N2 = ctx->N2, M = N1 * N2, pin = 1 * N2 + 1;
    const double dpin = A[ST_C * M + pin];
    const double ppin = b[pin] / dpin;
    double *r = malloc(sizeof(double) * M), *z = malloc(sizeof(double) * M);
    double *d = malloc(sizeof(double) * M), *q = malloc(sizeof(double) * M);
    or_mg mg;
    const int use_mg = ctx->P.precond == 0;
    if (use_mg) mg_setup(ctx, &mg, A_N, c);
    /* A symmetric Neumann matrix has zero column sums, so the pinned solution
     * is the solution of A_N x = b~ (b~ = b with the pin entry reduced by
     * sum(b)) shifted to x_pin = p_pin: the multigrid preconditioner then
     * sees the Neumann problem it is built for.  A non-symmetric A_N (nu < 1
     * with dx != dy, or K11 != K22) keeps the pinned row eliminated instead. */
    const int shift = is_symmetric9(N1, N2, A_N);
    double bsum = 0;
    if (shift) {
        for (int k = 0; k < M; k++) bsum += b[k];
    }
    x[pin] = ppin;
    or_matvec9(N1, N2, A_N, x, q);
    for (int k = 0; k < M; k++) r[k] = b[k] - q[k];
    if (shift) r[pin] -= bsum; else r[pin] = 0;
    const double bnorm = sqrt(dot(b, b, M));
    double rnorm = sqrt(dot(r, r, M));
    int its = 0, vcycles = 0, status = 0;
    double rho = 0, rho_old = 1;
    const int maxit = ctx->P.max_iter > 0 ? ctx->P.max_iter : 10000;
    while (!(rnorm <= tol * bnorm)) {
        if (its >= maxit) { status = 1; break; }
        if (use_mg) {
            memcpy(mg.L[0].b, r, sizeof(double) * M);
            project_mean(mg.L[0].b, M);
            vcycle(&mg, 0);
            vcycles++;
            memcpy(z, mg.L[0].x, sizeof(double) * M);
            project_mean(z, M);
        } else {
            for (int k = 0; k < M; k++) z[k] = r[k] / A_N[ST_C * M + k];
        }
        if (!shift) z[pin] = 0;
        rho = dot(r, z, M);
        if (its == 0) memcpy(d, z, sizeof(double) * M);
        else {
            double beta = rho / rho_old;
#pragma omp parallel for schedule(static)
            for (int k = 0; k < M; k++) d[k] = z[k] + beta * d[k];
        }
        or_matvec9(N1, N2, A_N, d, q);
        if (!shift) q[pin] = 0;
        double alpha = rho / dot(d, q, M);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < M; k++) { x[k] += alpha * d[k]; r[k] -= alpha * q[k]; }
        rho_old = rho;
        its++;
        rnorm = sqrt(dot(r, r, M));
    }
    if (shift) { double sh = ppin - x[pin]; for (int k = 0; k < M; k++) x[k] += sh; x[pin] = ppin; }
    /* true residual of the pinned system */
    or_matvec9(N1, N2, A, x, q);
    double s = 0;
    for (int k = 0; k < M; k++) s += (b[k] - q[k]) * (b[k] - q[k]);
    ctx->last_relres = bnorm > 0 ? sqrt(s) / bnorm : 0;
    ctx->last_its = its;
    ctx->last_vcycles = vcycles;
    if (use_mg) mg_free(&mg);
    free(r); free(z); free(d); free(q);
    return status;
}

/* BiCGStab with Jacobi preconditioning for the non-symmetric CN system;
 * stops on ||b - A x||_2 <= tol ||b||_2 like HYPRE's PCG two-norm test. */
static int bicgstab(or_ctx *ctx, const double *A, const double *b, double *x, double tol, int *its_out,
                    double *relres)
{
    const int N1 = ctx->N1, N2 = ctx->N2, M = N1 * N2;
    double *r = malloc(sizeof(double) * M), *r0 = malloc(sizeof(double) * M);
    double *p = calloc(M, sizeof(double)), *v = calloc(M, sizeof(double));
    double *s = malloc(sizeof(double) * M), *t = malloc(sizeof(double) * M);
    double *y = malloc(sizeof(double) * M), *zz = malloc(sizeof(double) * M);
    or_matvec9(N1, N2, A, x, t);
    for (int k = 0; k < M; k++) r[k] = b[k] - t[k];
    memcpy(r0, r, sizeof(double) * M);
    const double bnorm = sqrt(dot(b, b, M));
    double rnorm = sqrt(dot(r, r, M));
    double rho = 1, alpha = 1, omega = 1;
    int its = 0, status = 0;
    const int maxit = ctx->P.max_iter > 0 ? ctx->P.max_iter : 10000;
    while (!(rnorm <= tol * bnorm)) {
        if (its >= maxit) { status = 1; break; }
        double rho_new = dot(r0, r, M);
        if (rho_new == 0) { status = 2; break; }
        double beta = (rho_new / rho) * (alpha / omega);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < M; k++) {
            p[k] = r[k] + beta * (p[k] - omega * v[k]);
            y[k] = p[k] / A[ST_C * M + k];
        }
        or_matvec9(N1, N2, A, y, v);
        alpha = rho_new / dot(r0, v, M);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < M; k++) {
            s[k] = r[k] - alpha * v[k];
            zz[k] = s[k] / A[ST_C * M + k];
        }
        or_matvec9(N1, N2, A, zz, t);
        double tt = dot(t, t, M);
        omega = tt > 0 ? dot(t, s, M) / tt : 0;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < M; k++) { x[k] += alpha * y[k] + omega * zz[k]; r[k] = s[k] - omega * t[k]; }
        rho = rho_new;
        its++;
        rnorm = sqrt(dot(r, r, M));
        if (omega == 0) { status = 2; break; }
    }
    or_matvec9(N1, N2, A, x, t);
    double ss = 0;
    for (int k = 0; k < M; k++) ss += (b[k] - t[k]) * (b[k] - t[k]);
    *relres = bnorm > 0 ? sqrt(ss) / bnorm : 0;
    *its_out = its;
    free(r); free(r0); free(p); free(v); free(s); free(t); free(y); free(zz);
    return status;
}

/* Dense LU with partial pivoting of a 9-array stencil matrix (N1*N2 <= 4096).
 * Pins the solution of the assembled system on small grids. */
int or_dense_solve9(int N1, int N2, const double *A, const double *b, double *x)
{
    const int M = N1 * N2;
    if (M > 4096) return -1;
    double *D = calloc((size_t)M * M, sizeof(double));
    double *rhs = malloc(sizeof(double) * M);
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++)
            for (int k = 0; k < 9; k++) {
                int ii = i + ST_OFF[k][0], jj = j + ST_OFF[k][1];
                if (ii < 0 || jj < 0 || ii >= N1 || jj >= N2) continue;
                D[(size_t)(i * N2 + j) * M + ii * N2 + jj] += A[k * M + i * N2 + j];
            }
    memcpy(rhs, b, sizeof(double) * M);
    for (int col = 0; col < M; col++) {
        int piv = col;
        double best = fabs(D[(size_t)col * M + col]);
        for (int r = col + 1; r < M; r++)
            if (fabs(D[(size_t)r * M + col]) > best) { best = fabs(D[(size_t)r * M + col]); piv = r; }
        if (best == 0) { free(D); free(rhs); return -2; }
        if (piv != col) {
            for (int k = 0; k < M; k++) {
                double t = D[(size_t)col * M + k];
                D[(size_t)col * M + k] = D[(size_t)piv * M + k];
                D[(size_t)piv * M + k] = t;
            }
            double t = rhs[col]; rhs[col] = rhs[piv]; rhs[piv] = t;
        }
        for (int r = col + 1; r < M; r++) {
            double f = D[(size_t)r * M + col] / D[(size_t)col * M + col];
            if (f == 0) continue;
            for (int k = col; k < M; k++) D[(size_t)r * M + k] -= f * D[(size_t)col * M + k];
            rhs[r] -= f * rhs[col];
        }
    }
    for (int r = M - 1; r >= 0; r--) {
        double s = rhs[r];
        for (int k = r + 1; k < M; k++) s -= D[(size_t)r * M + k] * x[k];
        x[r] = s / D[(size_t)r * M + r];
    }
    free(D); free(rhs);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* pressure solve (MGSolver.cpp:145-149, 223-248, 573-590, 670-706)           */
/* ------------------------------------------------------------------------- */

/* Solve_Elliptic_cc_to_c: assemble A(c), b = -q, x0 = 0 (or the given p
 * interior when warm != 0, an extension), solve, write p's interior.
 * c must be ghost-filled.  Returns 0 ok, 1 no convergence. */
int or_pressure_solve(or_ctx *ctx, const double *c, double *p, const double *q, int warm)
{
    const int N1 = ctx->N1, N2 = ctx->N2, M = N1 * N2;
    double *A = malloc(sizeof(double) * 9 * M), *AN = malloc(sizeof(double) * 9 * M);
    double *b = malloc(sizeof(double) * M), *x = malloc(sizeof(double) * M);
    or_assemble_elliptic(ctx, c, A);
    /* the unpinned Neumann matrix: the same assembly without the pin row */
    double a[9], f[9];
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) {
            weighted_coeffs(ctx, c, i, j, a);
            fold_boundary(i, j, N1, N2, a, f);
            for (int k = 0; k < 9; k++) AN[k * M + i * N2 + j] = f[k];
        }
    for (int k = 0; k < M; k++) b[k] = -q[k];
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) x[i * N2 + j] = warm ? p[S_(ctx, i, j)] : 0.;
    int st = pcg_pinned(ctx, A, AN, c, b, x, ctx->P.eps);
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) p[S_(ctx, i, j)] = x[i * N2 + j];
    free(A); free(AN); free(b); free(x);
    return st;
}

/* Dense-LU variant of the pressure solve, for N1*N2 <= 4096. */
int or_pressure_solve_dense(or_ctx *ctx, const double *c, double *p, const double *q)
{
    const int N1 = ctx->N1, N2 = ctx->N2, M = N1 * N2;
    double *A = malloc(sizeof(double) * 9 * M), *b = malloc(sizeof(double) * M);
    double *x = malloc(sizeof(double) * M);
    or_assemble_elliptic(ctx, c, A);
    for (int k = 0; k < M; k++) b[k] = -q[k];
    int st = or_dense_solve9(N1, N2, A, b, x);
    if (st == 0)
        for (int i = 0; i < N1; i++)
            for (int j = 0; j < N2; j++) p[S_(ctx, i, j)] = x[i * N2 + j];
    free(A); free(b); free(x);
    return st;
}

/* ------------------------------------------------------------------------- */
/* velocities (EllipticSolver.cpp:541-683)                                    */
/* ------------------------------------------------------------------------- */
static double par_u1(const or_ctx *ctx, const double *c, const double *p, int i, int j)
{
    return -kom11_face(ctx, c, i, j) * (p[S_(ctx, i, j)] - p[S_(ctx, i - 1, j)]) / ctx->dx;
}
static double par_u2(const or_ctx *ctx, const double *c, const double *p, int i, int j)
{
    return -kom22_face(ctx, c, i, j) * (p[S_(ctx, i, j)] - p[S_(ctx, i, j - 1)]) / ctx->dy;
}

/* M_im12j + GetDiagonalW_im12j (EllipticSolver.cpp:554-574): vertex (i-1/2,j-1/2). */
static void diag_w(const or_ctx *ctx, const double *c, const double *p, int i, int j,
                   double *w1, double *w2)
{
    double d1dp = 1. / ctx->dxi * (p[S_(ctx, i, j)] - p[S_(ctx, i - 1, j - 1)]);
    double d2dp = 1. / ctx->deta * (p[S_(ctx, i - 1, j)] - p[S_(ctx, i, j - 1)]);
    const double K11 = kom_vertex(ctx, c, i - 1, j - 1, 0);
    const double K22 = kom_vertex(ctx, c, i - 1, j - 1, 1);
    double M11 = ctx->secamt2 * (K11 * ctx->cosa2 + K22 * ctx->sina2);
    double M12 = ctx->secamt2 * (K22 * ctx->costsina - K11 * ctx->cosasint);
    double M21 = ctx->secamt2 * (K22 * ctx->costsina - K11 * ctx->cosasint);
    double M22 = ctx->secamt2 * (K11 * ctx->sint2 + K22 * ctx->cost2);
    *w1 = -d1dp * M11 - d2dp * M12;
    *w2 = -d1dp * M21 - d2dp * M22;
}

/* GetDiagonalU1_im12j (EllipticSolver.cpp:600-621). */
static double diag_u1(const or_ctx *ctx, const double *c, const double *p, int i, int j)
{
    double w1a, w2a, w1b, w2b;
    diag_w(ctx, c, p, i, j, &w1a, &w2a);
    diag_w(ctx, c, p, i, j + 1, &w1b, &w2b);
    double u1m = ctx->cost * w1a - ctx->sina * w2a;
    double u1p = ctx->cost * w1b - ctx->sina * w2b;
    return 0.5 * (u1p + u1m);
}

/* GetDiagonalU2_ijm12 (EllipticSolver.cpp:624-644). */
static double diag_u2(const or_ctx *ctx, const double *c, const double *p, int i, int j)
{
    double w1a, w2a, w1b, w2b;
    diag_w(ctx, c, p, i, j, &w1a, &w2a);
    diag_w(ctx, c, p, i + 1, j, &w1b, &w2b);
    double u2m = ctx->sint * w1a + ctx->cosa * w2a;
    double u2p = ctx->sint * w1b + ctx->cosa * w2b;
    return 0.5 * (u2m + u2p);
}

/* EllipticSolver::CalcVelocities (EllipticSolver.cpp:660-683). */
void or_calc_velocities(const or_ctx *ctx, double *p, double *v, const double *c)
{
    or_fill_neumann_scalar(ctx, p);
    const double nu = ctx->P.nu;
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= ctx->N1; i++)
        for (int j = 0; j < ctx->N2; j++)
            v[V_(ctx, i, j, 0)] = nu * par_u1(ctx, c, p, i, j) + (1. - nu) * diag_u1(ctx, c, p, i, j);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ctx->N1; i++)
        for (int j = 0; j <= ctx->N2; j++)
            v[V_(ctx, i, j, 1)] = nu * par_u2(ctx, c, p, i, j) + (1. - nu) * diag_u2(ctx, c, p, i, j);
    or_fill_neumann_velocity(ctx, v);
}

/* ------------------------------------------------------------------------- */
/* time step (HyperbolicSolver.cpp:93-198)                                    */
/* ------------------------------------------------------------------------- */
double or_max_stable_advection_timestep(const or_ctx *ctx, const double *v)
{
    double dt_max = DBL_MAX;
    for (int i = 0; i < ctx->N1; i++)
        for (int j = 0; j < ctx->N2; j++) {
            double uc = fabs((double)(0.5 * (v[V_(ctx, i, j, 0)] + v[V_(ctx, i + 1, j, 0)])));
            double vc = fabs((double)(0.5 * (v[V_(ctx, i, j, 1)] + v[V_(ctx, i, j + 1, 1)])));
            if (uc < DBL_MIN) uc = DBL_MIN;
            if (vc < DBL_MIN) vc = DBL_MIN;
            double a = ctx->dx / uc, bb = ctx->dy / vc;
            double dt = a < bb ? a : bb;
            if (dt < dt_max) dt_max = dt;
        }
    dt_max *= ctx->P.cfl;
    return dt_max;
}

double or_max_stable_timestep_combined(const or_ctx *ctx, const double *v)
{
    double maxvel = 0;
    for (int i = 0; i < ctx->N1; i++)
        for (int j = 0; j < ctx->N2; j++) {
            double uc = fabs((double)(0.5 * (v[V_(ctx, i, j, 0)] + v[V_(ctx, i + 1, j, 0)])));
            double vc = fabs((double)(0.5 * (v[V_(ctx, i, j, 1)] + v[V_(ctx, i, j + 1, 1)])));
            if (uc < DBL_MIN) uc = DBL_MIN;
            if (vc < DBL_MIN) vc = DBL_MIN;
            if (maxvel < fabs(uc)) maxvel = fabs(uc);
            if (maxvel < fabs(vc)) maxvel = fabs(vc);
        }
    double a = ctx->dx / maxvel, bb = ctx->dy / maxvel;
    double dt_advec = a < bb ? a : bb;
    dt_advec *= ctx->P.cfl;
    double dt_diff = DBL_MAX;
    if (ctx->P.diffusion) {
        double mn = 1.e10;
        const double al = ctx->P.alpha_l, at = ctx->P.alpha_t, dx = ctx->dx;
        for (int i = 0; i < ctx->N1; i++)
            for (int j = 0; j < ctx->N2; j++) {
                double vij = 0.5 * (v[V_(ctx, i, j, 1)] + v[V_(ctx, i, j + 1, 1)]);
                double uij = 0.5 * (v[V_(ctx, i, j, 0)] + v[V_(ctx, i + 1, j, 0)]);
                double vij2 = vij * vij, uij2 = uij * uij;
                double s = vij2 + uij2;
                double absu = sqrt(s);
                double d11 = dx / absu * (al * uij2 + at * vij2);
                double d22 = dx / absu * (al * vij2 + at * uij2);
                double dt = 1. / 2 * dx * dx / (d11 + d22);
                if (dt < mn) mn = dt;
            }
        dt_diff = mn * ctx->P.dt_diff_limit;
    }
    if (ctx->P.para_cn) return dt_advec;
    return dt_diff < dt_advec ? dt_diff : dt_advec;
}

double or_max_stable_timestep(const or_ctx *ctx, const double *v)
{
    if (ctx->P.is_test) return or_max_stable_advection_timestep(ctx, v);
    return or_max_stable_timestep_combined(ctx, v);
}

/* ------------------------------------------------------------------------- */
/* source term (TRIntegrator.cpp:49-71; HyperbolicSolver.cpp:201-205)         */
/* ------------------------------------------------------------------------- */
double or_tr_integrate(const or_ctx *ctx, double *c, const double *q, double dt)
{
    double tmo = 0;
    for (int i = 0; i < ctx->N1; i++)
        for (int j = 0; j < ctx->N2; j++) {
            double qij = q[i * ctx->N2 + j];
            double *cij = &c[S_(ctx, i, j)];
            if (qij < 0) {
                double md = *cij;
                *cij = *cij * (1 + 0.5 * dt * qij) / (1 - 0.5 * dt * qij);
                md = *cij - md;
                tmo += md;
            } else if (qij > 0)
                *cij += dt * ctx->P.inj_fluid_conc * qij;
        }
    return tmo;
}

void or_source_split(or_ctx *ctx, double *c, const double *q, double dt)
{
    double tmo = or_tr_integrate(ctx, c, q, dt);
    ctx->total_mass_outflux += fabs(tmo);
}

/* ------------------------------------------------------------------------- */
/* MUSCL-Hancock / upwind fluxes (MUSCLSplitSolver.cpp:54-189;               */
/* UpwindSplitSolver.cpp:43-65; HyperbolicSolver.cpp:208-210)                 */
/* ------------------------------------------------------------------------- */
static inline double flux(const or_ctx *ctx, double vel, double conc)
{
    return vel * conc / ctx->P.porosity;
}

static double slope_limiter(double du_m, double du_p)
{
    double w = 0.;
    double r;
    if (du_m * du_p > 0) {
        r = du_m / du_p;
        if (r > DBL_MAX) r = DBL_MAX;
        double beta_plus = 1;
        double zr = 2 * beta_plus / (1 - w + (1 + w) * r);
        double zl = 2. * r / (1. + r);
        return (zr < zl) ? zr : zl;   /* std::min(a, b) returns a unless b < a */
    }
    return 0;
}

static double get_slope(double dU_prev, double dU_next)
{
    double w = 0.;
    return slope_limiter(dU_prev, dU_next) * 0.5 * ((1 + w) * dU_prev + (1 - w) * dU_next);
}

static void reconstruct(double U_i, double U_ip1, double U_im1, double *UL, double *UR)
{
    double dU_prev = U_i - U_im1;
    double dU_next = U_ip1 - U_i;
    double d_i = get_slope(dU_prev, dU_next);
    *UR = U_i + 0.5 * d_i;
    *UL = U_i - 0.5 * d_i;
}

static double upwind(const or_ctx *ctx, double speed, double UL, double UR)
{
    return (speed > 0) ? flux(ctx, speed, UL) : flux(ctx, speed, UR);
}

/* F(i,j) for i in [0,N1], j in [-b, N2+b); F is (N1+1) x (N2+2b), j base -b */
static void flux_F(const or_ctx *ctx, double *F, const double *c, const double *v, double dt)
{
    const int N1 = ctx->N1, N2 = ctx->N2, b = ctx->b, ldf = N2 + 2 * b;
#define F_(i, j) F[(i) * ldf + (j) + b]
    if (ctx->P.hyper == 1) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i <= N1; i++)
            for (int j = -b; j < N2 + b; j++)
                F_(i, j) = upwind(ctx, v[V_(ctx, i, j, 0)], c[S_(ctx, i - 1, j)], c[S_(ctx, i, j)]);
        return;
    }
    /* U_Lin (N1+3) x (N2+1+2b) x 2 with base (-1, -b) */
    const int ldu = N2 + 1 + 2 * b;
    double *U = calloc((size_t)(N1 + 3) * ldu * 2, sizeof(double));
#define U_(i, j, k) U[(((i) + 1) * ldu + ((j) + b)) * 2 + (k)]
#pragma omp parallel for schedule(static)
    for (int i = -1; i <= N1; i++)
        for (int j = -b; j < N2 + b; j++) {
            double UL, UR;
            reconstruct(c[S_(ctx, i, j)], c[S_(ctx, i + 1, j)], c[S_(ctx, i - 1, j)], &UL, &UR);
            double df = 0.5 * dt / ctx->dx * (flux(ctx, v[V_(ctx, i, j, 0)], UL) - flux(ctx, v[V_(ctx, i + 1, j, 0)], UR));
            U_(i + 1, j, 0) = UR + df;
            U_(i, j, 1) = UL + df;
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i <= N1; i++)
        for (int j = -b; j < N2 + b; j++)
            F_(i, j) = upwind(ctx, v[V_(ctx, i, j, 0)], U_(i, j, 0), U_(i, j, 1));
    free(U);
#undef U_
#undef F_
}

/* G(i,j) for i in [-b, N1+b), j in [0,N2]; G is (N1+2b) x (N2+1), i base -b */
static void flux_G(const or_ctx *ctx, double *G, const double *c, const double *v, double dt)
{
    const int N1 = ctx->N1, N2 = ctx->N2, b = ctx->b, ldg = N2 + 1;
#define G_(i, j) G[((i) + b) * ldg + (j)]
    if (ctx->P.hyper == 1) {
#pragma omp parallel for schedule(static)
        for (int i = -b; i < N1 + b; i++)
            for (int j = 0; j <= N2; j++)
                G_(i, j) = upwind(ctx, v[V_(ctx, i, j, 1)], c[S_(ctx, i, j - 1)], c[S_(ctx, i, j)]);
        return;
    }
    const int ldu = N2 + 3;
    double *U = calloc((size_t)(N1 + 1 + 2 * b) * ldu * 2, sizeof(double));
#define U_(i, j, k) U[(((i) + b) * ldu + ((j) + 1)) * 2 + (k)]
#pragma omp parallel for schedule(static)
    for (int i = -b; i < N1 + b; i++)
        for (int j = -1; j <= N2; j++) {
            double UL, UR;
            reconstruct(c[S_(ctx, i, j)], c[S_(ctx, i, j + 1)], c[S_(ctx, i, j - 1)], &UL, &UR);
            double df = 0.5 * dt / ctx->dy * (flux(ctx, v[V_(ctx, i, j, 1)], UL) - flux(ctx, v[V_(ctx, i, j + 1, 1)], UR));
            U_(i, j + 1, 0) = UR + df;
            U_(i, j, 1) = UL + df;
        }
#pragma omp parallel for schedule(static)
    for (int i = -b; i < N1 + b; i++)
        for (int j = 0; j <= N2; j++)
            G_(i, j) = upwind(ctx, v[V_(ctx, i, j, 1)], U_(i, j, 0), U_(i, j, 1));
    free(U);
#undef U_
#undef G_
}

/* Exported fluxes (CalculateIntercellFlux_F / _G of the selected scheme). */
void or_flux_F(const or_ctx *ctx, double *F, const double *c, const double *v, double dt) { flux_F(ctx, F, c, v, dt); }
void or_flux_G(const or_ctx *ctx, double *G, const double *c, const double *v, double dt) { flux_G(ctx, G, c, v, dt); }

/* Advect_X (HyperbolicSplitSolver.cpp:113-134). */
void or_advect_x(const or_ctx *ctx, double *c, const double *v, double dt)
{
    const int N1 = ctx->N1, N2 = ctx->N2, b = ctx->b, ldf = N2 + 2 * b;
    double *F = calloc((size_t)(N1 + 1) * ldf, sizeof(double));
    flux_F(ctx, F, c, v, dt);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N1; i++)
        for (int j = -b; j < N2 + b; j++)
            c[S_(ctx, i, j)] += dt / ctx->dx * (F[i * ldf + j + b] - F[(i + 1) * ldf + j + b]);
    free(F);
}

/* Advect_Y (HyperbolicSplitSolver.cpp:140-159). */
void or_advect_y(const or_ctx *ctx, double *c, const double *v, double dt)
{
    const int N1 = ctx->N1, N2 = ctx->N2, b = ctx->b, ldg = N2 + 1;
    double *G = calloc((size_t)(N1 + 2 * b) * ldg, sizeof(double));
    flux_G(ctx, G, c, v, dt);
#pragma omp parallel for schedule(static)
    for (int i = -b; i < N1 + b; i++)
        for (int j = 0; j < N2; j++)
            c[S_(ctx, i, j)] += dt / ctx->dy * (G[(i + b) * ldg + j] - G[(i + b) * ldg + j + 1]);
    free(G);
}

/* AdvectSplit (HyperbolicSplitSolver.cpp:51-106). */
void or_advect_split(or_ctx *ctx, double *c, const double *v, double dt)
{
    or_fill_boundaries(ctx, c);
    switch (ctx->P.dim_split) {
    case 0:
        if (ctx->prev_first_sweep == 1) {
            or_advect_x(ctx, c, v, dt); or_advect_y(ctx, c, v, dt);
            ctx->prev_first_sweep = or_mutation == MUT_SPLIT_ALT ? 1 : 0;
        } else {
            or_advect_y(ctx, c, v, dt); or_advect_x(ctx, c, v, dt); ctx->prev_first_sweep = 1;
        }
        break;
    case 1:
        if (or_mutation == MUT_SPLIT_XFIRST) { or_advect_y(ctx, c, v, dt); or_advect_x(ctx, c, v, dt); }
        else { or_advect_x(ctx, c, v, dt); or_advect_y(ctx, c, v, dt); }
        ctx->prev_first_sweep = 0;
        break;
    case 2:
        if (or_mutation == MUT_SPLIT_YFIRST) { or_advect_x(ctx, c, v, dt); or_advect_y(ctx, c, v, dt); }
        else { or_advect_y(ctx, c, v, dt); or_advect_x(ctx, c, v, dt); }
        ctx->prev_first_sweep = 1;
        break;
    case 3: {
        const double hx = or_mutation == MUT_SPLIT_STRANG ? dt : 0.5 * dt;
        or_advect_x(ctx, c, v, hx); or_advect_y(ctx, c, v, dt); or_advect_x(ctx, c, v, hx);
        ctx->prev_first_sweep = 0;
        break;
    }
    default: {
        const size_t n = (size_t)(ctx->N1 + 2 * ctx->b) * ctx->ld;
        double *ct = malloc(sizeof(double) * n);
        memcpy(ct, c, sizeof(double) * n);
        or_advect_x(ctx, c, v, dt); or_advect_y(ctx, c, v, dt);
        or_advect_y(ctx, ct, v, dt); or_advect_x(ctx, ct, v, dt);
        if (or_mutation != MUT_SPLIT_PARALLEL)
            for (size_t k = 0; k < n; k++) c[k] = 0.5 * (c[k] + ct[k]);
        free(ct);
        ctx->prev_first_sweep = 0;
    }
    }
}

/* ------------------------------------------------------------------------- */
/* Crank-Nicolson dispersion (MGSolver.cpp:846-1392)                          */
/* ------------------------------------------------------------------------- */
/* SetupDiffusionTensorF (MGSolver.cpp:846-883) at x-face (i-1/2, j). */
static void tensor_F(const or_ctx *ctx, const double *v, int i, int j, double *d11, double *d12)
{
    const double dx = ctx->dx, dy = ctx->dy;
    double u = v[V_(ctx, i, j, 0)];
    const int ia = or_mutation == MUT_CN_AVG_F ? i + 1 : i - 1;
    double vb = 1. / 4 * (v[V_(ctx, ia, j + 1, 1)] + v[V_(ctx, i, j + 1, 1)]
                          + v[V_(ctx, i - 1, j, 1)] + v[V_(ctx, i, j, 1)]);
    double u2 = u * u, v2 = vb * vb;
    double phi = or_mutation == MUT_CN_DROP_PHI ? 1. : ctx->P.porosity;
    double maxdx = dx > dy ? dx : dy;   /* std::max(dx, dy) */
    double al = ctx->P.alpha_abs ? ctx->P.alpha_l : ctx->P.alpha_l * maxdx;
    double at = ctx->P.alpha_abs ? ctx->P.alpha_t : ctx->P.alpha_t * maxdx;
    if (or_mutation == MUT_CN_SWAP_ALPHA) { double t_ = al; al = at; at = t_; }
    double s = sqrt(u2 + v2);
    double fl = phi * al / s, ft = phi * at / s;
    if (ctx->P.alpha_abs && !(s > 0)) { fl = 0; ft = 0; }
    double d11l = fl * u2, d12l = fl * (u * vb);
    double d11t = ft * v2, d12t = ft * (-u * vb);
    *d11 = d11l + d11t;
    *d12 = d12l + d12t;
    if (ctx->P.d_m != 0) *d11 = *d11 + phi * ctx->P.d_m;
    if (or_mutation == MUT_CN_DM_CROSS) *d12 = *d12 + phi * ctx->P.d_m;
}

/* SetupDiffusionTensorG (MGSolver.cpp:888-919) at y-face (i, j-1/2). */
static void tensor_G(const or_ctx *ctx, const double *v, int i, int j, double *d22, double *d21)
{
    const double dx = ctx->dx, dy = ctx->dy;
    const int ia = or_mutation == MUT_CN_AVG_G ? i - 1 : i + 1;
    double ub = 1. / 4 * (v[V_(ctx, ia, j - 1, 0)] + v[V_(ctx, i + 1, j, 0)]
                          + v[V_(ctx, i, j - 1, 0)] + v[V_(ctx, i, j, 0)]);
    double vv = v[V_(ctx, i, j, 1)];
    double u2 = ub * ub, v2 = vv * vv;
    double phi = or_mutation == MUT_CN_DROP_PHI ? 1. : ctx->P.porosity;
    double maxdx = dx > dy ? dx : dy;
    double al = ctx->P.alpha_abs ? ctx->P.alpha_l : ctx->P.alpha_l * maxdx;
    double at = ctx->P.alpha_abs ? ctx->P.alpha_t : ctx->P.alpha_t * maxdx;
    if (or_mutation == MUT_CN_SWAP_ALPHA) { double t_ = al; al = at; at = t_; }
    double s = sqrt(u2 + v2);
    double fl = phi * al / s, ft = phi * at / s;
    if (ctx->P.alpha_abs && !(s > 0)) { fl = 0; ft = 0; }
    double d21l = fl * (ub * vv), d22l = fl * v2;
    double d21t = ft * (-ub * vv), d22t = ft * u2;
    *d21 = d21l + d21t;
    *d22 = d22l + d22t;
    if (ctx->P.d_m != 0) *d22 = *d22 + phi * ctx->P.d_m;
    if (or_mutation == MUT_CN_DM_CROSS) *d21 = *d21 + phi * ctx->P.d_m;
}

/* GetCoefficients_Parabolic (MGSolver.cpp:923-948). */
static void para_coeffs(const or_ctx *ctx, const double *v, int i, int j, double dt, double a[9])
{
    const double dx = ctx->dx, dy = ctx->dy;
    double d11_im12j, d12_im12j, d22_ijm12, d21_ijm12, d21_ijp12, d22_ijp12, d12_ip12j, d11_ip12j;
    tensor_F(ctx, v, i, j, &d11_im12j, &d12_im12j);
    tensor_G(ctx, v, i, j, &d22_ijm12, &d21_ijm12);
    tensor_G(ctx, v, i, j + 1, &d22_ijp12, &d21_ijp12);
    tensor_F(ctx, v, i + 1, j, &d11_ip12j, &d12_ip12j);
    a[ST_NW] = dt / (8. * dx * dy) * (d12_im12j + d21_ijp12);
    a[ST_NE] = -dt / (8. * dx * dy) * (d12_ip12j + d21_ijp12);
    a[ST_N] = -dt / 8. * (4. / (dy * dy) * d22_ijp12 + 1. / (dx * dy) * (-d12_im12j + d12_ip12j));
    a[ST_SW] = -dt / (8. * dx * dy) * (d12_im12j + d21_ijm12);
    a[ST_SE] = dt / (8. * dx * dy) * (d12_ip12j + d21_ijm12);
    a[ST_S] = -dt / 8. * (4. / (dy * dy) * d22_ijm12 + 1. / (dx * dy) * (-d12_ip12j + d12_im12j));
    a[ST_W] = -dt / 8. * (4. / (dx * dx) * d11_im12j + 1. / (dx * dy) * (-d21_ijp12 + d21_ijm12));
    a[ST_C] = 1. + dt / (2. * dx * dx) * (d11_ip12j + d11_im12j) + dt / (2. * dy * dy) * (d22_ijp12 + d22_ijm12);
    a[ST_E] = -dt / 8. * (4. / (dx * dx) * d11_ip12j + 1. / (dx * dy) * (-d21_ijm12 + d21_ijp12));
}

/* SetMatrixCoefficients_Parabolic (MGSolver.cpp:997-1321); no Dirichlet
 * cells for the diagonal/parallel source types (MGSolver.cpp:117). */
void or_assemble_parabolic(const or_ctx *ctx, const double *v, double dt, double *A)
{
    const int N1 = ctx->N1, N2 = ctx->N2, M = N1 * N2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) {
            double a[9], f[9];
            para_coeffs(ctx, v, i, j, dt, a);
            fold_boundary(i, j, N1, N2, a, f);
            for (int k = 0; k < 9; k++) A[k * M + i * N2 + j] = f[k];
        }
}

/* SetRHS_Parabolic (MGSolver.cpp:1324-1370); c ghost-filled. */
void or_rhs_parabolic(const or_ctx *ctx, const double *c, const double *v, double dt, double *rhs)
{
    const double dx = ctx->dx, dy = ctx->dy;
#define C_(i, j) c[S_(ctx, i, j)]
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ctx->N1; i++)
        for (int j = 0; j < ctx->N2; j++) {
            double d11_im12j, d12_im12j, d22_ijm12, d21_ijm12, d21_ijp12, d22_ijp12, d12_ip12j, d11_ip12j;
            tensor_F(ctx, v, i, j, &d11_im12j, &d12_im12j);
            tensor_G(ctx, v, i, j, &d22_ijm12, &d21_ijm12);
            tensor_G(ctx, v, i, j + 1, &d22_ijp12, &d21_ijp12);
            tensor_F(ctx, v, i + 1, j, &d11_ip12j, &d12_ip12j);
            rhs[i * ctx->N2 + j] = C_(i, j) + dt / 8. * (
                4. / (dx * dx) * (d11_ip12j * (C_(i + 1, j) - C_(i, j))
                                  - d11_im12j * (C_(i, j) - C_(i - 1, j)))
                + 1. / (dx * dy) * (d12_ip12j * (C_(i, j + 1) + C_(i + 1, j + 1) - C_(i, j - 1) - C_(i + 1, j - 1))
                                    + (or_mutation == MUT_CN_RHS_CROSS ? 1. : -1.) * d12_im12j
                                          * (C_(i, j + 1) + C_(i - 1, j + 1) - C_(i - 1, j - 1) - C_(i, j - 1)))
                + 1. / (dx * dy) * (d21_ijp12 * (C_(i + 1, j) + C_(i + 1, j + 1) - C_(i - 1, j + 1) - C_(i - 1, j))
                                    - d21_ijm12 * (C_(i + 1, j) + C_(i + 1, j - 1) - C_(i - 1, j) - C_(i - 1, j - 1)))
                + 4. / (dy * dy) * (d22_ijp12 * (C_(i, j + 1) - C_(i, j))
                                    - d22_ijm12 * (C_(i, j) - C_(i, j - 1))));
        }
#undef C_
}

/* Solve_Parabolic_cc_to_c (MGSolver.cpp:952-976): x0 = 0
 * (AssignRHS_Parabolic, MGSolver.cpp:1375-1392), solution to c's interior. */
int or_parabolic_solve(or_ctx *ctx, double *c, const double *v, double dt)
{
    const int N1 = ctx->N1, N2 = ctx->N2, M = N1 * N2;
    double *A = malloc(sizeof(double) * 9 * M), *rhs = malloc(sizeof(double) * M);
    double *x = calloc(M, sizeof(double));
    or_assemble_parabolic(ctx, v, dt, A);
    or_rhs_parabolic(ctx, c, v, dt, rhs);
    int its = 0;
    double rr = 0;
    int st = bicgstab(ctx, A, rhs, x, ctx->P.para_eps, &its, &rr);
    ctx->last_para_its = its;
    ctx->last_para_relres = rr;
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) c[S_(ctx, i, j)] = x[i * N2 + j];
    free(A); free(rhs); free(x);
    return st;
}

/* ------------------------------------------------------------------------- */
/* HyperbolicSplitSolver::Solve with SPLIT2ND (HyperbolicSplitSolver.cpp:162-198) */
/* ------------------------------------------------------------------------- */
int or_hyperbolic_solve(or_ctx *ctx, double *c, const double *v, const double *q, double dt)
{
    int st = 0;
    or_source_split(ctx, c, q, 0.5 * dt);
    or_fill_boundaries(ctx, c);
    if (ctx->P.diffusion) st |= or_parabolic_solve(ctx, c, v, 0.5 * dt);
    or_advect_split(ctx, c, v, dt);
    or_source_split(ctx, c, q, 0.5 * dt);
    or_fill_boundaries(ctx, c);
    if (ctx->P.diffusion) st |= or_parabolic_solve(ctx, c, v, 0.5 * dt);
    return st;
}

/* ------------------------------------------------------------------------- */
/* diagnostics (Impec2DSolver.cpp:530-620)                                    */
/* ------------------------------------------------------------------------- */
double or_total_mass(const or_ctx *ctx, const double *c)
{
    double s = 0;
    for (int i = 0; i < ctx->N1; i++)
        for (int j = 0; j < ctx->N2; j++) s += c[S_(ctx, i, j)];
    return s;
}

double or_mass_error(const or_ctx *ctx, const double *c, double t)
{
    double mass = or_total_mass(ctx, c);
    double cpv = 1. / (ctx->dx * ctx->dy);
    double influx = (t - ctx->P.tstart) * ctx->P.I * cpv;
    if (ctx->P.stype == 0) influx *= 2.;
    return fabs((mass + ctx->total_mass_outflux - influx) / influx);
}

double or_divergence_error(const or_ctx *ctx, const double *v, const double *q)
{
    double err = 0;
    for (int i = 0; i < ctx->N1; i++)
        for (int j = 0; j < ctx->N2; j++) {
            double div = (v[V_(ctx, i, j + 1, 1)] - v[V_(ctx, i, j, 1)]) / ctx->dy
                         + (v[V_(ctx, i + 1, j, 0)] - v[V_(ctx, i, j, 0)]) / ctx->dx;
            err += fabs(div - q[i * ctx->N2 + j]);
        }
    return err * ctx->dx * ctx->dx;
}

double or_velocity_boundary_error(const or_ctx *ctx, const double *v)
{
    double e = 0;
    for (int i = 0; i < ctx->N1; i++) {
        e += fabs((double)v[V_(ctx, i, 0, 1)]);
        e += fabs((double)v[V_(ctx, i, ctx->N2, 1)]);
    }
    for (int j = 0; j < ctx->N2; j++) {
        e += fabs((double)v[V_(ctx, 0, j, 0)]);
        e += fabs((double)v[V_(ctx, ctx->N1, j, 0)]);
    }
    return e;
}

/* ------------------------------------------------------------------------- */
/* one IMPEC step: the body of Impec2DSolver::Solve (Impec2DSolver.cpp:339-442) */
/* `tnext` is the next output time (<= 0: none).  Returns the step status;   */
/* *dt_out receives the step size.                                           */
/* ------------------------------------------------------------------------- */
int or_impec_step(or_ctx *ctx, double *c, double *p, double *v, const double *q, double t, double tnext,
                  int warm, double *dt_out)
{
    or_fill_boundaries(ctx, c);
    int st = or_pressure_solve(ctx, c, p, q, warm);
    or_calc_velocities(ctx, p, v, c);
    double dt = or_max_stable_timestep(ctx, v);
    if (tnext > 0 && t + dt >= tnext) dt = tnext - t;
    st |= 2 * or_hyperbolic_solve(ctx, c, v, q, dt);
    *dt_out = dt;
    return st;
}

/* ------------------------------------------------------------------------- */
/* source set-up (Util.cpp:444-586; Impec2DSolver.cpp:241-280)                */
/* ------------------------------------------------------------------------- */
/* InitOneCellParallel / InitOneCellDiagonal (Util.cpp:565-586). */
void or_init_one_cell(const or_ctx *ctx, double *q)
{
    const int M = ctx->N1 * ctx->N2, N1 = ctx->N1, N2 = ctx->N2;
    double val = ctx->P.I / (ctx->dx * ctx->dy);
    for (int k = 0; k < M; k++) q[k] = 0;
    if (ctx->P.stype == 0) {
        q[0] = val; q[M - 1] = val;
        q[N2 - 1] = -val; q[(N1 - 1) * N2] = -val;
    } else {
        q[0] = val; q[M - 1] = -val;
    }
}

/* ExactVolFracCircle (Util.cpp:444-487). */
double or_exact_vol_frac_circle(double x1, double y1, double r, double dx, double dy)
{
    if (x1 * x1 + y1 * y1 < r * r) {
        double y2 = y1 + dy;
        double x2 = x1 + dx;
        if (x2 * x2 + y2 * y2 > r * r) {
            double xs = sqrt(r * r - y1 * y1);
            double ys = sqrt(r * r - x1 * x1);
            double area = 0;
            double xub, xlb, yub;
            if (ys <= y2 && xs <= x2) { xlb = x1; xub = xs; yub = ys; }
            else if (ys < y2 && xs > x2) { yub = ys; xlb = x1; xub = x2; }
            else if (ys > y2 && xs < x2) {
                yub = y2; xlb = sqrt(r * r - yub * yub); xub = xs; area += (xlb - x1) * dy;
            } else {
                yub = y2; xlb = sqrt(r * r - yub * yub); xub = x2; area += dy * (xlb - x1);
            }
            (void)yub;
            double t1 = asin(xlb / r);
            double t2 = asin(xub / r);
            area += r * r / 2 * (sin(2 * t2) / 2. + t2 - sin(2 * t1) / 2. - t1);
            area -= y1 * (xub - xlb);
            return area / (dx * dy);
        }
        return 1.;
    }
    return 0;
}

/* InitQuarterCircleDiagonal / Parallel (Util.cpp:494-559) with
 * r = well_rc (x_ub - x_lb), val = 4 I / (pi r^2) (Impec2DSolver.cpp:241-247). */
void or_init_circle(const or_ctx *ctx, double *q)
{
    const int N1 = ctx->N1, N2 = ctx->N2;
    double r = ctx->P.well_rc * (ctx->P.x_ub - ctx->P.x_lb);
    double val = ctx->P.I * 4. / (M_PI * r * r);
    double *qI = malloc(sizeof(double) * N1 * N2);
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++)
            qI[i * N2 + j] = val * or_exact_vol_frac_circle(ctx->P.x_lb + i * ctx->dx, ctx->P.y_lb + j * ctx->dy,
                                                            r, ctx->dx, ctx->dy);
    /* reverseSelf over the ghosted extent maps interior index k to N-1-k */
#define QI(i, j) qI[(i) * N2 + (j)]
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) {
            double v;
            if (ctx->P.stype == 1) {
                v = QI(i, j) - QI(N1 - 1 - i, N2 - 1 - j);
            } else {
                v = QI(i, j) + QI(N1 - 1 - i, N2 - 1 - j);
                v = v - QI(i, N2 - 1 - j);
                v = v - QI(N1 - 1 - i, j);
            }
            q[i * N2 + j] = v;
        }
#undef QI
    free(qI);
}

/* Stationary V-cycle iteration on the unpinned Neumann problem A_N x = b - mean(b)
 * for c (ghost-filled): x <- x + B(b - A_N x).  Writes the residual 2-norm
 * after each cycle into hist[0..ncyc-1].  Used to measure the V-cycle's
 * convergence factor. */
int or_mg_neumann_cycles(or_ctx *ctx, const double *c, const double *b_in, int ncyc, double *hist)
{
    const int N1 = ctx->N1, N2 = ctx->N2, M = N1 * N2;
    double *AN = malloc(sizeof(double) * 9 * M), *b = malloc(sizeof(double) * M);
    double *x = calloc(M, sizeof(double)), *r = malloc(sizeof(double) * M);
    double a[9], f[9];
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) {
            weighted_coeffs(ctx, c, i, j, a);
            fold_boundary(i, j, N1, N2, a, f);
            for (int k = 0; k < 9; k++) AN[k * M + i * N2 + j] = f[k];
        }
    memcpy(b, b_in, sizeof(double) * M);
    project_mean(b, M);
    or_mg mg;
    mg_setup(ctx, &mg, AN, c);
    for (int it = 0; it < ncyc; it++) {
        or_matvec9(N1, N2, AN, x, r);
        for (int k = 0; k < M; k++) r[k] = b[k] - r[k];
        memcpy(mg.L[0].b, r, sizeof(double) * M);
        project_mean(mg.L[0].b, M);
        vcycle(&mg, 0);
        for (int k = 0; k < M; k++) x[k] += mg.L[0].x[k];
        project_mean(x, M);
        or_matvec9(N1, N2, AN, x, r);
        for (int k = 0; k < M; k++) r[k] = b[k] - r[k];
        hist[it] = sqrt(dot(r, r, M));
    }
    mg_free(&mg);
    free(AN); free(b); free(x); free(r);
    return 0;
}
