/*
 * impec_hip.h -- C ABI of libimpec_hip.so, the gfx950 IMPEC hot path.
 *
 * Plain C: opaque handle, plain pointers and sizes, int status returns.
 * Host arrays passed across this boundary use the reference's host layouts
 * (pw359/Impec2D src/GenSimDefs.h:22-31, src/Impec2DSolver.cpp:151-182):
 *
 *   scalar field  double[(N1+2b)*(N2+2b)]        row-major, j fastest, (i,j) at
 *                                                 ((i+b)*(N2+2b) + (j+b)), i,j >= -b
 *   velocity      double[(N1+1+2b)*(N2+1+2b)*2]  component last: u = comp 0 on
 *                                                 x-faces, v = comp 1 on y-faces
 *   source q      double[N1*N2]                  index i*N2 + j
 *
 * with b = num_ghostcells (must be 2, the value every shipped config uses).
 * Device state is authoritative: transfers happen only in impec_set_* /
 * impec_get_*.  One host thread per context; a context owns one HIP stream.
 * Every function returns IMPEC_OK (0) or a negative IMPEC_E* code and leaves
 * a message for impec_last_error().  No function calls exit().
 *
 * Each stage entry point names the reference function it replaces.
 */
#ifndef IMPEC_HIP_H
#define IMPEC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define IMPEC_OK 0
#define IMPEC_EINVAL (-1)    /* bad argument or unsupported parameter        */
#define IMPEC_EHIP (-2)      /* HIP runtime error                            */
#define IMPEC_ENOCONV (-3)   /* linear solve hit max_iterations              */
#define IMPEC_ENAN (-4)      /* non-finite value detected                    */
#define IMPEC_ENOMEM (-5)    /* device allocation failed                     */
#define IMPEC_ENODEV (-6)    /* no usable GPU                                */

/* SimParams enums (pw359/Impec2D src/SimParams.h:14-27) */
#define IMPEC_SRC_PARALLEL 0
#define IMPEC_SRC_DIAGONAL 1
#define IMPEC_AVG_HARMONIC 0
#define IMPEC_AVG_MEANCONC 1
#define IMPEC_SPLIT_ALTERNATING 0
#define IMPEC_SPLIT_X_FIRST 1
#define IMPEC_SPLIT_Y_FIRST 2
#define IMPEC_SPLIT_STRANG 3
#define IMPEC_SPLIT_PARALLEL 4
#define IMPEC_HYPER_MUSCL 0
#define IMPEC_HYPER_UPWIND 1
#define IMPEC_PARA_EXPLICIT 0
#define IMPEC_PARA_CN 1
#define IMPEC_PREC_FP64 0    /* plain FP64 multigrid-preconditioned CG       */
#define IMPEC_PREC_MIXED 1   /* FP32 V-cycles inside FP64 defect correction  */
#define IMPEC_PREC_MIXED32 2 /* FP32 V-cycles and FP32 CG vectors, restarted  */
                             /* from the FP64 residual by reliable updates    */
#define IMPEC_GUESS_ZERO 0   /* reference semantics (MGSolver.cpp:573-590)   */
#define IMPEC_GUESS_PREVIOUS 1
#define IMPEC_GUESS_LEARNED 2   /* previous p plus the network's increment   */
                                /* and the last step's change, both scaled   */
#define IMPEC_GUESS_EXTRAPOLATE 3   /* previous p plus the last step's change */

/* Parameter block.  Field meanings follow SimParams (SimParams.h:29-164);
 * fields marked "ext" extend the reference. */
typedef struct impec_params {
    int N1, N2;                /* x_cells, y_cells                             */
    int num_ghostcells;        /* must be 2                                    */
    int source_type;           /* IMPEC_SRC_*                                  */
    int average_procedure;     /* IMPEC_AVG_*                                  */
    int dim_splitting;         /* IMPEC_SPLIT_*                                */
    int hyper_method;          /* IMPEC_HYPER_*                                */
    int diffusion;             /* hyperbolic.diffusion                         */
    int para_method;           /* IMPEC_PARA_*                                 */
    int is_test;               /* periodic BC, advection-only CFL              */
    int alpha_absolute;        /* ext: alpha_l/t are absolute dispersivities   */
    int max_iterations;        /* linear solver cap (reference: 10000)         */
    int precision;             /* ext: IMPEC_PREC_*                            */
    int initial_guess;         /* ext: IMPEC_GUESS_* of the pressure solve     */
    int mg_pre, mg_post;       /* red-black Gauss-Seidel sweeps per V-cycle    */
                               /* leg (default 2 and 2); the streaming cycle  */
                               /* runs 1/1 or 2/2, other counts the one-      */
                               /* kernel-per-pass cycle                        */
    int mg_variant;            /* 0: streaming fused V-cycle where available,  */
                               /* 1: one kernel per smoothing/transfer pass,   */
                               /* 2: streaming legs down to the coarsest level */
                               /* (no single-workgroup bottom kernel)          */
    int mg_rows;               /* rows per wavefront chunk of the fused finest */
                               /* level (0: automatic; rounded up to even)     */
    int cn_guess;              /* ext: initial guess of the dispersion solve:  */
                               /* 0 zero (AssignRHS_Parabolic), 1 current c    */
    int transport32;           /* ext: 1: the parallel-splitting transport     */
                               /* forms reconstructions, half steps and fluxes */
                               /* in FP32 (states, updates, mass sums FP64)    */
    double x_lb, x_ub, y_lb, y_ub;
    double nu;                 /* hyperbolic.nu, 5-pt/rotated stencil weight   */
    double mu_res, mu_inj;
    double porosity, cfl, inj_fluid_conc;
    double alpha_l, alpha_t, dt_diff_limiter;
    double d_m;                /* ext: molecular diffusion                     */
    double eps;                /* elliptic.eps (relative residual, 2-norm)     */
    double para_eps;           /* parabolic.eps                                */
    double I;                  /* general.I, injection rate                    */
    double tstart;             /* general.starttime                            */
    double inner_tol;          /* ext: FP32 inner tolerance (mixed precision)  */
    double cn_ahead;           /* ext: dispersion lookahead: an iteration's    */
                               /* p-launch is queued before the host tests     */
                               /* convergence while the last residual norm^2   */
                               /* exceeds cn_ahead times the target (0: 1e6;   */
                               /* negative: never); results do not depend on it */
    double cg_ahead;           /* ext: pressure lookahead: the next V-cycle's  */
                               /* first leg is queued likewise (0: 1e6;        */
                               /* negative: never); results do not depend on it */
    double reserved_d[1];
} impec_params;

typedef struct impec_stats {
    int last_its;              /* outer iterations of the last pressure solve  */
    int last_vcycles;          /* V-cycles in the last pressure solve          */
    int last_parabolic_its;    /* BiCGStab iterations, last CN solve           */
    int total_its, calls;      /* EllipticSolver::GetAvgIts bookkeeping        */
    int last_restarts;         /* ext: restarts from the true residual in the  */
                               /* last pressure solve                          */
    int last_elim_its;         /* ext: of last_its, those of the elimination   */
                               /* phase (0: the consistent phase converged)    */
    int cn_ahead_misses;       /* ext: dispersion p-launches queued ahead of   */
                               /* the host's convergence test for an iteration */
                               /* that ended its solve (cumulative)            */
    double last_relres;        /* true ||b - A p|| / ||b|| of the pinned system */
    double last_parabolic_relres;
    double total_mass_outflux; /* HyperbolicSolver::GetTotalMassOutflux        */
    double last_guess_alpha;   /* ext: step length given to the learned        */
                               /* increment in the last solve (0: not used)    */
    double last_guess_gamma;   /* ext: step length given to the last step's    */
                               /* change p_n - p_(n-1) (0: not used)           */
    double last_relres_floor;  /* the FP64 rounding floor of last_relres:      */
                               /* 8 u || |b| + |A| |p| || / ||b||; a solve     */
                               /* converges at max(eps, this)                  */
    int cg_ahead_misses;       /* ext: first legs of a V-cycle queued ahead of  */
                               /* the host's convergence test for an iteration */
                               /* that did not run (cumulative)                */
    int reserved_i2;
    double reserved_d[1];
} impec_stats;

typedef struct impec_ctx impec_ctx;

/* ---- lifecycle ---------------------------------------------------------- */
/* Fills p with the defaults of the shipped example configs. */
int impec_params_default(impec_params *p);
/* Replaces the constructors of Impec2DSolver/MGSolver/MUSCLSplitSolver
 * (Impec2DSolver.cpp:125-182, MGSolver.cpp:52-143): allocates device state,
 * c = p = v = 0, q = 0. */
int impec_create(const impec_params *p, impec_ctx **out);
int impec_destroy(impec_ctx *ctx);
/* the message of the context's last failure; NULL: of the last context-free
 * call on this thread */
const char *impec_last_error(const impec_ctx *ctx);
const char *impec_strerror(int code);
int impec_synchronize(impec_ctx *ctx);
/* Number of HIP devices visible (0 when none). */
int impec_device_count(void);

/* ---- host <-> device (layouts above) ----------------------------------- */
int impec_set_concentration(impec_ctx *ctx, const double *c);
int impec_get_concentration(impec_ctx *ctx, double *c);
int impec_set_pressure(impec_ctx *ctx, const double *p);
int impec_get_pressure(impec_ctx *ctx, double *p);
int impec_set_velocity(impec_ctx *ctx, const double *v);
int impec_get_velocity(impec_ctx *ctx, double *v);
/* Impec2DSolver::InitializeSource result (Impec2DSolver.cpp:241-280). */
int impec_set_source(impec_ctx *ctx, const double *q);
/* ext: per-cell K11, K22 on the scalar layout including ghosts (replaces the
 * analytic K_11/K_22 of EllipticSolver.cpp:418-470); NULL restores K = 1;
 * K22 == NULL with K11 != NULL means isotropic. */
int impec_set_permeability(impec_ctx *ctx, const double *K11, const double *K22);
/* Pin coefficient kappa of row (1,1) (MGSolver.cpp:528-548; 1 outside tests). */
int impec_set_pin_kappa(impec_ctx *ctx, double kappa);

/* ---- stages of one IMPEC step -------------------------------------------- */
/* HyperbolicSolver::FillBoundaries -> FillGhostCellsNeumannScalar /
 * FillGhostCellsPeriodicScalar (HyperbolicSolver.cpp:56-61, Util.cpp:215-324). */
int impec_fill_boundaries(impec_ctx *ctx);
/* EllipticSolver::Solve -> MGSolver::Solve_Elliptic_cc_to_c
 * (EllipticSolver.cpp:133-142, MGSolver.cpp:223-248): assembles the
 * nu-weighted 9-point operator of the ghost-filled c, solves for p (interior)
 * to ||b - A p||_2 <= eps ||b||_2, the residual recomputed from p; a residual
 * below the FP64 rounding floor of its own evaluation (last_relres_floor)
 * also counts as converged.  IMPEC_ENOCONV otherwise, with p the best
 * iterate.  its/relres may be NULL. */
int impec_pressure_solve(impec_ctx *ctx, int *its, double *relres);
/* EllipticSolver::CalcVelocities (EllipticSolver.cpp:660-683). */
int impec_velocities(impec_ctx *ctx);
/* HyperbolicSolver::MaxStableTimestep (HyperbolicSolver.cpp:93-198). */
int impec_max_stable_timestep(impec_ctx *ctx, double *dt);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the time step
 * of impec_max_stable_timestep for face velocities the caller holds on the
 * device in the host velocity layout above, written to the device word
 * *dt_dev on `stream` (a hipStream_t; NULL: the null stream) with no host
 * synchronisation; P gives the grid, domain, cfl and dispersion settings.
 * Allocates nothing: the reduction's partials live in a static device array
 * (3 x 2048 doubles per GPU) that every call on a GPU shares, so calls that
 * overlap on one GPU race.  Velocity arrays up to 2^31 - 1 doubles; larger
 * ones, bad grids and null pointers give IMPEC_EINVAL with a message in
 * impec_last_error(NULL). */
int impec_max_stable_timestep_noalloc(const impec_params *P, const double *v_dev, double *dt_dev, void *stream);
/* HyperbolicSolver::SourceSplit -> TRIntegrator::Integrate
 * (HyperbolicSolver.cpp:201-205, TRIntegrator.cpp:49-71). */
int impec_source_split(impec_ctx *ctx, double dt);
/* HyperbolicSplitSolver::AdvectSplit (HyperbolicSplitSolver.cpp:51-106). */
int impec_advect_split(impec_ctx *ctx, double dt);
/* HyperbolicSplitSolver::Advect_X / Advect_Y (HyperbolicSplitSolver.cpp:113-159)
 * with MUSCLSplitSolver or UpwindSplitSolver fluxes. */
int impec_advect_x(impec_ctx *ctx, double dt);
int impec_advect_y(impec_ctx *ctx, double dt);
/* HyperbolicSplitSolver::CalculateIntercellFlux_F / _G of the selected scheme
 * (MUSCLSplitSolver.cpp:102-189, UpwindSplitSolver.cpp:43-61) for the current c
 * (ghosts as filled) and face velocities.  dir 0: F(i,j), i in [0,N1],
 * j in [-2,N2+2), host array (N1+1) x (N2+4), F[i*(N2+4) + j+2]; dir 1: G(i,j),
 * i in [-2,N1+2), j in [0,N2], host array (N1+4) x (N2+1), G[(i+2)*(N2+1) + j]. */
int impec_intercell_flux(impec_ctx *ctx, int dir, double dt, double *F);
/* HyperbolicSolver::DiffusionSplitCN -> MGSolver::Solve_Parabolic_cc_to_c
 * (HyperbolicSolver.cpp:217-223, MGSolver.cpp:952-976); dt is the split step
 * (the caller passes dt/2 as HyperbolicSplitSolver::Solve does). */
int impec_dispersion_cn(impec_ctx *ctx, double dt, int *its);
/* HyperbolicSplitSolver::Solve (HyperbolicSplitSolver.cpp:162-198). */
int impec_hyperbolic_solve(impec_ctx *ctx, double dt);
/* One pass of the Impec2DSolver::Solve loop (Impec2DSolver.cpp:339-388):
 * fill, pressure, velocities, dt (clipped to t_next - t when t + dt >= t_next
 * and t_next > t), transport.  Writes the step size to *dt_out. */
int impec_step(impec_ctx *ctx, double t, double t_next, double *dt_out);

/* ---- diagnostics (Impec2DSolver.cpp:530-620) ------------------------------ */
int impec_total_mass(impec_ctx *ctx, double *mass);
int impec_mass_error(impec_ctx *ctx, double t, double *err);
int impec_divergence_error(impec_ctx *ctx, double *err);
int impec_velocity_boundary_error(impec_ctx *ctx, double *err);
/* max(c), min(c) over the whole ghosted array (Impec2DSolver.cpp:430-435). */
int impec_minmax_concentration(impec_ctx *ctx, double *cmax, double *cmin);
int impec_get_stats(impec_ctx *ctx, impec_stats *st);

/* ---- operator access for parity checks ----------------------------------- */
/* SetMatrixCoefficients_Elliptic (MGSolver.cpp:268-570) for the current c:
 * nine arrays of N1*N2 doubles, stencil order {C,W,SW,S,SE,E,NE,N,NW},
 * including the pinned row (1,1). */
int impec_pressure_operator(impec_ctx *ctx, double *A9);
/* SetMatrixCoefficients_Parabolic + SetRHS_Parabolic (MGSolver.cpp:997-1370)
 * for the current c, v and split step dt. */
int impec_parabolic_operator(impec_ctx *ctx, double dt, double *A9, double *rhs);

/* ---- measurement --------------------------------------------------------- */
/* Per-kernel-class HIP event timing on the context stream (off by default). */
int impec_profile_enable(impec_ctx *ctx, int on);
/* Accumulated launches and milliseconds of one kernel class by name; returns
 * IMPEC_EINVAL for an unknown name. */
int impec_profile_read(impec_ctx *ctx, const char *name, long long *launches, double *ms);
int impec_profile_reset(impec_ctx *ctx);
/* Plain device-to-device copy of nbytes (rounded down to doubles), repeated
 * reps times on a private stream; *ms gets the mean time of one copy.  The
 * copy kernel is the reference point that scales HBM byte counters. */
int impec_copy_calibration(long long nbytes, int reps, double *ms);

/* ---- learned initial guess of the pressure solve ------------------------- */
/* No counterpart in the reference (its SetRHS_Elliptic always starts from
 * zero, MGSolver.cpp:573-590): a CNN of 3x3 convolutions maps (log mobility,
 * previous pressure, source) to the pressure increment; with
 * initial_guess = IMPEC_GUESS_LEARNED the MG-CG solve starts from
 * p_prev + increment and still runs to the same tolerance.  The network is
 * trained on the GPU from pairs the solver generates. */
typedef struct impec_cnn impec_cnn;
typedef struct impec_cnn_config {
    int layers;                /* convolutions: 3->32, (layers-2) x 32->32, 32->1 */
    int channels;              /* hidden channels (32)                         */
    int capacity;              /* training pairs kept in device memory         */
    int crop;                  /* edge of a training crop (multiple of 32)     */
    unsigned long long seed;   /* weight initialisation and sampling           */
    int mse_loss;              /* 1: train on the mean squared error of the    */
                               /* increment; 0 (default): on its gradient      */
                               /* mismatch (H1 seminorm) + 0.01 x L2           */
    int reserved[3];
} impec_cnn_config;
int impec_cnn_config_default(impec_cnn_config *c);
int impec_cnn_create(const impec_cnn_config *c, impec_cnn **out);
int impec_cnn_destroy(impec_cnn *net);
const char *impec_cnn_last_error(const impec_cnn *net);
/* parameters: per layer W[tap][ci][co] (tap = 3*(dy+1) + (dx+1)) then b[co] */
int impec_cnn_num_params(const impec_cnn *net, long long *count);
int impec_cnn_get_params(impec_cnn *net, float *params);
int impec_cnn_set_params(impec_cnn *net, const float *params);
/* output scale: increment = output * std(p_prev) * scale */
int impec_cnn_get_scale(impec_cnn *net, float *scale);
int impec_cnn_set_scale(impec_cnn *net, float scale);
/* host batches, NHWC: x (n,h,w,3) -> y (n,h,w); gradients of the training
 * loss (mse_loss of the config) against t (n,h,w), in the parameter layout,
 * and the loss */
int impec_cnn_forward(impec_cnn *net, const float *x, int n, int h, int w, float *y);
int impec_cnn_gradients(impec_cnn *net, const float *x, const float *t, int n, int h, int w, float *grads,
                        double *loss);
/* training pairs from a context's latest pressure solve (features of the
 * previous pressure and the mobility of that solve; target: the increment),
 * `crops` random crops; then `iters` Adam steps on batches of `batch` pairs */
int impec_cnn_collect(impec_cnn *net, impec_ctx *ctx, int crops);
int impec_cnn_pairs(const impec_cnn *net, int *count);
int impec_cnn_train(impec_cnn *net, int iters, int batch, float lr, double *loss);
/* attach (or detach with NULL) the network used by IMPEC_GUESS_LEARNED; the
 * context keeps the previous pressure from then on */
int impec_set_warmstart(impec_ctx *ctx, impec_cnn *net);

/* ---- ensembles ------------------------------------------------------------ */
/* No counterpart in the reference (one Impec2DSolver runs one problem,
 * Impec2DSolver.cpp:293-466): `count` independent problems, each with its own
 * parameters and size, advanced together -- every pass of an IMPEC step is one
 * launch over all members, and a step replays captured HIP graphs.  Members
 * need nu = 1, no dispersion, parallel splitting, no test mode, initial_guess
 * zero or previous, isotropic K, and share precision, scheme and whether the
 * porosity is 1.  impec_ensemble_member lends member k's context: set its
 * source, permeability and concentration, and read its fields and
 * diagnostics, with the usual calls (do not destroy it).  A member's fields
 * after impec_ensemble_step equal, bit for bit, those impec_step computes on
 * a context of the same parameters (mg_rows included); dt[k] is member k's
 * time step (no output-time clipping) and its[k] its CG iterations. */
typedef struct impec_ensemble impec_ensemble;
int impec_ensemble_create(const impec_params *members, int count, impec_ensemble **out);
int impec_ensemble_destroy(impec_ensemble *e);
const char *impec_ensemble_last_error(const impec_ensemble *e);
int impec_ensemble_member(impec_ensemble *e, int k, impec_ctx **ctx);
/* graphs on (default) or off (eager launches); CG iterations per replay of
 * the iteration graph (1..64, 0 keeps the current value, default 4) */
int impec_ensemble_set_graphs(impec_ensemble *e, int on, int slots);
/* picks the row height of the finest-level multigrid chunks (8, 16, 32 or 64)
 * per member grid size by timing the ensemble's V-cycle `reps` times per
 * candidate; writes each member's choice to rows[k] (may be NULL) and to its
 * mg_rows */
int impec_ensemble_autotune(impec_ensemble *e, int reps, int *rows);
int impec_ensemble_step(impec_ensemble *e, double *dt, int *its);

#ifdef __cplusplus
}
#endif
#endif /* IMPEC_HIP_H */
