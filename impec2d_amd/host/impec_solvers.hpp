// impec_solvers.hpp -- the reference's solver class interfaces over the C ABI
// of impec_hip.h, so that an Impec2DSolver-style driver runs unchanged on one
// MI355X.  Class names, virtual methods, argument meaning and bool/int return
// conventions follow
//   EllipticSolver.h:15-110, MGSolver.h:11-80, HyperbolicSolver.h:14-43,
//   HyperbolicSplitSolver.h:10-25, MUSCLSplitSolver.h:10-23,
//   UpwindSplitSolver.h:8-17, TimeIntegrator.h:6-16, TRIntegrator.h:8-13,
//   Impec2DSolver.h (DiagState, Solve loop).
//
// Device state is authoritative.  Host arrays (Array2d, Array3d) keep the
// reference's Blitz layout (row-major, last index contiguous, base -b) and are
// bound to one device field each; a solver call uploads an argument only when
// its host copy was written since the last transfer, and host reads download
// lazily.  The per-step loop therefore moves no data over PCIe unless the
// caller reads or writes host arrays.
//
// Errors: methods return false / a negative status and print the library's
// message to std::cerr, as the reference's solvers do; nothing calls exit().
#pragma once

#include "impec_hip.h"

#include <memory>
#include <string>
#include <vector>

namespace impec_host {

enum SourceType { PARALLEL = IMPEC_SRC_PARALLEL, DIAGONAL = IMPEC_SRC_DIAGONAL };   // SimParams.h stype
enum AvgProcedure { HARMONIC = IMPEC_AVG_HARMONIC, MEANCONC = IMPEC_AVG_MEANCONC };
enum DimSplitting {   // SimParams.h:21, same order
    ALTERNATING_SWEEPS = IMPEC_SPLIT_ALTERNATING,
    X_SWEEP_FIRST = IMPEC_SPLIT_X_FIRST,
    Y_SWEEP_FIRST = IMPEC_SPLIT_Y_FIRST,
    STRANG = IMPEC_SPLIT_STRANG,
    SYMMETRIC_PARALLEL = IMPEC_SPLIT_PARALLEL
};
enum HyperbolicMethod { MUSCL = IMPEC_HYPER_MUSCL, UPWIND = IMPEC_HYPER_UPWIND };
enum ParabolicMethod { EXPLICIT = IMPEC_PARA_EXPLICIT, CRANKNICOLSON = IMPEC_PARA_CN };
enum WellProfile { ONECELL, CIRCULAR, GAUSSIAN };   // SimParams.h well_profile; GAUSSIAN leaves q = 0
enum SolverTest { ELLIPTIC, HYPERBOLIC };            // SimParams.h:17
enum HypTest { SINE, TOPHAT, GAUSS, HSINE };         // SimParams.h:18
enum HypNorm { L1, L2 };                             // SimParams.h:19

// The SimParams fields the hot path reads (SimParams.h:29-164), with the
// reference's names.
struct SimParams {
    int N1 = 64, N2 = 64, c_bounds = 2;
    double x_lb = 0., x_ub = 1., y_lb = 0., y_ub = 1.;
    double dx = 0., dy = 0.;                     // set by Finalize()
    double tstart = 0., tfinal = 1.;
    std::vector<double> times;                   // output times (the loop clips dt to them)
    int max_steps = 0;                           // 0: run to tfinal
    double cfl = 0.5, nu = 1., mu_res = 1., mu_inj = 0.1, porosity = 1., I = 1., inj_fluid_conc = 1.;
    double alpha_l = 0., alpha_t = 0., dt_diff_limit = 1., d_m = 0.;
    double eps = 1e-10, para_eps = 1e-10;
    SourceType stype = DIAGONAL;
    AvgProcedure avgproc = HARMONIC;
    DimSplitting dim_splitting = SYMMETRIC_PARALLEL;
    HyperbolicMethod hmethod = MUSCL;
    ParabolicMethod ptype = CRANKNICOLSON;
    bool diffusion = false, is_test = false, alpha_absolute = false;
    bool check_mass_conservation = false, check_noflow_boundary = false, check_divergence = false;
    bool verbose = true;                         // console line per step (Impec2DSolver::OutputState)
    int precision = IMPEC_PREC_FP64, initial_guess = IMPEC_GUESS_ZERO, max_iterations = 10000;
    // wells (Impec2DSolver::InitializeSource, Impec2DSolver.cpp:241-280)
    WellProfile well_profile = ONECELL;
    double well_rc = 0.05;                       // circle radius as a fraction of x_ub - x_lb
    // output (Impec2DSolver::LogResultsMatch / WriteDiagnostics)
    bool log_to_file = false;
    int log_skip_steps = 0;                      // 0: only the steps in counters
    std::vector<int> counters;                   // log_counter_list
    bool time_matching = false, log_mass_conservation = false;
    double max_mass_frac_error = 0.;
    std::string log_dir = ".", results_dir = ".";
    bool write_results = false;                  // final fields and diagnostics.dat into results_dir
    // test mode (SimParams.cpp:252-316): the convergence tests of
    // Impec2DSolver::TestComponent, run when is_test is set
    SolverTest test_component = ELLIPTIC;
    int testno = 1, init_x_cells = 0, init_y_cells = 0, levels = 1;
    HypTest hyp_test = SINE;
    HypNorm hyp_norm = L1;
    std::string num_result_file, ex_result_file, norm_file, num_u_file, num_v_file, ex_u_file, ex_v_file;

    void Finalize();                             // dx, dy from the domain and N1, N2
    impec_params ToImpec() const;
};

class Device;

// Index pair of Array2d::reindexSelf (blitz::TinyVector<int,2>, GenSimDefs.h:30).
struct Idx2d {
    int v[2] = {0, 0};
    int &operator[](int k) { return v[k]; }
    int operator[](int k) const { return v[k]; }
};

// Blitz-layout 2-D array, row-major with the second index contiguous.  A
// scalar field has extent (N1+2b) x (N2+2b) and base -b: element (i,j) for
// i in [-b, N1+b), j in [-b, N2+b).  As with Blitz, any extent and base can
// be set (resize(e1, e2) then reindexSelf), which the flux arrays of
// HyperbolicSplitSolver::Advect_X / _Y use (HyperbolicSplitSolver.cpp:113-159).
class Array2d {
  public:
    Array2d() = default;
    Array2d(int n1, int n2, int b) { resize(n1, n2, b); }
    Array2d(const Array2d &) = delete;
    Array2d &operator=(const Array2d &) = delete;
    ~Array2d();
    void resize(int n1, int n2, int b);   // a field: extent (n1+2b) x (n2+2b), base -b
    void resize(int e1, int e2);          // Blitz resize: extent e1 x e2, base 0, zeroed
    void reindexSelf(const Idx2d &base);  // Blitz reindexSelf: lower bounds of both indices
    Array2d &operator=(double x);         // Blitz "A = x": every element
    double &operator()(int i, int j) {
        for_write();
        return buf_[(size_t)(i - l1_) * ld_ + (j - l2_)];
    }
    double operator()(int i, int j) const {
        for_read();
        return buf_[(size_t)(i - l1_) * ld_ + (j - l2_)];
    }
    double *data() {            // element (lbound(0), lbound(1)); marks the host copy as written
        for_write();
        return buf_.data();
    }
    const double *data() const {
        for_read();
        return buf_.data();
    }
    int n1() const { return n1_; }
    int n2() const { return n2_; }
    int bounds() const { return b_; }
    int ld() const { return ld_; }   // row pitch of data()
    int lbound(int d) const { return d == 0 ? l1_ : l2_; }
    int ubound(int d) const { return d == 0 ? l1_ + e1_ - 1 : l2_ + e2_ - 1; }
    int extent(int d) const { return d == 0 ? e1_ : e2_; }

  private:
    friend class Device;
    void for_read() const;
    void for_write();
    void set_field_shape();
    std::vector<double> buf_;
    int n1_ = 0, n2_ = 0, b_ = 0, ld_ = 0;   // field view: cells and ghost width (b_ < 0: not a field)
    int l1_ = 0, l2_ = 0, e1_ = 0, e2_ = 0;  // lower bounds and extents
    Device *dev_ = nullptr;
    int field_ = -1;
    mutable bool host_stale_ = false;   // the device copy is newer
    bool host_dirty_ = true;            // the host copy was written since the last upload
};

// Blitz-layout face velocities: extent (N1+1+2b) x (N2+1+2b) x 2, component
// last (Array3d v of Impec2DSolver.cpp:151-182).
class Array3d {
  public:
    Array3d() = default;
    Array3d(const Array3d &) = delete;
    Array3d &operator=(const Array3d &) = delete;
    ~Array3d();
    void resize(int n1, int n2, int b);
    double &operator()(int i, int j, int k) {
        for_write();
        return buf_[((size_t)(i + b_) * ld_ + (j + b_)) * 2 + k];
    }
    double operator()(int i, int j, int k) const {
        for_read();
        return buf_[((size_t)(i + b_) * ld_ + (j + b_)) * 2 + k];
    }
    double *data() {
        for_write();
        return buf_.data();
    }
    const double *data() const {
        for_read();
        return buf_.data();
    }
    int n1() const { return n1_; }   // cells: the faces run to n1 and n2
    int n2() const { return n2_; }
    int bounds() const { return b_; }
    int ld() const { return ld_; }   // row pitch of data(), in (u, v) pairs

  private:
    friend class Device;
    void for_read() const;
    void for_write();
    std::vector<double> buf_;
    int n1_ = 0, n2_ = 0, b_ = 0, ld_ = 0;
    Device *dev_ = nullptr;
    mutable bool host_stale_ = false;
    bool host_dirty_ = true;
};

// One impec_ctx (one HIP stream, all fields resident in HBM) shared by the
// solver objects of one simulation.
class Device {
  public:
    enum Field { C = 0, P = 1 };
    explicit Device(const SimParams &sp);   // throws std::runtime_error when creation fails
    ~Device();
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    impec_ctx *ctx() const { return ctx_; }
    const impec_params &params() const { return params_; }
    // true on IMPEC_OK; otherwise prints "<where>: <message>" to std::cerr
    bool ok(int status, const char *where) const;
    // argument protocol of the solver shims
    bool use(Array2d &a, Field f);           // bind; upload if the host copy is newer
    bool use(const Array3d &v);
    bool use_source(double const *q);        // upload when the pointer changes
    void written(Array2d &a) { a.host_stale_ = true; }
    void written(const Array3d &v) { const_cast<Array3d &>(v).host_stale_ = true; }
    void produced(Array3d &v);               // bind v as the device's output (no upload)
    bool download(const Array2d &a) const;
    bool download(const Array3d &v) const;
    void unbind(Array2d *a);
    void unbind(Array3d *v);
    impec_stats stats() const;

  private:
    impec_ctx *ctx_ = nullptr;
    impec_params params_;
    Array2d *bound_[2] = {nullptr, nullptr};
    Array3d *bound_v_ = nullptr;
    double const *q_ = nullptr;
};

// ---------------------------------------------------------------------------
// EllipticSolver / MGSolver
// ---------------------------------------------------------------------------
class EllipticSolver {
  public:
    EllipticSolver(SimParams *sp, Device *dev);
    virtual ~EllipticSolver();
    virtual bool SolveLinearMultigrid(Array2d &c_c, int c_bounds, Array2d &p_c, int p_bounds,
                                      double const *const q_v) = 0;
    bool Solve(Array2d &p, Array2d &c, double *q);            // EllipticSolver.cpp:96-150
    void CalcVelocities(Array2d &p, Array3d &v, Array2d &c);  // EllipticSolver.cpp:660-683
    void SetNu(double nu);                                    // EllipticSolver.cpp:526-534
    int GetLastIts();
    int GetLastParabolicIts();
    double GetAvgIts();
    // the manufactured-solution convergence test (EllipticSolver.cpp:155-311):
    // levels resolutions from init_x_cells x init_y_cells, doubling; L1 errors
    // of p, u and v into norm_file, the fields of the first level into the
    // test's plot files.  No gnuplot window (the reference's GnuPlotter).
    bool TestNeumann();

  protected:
    SimParams *sp;
    Device *dev;
    int last_its = 0;
    int last_parabolic_its = 0;
    int total_its = 0;
    int calls = 0;
};

class MGSolver : public EllipticSolver {
  public:
    MGSolver(SimParams *sp, Device *dev);
    virtual ~MGSolver();
    // pressure: multigrid-preconditioned CG to sp->eps (MGSolver.cpp:268-667)
    virtual bool SolveLinearMultigrid(Array2d &c_c, int c_bounds, Array2d &p_c, int p_bounds,
                                      double const *const q_v);
    // Crank-Nicolson dispersion over dt into c (MGSolver.cpp:952-976); ls must be null
    bool Solve_Parabolic_cc_to_c(Array2d &c_c, int c_bounds, const Array3d &v, int v_bounds,
                                 Array2d const *const ls, int ls_bounds, const double dt);
    double GetLastRelres() const { return last_relres; }
    // solves that missed their tolerance (IMPEC_ENOCONV) and were kept
    int GetMissedSolves() const { return missed; }

  private:
    // the reference keeps whatever iterate its solver returns
    // (Solve_Elliptic_cc_to_c and Solve_Parabolic_cc_to_c return true,
    // MGSolver.cpp:223-248, 952-976): a tolerance below the FP64 rounding
    // floor (the shipped configurations ask for 1e-14) must not stop the run
    bool keep_unconverged(int status, const char *where);
    double last_relres = 0.;
    int missed = 0;
};

// ---------------------------------------------------------------------------
// TimeIntegrator / TRIntegrator
// ---------------------------------------------------------------------------
class TimeIntegrator {
  public:
    TimeIntegrator(SimParams *sp, Device *dev);
    virtual ~TimeIntegrator();
    double GetTotalMassOutflux();
    virtual int Integrate(Array2d &c, double const *const q, const double dt) = 0;

  protected:
    SimParams *sp;
    Device *dev;
};

class TRIntegrator : public TimeIntegrator {
  public:
    TRIntegrator(SimParams *sp, Device *dev);
    virtual ~TRIntegrator();
    // trapezoidal source step over dt (TRIntegrator.cpp:49-71); 0 on success
    virtual int Integrate(Array2d &c, double const *const q, const double dt);
};

// ---------------------------------------------------------------------------
// HyperbolicSolver / HyperbolicSplitSolver / MUSCLSplitSolver / UpwindSplitSolver
// ---------------------------------------------------------------------------
class HyperbolicSolver {
  public:
    HyperbolicSolver(SimParams *sp, Device *dev);
    virtual ~HyperbolicSolver();
    virtual bool Solve(Array2d &c, const Array3d &v, double const *const q, const double dt) = 0;
    double GetTotalMassOutflux();
    void FillBoundaries(Array2d &c);                                   // HyperbolicSolver.cpp:56-61
    double MaxStableTimestepCombined(const Array2d &c, const Array3d &v);   // :122-198
    double MaxStableAdvectionTimestep(const Array2d &c, const Array3d &v);  // :100-121
    double MaxStableTimestep(const Array2d &c, const Array3d &v);           // :93-98
    bool DiffusionSplitCN(Array2d &c, const Array3d &v, const double dt);   // :217-223
    void AssignMGSolver(MGSolver *mg);

  protected:
    bool SourceSplit(Array2d &c, double const *const q, const double dt);  // :201-205
    SimParams *sp;
    Device *dev;

  private:
    TimeIntegrator *ti;
    MGSolver *mg = nullptr;
};

class HyperbolicSplitSolver : public HyperbolicSolver {
  public:
    HyperbolicSplitSolver(SimParams *sp, Device *dev);
    virtual ~HyperbolicSplitSolver();
    // The reference's extension point (HyperbolicSplitSolver.h:19-20).  F(i,j):
    // faces i in [0,N1], j in [-b,N2+b) (extent (N1+1) x (N2+2b), base (0,-b));
    // G(i,j): i in [-b,N1+b), faces j in [0,N2] (extent (N1+2b) x (N2+1), base (-b,0)).
    virtual void CalculateIntercellFlux_F(Array2d &F, const Array2d &c, const Array3d &v, const double dt) = 0;
    virtual void CalculateIntercellFlux_G(Array2d &G, const Array2d &c, const Array3d &v, const double dt) = 0;
    // The sweeps.  With the library's own fluxes (MUSCLSplitSolver and
    // UpwindSplitSolver themselves) they run on the device; a subclass's
    // fluxes are honoured as the reference honours them: the flux array is
    // formed by the virtual and c is updated on the host
    // (HyperbolicSplitSolver.cpp:113-159).
    bool Advect_X(Array2d &c, const Array3d &v, const double dt);   // HyperbolicSplitSolver.cpp:113-134
    bool Advect_Y(Array2d &c, const Array3d &v, const double dt);   // :140-159
    // Source(dt/2), fill, [CN(dt/2)], AdvectSplit(dt), Source(dt/2), fill, [CN(dt/2)] (:162-198)
    virtual bool Solve(Array2d &c, const Array3d &v, double const *const q, const double dt);

  protected:
    bool AdvectSplit(Array2d &c, const Array3d &v, double const *const q, const double dt);   // :51-106
    // the library's device fluxes into F / G of the shapes above
    bool Flux(int dir, Array2d &F, const Array2d &c, const Array3d &v, double dt);
    // true when the flux virtuals are the library's own (the dynamic type is
    // MUSCLSplitSolver or UpwindSplitSolver): the sweeps then stay on the device
    virtual bool DeviceFluxes() const { return false; }

  private:
    int prev_first_sweep = 1;   // Y_SWEEP (HyperbolicSolver.cpp:84): the host sweeps' alternation
};

class MUSCLSplitSolver : public HyperbolicSplitSolver {
  public:
    MUSCLSplitSolver(SimParams *sp, Device *dev);   // the device must run the MUSCL scheme
    virtual ~MUSCLSplitSolver();
    virtual void CalculateIntercellFlux_F(Array2d &F, const Array2d &c, const Array3d &v, const double dt);
    virtual void CalculateIntercellFlux_G(Array2d &G, const Array2d &c, const Array3d &v, const double dt);

  protected:
    virtual bool DeviceFluxes() const;
};

class UpwindSplitSolver : public HyperbolicSplitSolver {
  public:
    UpwindSplitSolver(SimParams *sp, Device *dev);  // the device must run the upwind scheme
    virtual ~UpwindSplitSolver();
    virtual void CalculateIntercellFlux_F(Array2d &F, const Array2d &c, const Array3d &v, const double dt);
    virtual void CalculateIntercellFlux_G(Array2d &G, const Array2d &c, const Array3d &v, const double dt);

  protected:
    virtual bool DeviceFluxes() const;
};

// ---------------------------------------------------------------------------
// Impec2DSolver: the step loop of Impec2DSolver.cpp:293-466
// ---------------------------------------------------------------------------
struct DiagState {   // Impec2DSolver.h
    int ctr;
    int its;
    double t;
    double dt;
    double mass_err;
    double diff_err;
    double vel_err;
    double fillperdt;
    double outflux;   // HyperbolicSolver::GetTotalMassOutflux after the step
};

class Impec2DSolver {
  public:
    explicit Impec2DSolver(const SimParams &params);   // throws std::runtime_error without a device
    virtual ~Impec2DSolver();
    bool Run();                                        // TestComponent() when is_test, else Solve()
    bool Solve();
    bool TestComponent();                              // Impec2DSolver.cpp:719-745
    // the periodic advection convergence test (Impec2DSolver.cpp:773-843):
    // levels resolutions, Advection() each, log(dx) and log(norm) into norm_file
    bool TestAdvection();
    // one resolution of it (Impec2DSolver.cpp:845-990): the profile carried with
    // the constant velocity (-(x_ub - x_lb), y_ub - y_lb) from tstart to tfinal;
    // returns the L1 or L2 error against the initial profile
    double Advection();
    const std::vector<DiagState> &Diagnostics() const { return diagstates; }
    void WriteDiagnostics(const std::string &fname) const;   // diagnostics.dat columns
    // the long-run measures of every step (impec2d_amd/recovery.py defines
    // them): counter, t, dt, outflux, injected, recovered fraction of the
    // resident fluid, produced concentration
    void WriteRecovery(const std::string &fname) const;
    Array2d &Concentration() { return c; }
    Array2d &Pressure() { return p; }
    Array3d &Velocity() { return v; }
    std::vector<double> &Source() { return q; }
    SimParams sp;

  private:
    void Initialize();
    void InitializeSource();          // one-cell or circle wells (Impec2DSolver.cpp:241-280, Util.cpp:444-586)
    void InitializeConcentration();   // c = 0 (Impec2DSolver.cpp:191-233 without an init file)
    // plot files of c (and, at the end, p and v) into log_dir / results_dir;
    // parallel grids are projected onto the diagonal grid first
    // (LogResultsMatch, Impec2DSolver.cpp:622-717)
    void LogResultsMatch(double postfix, bool logTime);
    bool IsLoggingStep(int counter) const;   // Impec2DSolver.cpp:594-604
    bool IsLoggingTime(double t) const;
    double CalculateMassError(double t);
    double CalculateVelocityBoundaryError();
    double CalculateDivergenceError();
    void OutputState(const DiagState &ds);
    std::unique_ptr<Device> dev;
    Array2d p, c;
    Array3d v;
    std::vector<double> q;
    EllipticSolver *es = nullptr;
    HyperbolicSolver *hs = nullptr;
    std::vector<DiagState> diagstates;
};

}  // namespace impec_host
