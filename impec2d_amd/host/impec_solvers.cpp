// impec_solvers.cpp -- the reference's solver classes over the C ABI (see
// impec_solvers.hpp).  Every numerical operation is one call into
// libimpec_hip.so; this file only orders the calls and keeps host arrays
// coherent with the device.
#include "impec_solvers.hpp"
#include "impec_output.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <typeinfo>

namespace impec_host {

// ---------------------------------------------------------------------------
// SimParams
// ---------------------------------------------------------------------------
void SimParams::Finalize() {
    dx = (x_ub - x_lb) / N1;   // SimParams.cpp: dx = (x_ub - x_lb) / N1
    dy = (y_ub - y_lb) / N2;
}

impec_params SimParams::ToImpec() const {
    impec_params p;
    impec_params_default(&p);
    p.N1 = N1;
    p.N2 = N2;
    p.num_ghostcells = c_bounds;
    p.source_type = stype;
    p.average_procedure = avgproc;
    p.dim_splitting = dim_splitting;
    p.hyper_method = hmethod;
    p.diffusion = diffusion ? 1 : 0;
    p.para_method = ptype;
    p.is_test = is_test ? 1 : 0;
    p.alpha_absolute = alpha_absolute ? 1 : 0;
    p.max_iterations = max_iterations;
    p.precision = precision;
    p.initial_guess = initial_guess;
    p.x_lb = x_lb; p.x_ub = x_ub; p.y_lb = y_lb; p.y_ub = y_ub;
    p.nu = nu;
    p.mu_res = mu_res;
    p.mu_inj = mu_inj;
    p.porosity = porosity;
    p.cfl = cfl;
    p.inj_fluid_conc = inj_fluid_conc;
    p.alpha_l = alpha_l;
    p.alpha_t = alpha_t;
    p.dt_diff_limiter = dt_diff_limit;
    p.d_m = d_m;
    p.eps = eps;
    p.para_eps = para_eps;
    p.I = I;
    p.tstart = tstart;
    return p;
}

// ---------------------------------------------------------------------------
// host arrays
// ---------------------------------------------------------------------------
void Array2d::resize(int n1, int n2, int b) {
    resize(n1 + 2 * b, n2 + 2 * b);
    Idx2d base;
    base[0] = base[1] = -b;
    reindexSelf(base);
}

void Array2d::resize(int e1, int e2) {
    if (dev_) dev_->unbind(this);   // a new shape is no longer the device's field
    e1_ = e1 > 0 ? e1 : 0;
    e2_ = e2 > 0 ? e2 : 0;
    l1_ = l2_ = 0;
    ld_ = e2_;
    buf_.assign((size_t)e1_ * e2_, 0.);
    set_field_shape();
    host_dirty_ = true;
    host_stale_ = false;
}

void Array2d::reindexSelf(const Idx2d &base) {
    l1_ = base[0];
    l2_ = base[1];
    set_field_shape();
}

// the field view: equal ghost widths on both indices
void Array2d::set_field_shape() {
    if (l1_ == l2_ && l1_ <= 0 && e1_ + 2 * l1_ >= 0 && e2_ + 2 * l2_ >= 0) {
        b_ = -l1_;
        n1_ = e1_ - 2 * b_;
        n2_ = e2_ - 2 * b_;
    } else {
        b_ = -1;
        n1_ = n2_ = 0;
    }
}

Array2d &Array2d::operator=(double x) {
    for_write();
    std::fill(buf_.begin(), buf_.end(), x);
    return *this;
}

Array2d::~Array2d() {
    if (dev_) dev_->unbind(this);
}

void Array2d::for_read() const {
    if (host_stale_ && dev_) dev_->download(*this);
}

void Array2d::for_write() {
    for_read();
    host_dirty_ = true;
}

void Array3d::resize(int n1, int n2, int b) {
    n1_ = n1; n2_ = n2; b_ = b; ld_ = n2 + 1 + 2 * b;
    buf_.assign((size_t)(n1 + 1 + 2 * b) * ld_ * 2, 0.);
    host_dirty_ = true;
    host_stale_ = false;
}

Array3d::~Array3d() {
    if (dev_) dev_->unbind(this);
}

void Array3d::for_read() const {
    if (host_stale_ && dev_) dev_->download(*this);
}

void Array3d::for_write() {
    for_read();
    host_dirty_ = true;
}

// ---------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------
Device::Device(const SimParams &sp) {
    params_ = sp.ToImpec();
    const int st = impec_create(&params_, &ctx_);
    if (st != IMPEC_OK) {
        ctx_ = nullptr;
        throw std::runtime_error(std::string("impec_create: ") + impec_strerror(st));
    }
}

Device::~Device() {
    for (Array2d *a : bound_)
        if (a) a->dev_ = nullptr;
    if (bound_v_) bound_v_->dev_ = nullptr;
    if (ctx_) impec_destroy(ctx_);
}

bool Device::ok(int status, const char *where) const {
    if (status == IMPEC_OK) return true;
    std::cerr << "ERROR::" << where << ": " << impec_strerror(status) << " -- " << impec_last_error(ctx_)
              << std::endl;
    return false;
}

bool Device::use(Array2d &a, Field f) {
    if (a.n1_ != params_.N1 || a.n2_ != params_.N2 || a.b_ != params_.num_ghostcells) {
        std::cerr << "ERROR::Device: array extents do not match the simulation" << std::endl;
        return false;
    }
    if (bound_[f] != &a) {
        // a different host array becomes this field: the old one keeps a copy
        if (bound_[f]) {
            bound_[f]->for_read();
            bound_[f]->dev_ = nullptr;
        }
        if (a.dev_) a.dev_->unbind(&a);
        bound_[f] = &a;
        a.dev_ = this;
        a.field_ = f;
        a.host_dirty_ = true;
        a.host_stale_ = false;
    }
    if (a.host_dirty_) {
        const int st = f == C ? impec_set_concentration(ctx_, a.buf_.data()) : impec_set_pressure(ctx_, a.buf_.data());
        if (!ok(st, "Device::upload")) return false;
        a.host_dirty_ = false;
    }
    return true;
}

bool Device::use(const Array3d &vc) {
    Array3d &v = const_cast<Array3d &>(vc);
    if (v.n1_ != params_.N1 || v.n2_ != params_.N2) {
        std::cerr << "ERROR::Device: velocity extents do not match the simulation" << std::endl;
        return false;
    }
    if (bound_v_ != &v) {
        if (bound_v_) {
            bound_v_->for_read();
            bound_v_->dev_ = nullptr;
        }
        if (v.dev_) v.dev_->unbind(&v);
        bound_v_ = &v;
        v.dev_ = this;
        v.host_dirty_ = true;
        v.host_stale_ = false;
    }
    if (v.host_dirty_) {
        if (!ok(impec_set_velocity(ctx_, v.buf_.data()), "Device::upload")) return false;
        v.host_dirty_ = false;
    }
    return true;
}

bool Device::use_source(double const *q) {
    if (!q) return false;
    if (q != q_) {
        if (!ok(impec_set_source(ctx_, q), "Device::upload")) return false;
        q_ = q;
    }
    return true;
}

bool Device::download(const Array2d &ac) const {
    Array2d &a = const_cast<Array2d &>(ac);
    const int st = a.field_ == C ? impec_get_concentration(ctx_, a.buf_.data()) : impec_get_pressure(ctx_, a.buf_.data());
    if (!ok(st, "Device::download")) return false;
    a.host_stale_ = false;
    return true;
}

bool Device::download(const Array3d &vc) const {
    Array3d &v = const_cast<Array3d &>(vc);
    if (!ok(impec_get_velocity(ctx_, v.buf_.data()), "Device::download")) return false;
    v.host_stale_ = false;
    return true;
}

void Device::produced(Array3d &v) {
    if (bound_v_ != &v) {
        if (bound_v_) {
            bound_v_->for_read();
            bound_v_->dev_ = nullptr;
        }
        if (v.dev_) v.dev_->unbind(&v);
        bound_v_ = &v;
        v.dev_ = this;
    }
    v.host_dirty_ = false;
    v.host_stale_ = true;
}

void Device::unbind(Array2d *a) {
    for (Array2d *&b : bound_)
        if (b == a) b = nullptr;
    a->dev_ = nullptr;
}

void Device::unbind(Array3d *v) {
    if (bound_v_ == v) bound_v_ = nullptr;
    v->dev_ = nullptr;
}

impec_stats Device::stats() const {
    impec_stats s;
    impec_get_stats(ctx_, &s);
    return s;
}

// ---------------------------------------------------------------------------
// EllipticSolver / MGSolver
// ---------------------------------------------------------------------------
EllipticSolver::EllipticSolver(SimParams *sp_, Device *dev_) : sp(sp_), dev(dev_) {}
EllipticSolver::~EllipticSolver() {}

bool EllipticSolver::Solve(Array2d &p, Array2d &c, double *q) {
    const bool ok = SolveLinearMultigrid(c, sp->c_bounds, p, sp->c_bounds, q);
    if (ok) {
        total_its += last_its;
        calls++;
    }
    return ok;
}

void EllipticSolver::CalcVelocities(Array2d &p, Array3d &v, Array2d &c) {
    if (!dev->use(c, Device::C) || !dev->use(p, Device::P)) return;
    if (dev->ok(impec_velocities(dev->ctx()), "EllipticSolver::CalcVelocities")) dev->produced(v);
}

void EllipticSolver::SetNu(double nu) {
    if (nu != dev->params().nu)
        std::cerr << "ERROR::EllipticSolver::SetNu: the stencil weight is fixed when the device is created (nu="
                  << dev->params().nu << ")" << std::endl;
}

int EllipticSolver::GetLastIts() { return last_its; }
int EllipticSolver::GetLastParabolicIts() { return dev->stats().last_parabolic_its; }
double EllipticSolver::GetAvgIts() { return calls ? (double)total_its / calls : 0.; }

MGSolver::MGSolver(SimParams *sp_, Device *dev_) : EllipticSolver(sp_, dev_) {}
MGSolver::~MGSolver() {}

bool MGSolver::keep_unconverged(int status, const char *where) {
    if (status != IMPEC_ENOCONV) return dev->ok(status, where);
    if (missed++ == 0)
        std::cerr << "WARNING::" << where << ": " << impec_last_error(dev->ctx())
                  << " -- the iterate is kept, as the reference keeps its solver's (reported once)" << std::endl;
    return true;
}

bool MGSolver::SolveLinearMultigrid(Array2d &c_c, int, Array2d &p_c, int, double const *const q_v) {
    if (!dev->use(c_c, Device::C) || !dev->use_source(q_v)) return false;
    if (!dev->use(p_c, Device::P)) return false;   // the previous p (warm start) when the caller wrote one
    int its = 0;
    double relres = 0.;
    const bool ok = keep_unconverged(impec_pressure_solve(dev->ctx(), &its, &relres), "MGSolver::SolveLinearMultigrid");
    last_its = its;
    last_relres = relres;
    dev->written(p_c);
    return ok;
}

bool MGSolver::Solve_Parabolic_cc_to_c(Array2d &c_c, int, const Array3d &v, int, Array2d const *const ls, int,
                                       const double dt) {
    if (ls) {
        std::cerr << "ERROR::MGSolver::Solve_Parabolic_cc_to_c: level-set weighting is not on this path" << std::endl;
        return false;
    }
    if (!dev->use(c_c, Device::C) || !dev->use(v)) return false;
    int its = 0;
    const bool ok = keep_unconverged(impec_dispersion_cn(dev->ctx(), dt, &its), "MGSolver::Solve_Parabolic_cc_to_c");
    last_parabolic_its = its;
    dev->written(c_c);
    return ok;
}

// ---------------------------------------------------------------------------
// TimeIntegrator / TRIntegrator
// ---------------------------------------------------------------------------
TimeIntegrator::TimeIntegrator(SimParams *sp_, Device *dev_) : sp(sp_), dev(dev_) {}
TimeIntegrator::~TimeIntegrator() {}
double TimeIntegrator::GetTotalMassOutflux() { return dev->stats().total_mass_outflux; }

TRIntegrator::TRIntegrator(SimParams *sp_, Device *dev_) : TimeIntegrator(sp_, dev_) {}
TRIntegrator::~TRIntegrator() {}

int TRIntegrator::Integrate(Array2d &c, double const *const q, const double dt) {
    if (!dev->use(c, Device::C) || !dev->use_source(q)) return -1;
    const bool ok = dev->ok(impec_source_split(dev->ctx(), dt), "TRIntegrator::Integrate");
    dev->written(c);
    return ok ? 0 : -1;
}

// ---------------------------------------------------------------------------
// HyperbolicSolver
// ---------------------------------------------------------------------------
HyperbolicSolver::HyperbolicSolver(SimParams *sp_, Device *dev_) : sp(sp_), dev(dev_) {
    ti = new TRIntegrator(sp_, dev_);   // HyperbolicSolver.cpp:69-71
}

HyperbolicSolver::~HyperbolicSolver() { delete ti; }

double HyperbolicSolver::GetTotalMassOutflux() { return ti->GetTotalMassOutflux(); }

void HyperbolicSolver::FillBoundaries(Array2d &c) {
    if (!dev->use(c, Device::C)) return;
    if (dev->ok(impec_fill_boundaries(dev->ctx()), "HyperbolicSolver::FillBoundaries")) dev->written(c);
}

double HyperbolicSolver::MaxStableTimestep(const Array2d &c, const Array3d &v) {
    // the device applies the reference's selection: advection-only in test
    // mode, otherwise the combined advective (and, without CN, diffusive) limit
    if (!dev->use(v)) return NAN;
    (void)c;
    double dt = NAN;
    dev->ok(impec_max_stable_timestep(dev->ctx(), &dt), "HyperbolicSolver::MaxStableTimestep");
    return dt;
}

double HyperbolicSolver::MaxStableTimestepCombined(const Array2d &c, const Array3d &v) {
    if (sp->is_test)
        std::cerr << "ERROR::HyperbolicSolver: the device computes the advection-only limit in test mode" << std::endl;
    return MaxStableTimestep(c, v);
}

double HyperbolicSolver::MaxStableAdvectionTimestep(const Array2d &c, const Array3d &v) {
    if (!sp->is_test)
        std::cerr << "ERROR::HyperbolicSolver: the device computes the combined limit outside test mode" << std::endl;
    return MaxStableTimestep(c, v);
}

bool HyperbolicSolver::DiffusionSplitCN(Array2d &c, const Array3d &v, const double dt) {
    if (!mg) {
        std::cerr << "ERROR::HyperbolicSolver::DiffusionSplitCN: no MGSolver assigned" << std::endl;
        return false;
    }
    return mg->Solve_Parabolic_cc_to_c(c, sp->c_bounds, v, sp->c_bounds, nullptr, 0, dt);
}

void HyperbolicSolver::AssignMGSolver(MGSolver *m) { mg = m; }

bool HyperbolicSolver::SourceSplit(Array2d &c, double const *const q, const double dt) {
    return ti->Integrate(c, q, dt) == 0;
}

// ---------------------------------------------------------------------------
// HyperbolicSplitSolver
// ---------------------------------------------------------------------------
HyperbolicSplitSolver::HyperbolicSplitSolver(SimParams *sp_, Device *dev_) : HyperbolicSolver(sp_, dev_) {}
HyperbolicSplitSolver::~HyperbolicSplitSolver() {}

bool HyperbolicSplitSolver::Advect_X(Array2d &c, const Array3d &v, const double dt) {
    if (DeviceFluxes()) {
        if (!dev->use(c, Device::C) || !dev->use(v)) return false;
        const bool ok = dev->ok(impec_advect_x(dev->ctx(), dt), "HyperbolicSplitSolver::Advect_X");
        dev->written(c);
        return ok;
    }
    // a subclass's fluxes: the flux array of the reference's shape, then the
    // update of every row, ghost rows included
    const int b = sp->c_bounds;
    Array2d F;
    F.resize(sp->N1 + 1, sp->N2 + 2 * b);
    Idx2d idx;
    idx[0] = 0;
    idx[1] = -b;
    F.reindexSelf(idx);
    F = 0;
    CalculateIntercellFlux_F(F, c, v, dt);
    for (int i = 0; i < sp->N1; i++)
        for (int j = -b; j < sp->N2 + b; j++) c(i, j) += dt / sp->dx * (F(i, j) - F(i + 1, j));
    return true;
}

bool HyperbolicSplitSolver::Advect_Y(Array2d &c, const Array3d &v, const double dt) {
    if (DeviceFluxes()) {
        if (!dev->use(c, Device::C) || !dev->use(v)) return false;
        const bool ok = dev->ok(impec_advect_y(dev->ctx(), dt), "HyperbolicSplitSolver::Advect_Y");
        dev->written(c);
        return ok;
    }
    const int b = sp->c_bounds;
    Array2d G;
    G.resize(sp->N1 + 2 * b, sp->N2 + 1);
    Idx2d idx;
    idx[0] = -b;
    idx[1] = 0;
    G.reindexSelf(idx);
    G = 0;
    CalculateIntercellFlux_G(G, c, v, dt);
    for (int i = -b; i < sp->N1 + b; i++)
        for (int j = 0; j < sp->N2; j++) c(i, j) += dt / sp->dy * (G(i, j) - G(i, j + 1));
    return true;
}

bool HyperbolicSplitSolver::AdvectSplit(Array2d &c, const Array3d &v, double const *const, const double dt) {
    FillBoundaries(c);   // HyperbolicSplitSolver.cpp:55
    if (DeviceFluxes()) {
        if (!dev->use(c, Device::C) || !dev->use(v)) return false;
        const bool ok = dev->ok(impec_advect_split(dev->ctx(), dt), "HyperbolicSplitSolver::AdvectSplit");
        dev->written(c);
        return ok;
    }
    // the sweep orders of HyperbolicSplitSolver.cpp:58-104 over Advect_X / _Y
    bool ret = true;
    switch (sp->dim_splitting) {
    case ALTERNATING_SWEEPS:
        if (prev_first_sweep == 1) {
            ret = Advect_X(c, v, dt) && Advect_Y(c, v, dt);
            prev_first_sweep = 0;
        } else {
            ret = Advect_Y(c, v, dt) && Advect_X(c, v, dt);
            prev_first_sweep = 1;
        }
        break;
    case X_SWEEP_FIRST:
        ret = Advect_X(c, v, dt) && Advect_Y(c, v, dt);
        prev_first_sweep = 0;
        break;
    case Y_SWEEP_FIRST:
        ret = Advect_Y(c, v, dt) && Advect_X(c, v, dt);
        prev_first_sweep = 1;
        break;
    case STRANG:
        ret = Advect_X(c, v, 0.5 * dt) && Advect_Y(c, v, dt) && Advect_X(c, v, 0.5 * dt);
        prev_first_sweep = 0;
        break;
    default: {
        Array2d ct(sp->N1, sp->N2, sp->c_bounds);
        const Array2d &cr = c;
        std::copy(cr.data(), cr.data() + (size_t)cr.extent(0) * cr.extent(1), ct.data());
        ret = Advect_X(c, v, dt) && Advect_Y(c, v, dt) && Advect_Y(ct, v, dt) && Advect_X(ct, v, dt);
        const int b = sp->c_bounds;
        for (int i = -b; i < sp->N1 + b; i++)
            for (int j = -b; j < sp->N2 + b; j++) c(i, j) = 0.5 * (c(i, j) + ct(i, j));
        prev_first_sweep = 0;
    }
    }
    return ret;
}

bool HyperbolicSplitSolver::Solve(Array2d &c, const Array3d &v, double const *const q, const double dt) {
    bool state = SourceSplit(c, q, 0.5 * dt);
    FillBoundaries(c);
    if (state && sp->diffusion && sp->ptype == CRANKNICOLSON) state = DiffusionSplitCN(c, v, 0.5 * dt);
    state = state && AdvectSplit(c, v, q, dt);
    state = state && SourceSplit(c, q, 0.5 * dt);
    FillBoundaries(c);
    if (state && sp->diffusion && sp->ptype == CRANKNICOLSON) state = DiffusionSplitCN(c, v, 0.5 * dt);
    return state;
}

bool HyperbolicSplitSolver::Flux(int dir, Array2d &F, const Array2d &c, const Array3d &v, double dt) {
    const int b = sp->c_bounds;
    const int e1 = dir == 0 ? sp->N1 + 1 : sp->N1 + 2 * b, e2 = dir == 0 ? sp->N2 + 2 * b : sp->N2 + 1;
    const int l1 = dir == 0 ? 0 : -b, l2 = dir == 0 ? -b : 0;
    if (F.extent(0) != e1 || F.extent(1) != e2 || F.lbound(0) != l1 || F.lbound(1) != l2) {
        std::cerr << "ERROR::HyperbolicSplitSolver::CalculateIntercellFlux_" << (dir == 0 ? "F" : "G")
                  << ": the flux array must have extent " << e1 << " x " << e2 << " and base (" << l1 << "," << l2
                  << ")" << std::endl;
        return false;
    }
    Array2d &cc = const_cast<Array2d &>(c);
    if (!dev->use(cc, Device::C) || !dev->use(v)) return false;
    // the host layout of impec_intercell_flux is this array's storage
    return dev->ok(impec_intercell_flux(dev->ctx(), dir, dt, F.data()), "HyperbolicSplitSolver::Flux");
}

MUSCLSplitSolver::MUSCLSplitSolver(SimParams *sp_, Device *dev_) : HyperbolicSplitSolver(sp_, dev_) {
    if (dev_->params().hyper_method != IMPEC_HYPER_MUSCL)
        throw std::runtime_error("MUSCLSplitSolver: the device was created for another scheme");
}
MUSCLSplitSolver::~MUSCLSplitSolver() {}
void MUSCLSplitSolver::CalculateIntercellFlux_F(Array2d &F, const Array2d &c, const Array3d &v, const double dt) {
    Flux(0, F, c, v, dt);
}
void MUSCLSplitSolver::CalculateIntercellFlux_G(Array2d &G, const Array2d &c, const Array3d &v, const double dt) {
    Flux(1, G, c, v, dt);
}
bool MUSCLSplitSolver::DeviceFluxes() const { return typeid(*this) == typeid(MUSCLSplitSolver); }

UpwindSplitSolver::UpwindSplitSolver(SimParams *sp_, Device *dev_) : HyperbolicSplitSolver(sp_, dev_) {
    if (dev_->params().hyper_method != IMPEC_HYPER_UPWIND)
        throw std::runtime_error("UpwindSplitSolver: the device was created for another scheme");
}
UpwindSplitSolver::~UpwindSplitSolver() {}
void UpwindSplitSolver::CalculateIntercellFlux_F(Array2d &F, const Array2d &c, const Array3d &v, const double dt) {
    Flux(0, F, c, v, dt);
}
void UpwindSplitSolver::CalculateIntercellFlux_G(Array2d &G, const Array2d &c, const Array3d &v, const double dt) {
    Flux(1, G, c, v, dt);
}
bool UpwindSplitSolver::DeviceFluxes() const { return typeid(*this) == typeid(UpwindSplitSolver); }

// ---------------------------------------------------------------------------
// Impec2DSolver
// ---------------------------------------------------------------------------
Impec2DSolver::Impec2DSolver(const SimParams &params) : sp(params) {
    sp.Finalize();
    Initialize();
}

// members are destroyed in reverse order: the arrays unbind before dev goes
Impec2DSolver::~Impec2DSolver() {
    delete es;
    delete hs;
}

// Impec2DSolver::Initialize (Impec2DSolver.cpp:123-148): arrays, solvers, source, c
void Impec2DSolver::Initialize() {
    dev.reset(new Device(sp));
    p.resize(sp.N1, sp.N2, sp.c_bounds);
    c.resize(sp.N1, sp.N2, sp.c_bounds);
    v.resize(sp.N1, sp.N2, sp.c_bounds);
    MGSolver *mg = new MGSolver(&sp, dev.get());
    es = mg;
    es->SetNu(sp.nu);
    HyperbolicSplitSolver *h;
    if (sp.hmethod == MUSCL) h = new MUSCLSplitSolver(&sp, dev.get());
    else h = new UpwindSplitSolver(&sp, dev.get());
    h->AssignMGSolver(mg);
    hs = h;
    InitializeSource();
    InitializeConcentration();
}

// Impec2DSolver::InitializeSource (Impec2DSolver.cpp:241-280): one-cell wells
// of +-I/(dx dy) (InitOneCellDiagonal / InitOneCellParallel, Util.cpp:565-586)
// or exact quarter-circle wells of radius well_rc (x_ub - x_lb) and density
// 4 I / (pi r^2) (Util.cpp:444-559); a Gaussian profile leaves q = 0, as there
void Impec2DSolver::InitializeSource() {
    const int N1 = sp.N1, N2 = sp.N2;
    q.assign((size_t)N1 * N2, 0.);
    if (sp.well_profile == CIRCULAR) {
        const double r = sp.well_rc * (sp.x_ub - sp.x_lb);
        InitQuarterCircle(sp, q, r, sp.I * 4. / (M_PI * r * r));
        return;
    }
    if (sp.well_profile != ONECELL) return;
    const double val = sp.I / (sp.dx * sp.dy);
    if (sp.stype == DIAGONAL) {
        q[0] = val;
        q[(size_t)(N1 - 1) * N2 + (N2 - 1)] = -val;
    } else {
        q[0] = val;
        q[(size_t)(N1 - 1) * N2 + (N2 - 1)] = val;
        q[(size_t)(N2 - 1)] = -val;
        q[(size_t)(N1 - 1) * N2] = -val;
    }
}

bool Impec2DSolver::IsLoggingStep(int counter) const {
    return std::find(sp.counters.begin(), sp.counters.end(), counter) != sp.counters.end();
}

bool Impec2DSolver::IsLoggingTime(double t) const {
    return std::find(sp.times.begin(), sp.times.end(), t) != sp.times.end();
}

void Impec2DSolver::LogResultsMatch(double postfix, bool logTime) {
    std::string dir, suffix;
    if (postfix < 0) {   // the end of the simulation
        dir = sp.results_dir;
        suffix = ".dat";
    } else {
        dir = sp.log_dir;
        std::stringstream ss;
        ss << "_";
        if (logTime) ss << "t_";
        ss << postfix;
        suffix = ss.str() + ".dat";
    }
    const std::string cfile = dir + "/concentration" + suffix, pfile = dir + "/pressure" + suffix;
    const bool final = postfix < 0;
    if (sp.stype == PARALLEL) {
        PrintProjectionToPlotFile(cfile, ProjectParallelToDiagonal(c, sp.x_lb, sp.y_lb, sp.x_ub, sp.y_ub), sp.x_lb,
                                  sp.y_lb);
        if (final)
            PrintProjectionToPlotFile(pfile, ProjectParallelToDiagonal(p, sp.x_lb, sp.y_lb, sp.x_ub, sp.y_ub),
                                      sp.x_lb, sp.y_lb);
    } else {
        PrintDToPlotFile(cfile, c, sp.x_lb, sp.y_lb, sp.dx, sp.dy);
        if (final) {
            PrintDToPlotFile(pfile, p, sp.x_lb, sp.y_lb, sp.dx, sp.dy);
            PrintVelToPlotFile(dir + "/velocity_u" + suffix, dir + "/velocity_v" + suffix, v, sp.x_lb, sp.y_lb,
                               sp.dx, sp.dy);
        }
    }
}

void Impec2DSolver::InitializeConcentration() {
    for (int i = -sp.c_bounds; i < sp.N1 + sp.c_bounds; i++)
        for (int j = -sp.c_bounds; j < sp.N2 + sp.c_bounds; j++) c(i, j) = 0.;
}

double Impec2DSolver::CalculateMassError(double t) {
    double e = -1;
    dev->ok(impec_mass_error(dev->ctx(), t, &e), "Impec2DSolver::CalculateMassError");
    return e;
}

double Impec2DSolver::CalculateVelocityBoundaryError() {
    double e = -1;
    dev->ok(impec_velocity_boundary_error(dev->ctx(), &e), "Impec2DSolver::CalculateVelocityBoundaryError");
    return e;
}

double Impec2DSolver::CalculateDivergenceError() {
    double e = -1;
    dev->ok(impec_divergence_error(dev->ctx(), &e), "Impec2DSolver::CalculateDivergenceError");
    return e;
}

// console line of Impec2DSolver::OutputState (Impec2DSolver.cpp:497-527)
void Impec2DSolver::OutputState(const DiagState &ds) {
    std::stringstream ss;
    ss.setf(std::ios_base::fixed, std::ios_base::floatfield);
    ss.precision(6);
    ss << "[" << ds.ctr << "]\t" << "ITS_e: " << ds.its;
    ss << "\tITS_p: " << es->GetLastParabolicIts();
    ss << "\t t: " << ds.t << "\tdt: " << ds.dt << "\tcells/dt: " << ds.fillperdt;
    if (sp.check_mass_conservation) ss << "\tE_m: " << ds.mass_err * 100;
    if (sp.check_noflow_boundary) ss << "\tE_v: " << ds.vel_err;
    if (sp.check_divergence) ss << "\tE_div: " << ds.diff_err;
    std::cout << ss.str() << std::endl;
}

// diagnostics.dat rows (Impec2DSolver.cpp:469-495)
void Impec2DSolver::WriteDiagnostics(const std::string &fname) const {
    std::ofstream file(fname.c_str());
    if (!file.is_open()) {
        std::cerr << "Unable to write to diagnostics-file: " << fname << std::endl;
        return;
    }
    const std::string SEP = "    ";
    for (const DiagState &ds : diagstates) {
        file << ds.ctr << SEP << ds.its << SEP << ds.t << SEP << ds.dt << SEP << ds.fillperdt << SEP;
        if (sp.check_mass_conservation) file << ds.mass_err << SEP;
        if (sp.check_noflow_boundary) file << ds.vel_err << SEP;
        if (sp.check_divergence) file << ds.diff_err << SEP;
        file << std::endl;
    }
}

void Impec2DSolver::WriteRecovery(const std::string &fname) const {
    std::ofstream file(fname.c_str());
    if (!file.is_open()) {
        std::cerr << "Unable to write to recovery-file: " << fname << std::endl;
        return;
    }
    file.precision(17);
    // injected mass per unit time in the mass-balance units of CalculateMassError
    // (Impec2DSolver.cpp:530-540): I / (dx dy), twice for two injectors
    const double rate = sp.I / (sp.dx * sp.dy) * (sp.stype == PARALLEL ? 2. : 1.);
    const double pv = (double)sp.N1 * sp.N2 * sp.porosity;
    file << "# counter t dt outflux injected recovered produced_conc\n";
    double prev_out = 0., prev_inj = 0.;
    for (const DiagState &ds : diagstates) {
        const double inj = (ds.t - sp.tstart) * rate;
        const double d_inj = inj - prev_inj;
        const double cp = d_inj > 0 ? (ds.outflux - prev_out) / d_inj : 0.;
        file << ds.ctr << " " << ds.t << " " << ds.dt << " " << ds.outflux << " " << inj << " "
             << (inj - ds.outflux) / pv << " " << cp << "\n";
        prev_out = ds.outflux;
        prev_inj = inj;
    }
}

// Impec2DSolver::Run (Impec2DSolver.cpp:747-760), without the gnuplot pause
bool Impec2DSolver::Run() { return sp.is_test ? TestComponent() : Solve(); }

// the step loop (Impec2DSolver.cpp:293-466)
bool Impec2DSolver::Solve() {
    bool state = true;
    double t = sp.tstart;
    double dt;
    int counter = 0;
    if (std::find(sp.times.begin(), sp.times.end(), sp.tfinal) == sp.times.end()) sp.times.push_back(sp.tfinal);
    std::sort(sp.times.begin(), sp.times.end());
    size_t tidx = 0;
    double maxc = 0, minc = 10e10;
    while (state && t < sp.tfinal && (sp.max_steps <= 0 || counter < sp.max_steps)) {
        counter++;
        hs->FillBoundaries(c);
        state = state && es->Solve(p, c, q.data());
        if (!state) {
            std::cerr << "ERROR::EllipticSolver: [" << counter << "]\tt:\t" << t << std::endl;
            break;
        }
        es->CalcVelocities(p, v, c);
        dt = hs->MaxStableTimestep(c, v);
        if (tidx < sp.times.size() && t + dt >= sp.times[tidx]) {
            dt = sp.times[tidx] - t;
            tidx++;
        }
        state = state && hs->Solve(c, v, q.data(), dt);
        if (!state) {
            std::cerr << "ERROR::HyperbolicSolver: [" << counter << "]\tt:\t\tdt:\t" << dt << std::endl;
            break;
        }
        t += dt;
        DiagState ds;
        ds.ctr = counter;
        ds.t = t;
        ds.dt = dt;
        ds.its = es->GetLastIts();
        ds.fillperdt = dt * sp.I / (sp.dx * sp.dy);
        ds.outflux = hs->GetTotalMassOutflux();
        ds.mass_err = sp.check_mass_conservation ? CalculateMassError(t) : -1;
        ds.vel_err = sp.check_noflow_boundary ? CalculateVelocityBoundaryError() : -1;
        ds.diff_err = sp.check_divergence ? CalculateDivergenceError() : -1;
        diagstates.push_back(ds);
        if (sp.verbose) OutputState(ds);
        if (sp.log_to_file) {   // Impec2DSolver.cpp:420-428
            if (IsLoggingTime(t)) LogResultsMatch(t, true);
            // log_skip_steps = 0: only the steps of the list
            if (IsLoggingStep(counter) ||
                (sp.log_skip_steps != 0 && (counter == 1 || counter % sp.log_skip_steps == 0)))
                LogResultsMatch(counter, false);
        }
        double cmax = 0, cmin = 0;
        if (dev->ok(impec_minmax_concentration(dev->ctx(), &cmax, &cmin), "Impec2DSolver::Solve")) {
            maxc = std::max(maxc, cmax);
            minc = std::min(minc, cmin);
        }
    }
    if (sp.verbose && counter > 0) {
        std::cout << "overall maximum: " << maxc << std::endl;
        std::cout << "overall minimum: " << minc << std::endl;
        std::cout << "Average no. of Iterations: " << es->GetAvgIts() << std::endl;
    }
    if (sp.write_results) {   // Impec2DSolver.cpp:461-462
        LogResultsMatch(-1, false);
        WriteDiagnostics(sp.results_dir + "/diagnostics.dat");
    }
    return state;
}

}  // namespace impec_host
