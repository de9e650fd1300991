// impec_tests.cpp -- the reference's test mode (is_test = true): the
// manufactured-solution convergence test of the pressure solve
// (EllipticSolver::TestNeumann, EllipticSolver.cpp:155-470) and the periodic
// advection test of the transport (Impec2DSolver::TestAdvection / Advection,
// Impec2DSolver.cpp:719-1015, profiles and norms of Util.cpp:51-147).  Every
// solve and sweep runs in libimpec_hip.so; this file sets up the analytic
// data, loops over the resolutions and writes the norm and plot files.
#include "impec_output.hpp"
#include "impec_solvers.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <sys/time.h>

namespace impec_host {

namespace {

// K_11 / K_22 of the test numbers 1-4 (EllipticSolver.cpp:418-470)
double K11(int testno, double x, double y) {
    const double t = 2.;
    switch (testno) {
    case 1: return 1.;
    case 2: return std::exp(-0.3 * x);
    case 3: return 2. + std::sin(t * x) * std::sin(y);
    default: return 2. + std::sin(t * x) * std::sin(y);
    }
}
double K22(int testno, double x, double y) {
    const double t = 2.;
    switch (testno) {
    case 1: return 1.;
    case 2: return std::exp(-0.3 * y);
    case 3: return 2. + std::sin(t * x) * std::sin(y);
    default: return 2. + std::sin(x) * std::sin(t * y);
    }
}

// qTest1-4 (EllipticSolver.cpp:319-350): -div(K/mu grad p) for p = cos x cos y
double QTest(int testno, double mu_res, double x, double y) {
    const double t = 2.;
    double a1, a2, a1x, a2y;
    switch (testno) {
    case 1:
        return 2. / mu_res * std::cos(x) * std::cos(y);
    case 2:
        a1 = std::exp(-0.3 * x);
        a2 = std::exp(-0.3 * y);
        a1x = -0.3 * std::exp(-0.3 * x);
        a2y = -0.3 * std::exp(-0.3 * y);
        break;
    case 3:
        a1 = 2. + std::sin(t * x) * std::sin(y);
        a2 = 2. + std::sin(t * x) * std::sin(y);
        a1x = t * std::cos(t * x) * std::sin(y);
        a2y = std::cos(y) * std::sin(t * x);
        break;
    default:
        a1 = 2. + std::sin(t * x) * std::sin(y);
        a2 = 2. + std::sin(x) * std::sin(t * y);
        a1x = t * std::cos(t * x) * std::sin(y);
        a2y = t * std::cos(t * y) * std::sin(x);
        break;
    }
    return 1. / mu_res *
           ((a1 + a2) * std::cos(x) * std::cos(y) + a1x * std::sin(x) * std::cos(y) + a2y * std::cos(x) * std::sin(y));
}

double PTest(double x, double y) { return std::cos(x) * std::cos(y); }

// advection profiles (Util.cpp:51-105): cell values at the centre (x, y)
double Profile(const SimParams &sp, double x, double y) {
    switch (sp.hyp_test) {
    case TOPHAT: {
        const double width = 0.3 * (sp.x_ub - sp.x_lb), height = 0.3 * (sp.y_ub - sp.y_lb);
        const double xc = (sp.x_ub - sp.x_lb) / 2., yc = (sp.y_ub - sp.y_lb) / 2.;
        return (std::fabs(x - xc) <= width / 2. && std::fabs(y - yc) <= height / 2.) ? 1. : 0.;
    }
    case GAUSS: {   // cell average of exp(-64 r^2) through erf
        const double xc = (sp.x_ub - sp.x_lb) / 2., yc = (sp.y_ub - sp.y_lb) / 2.;
        const double ix = 1. / 16 * std::sqrt(M_PI) *
                          (-std::erf(8. * xc - 4. * sp.dx - 8. * x) + std::erf(8. * xc + 4. * sp.dx - 8. * x));
        const double iy = 1. / 16 * std::sqrt(M_PI) *
                          (-std::erf(8. * yc - 4. * sp.dy - 8. * y) + std::erf(8. * yc + 4. * sp.dy - 8. * y));
        return 1. / (sp.dx * sp.dy) * ix * iy;
    }
    case HSINE:   // VerticalSine: cell average of a sine in y
        return 1. + 1. / sp.dy * (-1. / (2. * M_PI)) *
                        (std::cos(2 * M_PI * (y + sp.dy / 2.)) - std::cos(2 * M_PI * (y - sp.dy / 2.)));
    default:      // Sinewave: cell average of a product sine wave
        return 1. + 1. / (sp.dx * sp.dy * 4. * M_PI * M_PI) *
                        (std::cos(2. * M_PI * (x + sp.dx / 2.)) - std::cos(2. * M_PI * (x - sp.dx / 2.))) *
                        (std::cos(2. * M_PI * (y + sp.dy / 2.)) - std::cos(2. * M_PI * (y - sp.dy / 2.)));
    }
}

// CalcAdvectionL1Norm / CalcAdvectionL2Norm (Util.cpp:107-147)
double AdvectionNorm(const SimParams &sp, const Array2d &c) {
    double sum = 0.;
    for (int i = 0; i < sp.N1; i++)
        for (int j = 0; j < sp.N2; j++) {
            const double x = sp.x_lb + (i + 0.5) * sp.dx, y = sp.y_lb + (j + 0.5) * sp.dy;
            const double e = Profile(sp, x, y) - c(i, j);
            sum += sp.hyp_norm == L2 ? e * e : std::fabs(e);
        }
    return sp.hyp_norm == L2 ? std::sqrt(sp.dx * sp.dy * sum) : sum * sp.dx * sp.dy;
}

}  // namespace

// ---------------------------------------------------------------------------
// EllipticSolver::TestNeumann
// ---------------------------------------------------------------------------
bool EllipticSolver::TestNeumann() {
    bool stat = true;
    sp->N1 = sp->init_x_cells;
    sp->N2 = sp->init_y_cells;
    std::ofstream norm_stream(sp->norm_file.c_str());
    if (!norm_stream.is_open()) {
        std::cerr << "Cannot write norm-file! Convergence test aborted..." << std::endl;
        return false;
    }
    norm_stream.precision(30);
    const int tno = sp->testno;
    if (tno < 1 || tno > 4) {
        std::cerr << "EllipticSolver::TestNeumann::Initialization> Unknown test no.!" << std::endl;
        return false;
    }
    for (int k = 0; k < sp->levels; k++) {
        std::cout << "-> (N1,N2):\t(" << sp->N1 << " , " << sp->N2 << ")" << std::flush;
        sp->Finalize();
        // the pressure problem of this resolution on its own context: Neumann
        // ghost fills (the test's c is zero), the analytic K, the exact pin
        SimParams lp = *sp;
        lp.is_test = false;
        const int N1 = lp.N1, N2 = lp.N2, b = lp.c_bounds;
        const double dx = lp.dx, dy = lp.dy;
        Device d(lp);
        MGSolver mg(&lp, &d);
        const int ng1 = N1 + 2 * b, ng2 = N2 + 2 * b;
        std::vector<double> k11((size_t)ng1 * ng2), k22((size_t)ng1 * ng2);
        for (int i = -b; i < N1 + b; i++)
            for (int j = -b; j < N2 + b; j++) {
                const double x = lp.x_lb + (i + 0.5) * dx, y = lp.y_lb + (j + 0.5) * dy;
                k11[(size_t)(i + b) * ng2 + (j + b)] = K11(tno, x, y);
                k22[(size_t)(i + b) * ng2 + (j + b)] = K22(tno, x, y);
            }
        const bool iso = tno == 1 || tno == 3;   // K_22 = K_11
        if (!d.ok(impec_set_permeability(d.ctx(), k11.data(), iso ? nullptr : k22.data()), "TestNeumann")) return false;
        {   // MGSolver.cpp:528-548: kappa such that the pinned cell (1,1) takes the exact pressure
            const double x = lp.x_lb + 1.5 * dx, y = lp.y_lb + 1.5 * dy;
            const double kappa = -dx * dy * QTest(tno, lp.mu_res, x, y) / PTest(x, y);
            if (!d.ok(impec_set_pin_kappa(d.ctx(), kappa), "TestNeumann")) return false;
        }
        // InitializeEllipticNeumannTest (EllipticSolver.cpp:357-415)
        Array2d p(N1, N2, b), p_ex(N1, N2, b), c(N1, N2, b);
        Array3d v, v_ex;
        v.resize(N1, N2, b);
        v_ex.resize(N1, N2, b);
        std::vector<double> q((size_t)N1 * N2);
        for (int i = 0; i <= N1; i++)
            for (int j = 0; j < N2; j++) {
                const double xu = lp.x_lb + i * dx, yu = lp.y_lb + (j + 0.5) * dy;
                v_ex(i, j, 0) = K11(tno, xu, yu) / lp.mu_res * std::sin(xu) * std::cos(yu);
            }
        for (int i = 0; i < N1; i++)
            for (int j = 0; j <= N2; j++) {
                const double xv = lp.x_lb + (i + 0.5) * dx, yv = lp.y_lb + j * dy;
                v_ex(i, j, 1) = K22(tno, xv, yv) / lp.mu_res * std::cos(xv) * std::sin(yv);
            }
        for (int i = 0; i < N1; i++)
            for (int j = 0; j < N2; j++) {
                const double x = lp.x_lb + (i + 0.5) * dx, y = lp.y_lb + (j + 0.5) * dy;
                p_ex(i, j) = PTest(x, y);
                q[(size_t)i * N2 + j] = QTest(tno, lp.mu_res, x, y);
            }
        stat = stat && mg.Solve(p, c, q.data());
        std::cout << "\tITS:\t" << mg.GetLastIts() << std::endl;
        mg.CalcVelocities(p, v, c);
        double p_sum = 0., u_sum = 0., v_sum = 0.;
        for (int i = 0; i < N1; i++)
            for (int j = 0; j < N2; j++) p_sum += std::fabs(p(i, j) - p_ex(i, j));
        for (int i = 0; i <= N1; i++)
            for (int j = 0; j < N2; j++) u_sum += std::fabs(v_ex(i, j, 0) - v(i, j, 0));
        for (int i = 0; i < N1; i++)
            for (int j = 0; j <= N2; j++) v_sum += std::fabs(v_ex(i, j, 1) - v(i, j, 1));
        p_sum *= dx * dy;
        u_sum *= dx * dy;
        v_sum *= dx * dy;
        if (k == 0) {
            PrintDToPlotFile(sp->num_result_file, p, lp.x_lb, lp.y_lb, dx, dy);
            PrintDToPlotFile(sp->ex_result_file, p_ex, lp.x_lb, lp.y_lb, dx, dy);
            PrintVelToPlotFile(sp->num_u_file, sp->num_v_file, v, lp.x_lb, lp.y_lb, dx, dy);
            PrintVelToPlotFile(sp->ex_u_file, sp->ex_v_file, v_ex, lp.x_lb, lp.y_lb, dx, dy);
            std::cout << "\t(Writing files " << sp->num_result_file << ", " << sp->ex_result_file << ", "
                      << sp->num_u_file << ", " << sp->ex_u_file << ")" << std::endl;
        }
        norm_stream << std::log(dx) << "\t" << std::log(p_sum) << "\t" << std::log(u_sum) << "\t" << std::log(v_sum)
                    << std::endl;
        sp->N1 *= 2;
        sp->N2 *= 2;
    }
    return stat;
}

// ---------------------------------------------------------------------------
// Impec2DSolver::TestComponent / TestAdvection / Advection
// ---------------------------------------------------------------------------
bool Impec2DSolver::TestComponent() {
    std::cout << "Executing test no. " << sp.testno << " for the ";
    if (sp.test_component == ELLIPTIC) {
        std::cout << "ELLIPTIC Solver " << std::endl;
        return es->TestNeumann();
    }
    std::cout << "Hyperbolic Solver " << std::endl;
    if (sp.testno != 1) {
        std::cerr << "Hyperbolic TEST:: Unknown test number: " << sp.testno << std::endl;
        return false;
    }
    return TestAdvection();
}

bool Impec2DSolver::TestAdvection() {
    bool stat = true;
    sp.N1 = sp.init_x_cells;
    sp.N2 = sp.init_y_cells;
    std::ofstream norm_stream(sp.norm_file.c_str());
    if (!norm_stream.is_open()) {
        std::cerr << "Cannot write norm-file! Convergence test aborted..." << std::endl;
        return false;
    }
    norm_stream.precision(30);
    for (int k = 0; k < sp.levels; k++) {
        sp.Finalize();
        // a context, solvers and arrays of this resolution (the reference
        // resizes its arrays in place)
        delete es;
        delete hs;
        es = nullptr;
        hs = nullptr;
        dev.reset();
        dev.reset(new Device(sp));
        p.resize(sp.N1, sp.N2, sp.c_bounds);
        c.resize(sp.N1, sp.N2, sp.c_bounds);
        v.resize(sp.N1, sp.N2, sp.c_bounds);
        MGSolver *mg = new MGSolver(&sp, dev.get());
        es = mg;
        HyperbolicSplitSolver *h;
        if (sp.hmethod == MUSCL) h = new MUSCLSplitSolver(&sp, dev.get());
        else h = new UpwindSplitSolver(&sp, dev.get());
        h->AssignMGSolver(mg);
        hs = h;
        q.assign((size_t)sp.N1 * sp.N2, 0.);
        std::cout << "->Resolution:\t" << sp.N1 << "  x  " << sp.N2 << std::flush;
        const double lpnorm = Advection();
        std::cout << "\t\tL" << sp.hyp_norm + 1 << ":\t" << lpnorm << std::endl;
        if (!(lpnorm >= 0.)) stat = false;
        norm_stream << std::log(sp.dx) << "\t" << std::log(lpnorm) << std::endl;
        sp.N1 *= 2;
        sp.N2 *= 2;
    }
    return stat;
}

double Impec2DSolver::Advection() {
    // initial profile at the cell centres, ghosts left zero (InitAdvection)
    for (int i = -sp.c_bounds; i < sp.N1 + sp.c_bounds; i++)
        for (int j = -sp.c_bounds; j < sp.N2 + sp.c_bounds; j++) c(i, j) = 0.;
    for (int i = 0; i < sp.N1; i++)
        for (int j = 0; j < sp.N2; j++) c(i, j) = Profile(sp, sp.x_lb + (i + 0.5) * sp.dx, sp.y_lb + (j + 0.5) * sp.dy);
    {
        std::stringstream ss;
        ss << sp.num_result_file << "init" << sp.N1;
        PrintDToPlotFile(ss.str(), c, sp.x_lb, sp.y_lb, sp.dx, sp.dy);
    }
    // CalcAdvectionVelocities: the domain crossed diagonally once per time unit
    const int vb = sp.c_bounds;
    for (int i = -vb; i <= sp.N1 + vb; i++)
        for (int j = -vb; j <= sp.N2 + vb; j++) {
            v(i, j, 0) = -(sp.x_ub - sp.x_lb);
            v(i, j, 1) = sp.y_ub - sp.y_lb;
        }
    {
        std::stringstream su, sv;
        su << sp.ex_u_file << sp.N1;
        sv << sp.ex_v_file << sp.N1;
        PrintVelToPlotFile(su.str(), sv.str(), v, sp.x_lb, sp.y_lb, sp.dx, sp.dy);
    }
    const double dt_max = hs->MaxStableAdvectionTimestep(c, v);
    bool state = dt_max > 0.;
    double t = sp.tstart;
    int counter = 0;
    std::vector<double> times = sp.times;
    if (std::find(times.begin(), times.end(), sp.tfinal) == times.end()) times.push_back(sp.tfinal);
    std::sort(times.begin(), times.end());
    size_t tidx = 0;
    struct timeval start, end;
    gettimeofday(&start, nullptr);
    double dt = dt_max;
    while (state && t < sp.tfinal) {
        counter++;
        if (tidx < times.size() && t + dt >= times[tidx]) {
            dt = times[tidx] - t;
            tidx++;
        } else {
            dt = dt_max;
        }
        state = state && hs->Solve(c, v, q.data(), dt);
        if (!state) {
            std::cerr << "ERROR::HyperbolicSolver: [" << counter << "]\tt:\t\tdt:\t" << dt << std::endl;
            break;
        }
        t += dt;
        DiagState ds{};
        ds.ctr = counter;
        ds.dt = dt;
        ds.t = t;
        ds.its = 0;
        if (sp.verbose && (counter % 100)) OutputState(ds);
        if (sp.log_to_file) {
            if (IsLoggingTime(t)) LogResultsMatch(t, true);
            if (IsLoggingStep(counter) ||
                (sp.log_skip_steps != 0 && (counter == 1 || counter % sp.log_skip_steps == 0)))
                LogResultsMatch(counter, false);
        }
    }
    gettimeofday(&end, nullptr);
    const long mtime = (long)(((end.tv_sec - start.tv_sec) * 1000 + (end.tv_usec - start.tv_usec) / 1000.0) + 0.5);
    std::cout << "\ttime [ms]: " << mtime << std::flush;
    {
        std::stringstream ss;
        ss << sp.num_result_file << sp.N1;
        PrintDToPlotFile(ss.str(), c, sp.x_lb, sp.y_lb, sp.dx, sp.dy);
    }
    if (!state) return -1.;
    return AdvectionNorm(sp, c);
}

}  // namespace impec_host
