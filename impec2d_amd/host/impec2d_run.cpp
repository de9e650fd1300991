// impec2d_run.cpp -- command-line driver over the C++ solver classes, in the
// role of the reference's Driver.cpp + Impec2DSolver::Run: builds SimParams,
// runs the step loop of Impec2DSolver::Solve and optionally writes the final
// fields and diagnostics.
//
//   impec2d_run --config file.cfg [options below, which override the file]
//   impec2d_run [--N n] [--N1 n --N2 n] [--steps k] [--tfinal t] [--stype diagonal|parallel]
//               [--nu x] [--mu-res x] [--mu-inj x] [--cfl x] [--eps x] [--para-eps x]
//               [--diffusion] [--alpha-l x] [--alpha-t x] [--d-m x] [--alpha-absolute]
//               [--upwind] [--split alternating|x_first|y_first|strang|parallel]
//               [--meanconc] [--porosity x] [--guess zero|previous] [--quiet]
//               [--checks] [--dump file] [--diagnostics file] [--recovery file]
//
// --config reads an Impec2D configuration file (libconfig syntax, the keys of
// SimParams.cpp:45-463) and, as the reference does, writes the final fields
// and diagnostics.dat into its results_dir and, with log_to_file, the
// concentration every log_skip_steps steps into log_dir.  A file with
// is_test = true runs the convergence test its test section names
// (Impec2DSolver::TestComponent) and writes the norm and plot files there.
// --dump writes c, p (each (N1+4) x (N2+4)) and v ((N1+5) x (N2+5) x 2) as raw
// little-endian doubles, in that order.  --recovery writes the long-run
// measures of every step (t, produced mass, recovered fraction, produced
// concentration; impec2d_amd/recovery.py defines them) as text.  Exit status
// 0 on success, 1 on a solver error, 2 on a usage or configuration error, 3
// when no GPU can be used.
#include "impec_config.hpp"
#include "impec_solvers.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

using namespace impec_host;

static void usage() {
    std::cerr << "usage: impec2d_run [--config file.cfg] [--N n] [--steps k] [--tfinal t] [--stype diagonal|parallel] [--nu x]\n"
                 "                   [--diffusion] [--alpha-l x] [--alpha-t x] [--d-m x] [--alpha-absolute]\n"
                 "                   [--upwind] [--split mode] [--meanconc] [--guess zero|previous]\n"
                 "                   [--quiet] [--checks] [--dump file] [--diagnostics file] [--recovery file]\n";
}

int main(int argc, char **argv) {
    SimParams sp;
    sp.tfinal = 1e30;
    sp.max_steps = 10;
    std::string dump, diag, recovery;
    bool print_params = false;
    for (int a = 1; a + 1 < argc; a++)   // a configuration file first; the options override it
        if (std::string(argv[a]) == "--config") {
            std::string err;
            if (!ReadConfigFile(argv[a + 1], sp, err)) {
                std::cerr << argv[a + 1] << ": " << err << std::endl;
                return 2;
            }
            sp.max_steps = 0;
            sp.write_results = true;
        }
    for (int a = 1; a < argc; a++) {
        const std::string k = argv[a];
        auto val = [&]() -> const char * {
            if (a + 1 >= argc) {
                usage();
                std::exit(2);
            }
            return argv[++a];
        };
        if (k == "--config") val();
        else if (k == "--N") sp.N1 = sp.N2 = std::atoi(val());
        else if (k == "--N1") sp.N1 = std::atoi(val());
        else if (k == "--N2") sp.N2 = std::atoi(val());
        else if (k == "--steps") sp.max_steps = std::atoi(val());
        else if (k == "--tfinal") sp.tfinal = std::atof(val());
        else if (k == "--stype") sp.stype = std::strcmp(val(), "parallel") == 0 ? PARALLEL : DIAGONAL;
        else if (k == "--nu") sp.nu = std::atof(val());
        else if (k == "--mu-res") sp.mu_res = std::atof(val());
        else if (k == "--mu-inj") sp.mu_inj = std::atof(val());
        else if (k == "--cfl") sp.cfl = std::atof(val());
        else if (k == "--eps") sp.eps = std::atof(val());
        else if (k == "--para-eps") sp.para_eps = std::atof(val());
        else if (k == "--porosity") sp.porosity = std::atof(val());
        else if (k == "--diffusion") sp.diffusion = true;
        else if (k == "--alpha-l") sp.alpha_l = std::atof(val());
        else if (k == "--alpha-t") sp.alpha_t = std::atof(val());
        else if (k == "--d-m") sp.d_m = std::atof(val());
        else if (k == "--alpha-absolute") sp.alpha_absolute = true;
        else if (k == "--upwind") sp.hmethod = UPWIND;
        else if (k == "--meanconc") sp.avgproc = MEANCONC;
        else if (k == "--split") {
            const std::string m = val();
            if (m == "alternating") sp.dim_splitting = ALTERNATING_SWEEPS;
            else if (m == "x_first") sp.dim_splitting = X_SWEEP_FIRST;
            else if (m == "y_first") sp.dim_splitting = Y_SWEEP_FIRST;
            else if (m == "strang") sp.dim_splitting = STRANG;
            else sp.dim_splitting = SYMMETRIC_PARALLEL;
        } else if (k == "--guess") sp.initial_guess = std::strcmp(val(), "previous") == 0 ? IMPEC_GUESS_PREVIOUS
                                                                                          : IMPEC_GUESS_ZERO;
        else if (k == "--quiet") sp.verbose = false;
        else if (k == "--checks") sp.check_mass_conservation = sp.check_noflow_boundary = sp.check_divergence = true;
        else if (k == "--dump") dump = val();
        else if (k == "--results-dir") sp.results_dir = val();
        else if (k == "--log-dir") sp.log_dir = val();
        else if (k == "--print-params") print_params = true;
        else if (k == "--diagnostics") diag = val();
        else if (k == "--recovery") recovery = val();
        else {
            usage();
            return 2;
        }
    }
    if (sp.N1 < 2 || sp.N2 < 2) {
        usage();
        return 2;
    }
    if (print_params) {   // the parameters as read, one "name value" line each; no device needed
        sp.Finalize();
        std::cout.precision(17);
        std::cout << "N1 " << sp.N1 << "\nN2 " << sp.N2 << "\nc_bounds " << sp.c_bounds << "\nx_lb " << sp.x_lb
                  << "\nx_ub " << sp.x_ub << "\ny_lb " << sp.y_lb << "\ny_ub " << sp.y_ub << "\ndx " << sp.dx
                  << "\ndy " << sp.dy << "\ntstart " << sp.tstart << "\ntfinal " << sp.tfinal << "\nI " << sp.I
                  << "\nstype " << (sp.stype == PARALLEL ? "parallel" : "diagonal") << "\neps " << sp.eps
                  << "\npara_eps " << sp.para_eps << "\nmu_res " << sp.mu_res << "\nmu_inj " << sp.mu_inj
                  << "\navgproc " << (sp.avgproc == HARMONIC ? "harmonic" : "mean_concentration")
                  << "\nnu " << sp.nu << "\ncfl " << sp.cfl << "\nporosity " << sp.porosity
                  << "\ninj_fluid_conc " << sp.inj_fluid_conc << "\ndiffusion " << sp.diffusion
                  << "\nalpha_l " << sp.alpha_l << "\nalpha_t " << sp.alpha_t << "\ndim_splitting "
                  << (int)sp.dim_splitting << "\nwell_profile " << (int)sp.well_profile << "\nwell_rc "
                  << sp.well_rc << "\nlog_to_file " << sp.log_to_file << "\nlog_skip_steps " << sp.log_skip_steps
                  << "\nlog_dir " << sp.log_dir << "\nresults_dir " << sp.results_dir
                  << "\ncheck_mass_conservation " << sp.check_mass_conservation << "\ncheck_noflow_boundary "
                  << sp.check_noflow_boundary << "\ncheck_divergence " << sp.check_divergence << "\ntimes";
        for (double t : sp.times) std::cout << " " << t;
        std::cout << "\ncounters";
        for (int c : sp.counters) std::cout << " " << c;
        std::cout << "\nis_test " << sp.is_test;
        if (sp.is_test)
            std::cout << "\ntest_component " << (sp.test_component == ELLIPTIC ? "elliptic" : "hyperbolic")
                      << "\ntestno " << sp.testno << "\ninit_x_cells " << sp.init_x_cells << "\ninit_y_cells "
                      << sp.init_y_cells << "\nlevels " << sp.levels << "\nhyp_test " << (int)sp.hyp_test
                      << "\nhyp_norm " << (int)sp.hyp_norm << "\nnorm_file " << sp.norm_file;
        std::cout << std::endl;
        return 0;
    }
    if (sp.is_test && !dump.empty()) {
        std::cerr << "--dump is not available in test mode (the fields change size between levels)" << std::endl;
        return 2;
    }
    try {
        Impec2DSolver solver(sp);
        const bool ok = solver.Run();
        if (!diag.empty()) solver.WriteDiagnostics(diag);
        if (!recovery.empty()) solver.WriteRecovery(recovery);
        if (!dump.empty()) {
            FILE *f = std::fopen(dump.c_str(), "wb");
            if (!f) {
                std::cerr << "cannot write " << dump << std::endl;
                return 1;
            }
            const Array2d &c = solver.Concentration(), &p = solver.Pressure();
            const Array3d &v = solver.Velocity();
            const size_t ns = (size_t)(sp.N1 + 4) * (sp.N2 + 4), nv = (size_t)(sp.N1 + 5) * (sp.N2 + 5) * 2;
            std::fwrite(c.data(), sizeof(double), ns, f);
            std::fwrite(p.data(), sizeof(double), ns, f);
            std::fwrite(v.data(), sizeof(double), nv, f);
            std::fclose(f);
        }
        return ok ? 0 : 1;
    } catch (const std::runtime_error &e) {
        std::cerr << e.what() << std::endl;
        return 3;
    }
}
