// A HyperbolicSplitSolver subclass written against the reference's extension
// point (HyperbolicSplitSolver.h:19-20): first-order upwind fluxes restated
// from UpwindSplitSolver.cpp:43-65 with HyperbolicSolver::Flux
// (HyperbolicSolver.cpp:208-210).  The library must honour the override: its
// sweeps with this subclass (host updates through the virtuals) equal the
// library's own UpwindSplitSolver (device sweeps) bit for bit.
//
//   flux_subclass            prints "max_abs_diff <x> <y> <split_parallel> <split_alternating>"
#include "impec_solvers.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>

using namespace impec_host;

class ReferenceUpwind : public HyperbolicSplitSolver {
  public:
    ReferenceUpwind(SimParams *sp_, Device *dev_) : HyperbolicSplitSolver(sp_, dev_) {}
    void CalculateIntercellFlux_F(Array2d &F, const Array2d &c, const Array3d &v, const double dt) override {
        for (int i = 0; i <= sp->N1; i++)
            for (int j = -sp->c_bounds; j < sp->N2 + sp->c_bounds; j++)
                F(i, j) = Upwind(v(i, j, 0), c(i - 1, j), c(i, j));
    }
    void CalculateIntercellFlux_G(Array2d &G, const Array2d &c, const Array3d &v, const double dt) override {
        for (int i = -sp->c_bounds; i < sp->N1 + sp->c_bounds; i++)
            for (int j = 0; j <= sp->N2; j++) G(i, j) = Upwind(v(i, j, 1), c(i, j - 1), c(i, j));
    }

  private:
    double Upwind(double speed, double UL, double UR) { return speed > 0 ? Flux(speed, UL) : Flux(speed, UR); }
    double Flux(double vel, double conc) { return vel * conc / sp->porosity; }
};

static double lcg(unsigned long long &s) {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) / 9007199254740992.0;
}

static void fill(Array2d &c, Array3d &v, const SimParams &sp, unsigned long long seed) {
    const int b = sp.c_bounds;
    for (int i = -b; i < sp.N1 + b; i++)
        for (int j = -b; j < sp.N2 + b; j++) c(i, j) = lcg(seed);
    for (int i = -b; i < sp.N1 + 1 + b; i++)
        for (int j = -b; j < sp.N2 + 1 + b; j++) {
            v(i, j, 0) = 2 * lcg(seed) - 1;
            v(i, j, 1) = 2 * lcg(seed) - 1;
        }
}

static double maxdiff(const Array2d &a, const Array2d &b, const SimParams &sp) {
    double m = 0;
    for (int i = -sp.c_bounds; i < sp.N1 + sp.c_bounds; i++)
        for (int j = -sp.c_bounds; j < sp.N2 + sp.c_bounds; j++) m = std::fmax(m, std::fabs(a(i, j) - b(i, j)));
    return m;
}

int main() {
    SimParams sp;
    sp.N1 = 40;
    sp.N2 = 28;
    sp.hmethod = UPWIND;
    sp.porosity = 0.8;
    sp.Finalize();
    const double dt = 0.2 * sp.dx;
    double d[4];
    for (int k = 0; k < 4; k++) {
        sp.dim_splitting = k == 3 ? ALTERNATING_SWEEPS : SYMMETRIC_PARALLEL;
        Device dl(sp), du(sp);
        UpwindSplitSolver lib(&sp, &dl);
        ReferenceUpwind user(&sp, &du);
        Array2d cl(sp.N1, sp.N2, sp.c_bounds), cu(sp.N1, sp.N2, sp.c_bounds);
        Array3d vl, vu;
        vl.resize(sp.N1, sp.N2, sp.c_bounds);
        vu.resize(sp.N1, sp.N2, sp.c_bounds);
        fill(cl, vl, sp, 7 + k);
        fill(cu, vu, sp, 7 + k);
        std::vector<double> q((size_t)sp.N1 * sp.N2, 0.);
        bool ok = true;
        if (k == 0) ok = lib.Advect_X(cl, vl, dt) && user.Advect_X(cu, vu, dt);
        else if (k == 1) ok = lib.Advect_Y(cl, vl, dt) && user.Advect_Y(cu, vu, dt);
        else   // two steps: the alternating mode changes its first sweep
            for (int s = 0; s < 2 && ok; s++) ok = lib.Solve(cl, vl, q.data(), dt) && user.Solve(cu, vu, q.data(), dt);
        if (!ok) {
            std::fprintf(stderr, "a sweep failed (case %d)\n", k);
            return 1;
        }
        d[k] = maxdiff(cl, cu, sp);
    }
    std::printf("max_abs_diff %.17g %.17g %.17g %.17g\n", d[0], d[1], d[2], d[3]);
    return 0;
}
