// impec_output.cpp -- plot files, the parallel-to-diagonal projection and the
// quarter-circle wells (see impec_output.hpp).
#include "impec_output.hpp"

#include <cmath>
#include <fstream>
#include <iostream>

namespace impec_host {

bool PrintDToPlotFile(const std::string &fname, const Array2d &a, double x_lb, double y_lb, double dx, double dy) {
    std::ofstream file(fname.c_str());
    if (!file.is_open()) {
        std::cerr << "PRINTER-ERROR: Unable to write to file: " << fname << std::endl;
        return false;
    }
    const double *d = a.data();
    const int b = a.bounds(), ld = a.ld();
    const std::string SEP = " ";
    for (int i = 0; i < a.n1(); i++) {
        for (int j = 0; j < a.n2(); j++) {
            const double x = x_lb + (i + 0.5) * dx;
            const double y = y_lb + (j + 0.5) * dy;
            file << x << SEP << y << SEP << d[(size_t)(i + b) * ld + (j + b)] << std::endl;
        }
        file << std::endl;
    }
    return true;
}

bool PrintVelToPlotFile(const std::string &ufile, const std::string &vfile, const Array3d &v, double x_lb,
                        double y_lb, double dx, double dy) {
    std::ofstream uf(ufile.c_str()), vf(vfile.c_str());
    if (!uf.is_open() || !vf.is_open()) {
        std::cerr << "PRINTER-ERROR: Unable to write to file: " << ufile << " or " << vfile << std::endl;
        return false;
    }
    const double *d = v.data();
    const int b = v.bounds(), ld = v.ld(), N1 = v.n1(), N2 = v.n2();
    auto at = [&](int i, int j, int k) { return d[((size_t)(i + b) * ld + (j + b)) * 2 + k]; };
    const std::string SEP = " ";
    for (int i = 0; i <= N1; i++) {
        for (int j = 0; j < N2; j++) uf << x_lb + i * dx << SEP << y_lb + (j + 0.5) * dy << SEP << at(i, j, 0) << std::endl;
        uf << std::endl;
    }
    for (int i = 0; i < N1; i++) {
        for (int j = 0; j <= N2; j++) vf << x_lb + (i + 0.5) * dx << SEP << y_lb + j * dy << SEP << at(i, j, 1) << std::endl;
        vf << std::endl;
    }
    return true;
}

// ParallelProjector (ParallelProjector.cpp:51-199): rotate the diagonal grid's
// cell centres by -pi/4 into the parallel grid, fold y < 0 by symmetry, and
// interpolate bilinearly between the four surrounding parallel cell centres
// (ghost cells included at the upper edges)
Projection ProjectParallelToDiagonal(const Array2d &par, double x_lb, double y_lb, double x_ub, double y_ub) {
    Projection P;
    const int p_N1 = par.n1(), p_N2 = par.n2(), b = par.bounds(), ld = par.ld();
    P.d_N1 = (int)std::round(p_N1 / std::sqrt(2.));
    P.d_N2 = (int)std::round(p_N2 / std::sqrt(2.));
    P.d_dx = (x_ub - x_lb) / (std::sqrt(2.) * P.d_N1);
    P.d_dy = (y_ub - y_lb) / (std::sqrt(2.) * P.d_N2);
    const double p_dx = (x_ub - x_lb) / p_N1, p_dy = (y_ub - y_lb) / p_N2;
    const double angle = -M_PI / 4.;
    const double *d = par.data();
    auto at = [&](int i, int j) { return d[(size_t)(i + b) * ld + (j + b)]; };
    auto centred = [](double lb, double h, int idx) { return lb + (idx + 0.5) * h; };
    P.values.assign((size_t)P.d_N1 * P.d_N2, 0.);
    for (int i = 0; i < P.d_N1; i++)
        for (int j = 0; j < P.d_N2; j++) {
            const double dxc = centred(x_lb, P.d_dx, i), dyc = centred(y_lb, P.d_dy, j);
            const double x = std::cos(angle) * dxc - std::sin(angle) * dyc;
            double y = std::sin(angle) * dxc + std::cos(angle) * dyc;
            y = std::fabs(y);
            if (x > x_ub || x < x_lb || y > y_ub || y < y_lb)
                std::cerr << "ParallelProjector: OUT OF RANGE: (" << x << " , " << y << ")" << std::endl;
            const int x00 = (int)std::floor(x / (x_ub - x_lb) * p_N1), y00 = (int)std::floor(y / (y_ub - y_lb) * p_N2);
            const int x11 = x00 + 1, y11 = y00 + 1;
            const double x_00 = centred(x_lb, p_dx, x00), x_11 = centred(x_lb, p_dx, x11);
            const double y_00 = centred(y_lb, p_dy, y00), y_11 = centred(y_lb, p_dy, y11);
            const double w = 1. / ((x_11 - x_00) * (y_11 - y_00));
            P.values[(size_t)i * P.d_N2 + j] = w * (x_11 - x) * (y_11 - y) * at(x00, y00) +
                                              w * (x - x_00) * (y_11 - y) * at(x11, y00) +
                                              w * (x_11 - x) * (y - y_00) * at(x00, y11) +
                                              w * (x - x_00) * (y - y_00) * at(x11, y11);
        }
    return P;
}

bool PrintProjectionToPlotFile(const std::string &fname, const Projection &p, double x_lb, double y_lb) {
    std::ofstream file(fname.c_str());
    if (!file.is_open()) {
        std::cerr << "PRINTER-ERROR: Unable to write to file: " << fname << std::endl;
        return false;
    }
    const std::string SEP = " ";
    for (int i = 0; i < p.d_N1; i++) {
        for (int j = 0; j < p.d_N2; j++)
            file << x_lb + (i + 0.5) * p.d_dx << SEP << y_lb + (j + 0.5) * p.d_dy << SEP
                 << p.values[(size_t)i * p.d_N2 + j] << std::endl;
        file << std::endl;
    }
    return true;
}

// ExactVolFracCircle (Util.cpp:444-487): the cell's area inside the circle is
// the circular segment between the two crossing abscissae (by the integral of
// sqrt(r^2 - x^2), expressed through asin) above y1, plus the rectangle the
// circle covers fully when it leaves through the top edge
double ExactVolFracCircle(double x1, double y1, double r, double dx, double dy) {
    if (!(x1 * x1 + y1 * y1 < r * r)) return 0.;
    const double x2 = x1 + dx, y2 = y1 + dy;
    if (!(x2 * x2 + y2 * y2 > r * r)) return 1.;
    const double xs = std::sqrt(r * r - y1 * y1);   // crossing of y = y1
    const double ys = std::sqrt(r * r - x1 * x1);   // crossing of x = x1
    double area = 0., xlb, xub;
    if (ys <= y2 && xs <= x2) {          // enters through the left edge, leaves through the bottom
        xlb = x1;
        xub = xs;
    } else if (ys < y2 && xs > x2) {     // left edge to right edge
        xlb = x1;
        xub = x2;
    } else if (ys > y2 && xs < x2) {     // top edge to bottom edge
        xlb = std::sqrt(r * r - y2 * y2);
        xub = xs;
        area += (xlb - x1) * dy;
    } else {                             // top edge to right edge
        xlb = std::sqrt(r * r - y2 * y2);
        xub = x2;
        area += dy * (xlb - x1);
    }
    const double t1 = std::asin(xlb / r), t2 = std::asin(xub / r);
    area += r * r / 2 * (std::sin(2 * t2) / 2. + t2 - std::sin(2 * t1) / 2. - t1);
    area -= y1 * (xub - xlb);
    return area / (dx * dy);
}

// InitQuarterCircleDiagonal / Parallel (Util.cpp:494-559): the exact quarter
// disc at the origin, mirrored to the other wells (the reversal of the ghosted
// array maps interior index k to N - 1 - k)
void InitQuarterCircle(const SimParams &sp, std::vector<double> &q, double r, double val) {
    const int N1 = sp.N1, N2 = sp.N2;
    std::vector<double> qI((size_t)N1 * N2);
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++)
            qI[(size_t)i * N2 + j] = val * ExactVolFracCircle(sp.x_lb + i * sp.dx, sp.y_lb + j * sp.dy, r, sp.dx, sp.dy);
    auto QI = [&](int i, int j) { return qI[(size_t)i * N2 + j]; };
    q.assign((size_t)N1 * N2, 0.);
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < N2; j++) {
            double v;
            if (sp.stype == DIAGONAL) {
                v = QI(i, j) - QI(N1 - 1 - i, N2 - 1 - j);
            } else {
                v = QI(i, j) + QI(N1 - 1 - i, N2 - 1 - j);
                v = v - QI(i, N2 - 1 - j);
                v = v - QI(N1 - 1 - i, j);
            }
            q[(size_t)i * N2 + j] = v;
        }
}

}  // namespace impec_host
