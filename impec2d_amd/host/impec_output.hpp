// impec_output.hpp -- the output path of Impec2DSolver::LogResultsMatch
// (Impec2DSolver.cpp:622-717): "x y value" plot files (Printer.cpp:98-188) and
// the bilinear projection of a parallel-grid field onto the diagonal grid
// (ParallelProjector.cpp:91-199), and the quarter-circle well sources
// (Util.cpp:444-559).  Host-side; run on downloaded fields at logging time.
#pragma once

#include "impec_solvers.hpp"

#include <string>
#include <vector>

namespace impec_host {

// Printer::PrintDToPlotFile: one "x y value" line per interior cell centre, a
// blank line after every i
bool PrintDToPlotFile(const std::string &file, const Array2d &a, double x_lb, double y_lb, double dx, double dy);
// Printer::PrintVelToPlotFile: u at x-faces (i in [0, N1], j in [0, N2)) and v at
// y-faces (i in [0, N1), j in [0, N2]) into two files
bool PrintVelToPlotFile(const std::string &ufile, const std::string &vfile, const Array3d &v, double x_lb,
                        double y_lb, double dx, double dy);

// ParallelProjector: the diagonal grid (d_N1 = round(N1 / sqrt 2) cells a side,
// spacing (x_ub - x_lb) / (sqrt 2 d_N1)) filled by bilinear interpolation of
// the parallel-grid field `par` at the rotated cell centres; out holds
// d_N1 x d_N2 values, row-major
struct Projection {
    int d_N1 = 0, d_N2 = 0;
    double d_dx = 0., d_dy = 0.;
    std::vector<double> values;
};
Projection ProjectParallelToDiagonal(const Array2d &par, double x_lb, double y_lb, double x_ub, double y_ub);
bool PrintProjectionToPlotFile(const std::string &file, const Projection &p, double x_lb, double y_lb);

// ExactVolFracCircle: the fraction of cell [x1, x1+dx] x [y1, y1+dy] inside the
// circle of radius r about the origin (first quadrant)
double ExactVolFracCircle(double x1, double y1, double r, double dx, double dy);
// InitQuarterCircleDiagonal / InitQuarterCircleParallel into q (N1 * N2, i * N2 + j)
void InitQuarterCircle(const SimParams &sp, std::vector<double> &q, double r, double val);

}  // namespace impec_host
