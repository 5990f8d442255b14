#pragma once

#include <vector>

namespace tofe {

// 24x24 Q1 hexahedron stiffness, row-major, DOF 3*localnode + component
// (TopOpt::LocalK, FEAnalysis.cpp:869-1050).
void compute_ke(double E0, double nu, double hx, double hy, double hz, double KE[576]);

// Density-filter neighbourhood: offsets (i,j,k) walked k, j, i ascending with
// dist < R, and weights 1 - dist/R (TopOpt::RecFilter, Filter.cpp:21-94).
int filter_template(double dx, double dy, double dz, double R, std::vector<int> &off,
                    std::vector<double> &w);

// Interpolation forms of a fine element inside its 2x2x2 coarse window: per direction
// q[l][L] for the even element, the odd element and the last element of a grid with an
// even node count; Qforms[f][l][L] (8x8, f = fx + 3 fy + 9 fz) and
// KEC[f] = Qd^T KE Qd with Qd = Q (x) I_3 (24x24).
void galerkin_forms(const double KE[576], double Qforms[27 * 64], double KEC[27 * 576]);

// Largest eigenvalue of a symmetric tridiagonal matrix (diagonal a, off-diagonal b) by
// Sturm-sequence bisection.
double tridiag_max_eig(int k, const double *a, const double *b);

}  // namespace tofe
