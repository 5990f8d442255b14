#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace tofe {

// 24x24 Q1 hexahedron stiffness, row-major, DOF 3*localnode + component
// (TopOpt::LocalK, FEAnalysis.cpp:869-1050).
void compute_ke(double E0, double nu, double hx, double hy, double hz, double KE[576]);

// Isotropic D (6x6 row-major, Voigt xx,yy,zz,xy,yz,xz) and the shape-function gradients
// dN_n/dx_a at the 8 Gauss points of the box element with det(J) per point (unit weights).
void dmatrix(double E0, double nu, double D[36]);
void shape_gradients(double hx, double hy, double hz, double dNdx[8][3][8], double detJ[8]);

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

// DOF-level interpolation from a coarse node grid (Nf+1)/2 to the fine grid Nf:
// weights (0.5, 1, 0.5) per direction for coarse nodes at even fine indices, each row
// divided by its sum (Interpolation.cpp:123-169, 322-347); CSR with ascending columns.
void build_interpolation(int Nfx, int Nfy, int Nfz, std::vector<int64_t> &rowptr,
                         std::vector<int32_t> &col, std::vector<double> &val);

// Raw mesh files of TopOpt::MeshOut for one rank (TopOpt.cpp:97-360).
int write_mesh_files(const std::string &dir, int nx, int ny, int nz, double hx, double hy,
                     double hz, const double origin[3], int index_bytes);

}  // namespace tofe
