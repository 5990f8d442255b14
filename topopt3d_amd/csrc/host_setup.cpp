// Host-side set-up computations that run once per context: the element stiffness
// matrix and the density-filter neighbourhood template.
#include "host_setup.hpp"

#include <cmath>
#include <algorithm>
#include <cstring>

namespace tofe {

// Element-local node order (RecMesh.cpp:486-512).
static const int kLN[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

void dmatrix(double E0, double nu, double D[36])
{
  // Isotropic constitutive matrix, Voigt [xx,yy,zz,xy,yz,xz] (RecMesh.cpp:453-462).
  const double c = E0 / ((1 + nu) * (1 - 2 * nu));
  const double G = E0 / (2 * (1 + nu));
  const double t1 = (1 - nu) * c, t2 = G, t3 = nu * c;
  const double v[36] = {t1, t3, t3, 0, 0, 0, t3, t1, t3, 0, 0, 0, t3, t3, t1, 0, 0, 0,
                        0,  0,  0,  t2, 0, 0, 0,  0,  0,  0, t2, 0, 0,  0,  0,  0, 0, t2};
  std::memcpy(D, v, sizeof(v));
}

void shape_gradients(double hx, double hy, double hz, double dNdx[8][3][8], double detJ[8])
{
  const double g = 1 / std::sqrt(3.0);
  const double h[3] = {hx, hy, hz};
  // 2x2x2 Gauss points at +-1/sqrt(3), unit weights, in local-node sign order
  // (GaussPoints FEAnalysis.cpp:900-926; LocalK 869-892).
  for (int q = 0; q < 8; q++) {
    const double xi = (2 * kLN[q][0] - 1) * g, eta = (2 * kLN[q][1] - 1) * g,
                 zeta = (2 * kLN[q][2] - 1) * g;
    double dNdxi[3][8];  // trilinear shape-function derivatives (dN, 936-990)
    for (int n = 0; n < 8; n++) {
      const double sx = 2 * kLN[n][0] - 1, sy = 2 * kLN[n][1] - 1, sz = 2 * kLN[n][2] - 1;
      dNdxi[0][n] = sx / 8 * (1 + sy * eta) * (1 + sz * zeta);
      dNdxi[1][n] = sy / 8 * (1 + sx * xi) * (1 + sz * zeta);
      dNdxi[2][n] = sz / 8 * (1 + sx * xi) * (1 + sy * eta);
    }
    // The element is an axis-aligned box, so J = diag(h/2).
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++)
        for (int n = 0; n < 8; n++) J[a][b] += dNdxi[a][n] * (h[b] * kLN[n][b]);
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double inv[3][3];
    inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int a = 0; a < 3; a++)
      for (int n = 0; n < 8; n++) {
        double s = 0;
        for (int b = 0; b < 3; b++) s += inv[a][b] * dNdxi[b][n];
        dNdx[q][a][n] = s;
      }
    detJ[q] = det;  // rounding of J differs between Gauss points
  }
}

void compute_ke(double E0, double nu, double hx, double hy, double hz, double KE[576])
{
  double Dm[36];
  dmatrix(E0, nu, Dm);
  double dN[8][3][8], det[8];
  shape_gradients(hx, hy, hz, dN, det);
  std::memset(KE, 0, 576 * sizeof(double));
  for (int q = 0; q < 8; q++) {
    const double(&dNdx)[3][8] = dN[q];
    // Strain-displacement matrix B (AssignB, 1024-1046)
    double B[6][24];
    std::memset(B, 0, sizeof(B));
    for (int n = 0; n < 8; n++) {
      B[0][3 * n + 0] = dNdx[0][n];
      B[1][3 * n + 1] = dNdx[1][n];
      B[2][3 * n + 2] = dNdx[2][n];
      B[3][3 * n + 0] = dNdx[1][n];
      B[3][3 * n + 1] = dNdx[0][n];
      B[4][3 * n + 1] = dNdx[2][n];
      B[4][3 * n + 2] = dNdx[1][n];
      B[5][3 * n + 0] = dNdx[2][n];
      B[5][3 * n + 2] = dNdx[0][n];
    }
    double DB[6][24];
    for (int a = 0; a < 6; a++)
      for (int cc = 0; cc < 24; cc++) {
        double s = 0;
        for (int b = 0; b < 6; b++) s += Dm[a * 6 + b] * B[b][cc];
        DB[a][cc] = s;
      }
    for (int r = 0; r < 24; r++)
      for (int cc = 0; cc < 24; cc++) {
        double s = 0;
        for (int a = 0; a < 6; a++) s += B[a][r] * DB[a][cc];
        KE[r * 24 + cc] += s * det[q];
      }
  }
}

int filter_template(double dx, double dy, double dz, double R, std::vector<int> &off,
                    std::vector<double> &w)
{
  // N = (short)(R/dx) per direction (Filter.cpp:32-35)
  const short N[3] = {(short)(R / dx), (short)(R / dy), (short)(R / dz)};
  off.clear();
  w.clear();
  for (int k = -N[2]; k < N[2] + 1; k++)
    for (int j = -N[1]; j < N[1] + 1; j++)
      for (int i = -N[0]; i < N[0] + 1; i++) {
        // dist = sqrt(pow(i dx,2) + pow(j dx,2) + pow(k dx,2)) (Filter.cpp:47), pow(v,2)
        // being the correctly rounded v*v; strict dist < R (48); weight 1 - dist/R (80)
        const double vi = i * dx, vj = j * dy, vk = k * dz;
        const double dist = std::sqrt(vi * vi + vj * vj + vk * vk);
        if (dist < R) {
          off.push_back(i);
          off.push_back(j);
          off.push_back(k);
          w.push_back(1 - dist / R);
        }
      }
  return (int)w.size();
}

void galerkin_forms(const double KE[576], double Qforms[27 * 64], double KEC[27 * 576])
{
  // q[form][fine node l of the element edge][coarse node L of the window]
  static const double q[3][2][2] = {{{1.0, 0.0}, {0.5, 0.5}},   // even element
                                    {{0.5, 0.5}, {0.0, 1.0}},   // odd element
                                    {{1.0, 0.0}, {1.0, 0.0}}};  // last element, single parent
  for (int f = 0; f < 27; ++f) {
    const int fx = f % 3, fy = (f / 3) % 3, fz = f / 9;
    double *Q = Qforms + f * 64;
    for (int l = 0; l < 8; ++l)
      for (int L = 0; L < 8; ++L)
        Q[l * 8 + L] = q[fx][kLN[l][0]][kLN[L][0]] * q[fy][kLN[l][1]][kLN[L][1]] *
                       q[fz][kLN[l][2]][kLN[L][2]];
    double X[576];
    for (int r = 0; r < 24; ++r)
      for (int c = 0; c < 24; ++c) {
        const int L = c / 3, b = c % 3;
        double s = 0;
        for (int l = 0; l < 8; ++l) s += KE[r * 24 + 3 * l + b] * Q[l * 8 + L];
        X[r * 24 + c] = s;
      }
    for (int r = 0; r < 24; ++r)
      for (int c = 0; c < 24; ++c) {
        const int L = r / 3, a = r % 3;
        double s = 0;
        for (int l = 0; l < 8; ++l) s += Q[l * 8 + L] * X[(3 * l + a) * 24 + c];
        KEC[f * 576 + r * 24 + c] = s;
      }
  }
}

double tridiag_max_eig(int k, const double *a, const double *b)
{
  double lo = a[0], hi = a[0];
  for (int j = 0; j < k; j++) {
    const double r = (j > 0 ? std::fabs(b[j - 1]) : 0) + (j + 1 < k ? std::fabs(b[j]) : 0);
    lo = std::min(lo, a[j] - r);
    hi = std::max(hi, a[j] + r);
  }
  for (int it = 0; it < 200; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    int cnt = 0;  // eigenvalues below mid
    double q = a[0] - mid;
    if (q < 0) cnt++;
    for (int j = 1; j < k; j++) {
      if (q == 0) q = 1e-300;
      q = a[j] - mid - b[j - 1] * b[j - 1] / q;
      if (q < 0) cnt++;
    }
    if (cnt == k) hi = mid;
    else lo = mid;
  }
  return hi;
}

void build_interpolation(int Nfx, int Nfy, int Nfz, std::vector<int64_t> &rowptr,
                         std::vector<int32_t> &col, std::vector<double> &val)
{
  static const double vals[3] = {0.5, 1.0, 0.5};
  const int64_t Ncx = (Nfx + 1) / 2, Ncy = (Nfy + 1) / 2, Ncz = (Nfz + 1) / 2;
  const int64_t nf = (int64_t)Nfx * Nfy * Nfz;
  rowptr.assign(3 * nf + 1, 0);
  col.clear();
  val.clear();
  col.reserve(3 * nf * 8);
  val.reserve(3 * nf * 8);
  auto parents = [](int64_t i, int64_t Nc, int64_t *J, int *t) {
    int n = 0;
    for (int64_t j = i / 2 - 1; j <= i / 2 + 1; ++j) {
      const int64_t d = i - 2 * j;
      if (j < 0 || j >= Nc || d < -1 || d > 1) continue;
      J[n] = j;
      t[n] = (int)d + 1;
      ++n;
    }
    return n;
  };
  for (int64_t f = 0; f < nf; ++f) {
    const int64_t ix = f % Nfx, iy = (f / Nfx) % Nfy, iz = f / ((int64_t)Nfx * Nfy);
    int64_t Jx[3], Jy[3], Jz[3];
    int tx[3], ty[3], tz[3];
    const int nxp = parents(ix, Ncx, Jx, tx), nyp = parents(iy, Ncy, Jy, ty),
              nzp = parents(iz, Ncz, Jz, tz);
    int64_t cn[8];
    double w[8];
    int c = 0;
    for (int a = 0; a < nzp; ++a)
      for (int b = 0; b < nyp; ++b)
        for (int d = 0; d < nxp; ++d) {
          cn[c] = Jx[d] + Ncx * (Jy[b] + Ncy * Jz[a]);
          w[c] = vals[tx[d]] * vals[ty[b]] * vals[tz[a]];
          ++c;
        }
    // MatGetRowSum in column order, then the row times the reciprocal
    double s = 0.0;
    for (int q = 0; q < c; ++q) s += w[q];
    const double inv = 1.0 / s;
    for (int comp = 0; comp < 3; ++comp) {
      for (int q = 0; q < c; ++q) {
        col.push_back((int32_t)(3 * cn[q] + comp));
        val.push_back(w[q] * inv);
      }
      rowptr[3 * f + comp + 1] = (int64_t)col.size();
    }
  }
}

}  // namespace tofe
