// Linearised buckling on the structured mesh: the stress stiffness matrix of a displacement
// state and the element terms of its eigenvalue sensitivities (Stability::Function,
// StressFnc and sigtos, Buckling.cpp:15-414).
//
// ks_e(u) = sum_q W G_q^T S(D B_q u_e) G_q detJ is block diagonal over the three
// displacement components with the scalar 8x8 matrix
//   H_e[k][m] = sum_q W detJ grad N_k(q)^T s_q grad N_m(q),   s_q = stress tensor of D B_q u_e,
// so a state is stored as H_e (64 values per element, [(k*8+m)*nelem + e]) and
// Ks = sum_e (-Es_e) H_e (x) I_3 is applied node by node.  The adjoint right-hand side uses
// phi^T ks(e_d) phi = (D B_q e_d) . tau_q summed over q, with tau the Voigt vector of
// T = sum_i grad phi_i grad phi_i^T (shear entries doubled).
#include "k_common.hpp"

namespace tofe {
namespace {

constexpr int NT = 256;

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

__device__ inline void element_nodes(int64_t e, int nx, int ny, int64_t (&node)[8])
{
  const int Nx = nx + 1, Ny = ny + 1;
  const int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((int64_t)nx * ny));
#pragma unroll
  for (int l = 0; l < 8; ++l)
    node[l] = (ex + lnx(l)) + (int64_t)Nx * ((ey + lny(l)) + (int64_t)Ny * (ez + lnz(l)));
}

// H_e of the displacement state U (fixed DOFs taken as given: U is a solution, zero there).
__global__ __launch_bounds__(NT) void k_stress_h(const double *__restrict__ U, int nx, int ny,
                                                 int nz, ShapeData sd, double *H)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  const int64_t e = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (e >= ne) return;
  int64_t node[8];
  element_nodes(e, nx, ny, node);
  double u[8][3];
#pragma unroll
  for (int l = 0; l < 8; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c) u[l][c] = U[3 * node[l] + c];
  double h[36];  // upper triangle k <= m, packed row by row
#pragma unroll
  for (int v = 0; v < 36; ++v) h[v] = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    // displacement gradient g[c][a] = du_c/dx_a
    double g[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double t = 0.0;
#pragma unroll
        for (int l = 0; l < 8; ++l) t += sd.dN[q][a][l] * u[l][c];
        g[c][a] = t;
      }
    const double eps[6] = {g[0][0], g[1][1], g[2][2], g[0][1] + g[1][0], g[1][2] + g[2][1],
                           g[0][2] + g[2][0]};
    double sig[6];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      double t = 0.0;
#pragma unroll
      for (int b = 0; b < 6; ++b) t += sd.D[a * 6 + b] * eps[b];
      sig[a] = t * sd.detJ[q];  // unit Gauss weight
    }
    const double s[3][3] = {{sig[0], sig[3], sig[5]}, {sig[3], sig[1], sig[4]}, {sig[5], sig[4], sig[2]}};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double t[3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
        t[a] = s[a][0] * sd.dN[q][0][k] + s[a][1] * sd.dN[q][1][k] + s[a][2] * sd.dN[q][2][k];
#pragma unroll
      for (int m = k; m < 8; ++m)
        h[k * 8 - k * (k - 1) / 2 + (m - k)] +=
            t[0] * sd.dN[q][0][m] + t[1] * sd.dN[q][1][m] + t[2] * sd.dN[q][2][m];
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int m = k; m < 8; ++m) {
      const double v = h[k * 8 - k * (k - 1) / 2 + (m - k)];
      H[(int64_t)(k * 8 + m) * ne + e] = v;
      H[(int64_t)(m * 8 + k) * ne + e] = v;
    }
}

// y = Ks x with Ks = sum_e (-Es_e) H_e (x) I_3 and the Dirichlet rows and columns of
// MatZeroRowsColumns(Ks, fixed, 0.0) (Buckling.cpp:28-29): fixed x entries are ignored
// and fixed rows return 0.  One thread per node; elements in ascending order.
// V vectors per thread (block row blockIdx.y holds vectors V*blockIdx.y ...), so each H value
// read serves all of them; every vector's sums run in the same order as for one vector.
template <int V>
__global__ __launch_bounds__(NT) void k_apply_ks(const double *__restrict__ H,
                                                 const double *__restrict__ Es,
                                                 const uint8_t *__restrict__ fix, int nx, int ny,
                                                 int nz, const double *__restrict__ x, double *y,
                                                 int nvec)
{
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  const int64_t nn = (int64_t)Nx * Ny * Nz, ne = (int64_t)nx * ny * nz;
  const int v0 = blockIdx.y * V, nv = min(V, nvec - v0);
  x += (int64_t)v0 * 3 * nn;
  y += (int64_t)v0 * 3 * nn;
  const int64_t n = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (n >= nn) return;
  int i, j, k;
  node_ijk(n, Nx, Ny, i, j, k);
  double acc[V][3];
#pragma unroll
  for (int v = 0; v < V; ++v) acc[v][0] = acc[v][1] = acc[v][2] = 0.0;
  for (int dz = -1; dz <= 0; ++dz)
    for (int dy = -1; dy <= 0; ++dy)
      for (int dx = -1; dx <= 0; ++dx) {
        const int ex = i + dx, ey = j + dy, ez = k + dz;
        if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
        const int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
        const int kl = lidx(-dx, -dy, -dz);
        int64_t node[8];
        element_nodes(e, nx, ny, node);
        double t[V][3];
#pragma unroll
        for (int v = 0; v < V; ++v) t[v][0] = t[v][1] = t[v][2] = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const double hkm = H[(int64_t)(kl * 8 + m) * ne + e];
          const uint8_t fm = fix ? fix[node[m]] : 0;
#pragma unroll
          for (int v = 0; v < V; ++v) {
            if (v >= nv) break;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const double xv = ((fm >> c) & 1) ? 0.0 : x[v * 3 * nn + 3 * node[m] + c];
              t[v][c] += hkm * xv;
            }
          }
        }
        const double s = -Es[e];
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[v][c] += s * t[v][c];
      }
  const uint8_t fn = fix ? fix[n] : 0;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    if (v >= nv) break;
#pragma unroll
    for (int c = 0; c < 3; ++c) y[v * 3 * nn + 3 * n + c] = ((fn >> c) & 1) ? 0.0 : acc[v][c];
  }
}

// q_e = sum_c phi_c^T H_e phi_c for each vector (the unscaled element term of
// phi^T dKs/dy phi; Buckling.cpp:239-242 with dKsdy = -dEsdz ks, 324).
__global__ __launch_bounds__(NT) void k_ks_quad(const double *__restrict__ H,
                                                const double *__restrict__ phi, int nx, int ny,
                                                int nz, int64_t ndof, int nvec, double *q)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= ne * nvec) return;
  const int64_t e = t % ne;
  const double *p = phi + (t / ne) * ndof;
  int64_t node[8];
  element_nodes(e, nx, ny, node);
  double f[8][3];
#pragma unroll
  for (int l = 0; l < 8; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c) f[l][c] = p[3 * node[l] + c];
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const double hkm = H[(int64_t)(k * 8 + m) * ne + e];
      s += hkm * (f[k][0] * f[m][0] + f[k][1] * f[m][1] + f[k][2] * f[m][2]);
    }
  q[t] = s;
}

// Adjoint right-hand side of one eigenvector (Buckling.cpp:150-191):
// r_d = sum_{e holding d} Es_e sum_q (D B_q e_d) . tau_q(phi_e), fixed rows zero.
// One thread per node; the node's elements in ascending order, each recomputing tau.
__global__ __launch_bounds__(NT) void k_stab_rhs(const double *__restrict__ phi,
                                                 const double *__restrict__ Es,
                                                 const uint8_t *__restrict__ fix, int nx, int ny,
                                                 int nz, ShapeData sd, double *r)
{
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  const int64_t nn = (int64_t)Nx * Ny * Nz;
  const int64_t n = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (n >= nn) return;
  int i, j, k;
  node_ijk(n, Nx, Ny, i, j, k);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int dz = -1; dz <= 0; ++dz)
    for (int dy = -1; dy <= 0; ++dy)
      for (int dx = -1; dx <= 0; ++dx) {
        const int ex = i + dx, ey = j + dy, ez = k + dz;
        if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
        const int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
        const int kl = lidx(-dx, -dy, -dz);
        int64_t node[8];
        element_nodes(e, nx, ny, node);
        double f[8][3];
#pragma unroll
        for (int l = 0; l < 8; ++l)
#pragma unroll
          for (int c = 0; c < 3; ++c) f[l][c] = phi[3 * node[l] + c];
        double t[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < 8; ++q) {
          double g[3][3];  // g[c][a] = d phi_c / dx_a
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              double v = 0.0;
#pragma unroll
              for (int l = 0; l < 8; ++l) v += sd.dN[q][a][l] * f[l][c];
              g[c][a] = v;
            }
          double T[3][3];
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) T[a][b] = g[0][a] * g[0][b] + g[1][a] * g[1][b] + g[2][a] * g[2][b];
          const double tau[6] = {T[0][0], T[1][1], T[2][2], 2 * T[0][1], 2 * T[1][2], 2 * T[0][2]};
          double Dt[6];
#pragma unroll
          for (int a = 0; a < 6; ++a) {
            double v = 0.0;
#pragma unroll
            for (int b = 0; b < 6; ++b) v += sd.D[a * 6 + b] * tau[b];
            Dt[a] = v * sd.detJ[q];
          }
          // B^T rows of the node's three DOFs (AssignB row pattern)
          const double nx_ = sd.dN[q][0][kl], ny_ = sd.dN[q][1][kl], nz_ = sd.dN[q][2][kl];
          t[0] += nx_ * Dt[0] + ny_ * Dt[3] + nz_ * Dt[5];
          t[1] += ny_ * Dt[1] + nx_ * Dt[3] + nz_ * Dt[4];
          t[2] += nz_ * Dt[2] + ny_ * Dt[4] + nx_ * Dt[5];
        }
        const double s = Es[e];
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] += s * t[c];
      }
  const uint8_t fn = fix ? fix[n] : 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) r[3 * n + c] = ((fn >> c) & 1) ? 0.0 : acc[c];
}

// out_e = a_e^T KE b_e for nvec vectors a against one vector b (the v^T dK/dy U term of
// the buckling sensitivity, Buckling.cpp:245-246, without dEdz).
__global__ __launch_bounds__(NT) void k_elem_bilinear(const double *__restrict__ A,
                                                      const double *__restrict__ Bv,
                                                      const double *__restrict__ KE, int nx,
                                                      int ny, int nz, int64_t ndof, int nvec,
                                                      double *out)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= ne * nvec) return;
  const int64_t e = t % ne;
  const double *a = A + (t / ne) * ndof;
  int64_t node[8];
  element_nodes(e, nx, ny, node);
  double ua[24], ub[24];
#pragma unroll
  for (int l = 0; l < 8; ++l)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ua[3 * l + c] = a[3 * node[l] + c];
      ub[3 * l + c] = Bv[3 * node[l] + c];
    }
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < 24; ++r) {
    double tr = 0.0;
#pragma unroll
    for (int c = 0; c < 24; ++c) tr += KE[r * 24 + c] * ub[c];
    s += ua[r] * tr;
  }
  out[t] = s;
}

}  // namespace

void stress_h(hipStream_t s, const double *U, int nx, int ny, int nz, const ShapeData &sd,
              double *H)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  k_stress_h<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(U, nx, ny, nz, sd, H);
  TOFE_LAUNCH_CHECK();
}

void apply_ks(hipStream_t s, const double *H, const double *Es, const uint8_t *fix, int nx,
              int ny, int nz, const double *x, double *y, int nvec)
{
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  if (nvec == 1)
    k_apply_ks<1><<<dim3(blocks_for(nn)), dim3(NT), 0, s>>>(H, Es, fix, nx, ny, nz, x, y, 1);
  else
    k_apply_ks<4><<<dim3(blocks_for(nn), (nvec + 3) / 4), dim3(NT), 0, s>>>(H, Es, fix, nx, ny, nz,
                                                                           x, y, nvec);
  TOFE_LAUNCH_CHECK();
}

void ks_quad(hipStream_t s, const double *H, const double *phi, int nx, int ny, int nz,
             int64_t ndof, int nvec, double *q)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  k_ks_quad<<<dim3(blocks_for(ne * nvec)), dim3(NT), 0, s>>>(H, phi, nx, ny, nz, ndof, nvec, q);
  TOFE_LAUNCH_CHECK();
}

void stab_rhs(hipStream_t s, const double *phi, const double *Es, const uint8_t *fix, int nx,
              int ny, int nz, const ShapeData &sd, double *r)
{
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  k_stab_rhs<<<dim3(blocks_for(nn)), dim3(NT), 0, s>>>(phi, Es, fix, nx, ny, nz, sd, r);
  TOFE_LAUNCH_CHECK();
}

void elem_bilinear(hipStream_t s, const double *A, const double *B, const double *KE, int nx,
                   int ny, int nz, int64_t ndof, int nvec, double *out)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  k_elem_bilinear<<<dim3(blocks_for(ne * nvec)), dim3(NT), 0, s>>>(A, B, KE, nx, ny, nz, ndof,
                                                                   nvec, out);
  TOFE_LAUNCH_CHECK();
}

}  // namespace tofe
