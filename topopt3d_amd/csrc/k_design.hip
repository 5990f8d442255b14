// Design-side kernels: density and maximum-length-scale filters (Filter.cpp:21-150,
// RecMesh.cpp:544,706-718), the material interpolation with its SIMP / SIMP_CUT /
// SIMP_LOGISTIC branches and the filter chain rule (Material_Interpolation.cpp:20-187), and
// compliance sensitivities (Compliance.cpp:37-46).
//
// These are HBM-bound element loops.  They reproduce the reference's floating-point
// operation order exactly (file-wide fp contract(off): no fused multiply-add), so the filter
// weights, rho, E, dE/dz and the element sensitivities are bit-identical to a
// sequential non-FMA evaluation of the reference formulas.
#include "k_common.hpp"

// Keep every multiply and add separately rounded (no fused multiply-add) in this file.
#pragma clang fp contract(off)

namespace tofe {
namespace {

constexpr int NT = 256;
constexpr int kMaxTpl = 1024;

__device__ inline bool tpl_valid(const FilterTemplate &t, int ex, int ey, int ez, int q,
                                 const int *off, int &nb)
{
  const int i = ex + off[3 * q], j = ey + off[3 * q + 1], k = ez + off[3 * q + 2];
  if (i < 0 || i >= t.nx || j < 0 || j >= t.ny || k < 0 || k >= t.nz) return false;
  nb = i + t.nx * (j + t.ny * k);
  return true;
}

#define TOFE_STAGE_TEMPLATE()                                              \
  __shared__ int soff[3 * kMaxTpl];                                        \
  __shared__ double sw[kMaxTpl];                                           \
  for (int v = threadIdx.x; v < t.ntpl; v += NT) {                         \
    soff[3 * v] = t.off[3 * v];                                            \
    soff[3 * v + 1] = t.off[3 * v + 1];                                    \
    soff[3 * v + 2] = t.off[3 * v + 2];                                    \
    sw[v] = t.w[v];                                                        \
  }                                                                        \
  __syncthreads();

// inv_e = 1 / (sequential sum of the row's weights in ascending column order)
// (MatGetRowSum + VecPointwiseDivide, Filter.cpp:136-147)
__global__ __launch_bounds__(NT) void k_filter_inv(FilterTemplate t, double *inv, int64_t ne)
{
  TOFE_STAGE_TEMPLATE();
  const int64_t el = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (el >= ne) return;
  const int ex = (int)(el % t.nx), ey = (int)((el / t.nx) % t.ny), ez = (int)(el / ((int64_t)t.nx * t.ny));
  double s = 0.0;
  for (int q = 0; q < t.ntpl; ++q) {
    int nb;
    if (tpl_valid(t, ex, ey, ez, q, soff, nb)) s = (s + sw[q]);
  }
  inv[el] = 1.0 / s;
}

// rho_e = sum_j P_ej x_j, P_ej = w_q * inv_e, ascending columns (MatMult, Material_Interpolation.cpp:35)
__global__ __launch_bounds__(NT) void k_filter_apply(FilterTemplate t, const double *__restrict__ inv,
                                                     const double *__restrict__ x, double *rho,
                                                     int64_t ne)
{
  TOFE_STAGE_TEMPLATE();
  const int64_t el = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (el >= ne) return;
  const int ex = (int)(el % t.nx), ey = (int)((el / t.nx) % t.ny), ez = (int)(el / ((int64_t)t.nx * t.ny));
  const double ie = inv[el];
  double s = 0.0;
  for (int q = 0; q < t.ntpl; ++q) {
    int nb;
    if (tpl_valid(t, ex, ey, ez, q, soff, nb)) s = s + (sw[q] * ie) * x[nb];
  }
  rho[el] = s;
}

// gx_j = sum_i P_ij g_i over rows i = j - off_q in ascending order, i.e. the template
// walked backwards (MatMultTranspose, Material_Interpolation.cpp:159-162, 180).
__global__ __launch_bounds__(NT) void k_filter_apply_T(FilterTemplate t,
                                                       const double *__restrict__ inv,
                                                       const double *__restrict__ g, double *gx,
                                                       int64_t ne)
{
  TOFE_STAGE_TEMPLATE();
  const int64_t el = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (el >= ne) return;
  const int ex = (int)(el % t.nx), ey = (int)((el / t.nx) % t.ny), ez = (int)(el / ((int64_t)t.nx * t.ny));
  double s = 0.0;
  for (int q = t.ntpl - 1; q >= 0; --q) {
    const int i = ex - soff[3 * q], j = ey - soff[3 * q + 1], k = ez - soff[3 * q + 2];
    if (i < 0 || i >= t.nx || j < 0 || j >= t.ny || k < 0 || k >= t.nz) continue;
    const int64_t row = i + (int64_t)t.nx * (j + (int64_t)t.ny * k);
    s = s + (sw[q] * inv[row]) * g[row];
  }
  gx[el] = s;
}

__global__ __launch_bounds__(NT) void k_filter_count(FilterTemplate t, int64_t *cnt, int64_t ne)
{
  TOFE_STAGE_TEMPLATE();
  const int64_t el = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (el >= ne) return;
  const int ex = (int)(el % t.nx), ey = (int)((el / t.nx) % t.ny), ez = (int)(el / ((int64_t)t.nx * t.ny));
  int64_t c = 0;
  for (int q = 0; q < t.ntpl; ++q) {
    int nb;
    if (tpl_valid(t, ex, ey, ez, q, soff, nb)) ++c;
  }
  cnt[el] = c;
}

__global__ __launch_bounds__(NT) void k_filter_fill(FilterTemplate t, const double *__restrict__ inv,
                                                    const int64_t *__restrict__ rowptr,
                                                    int32_t *col, double *w, int64_t ne)
{
  TOFE_STAGE_TEMPLATE();
  const int64_t el = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (el >= ne) return;
  const int ex = (int)(el % t.nx), ey = (int)((el / t.nx) % t.ny), ez = (int)(el / ((int64_t)t.nx * t.ny));
  const double ie = inv[el];
  int64_t p = rowptr[el];
  for (int q = 0; q < t.ntpl; ++q) {
    int nb;
    if (tpl_valid(t, ex, ey, ez, q, soff, nb)) {
      col[p] = nb;
      w[p] = (sw[q] * ie);
      ++p;
    }
  }
}

// VecPow with PETSc's special exponents; pow() otherwise.
__device__ inline double vecpow1(double x, double p)
{
  if (p == 1.0) return x;
  if (p == 0.0) return 1.0;
  if (p == -1.0) return 1.0 / x;
  if (p == 0.5) return sqrt(x);
  if (p == -0.5) return 1.0 / sqrt(x);
  if (p == 2.0) return (x * x);
  if (p == -2.0) return 1.0 / (x * x);
  return pow(x, p);
}

// z = (1 - rho)^q after rho = min(rho, 1), in place on rho (Material_Interpolation.cpp:42-45)
__global__ void k_void_pre(double *rho, int64_t ne, double q, double *zt)
{
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < ne; i += (int64_t)gridDim.x * NT) {
    double r = rho[i];
    if (r > 1.0) r = 1.0;
    rho[i] = r;
    zt[i] = vecpow1(1.0 + -1 * r, q);
  }
}

// y = min(1, (REdge + R z) / vdMin), R with unit weights summed in ascending columns
// after REdge (MatMultAdd, VecScale(1/vdMin), VecPointwiseMin; Material_Interpolation.cpp:46-52)
__global__ __launch_bounds__(NT) void k_rfilter_madd(FilterTemplate t, const double *__restrict__ REdge,
                                                     const double *__restrict__ zt, double inv_vdmin,
                                                     double *y, int64_t ne)
{
  TOFE_STAGE_TEMPLATE();
  const int64_t el = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (el >= ne) return;
  const int ex = (int)(el % t.nx), ey = (int)((el / t.nx) % t.ny), ez = (int)(el / ((int64_t)t.nx * t.ny));
  double s = REdge[el];
  for (int q = 0; q < t.ntpl; ++q) {
    int nb;
    if (tpl_valid(t, ex, ey, ez, q, soff, nb)) s = s + 1.0 * zt[nb];
  }
  s = s * inv_vdmin;
  y[el] = s < 1.0 ? s : 1.0;
}

// R^T g with unit weights: rows i = j - off_q in ascending order (template walked backwards)
__global__ __launch_bounds__(NT) void k_rfilter_T(FilterTemplate t, const double *__restrict__ g,
                                                  double *gx, int64_t ne)
{
  TOFE_STAGE_TEMPLATE();
  const int64_t el = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (el >= ne) return;
  const int ex = (int)(el % t.nx), ey = (int)((el / t.nx) % t.ny), ez = (int)(el / ((int64_t)t.nx * t.ny));
  double s = 0.0;
  for (int q = t.ntpl - 1; q >= 0; --q) {
    const int i = ex - soff[3 * q], j = ey - soff[3 * q + 1], k = ez - soff[3 * q + 2];
    if (i < 0 || i >= t.nx || j < 0 || j >= t.ny || k < 0 || k >= t.nz) continue;
    s = s + 1.0 * g[i + (int64_t)t.nx * (j + (int64_t)t.ny * k)];
  }
  gx[el] = s;
}

// MatIntFnc after the filters (Material_Interpolation.cpp:54-116): rhoq = (q/vdMin)(1-rho)^(q-1);
// z = rho y; dEsdz = z^(p-1); Es = dEsdz z; dEsdz *= p; dEdz = (1-eps) dEsdz;
// E = eps + (1-eps) Es; then SIMP_CUT / SIMP_LOGISTIC modify Es and dEsdz only.
// y == nullptr means the plain density filter (y = 1, written to yout).
__global__ void k_simp(const double *__restrict__ rho, const double *y_in, int64_t ne,
                       SimpParams sp, double *yout, double *rhoq, double *E, double *dEdz,
                       double *Es, double *dEsdz, float *E32)
{
  const double sc = sp.vd_penal / sp.vd_min;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < ne; i += (int64_t)gridDim.x * NT) {
    const double r = rho[i];
    const double y = y_in ? y_in[i] : 1.0;
    if (!y_in) yout[i] = 1.0;
    rhoq[i] = vecpow1(1.0 + -1 * r, sp.vd_penal - 1) * sc;
    const double z = r * y;
    double d = vecpow1(z, sp.penal - 1);
    double es = d * z;
    d = d * sp.penal;
    const double e = sp.eps + (1 - sp.eps) * es;
    dEdz[i] = d * (1 - sp.eps);
    E[i] = e;
    if (E32) E32[i] = (float)e;
    if (sp.interp == TOFE_INTERP_SIMP_CUT) {
      const double cut = (z > sp.p0) ? 1.0 : 0.0;
      es = es * cut;
      d = d * cut;
    } else if (sp.interp == TOFE_INTERP_SIMP_LOGISTIC) {
      const double denom = 1 + exp(sp.p0 * (sp.p1 - z));
      es = es / denom;
      d = (d + (es * sp.p0) * (1 - 1 / denom)) / denom;
    }
    Es[i] = es;
    dEsdz[i] = d;
  }
}

// Chain_Filter middle step (Material_Interpolation.cpp:168-179):
// out = dfdE .* y - rhoq .* ((y < 1) .* RtT), RtT = R^T(dfdE .* rho) (or nullptr: no R term)
__global__ void k_chain_mid(const double *__restrict__ dfdE, const double *__restrict__ y,
                            const double *__restrict__ rhoq, const double *__restrict__ RtT,
                            double *out, int64_t ne)
{
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < ne; i += (int64_t)gridDim.x * NT) {
    const double b = dfdE[i] * y[i];
    if (RtT) {
      const double mask = (y[i] < 1.0) ? 1.0 : 0.0;
      double a = RtT[i] * mask;
      a = rhoq[i] * a;
      out[i] = b - a;
    } else {
      out[i] = b;
    }
  }
}

__global__ void k_mul(const double *__restrict__ a, const double *__restrict__ b, double *c, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
    c[i] = a[i] * b[i];
}

// g_e = -dEdz_e * u_e^T (KE u_e), sequential sums without contraction (Compliance.cpp:37-43)
__global__ __launch_bounds__(NT) void k_sensitivity(const double *__restrict__ U,
                                                    const double *__restrict__ dEdz,
                                                    const double *__restrict__ KE, int nx, int ny,
                                                    int nz, double *g)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  const int64_t el = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (el >= ne) return;
  const int ex = (int)(el % nx), ey = (int)((el / nx) % ny), ez = (int)(el / ((int64_t)nx * ny));
  const int Nx = nx + 1, Ny = ny + 1;
  double u[24];
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const int64_t n = (ex + lnx(l)) + (int64_t)Nx * ((ey + lny(l)) + (int64_t)Ny * (ez + lnz(l)));
#pragma unroll
    for (int c = 0; c < 3; ++c) u[3 * l + c] = U[3 * n + c];
  }
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < 24; ++r) {
    double tr = 0.0;
#pragma unroll
    for (int c = 0; c < 24; ++c) tr = tr + KE[r * 24 + c] * u[c];
    s = s + u[r] * tr;
  }
  g[el] = (-dEdz[el] * s);
}

// Lumped mass diagonal (Frequency::DiagMassFnc, Dynamic.cpp:181-236): the element blocks
// V_e m0 I are added in ascending element order, then the lumped point masses
// (MatDiagonalSet ADD_VALUES, 233); Dirichlet DOFs get 1e-8 (MatZeroRowsColumns, 44-45).
__global__ __launch_bounds__(NT) void k_mass_diag(const double *__restrict__ V, int nx, int ny,
                                                  int nz, double m0, const double *__restrict__ mlump,
                                                  const uint8_t *__restrict__ fix, double *M)
{
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  const int64_t nn = (int64_t)Nx * Ny * Nz;
  const int64_t n = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (n >= nn) return;
  int i, j, k;
  node_ijk(n, Nx, Ny, i, j, k);
  double s = 0.0;
  for (int dz = -1; dz <= 0; ++dz)
    for (int dy = -1; dy <= 0; ++dy)
      for (int dx = -1; dx <= 0; ++dx) {
        const int ex = i + dx, ey = j + dy, ez = k + dz;
        if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
        s = s + V[ex + (int64_t)nx * (ey + (int64_t)ny * ez)] * m0;
      }
  const uint8_t fm = fix ? fix[n] : 0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double v = mlump ? s + mlump[3 * n + c] : s;
    if ((fm >> c) & 1) v = 1e-8;
    M[3 * n + c] = v;
  }
}

__global__ void k_diag_mul(const double *__restrict__ d, const double *__restrict__ x, double *y,
                           int64_t n, int nrhs)
{
  const int64_t tot = n * nrhs;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < tot; i += (int64_t)gridDim.x * NT)
    y[i] = d[i % n] * x[i];
}

// Element quadratic forms of nvec vectors (the element terms of the eigenvalue
// sensitivities, Dynamic.cpp:113-154): kq = phi_e^T KE phi_e, mq = sum over the element's
// 24 DOFs of phi^2.  One thread per (element, vector).
// Consistent element mass over m0 = elemSize * density / 24 (the lumped matrix's per-DOF
// element mass): Q (x) I_3 with Q_ab = 24 prod_d (a_d == b_d ? 1/3 : 1/6) for element-local
// nodes a, b (the exact integral of N_a N_b over the box element, over elemSize / 24).
__device__ inline double mass_q(int a, int b)
{
  const int same = (lnx(a) == lnx(b)) + (lny(a) == lny(b)) + (lnz(a) == lnz(b));
  // 24 / 216 * 2^same
  return same == 3 ? 8.0 / 9.0 : same == 2 ? 4.0 / 9.0 : same == 1 ? 2.0 / 9.0 : 1.0 / 9.0;
}

// y = M x with the consistent mass M = sum_e V_e m0 Q (x) I_3, plus mlump on the diagonal,
// and the Dirichlet rows and columns of MatZeroRowsColumns(M, fixed, 1e-8): one thread per
// node gathers its eight elements; V vectors per thread (block row blockIdx.y holds vectors
// V * blockIdx.y ...).
template <int V>
__global__ __launch_bounds__(NT) void k_apply_mc(const double *__restrict__ Ve, double m0,
                                                 const double *__restrict__ mlump,
                                                 const uint8_t *__restrict__ fix, int nx, int ny,
                                                 int nz, const double *__restrict__ x, double *y,
                                                 int nvec)
{
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  const int64_t nn = (int64_t)Nx * Ny * Nz;
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
        const int kl = lidx(-dx, -dy, -dz);
        double t[V][3];
#pragma unroll
        for (int v = 0; v < V; ++v) t[v][0] = t[v][1] = t[v][2] = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const int64_t nm = (ex + lnx(m)) + (int64_t)Nx * ((ey + lny(m)) + (int64_t)Ny * (ez + lnz(m)));
          const double w = mass_q(kl, m);
          const uint8_t fm = fix ? fix[nm] : 0;
#pragma unroll
          for (int v = 0; v < V; ++v) {
            if (v >= nv) break;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const double xv = ((fm >> c) & 1) ? 0.0 : x[v * 3 * nn + 3 * nm + c];
              t[v][c] += w * xv;
            }
          }
        }
        const double se = Ve[ex + (int64_t)nx * (ey + (int64_t)ny * ez)] * m0;
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[v][c] += se * t[v][c];
      }
  const uint8_t fn = fix ? fix[n] : 0;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    if (v >= nv) break;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t d = v * 3 * nn + 3 * n + c;
      y[d] = ((fn >> c) & 1) ? 1e-8 * x[d] : acc[v][c] + (mlump ? mlump[3 * n + c] * x[d] : 0.0);
    }
  }
}

__global__ __launch_bounds__(NT) void k_elem_quad(const double *__restrict__ phi,
                                                  const double *__restrict__ KE, int nx, int ny,
                                                  int nz, int64_t ndof, int nvec, double *kq,
                                                  double *mq, int consistent)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= ne * nvec) return;
  const int64_t el = t % ne;
  const int v = (int)(t / ne);
  const double *U = phi + (int64_t)v * ndof;
  const int ex = (int)(el % nx), ey = (int)((el / nx) % ny), ez = (int)(el / ((int64_t)nx * ny));
  const int Nx = nx + 1, Ny = ny + 1;
  double u[24];
#pragma unroll
  for (int l = 0; l < 8; ++l) {
    const int64_t n = (ex + lnx(l)) + (int64_t)Nx * ((ey + lny(l)) + (int64_t)Ny * (ez + lnz(l)));
#pragma unroll
    for (int c = 0; c < 3; ++c) u[3 * l + c] = U[3 * n + c];
  }
  double s = 0.0, m = 0.0;
#pragma unroll
  for (int r = 0; r < 24; ++r) {
    double tr = 0.0;
#pragma unroll
    for (int c = 0; c < 24; ++c) tr = tr + KE[r * 24 + c] * u[c];
    s = s + u[r] * tr;
    m = m + u[r] * u[r];
  }
  kq[t] = s;
  if (mq && consistent) {  // phi_e^T (Q (x) I_3) phi_e
    m = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const double w = mass_q(a, b);
#pragma unroll
        for (int c = 0; c < 3; ++c) m += w * u[3 * a + c] * u[3 * b + c];
      }
  }
  if (mq) mq[t] = m;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

}  // namespace

void filter_rowsum_inv(hipStream_t s, const FilterTemplate &t, double *inv)
{
  const int64_t ne = (int64_t)t.nx * t.ny * t.nz;
  k_filter_inv<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(t, inv, ne);
  TOFE_LAUNCH_CHECK();
}

void filter_apply(hipStream_t s, const FilterTemplate &t, const double *inv, const double *x,
                  double *rho)
{
  const int64_t ne = (int64_t)t.nx * t.ny * t.nz;
  k_filter_apply<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(t, inv, x, rho, ne);
  TOFE_LAUNCH_CHECK();
}

void filter_apply_T(hipStream_t s, const FilterTemplate &t, const double *inv, const double *g,
                    double *gx)
{
  const int64_t ne = (int64_t)t.nx * t.ny * t.nz;
  k_filter_apply_T<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(t, inv, g, gx, ne);
  TOFE_LAUNCH_CHECK();
}

void filter_count(hipStream_t s, const FilterTemplate &t, int64_t *cnt)
{
  const int64_t ne = (int64_t)t.nx * t.ny * t.nz;
  k_filter_count<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(t, cnt, ne);
  TOFE_LAUNCH_CHECK();
}

void filter_fill(hipStream_t s, const FilterTemplate &t, const double *inv, const int64_t *rowptr,
                 int32_t *col, double *w)
{
  const int64_t ne = (int64_t)t.nx * t.ny * t.nz;
  k_filter_fill<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(t, inv, rowptr, col, w, ne);
  TOFE_LAUNCH_CHECK();
}

static unsigned stride_blocks(int64_t n)
{
  unsigned nb = blocks_for(n);
  return nb > 4096 ? 4096 : (nb ? nb : 1);
}

void void_pre(hipStream_t s, double *rho, int64_t ne, double q, double *zt)
{
  k_void_pre<<<dim3(stride_blocks(ne)), dim3(NT), 0, s>>>(rho, ne, q, zt);
  TOFE_LAUNCH_CHECK();
}

void rfilter_madd(hipStream_t s, const FilterTemplate &t, const double *REdge, const double *zt,
                  double inv_vdmin, double *y)
{
  const int64_t ne = (int64_t)t.nx * t.ny * t.nz;
  k_rfilter_madd<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(t, REdge, zt, inv_vdmin, y, ne);
  TOFE_LAUNCH_CHECK();
}

void rfilter_T(hipStream_t s, const FilterTemplate &t, const double *g, double *gx)
{
  const int64_t ne = (int64_t)t.nx * t.ny * t.nz;
  k_rfilter_T<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(t, g, gx, ne);
  TOFE_LAUNCH_CHECK();
}

void simp(hipStream_t s, const double *rho, const double *y, int64_t ne, const SimpParams &sp,
          double *yout, double *rhoq, double *E, double *dEdz, double *Es, double *dEsdz, float *E32)
{
  k_simp<<<dim3(stride_blocks(ne)), dim3(NT), 0, s>>>(rho, y, ne, sp, yout, rhoq, E, dEdz, Es, dEsdz,
                                                     E32);
  TOFE_LAUNCH_CHECK();
}

void chain_mid(hipStream_t s, const double *dfdE, const double *y, const double *rhoq,
               const double *RtT, double *out, int64_t ne)
{
  k_chain_mid<<<dim3(stride_blocks(ne)), dim3(NT), 0, s>>>(dfdE, y, rhoq, RtT, out, ne);
  TOFE_LAUNCH_CHECK();
}

void mul(hipStream_t s, const double *a, const double *b, double *c, int64_t n)
{
  k_mul<<<dim3(stride_blocks(n)), dim3(NT), 0, s>>>(a, b, c, n);
  TOFE_LAUNCH_CHECK();
}

void sensitivity(hipStream_t s, const double *U, const double *dEdz, const double *KE, int nx,
                 int ny, int nz, double *g)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  k_sensitivity<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(U, dEdz, KE, nx, ny, nz, g);
  TOFE_LAUNCH_CHECK();
}

void mass_diag(hipStream_t s, const double *V, int nx, int ny, int nz, double m0,
               const double *mlump, const uint8_t *fix, double *M)
{
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  k_mass_diag<<<dim3(blocks_for(nn)), dim3(NT), 0, s>>>(V, nx, ny, nz, m0, mlump, fix, M);
  TOFE_LAUNCH_CHECK();
}

void apply_mass_consistent(hipStream_t s, const double *V, double m0, const double *mlump,
                           const uint8_t *fix, int nx, int ny, int nz, const double *x, double *y,
                           int nvec)
{
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  if (nvec == 1)
    k_apply_mc<1><<<dim3(blocks_for(nn)), dim3(NT), 0, s>>>(V, m0, mlump, fix, nx, ny, nz, x, y, 1);
  else
    k_apply_mc<4><<<dim3(blocks_for(nn), (nvec + 3) / 4), dim3(NT), 0, s>>>(V, m0, mlump, fix, nx,
                                                                           ny, nz, x, y, nvec);
  TOFE_LAUNCH_CHECK();
}

void diag_mul(hipStream_t s, const double *d, const double *x, double *y, int64_t n, int nrhs)
{
  k_diag_mul<<<dim3(stride_blocks(n * nrhs)), dim3(NT), 0, s>>>(d, x, y, n, nrhs);
  TOFE_LAUNCH_CHECK();
}

void elem_quad(hipStream_t s, const double *phi, const double *KE, int nx, int ny, int nz,
               int64_t ndof, int nvec, double *kq, double *mq, bool consistent)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  k_elem_quad<<<dim3(blocks_for(ne * nvec)), dim3(NT), 0, s>>>(phi, KE, nx, ny, nz, ndof, nvec,
                                                              kq, mq, consistent ? 1 : 0);
  TOFE_LAUNCH_CHECK();
}

}  // namespace tofe
