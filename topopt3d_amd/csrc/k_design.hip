// Design-side kernels: density filter (Filter.cpp:21-150), SIMP interpolation
// (Material_Interpolation.cpp:20-75) and compliance sensitivities (Compliance.cpp:37-46).
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

// Plain-density SIMP (y = 1): z = rho; dEsdz = z^(p-1); Es = dEsdz z; dEsdz *= p;
// dEdz = (1-eps) dEsdz; E = eps + (1-eps) Es  (Material_Interpolation.cpp:61-75)
__global__ void k_simp(const double *__restrict__ rho, int64_t ne, double penal, double eps,
                       double *E, double *dEdz, double *Es, double *dEsdz, float *E32)
{
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < ne; i += (int64_t)gridDim.x * NT) {
    const double z = (rho[i] * 1.0);
    double d = vecpow1(z, penal - 1);
    const double es = (d * z);
    d = (d * penal);
    const double e = eps + (1 - eps) * es;
    Es[i] = es;
    dEsdz[i] = d;
    dEdz[i] = d * (1 - eps);
    E[i] = e;
    if (E32) E32[i] = (float)e;
  }
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

void simp(hipStream_t s, const double *rho, int64_t ne, double penal, double eps, double *E,
          double *dEdz, double *Es, double *dEsdz, float *E32)
{
  unsigned nb = blocks_for(ne);
  if (nb > 4096) nb = 4096;
  k_simp<<<dim3(nb), dim3(NT), 0, s>>>(rho, ne, penal, eps, E, dEdz, Es, dEsdz, E32);
  TOFE_LAUNCH_CHECK();
}

void sensitivity(hipStream_t s, const double *U, const double *dEdz, const double *KE, int nx,
                 int ny, int nz, double *g)
{
  const int64_t ne = (int64_t)nx * ny * nz;
  k_sensitivity<<<dim3(blocks_for(ne)), dim3(NT), 0, s>>>(U, dEdz, KE, nx, ny, nz, g);
  TOFE_LAUNCH_CHECK();
}

}  // namespace tofe
