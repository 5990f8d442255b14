// Coarse-level operators of the geometric multigrid hierarchy.
//
// Levels >= 1 hold the Galerkin operator A_{l+1} = P_l^T A_l P_l (PCMGSetGalerkin
// PC_MG_GALERKIN_BOTH, FEAnalysis.cpp:412) as a 27-point 3x3-block stencil on the
// coarse node grid.  With the row-scaled trilinear transfers of Interpolation.cpp the
// coarse operator couples each coarse node only with its 26 neighbours, so the
// stencil is exact, including the Dirichlet identity rows carried up from the fine
// level and the degenerate last cell of grids with an odd element count.
#include "k_common.hpp"

namespace tofe {
namespace {

constexpr int NT = 256;

template <typename T, int EPI>
__global__ __launch_bounds__(NT) void k_stencil_op(StencilOp<T> op, const T *__restrict__ u,
                                                   EpiArgs<T> ep, int64_t nn)
{
  __shared__ double sred[NT / 64];
  const int64_t n = (int64_t)blockIdx.x * NT + threadIdx.x;
  const T *__restrict__ S = op.S;
  double dsum = 0, dsum2 = 0;
  if (n < nn) {
    const int Nx = op.Nx, Ny = op.Ny, Nz = op.Nz;
    const int i = (int)(n % Nx), j = (int)((n / Nx) % Ny), k = (int)(n / ((int64_t)Nx * Ny));
    T acc[3] = {0, 0, 0};
#pragma unroll
    for (int o = 0; o < 27; ++o) {
      const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
      const int mi = i + dx, mj = j + dy, mk = k + dz;
      if (mi >= 0 && mi < Nx && mj >= 0 && mj < Ny && mk >= 0 && mk < Nz) {
        const int64_t m = n + dx + (int64_t)Nx * (dy + (int64_t)Ny * dz);
        const T u0 = u[3 * m], u1 = u[3 * m + 1], u2 = u[3 * m + 2];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          acc[a] += S[(int64_t)(o * 9 + a * 3 + 0) * nn + n] * u0;
          acc[a] += S[(int64_t)(o * 9 + a * 3 + 1) * nn + n] * u1;
          acc[a] += S[(int64_t)(o * 9 + a * 3 + 2) * nn + n] * u2;
        }
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int64_t dof = 3 * n + a;
      apply_epilogue<T, EPI>(ep, dof, acc[a], u[dof], dsum, dsum2);
    }
  }
  if constexpr (EPI == EPI_APPLY_DOT || EPI == EPI_APPLY_DOT2) {
    const double t = block_sum<NT>(dsum, sred);
    if (threadIdx.x == 0) ep.partial[blockIdx.x] = t;
    if constexpr (EPI == EPI_APPLY_DOT2) {
      __syncthreads();
      const double t2 = block_sum<NT>(dsum2, sred);
      if (threadIdx.x == 0) ep.partial2[blockIdx.x] = t2;
    }
  }
}

// One thread per (coarse node I, coarse offset o >= 13, block entry q = 3a + b): the
// entry of P^T A P coupling I with J = I + off(o), summed over fine pairs (i, i+d) with
// nonzero transfer weights; the transposed entry goes to (J, 26-o), so the operator is
// exactly symmetric.  Nine threads per block keep the small coarse levels busy.
__global__ __launch_bounds__(NT) void k_galerkin(const double *__restrict__ Sf, int Nfx, int Nfy,
                                                 int Nfz, double *__restrict__ Sc, int Ncx,
                                                 int Ncy, int Ncz)
{
  const int64_t ncn = (int64_t)Ncx * Ncy * Ncz, nfn = (int64_t)Nfx * Nfy * Nfz;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= 126 * ncn) return;
  const int oq = (int)(t / ncn);
  const int64_t I = t - (int64_t)oq * ncn;
  const int o = 13 + oq / 9, q = oq % 9;
  const int Dx = o % 3 - 1, Dy = (o / 3) % 3 - 1, Dz = o / 9 - 1;
  const int Ix = (int)(I % Ncx), Iy = (int)((I / Ncx) % Ncy), Iz = (int)(I / ((int64_t)Ncx * Ncy));
  const int Jx = Ix + Dx, Jy = Iy + Dy, Jz = Iz + Dz;
  if (Jx < 0 || Jx >= Ncx || Jy < 0 || Jy >= Ncy || Jz < 0 || Jz >= Ncz) return;

  int li[3][9], ld[3][9];
  double lw[3][9];
  int cnt[3];
  const int Ic[3] = {Ix, Iy, Iz}, Jc[3] = {Jx, Jy, Jz}, Nf[3] = {Nfx, Nfy, Nfz};
#pragma unroll
  for (int dim = 0; dim < 3; ++dim) {
    int c = 0;
    for (int i = 2 * Ic[dim] - 1; i <= 2 * Ic[dim] + 1; ++i) {
      if (i < 0 || i >= Nf[dim]) continue;
      const double wi = w1d(i, Ic[dim], Nf[dim]);
      if (wi == 0.0) continue;
      for (int d = -1; d <= 1; ++d) {
        const int jj = i + d;
        if (jj < 0 || jj >= Nf[dim]) continue;
        const double wj = w1d(jj, Jc[dim], Nf[dim]);
        if (wj == 0.0) continue;
        li[dim][c] = i;
        ld[dim][c] = d;
        lw[dim][c] = wi * wj;
        ++c;
      }
    }
    cnt[dim] = c;
  }
  const double *Sq = Sf + (int64_t)q * nfn;
  double acc = 0.0;
  for (int cz = 0; cz < cnt[2]; ++cz)
    for (int cy = 0; cy < cnt[1]; ++cy) {
      const double wyz = lw[1][cy] * lw[2][cz];
      const int64_t rowyz = (int64_t)Nfx * (li[1][cy] + (int64_t)Nfy * li[2][cz]);
      const int odyz = 3 * (ld[1][cy] + 1) + 9 * (ld[2][cz] + 1);
      for (int cx = 0; cx < cnt[0]; ++cx) {
        const int od = (ld[0][cx] + 1) + odyz;
        acc += lw[0][cx] * wyz * Sq[(int64_t)(od * 9) * nfn + li[0][cx] + rowyz];
      }
    }
  Sc[(int64_t)(o * 9 + q) * ncn + I] = acc;
  if (o != 13) {
    const int64_t J = Jx + (int64_t)Ncx * (Jy + (int64_t)Ncy * Jz);
    const int a = q / 3, b = q % 3;
    Sc[(int64_t)((26 - o) * 9 + b * 3 + a) * ncn + J] = acc;
  }
}

__global__ void k_stencil_diag(const double *__restrict__ S, int64_t nn, double *diag,
                               double *dinv)
{
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= 3 * nn) return;
  const int64_t n = t / 3;
  const int a = (int)(t - 3 * n);
  const double d = S[(int64_t)(13 * 9 + a * 4) * nn + n];
  if (diag) diag[t] = d;
  dinv[t] = 1.0 / d;
}

__global__ void k_stencil_dense(const double *__restrict__ S, int Nx, int Ny, int Nz,
                                double *__restrict__ A)
{
  const int64_t nn = (int64_t)Nx * Ny * Nz, n = 3 * nn;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= 27 * nn) return;
  const int o = (int)(t / nn);
  const int64_t I = t - (int64_t)o * nn;
  const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
  const int i = (int)(I % Nx), j = (int)((I / Nx) % Ny), k = (int)(I / ((int64_t)Nx * Ny));
  if (i + dx < 0 || i + dx >= Nx || j + dy < 0 || j + dy >= Ny || k + dz < 0 || k + dz >= Nz) return;
  const int64_t J = I + dx + (int64_t)Nx * (dy + (int64_t)Ny * dz);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) A[(3 * I + a) * n + 3 * J + b] = S[(int64_t)(o * 9 + a * 3 + b) * nn + I];
}

// column-major lower triangle -> full symmetric
__global__ void k_symmetrize(double *A, int64_t n)
{
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= n * n) return;
  const int64_t c = t / n, r = t - c * n;
  if (r < c) A[r + c * n] = A[c + r * n];
}

// y = A x for a symmetric dense A: one wavefront per row, fixed-order reduction.
template <typename Tin, typename Tout>
__global__ __launch_bounds__(NT) void k_gemv(const double *__restrict__ A, int64_t n,
                                             const Tin *__restrict__ x, Tout *__restrict__ y)
{
  const int64_t row = (int64_t)blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const double *a = A + row * n;
  double s = 0;
  for (int64_t c = lane; c < n; c += 64) s += a[c] * (double)x[c];
  s = wave_sum(s);
  if (lane == 0) y[row] = (Tout)s;
}

// Cholesky factor of one kb x kb diagonal block held in LDS (column-major A, leading
// dimension lda, kb <= kCholNB): L_kk overwrites the lower triangle of the block and its
// inverse goes to W (leading dimension kCholNB).  info[0] = global pivot index + 1 of the
// first non-positive pivot (0 if none).
// One wavefront: lane i owns row i of the block during the factorisation and column i of
// W = L^-1 during the inversion, so the inversion needs no barrier at all and the
// factorisation one per column.  a is column-major (a[i + 64 j]); w is row-major
// (w[c + 64 i]) so that the 64 lanes touch consecutive words.
__global__ __launch_bounds__(64) void k_chol_block(double *A, int lda, int kb, int k0, double *W,
                                                   int *info)
{
  __shared__ double a[kCholNB * kCholNB];
  __shared__ double w[kCholNB * kCholNB];
  const int i = threadIdx.x;
  for (int j = 0; j < kb; ++j) a[i + kCholNB * j] = (i < kb && i >= j) ? A[i + (int64_t)j * lda] : 0.0;
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    double d = a[j + kCholNB * j];
    if (!(d > 0)) {
      if (i == 0 && info[0] == 0) info[0] = k0 + j + 1;
      d = 1.0;
    }
    d = sqrt(d);
    double lij = 0.0;
    if (i > j && i < kb) lij = a[i + kCholNB * j] / d;
    __syncthreads();  // every lane has read column j and the pivot
    if (i == j) a[j + kCholNB * j] = d;
    if (i > j && i < kb) a[i + kCholNB * j] = lij;
    // trailing update of row i: a(i, l) -= l(i, j) l(l, j) for j < l <= i
    for (int l = j + 1; l < kb; ++l) {
      const double llj = __shfl(lij, l, 64);
      if (i >= l && i < kb) a[i + kCholNB * l] -= lij * llj;
    }
    __syncthreads();
  }
  for (int j = 0; j < kb; ++j)
    if (i < kb && i >= j) A[i + (int64_t)j * lda] = a[i + kCholNB * j];
  // W = L^-1: lane c solves L w = e_c by forward substitution on its own column
  const int c = i;
  for (int r = 0; r < kb; ++r) w[c + kCholNB * r] = (r == c) ? 1.0 : 0.0;
  for (int q = 0; q < kb; ++q) {
    const double wq = w[c + kCholNB * q] / a[q + kCholNB * q];
    w[c + kCholNB * q] = wq;
    for (int r = q + 1; r < kb; ++r) w[c + kCholNB * r] -= a[r + kCholNB * q] * wq;
  }
  if (c < kCholNB)
    for (int r = 0; r < kCholNB; ++r)
      W[r + kCholNB * c] = (r < kb && c < kb && r >= c) ? w[c + kCholNB * r] : 0.0;
}

}  // namespace

void chol_block(hipStream_t s, double *A, int lda, int kb, int k0, double *W, int *info)
{
  k_chol_block<<<dim3(1), dim3(kCholNB), 0, s>>>(A, lda, kb, k0, W, info);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
int stencil_op_blocks(const StencilOp<T> &op)
{
  const int64_t nn = (int64_t)op.Nx * op.Ny * op.Nz;
  return (int)((nn + NT - 1) / NT);
}

template <typename T>
void stencil_op(hipStream_t s, const StencilOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep)
{
  const int64_t nn = (int64_t)op.Nx * op.Ny * op.Nz;
  dim3 grid((unsigned)((nn + NT - 1) / NT)), block(NT);
  switch (epi) {
    case EPI_APPLY: k_stencil_op<T, EPI_APPLY><<<grid, block, 0, s>>>(op, u, ep, nn); break;
    case EPI_APPLY_DOT: k_stencil_op<T, EPI_APPLY_DOT><<<grid, block, 0, s>>>(op, u, ep, nn); break;
    case EPI_APPLY_DOT2: k_stencil_op<T, EPI_APPLY_DOT2><<<grid, block, 0, s>>>(op, u, ep, nn); break;
    case EPI_RESID: k_stencil_op<T, EPI_RESID><<<grid, block, 0, s>>>(op, u, ep, nn); break;
    case EPI_CHEB: k_stencil_op<T, EPI_CHEB><<<grid, block, 0, s>>>(op, u, ep, nn); break;
  }
  TOFE_LAUNCH_CHECK();
}

void galerkin(hipStream_t s, const double *Sf, int Nfx, int Nfy, int Nfz, double *Sc, int Ncx,
              int Ncy, int Ncz)
{
  const int64_t ncn = (int64_t)Ncx * Ncy * Ncz;
  TOFE_HIP(hipMemsetAsync(Sc, 0, sizeof(double) * 243 * ncn, s));
  const int64_t nt = 126 * ncn;
  k_galerkin<<<dim3((unsigned)((nt + NT - 1) / NT)), dim3(NT), 0, s>>>(Sf, Nfx, Nfy, Nfz, Sc, Ncx,
                                                                      Ncy, Ncz);
  TOFE_LAUNCH_CHECK();
}

void stencil_diag(hipStream_t s, const double *S, int64_t nn, double *diag, double *dinv)
{
  k_stencil_diag<<<dim3((unsigned)((3 * nn + NT - 1) / NT)), dim3(NT), 0, s>>>(S, nn, diag, dinv);
  TOFE_LAUNCH_CHECK();
}

void stencil_to_dense(hipStream_t s, const double *S, int Nx, int Ny, int Nz, double *A)
{
  const int64_t nn = (int64_t)Nx * Ny * Nz;
  TOFE_HIP(hipMemsetAsync(A, 0, sizeof(double) * 9 * nn * nn, s));
  k_stencil_dense<<<dim3((unsigned)((27 * nn + NT - 1) / NT)), dim3(NT), 0, s>>>(S, Nx, Ny, Nz, A);
  TOFE_LAUNCH_CHECK();
}

void symmetrize_lower(hipStream_t s, double *A, int64_t n)
{
  k_symmetrize<<<dim3((unsigned)((n * n + NT - 1) / NT)), dim3(NT), 0, s>>>(A, n);
  TOFE_LAUNCH_CHECK();
}

template <typename Tin, typename Tout>
void dense_gemv(hipStream_t s, const double *A, int64_t n, const Tin *x, Tout *y)
{
  k_gemv<Tin, Tout><<<dim3((unsigned)((n + 3) / 4)), dim3(NT), 0, s>>>(A, n, x, y);
  TOFE_LAUNCH_CHECK();
}

template int stencil_op_blocks<double>(const StencilOp<double> &);
template int stencil_op_blocks<float>(const StencilOp<float> &);
template void stencil_op<double>(hipStream_t, const StencilOp<double> &, const double *, Epi,
                                 const EpiArgs<double> &);
template void stencil_op<float>(hipStream_t, const StencilOp<float> &, const float *, Epi,
                                const EpiArgs<float> &);
template void dense_gemv<double, double>(hipStream_t, const double *, int64_t, const double *, double *);
template void dense_gemv<float, float>(hipStream_t, const double *, int64_t, const float *, float *);

}  // namespace tofe
