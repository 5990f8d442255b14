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

// P^T A P for a 27-point 3x3-block stencil field, one axis at a time.  P is a tensor
// product of the 1-D row-scaled interpolations w1d, so the triple sum over fine nodes and
// fine offsets factorises: restricting x, then y, then z, each step is a 1-D Galerkin
// product of a stencil field on a grid (n0, n1, n2) whose `axis` has Nf nodes into one
// with Nc = (Nf + 1) / 2 nodes along that axis:
//   out(node', o', q) = sum_{i, d} w(i, I) w(i + d, I + D) in(node' with i, o' with d, q)
// where I, D are node' and o' along the axis and the other coordinates are unchanged.
// One thread per output value; every output value is written exactly once.  In the last
// step (upper != 0) offsets o' >= 13 are computed and mirrored to (J, 26 - o', b, a), so
// the coarse operator is exactly symmetric; o' < 13 threads only write the zeros of
// couplings that leave the grid.
__global__ __launch_bounds__(NT) void k_galerkin_axis(const double *__restrict__ in, int n0,
                                                      int n1, int n2, int axis,
                                                      double *__restrict__ out, int upper)
{
  int m[3] = {n0, n1, n2};
  const int Nf = m[axis];
  const int Nc = (Nf + 1) / 2;
  m[axis] = Nc;
  const int64_t nin = (int64_t)n0 * n1 * n2, nout = (int64_t)m[0] * m[1] * m[2];
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= 243 * nout) return;
  const int oq = (int)(t / nout);
  const int64_t node = t - (int64_t)oq * nout;
  const int o = oq / 9, q = oq % 9;
  int c[3];
  c[0] = (int)(node % m[0]);
  c[1] = (int)((node / m[0]) % m[1]);
  c[2] = (int)(node / ((int64_t)m[0] * m[1]));
  const int off[3] = {o % 3 - 1, (o / 3) % 3 - 1, o / 9 - 1};
  // the coupled output node must exist
  bool inside = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int cc = c[k] + off[k];
    if (cc < 0 || cc >= m[k]) inside = false;
  }
  if (!inside) {
    out[t] = 0.0;
    return;
  }
  if (upper && o < 13) return;  // written by the mirror of (J, 26 - o)
  const int I = c[axis], D = off[axis];
  const int sa = axis == 0 ? 1 : (axis == 1 ? n0 : n0 * n1);  // input node stride along axis
  const int so = axis == 0 ? 1 : (axis == 1 ? 3 : 9);         // offset-index stride of axis
  int cin[3] = {c[0], c[1], c[2]};
  cin[axis] = 0;
  const int64_t base_node = cin[0] + (int64_t)n0 * (cin[1] + (int64_t)n1 * cin[2]);
  const int o_base = o - (D + 1) * so;
  double acc = 0.0;
  for (int i = 2 * I - 1; i <= 2 * I + 1; ++i) {
    if (i < 0 || i >= Nf) continue;
    const double wi = w1d(i, I, Nf);
    if (wi == 0.0) continue;
    for (int d = -1; d <= 1; ++d) {
      const int j = i + d;
      if (j < 0 || j >= Nf) continue;
      const double wj = w1d(j, I + D, Nf);
      if (wj == 0.0) continue;
      const int oin = o_base + (d + 1) * so;
      acc += wi * wj * in[(int64_t)(oin * 9 + q) * nin + base_node + (int64_t)i * sa];
    }
  }
  out[t] = acc;
  if (upper && o != 13) {
    const int64_t J = (c[0] + off[0]) + (int64_t)m[0] * ((c[1] + off[1]) + (int64_t)m[1] * (c[2] + off[2]));
    const int a = q / 3, b = q % 3;
    out[(int64_t)((26 - o) * 9 + b * 3 + a) * nout + J] = acc;
  }
}

// The same operator from the upper half of an exactly symmetric stencil: offsets
// o >= 13 are stored at the node (Su[((o-13)*9 + a*3 + b)*nn + n], 126 values per node) and
// the block of a lower offset o < 13 is the transpose of the neighbour's block 26 - o.  Each
// stored block is read by its node and by one neighbour, so the HBM traffic is about half of
// the full stencil's.
template <typename T, int EPI>
__global__ __launch_bounds__(NT) void k_stencil_op_sym(StencilOp<T> op, const T *__restrict__ u,
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
        if (o >= 13) {
          const T *B = S + (int64_t)((o - 13) * 9) * nn + n;
#pragma unroll
          for (int a = 0; a < 3; ++a)
            acc[a] += B[(a * 3 + 0) * nn] * u0 + B[(a * 3 + 1) * nn] * u1 + B[(a * 3 + 2) * nn] * u2;
        } else {
          const T *B = S + (int64_t)((13 - o) * 9) * nn + m;  // block (m, 26 - o), transposed
#pragma unroll
          for (int a = 0; a < 3; ++a)
            acc[a] += B[(0 * 3 + a) * nn] * u0 + B[(1 * 3 + a) * nn] * u1 + B[(2 * 3 + a) * nn] * u2;
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

// FP64 full stencil -> FP32 upper half (offsets 13..26).
__global__ void k_stencil_upper32(const double *__restrict__ S, int64_t nn, float *Su)
{
  const int64_t tot = 126 * nn;
  for (int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x; t < tot; t += (int64_t)gridDim.x * NT)
    Su[t] = (float)S[13 * 9 * nn + t];
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
// Value of a double held by lane `lane` (uniform) of the wavefront.
__device__ inline double lane_value(double v, int lane)
{
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// One wavefront, everything in registers: lane i holds row i of the block while it is
// factorised and column i of W = L^-1 while that is formed; the other lanes' values come
// through v_readlane with compile-time lane numbers, so there are no LDS round trips and
// no barriers.  A partial block (kb < 64) is padded with the identity, which factors and
// inverts to itself and leaves the leading kb x kb part unchanged.
__global__ __launch_bounds__(64) void k_chol_block(double *A, int lda, int kb, int k0, double *W,
                                                   int *info)
{
  const int i = threadIdx.x;
  double r[kCholNB];
#pragma unroll
  for (int j = 0; j < kCholNB; ++j)
    r[j] = (i < kb && j < kb) ? (i >= j ? A[i + (int64_t)j * lda] : 0.0) : (i == j ? 1.0 : 0.0);
  int bad = 0;
#pragma unroll
  for (int j = 0; j < kCholNB; ++j) {
    double d = lane_value(r[j], j);
    if (!(d > 0)) {
      if (bad == 0) bad = k0 + j + 1;
      d = 1.0;
    }
    d = sqrt(d);
    const double lij = (i > j) ? r[j] / d : (i == j ? d : 0.0);
    r[j] = lij;
    // trailing update of row i; entries above the diagonal are never read as L values
#pragma unroll
    for (int l = j + 1; l < kCholNB; ++l) r[l] -= lij * lane_value(lij, l);
  }
  if (bad != 0 && i == 0 && info[0] == 0) info[0] = bad;
#pragma unroll
  for (int j = 0; j < kCholNB; ++j)
    if (i < kb && j < kb && i >= j) A[i + (int64_t)j * lda] = r[j];
  // W = L^-1: lane c solves L w = e_c; L(q, p) is r[p] of lane q
  const int c = i;
  double w[kCholNB];
#pragma unroll
  for (int q = 0; q < kCholNB; ++q) w[q] = (q == c) ? 1.0 : 0.0;
#pragma unroll
  for (int q = 0; q < kCholNB; ++q) {
    w[q] = w[q] / lane_value(r[q], q);
#pragma unroll
    for (int p = q + 1; p < kCholNB; ++p) w[p] -= lane_value(r[q], p) * w[q];
  }
#pragma unroll
  for (int q = 0; q < kCholNB; ++q)
    W[q + kCholNB * c] = (q < kb && c < kb && q >= c) ? w[q] : 0.0;
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

template <typename T>
void stencil_op_sym(hipStream_t s, const StencilOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep)
{
  const int64_t nn = (int64_t)op.Nx * op.Ny * op.Nz;
  dim3 grid((unsigned)((nn + NT - 1) / NT)), block(NT);
  switch (epi) {
    case EPI_APPLY: k_stencil_op_sym<T, EPI_APPLY><<<grid, block, 0, s>>>(op, u, ep, nn); break;
    case EPI_APPLY_DOT: k_stencil_op_sym<T, EPI_APPLY_DOT><<<grid, block, 0, s>>>(op, u, ep, nn); break;
    case EPI_APPLY_DOT2: k_stencil_op_sym<T, EPI_APPLY_DOT2><<<grid, block, 0, s>>>(op, u, ep, nn); break;
    case EPI_RESID: k_stencil_op_sym<T, EPI_RESID><<<grid, block, 0, s>>>(op, u, ep, nn); break;
    case EPI_CHEB: k_stencil_op_sym<T, EPI_CHEB><<<grid, block, 0, s>>>(op, u, ep, nn); break;
  }
  TOFE_LAUNCH_CHECK();
}

void stencil_upper32(hipStream_t s, const double *S, int64_t nn, float *Su)
{
  unsigned nb = (unsigned)((126 * nn + NT - 1) / NT);
  if (nb > 4096) nb = 4096;
  k_stencil_upper32<<<dim3(nb), dim3(NT), 0, s>>>(S, nn, Su);
  TOFE_LAUNCH_CHECK();
}

void galerkin(hipStream_t s, const double *Sf, int Nfx, int Nfy, int Nfz, double *Sc, double *T1,
              double *T2)
{
  // x: (Nfx, Nfy, Nfz) -> (Ncx, Nfy, Nfz); y: -> (Ncx, Ncy, Nfz); z: -> (Ncx, Ncy, Ncz)
  const int Ncx = (Nfx + 1) / 2, Ncy = (Nfy + 1) / 2, Ncz = (Nfz + 1) / 2;
  const int64_t n1 = (int64_t)Ncx * Nfy * Nfz, n2 = (int64_t)Ncx * Ncy * Nfz,
                n3 = (int64_t)Ncx * Ncy * Ncz;
  k_galerkin_axis<<<dim3((unsigned)((243 * n1 + NT - 1) / NT)), dim3(NT), 0, s>>>(Sf, Nfx, Nfy, Nfz,
                                                                                  0, T1, 0);
  TOFE_LAUNCH_CHECK();
  k_galerkin_axis<<<dim3((unsigned)((243 * n2 + NT - 1) / NT)), dim3(NT), 0, s>>>(T1, Ncx, Nfy, Nfz,
                                                                                  1, T2, 0);
  TOFE_LAUNCH_CHECK();
  k_galerkin_axis<<<dim3((unsigned)((243 * n3 + NT - 1) / NT)), dim3(NT), 0, s>>>(T2, Ncx, Ncy, Nfz,
                                                                                  2, Sc, 1);
  TOFE_LAUNCH_CHECK();
}

int64_t galerkin_scratch(int Nfx, int Nfy, int Nfz, int64_t *n2)
{
  const int Ncx = (Nfx + 1) / 2, Ncy = (Nfy + 1) / 2;
  *n2 = (int64_t)243 * Ncx * Ncy * Nfz;
  return (int64_t)243 * Ncx * Nfy * Nfz;
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

template void stencil_op_sym<float>(hipStream_t, const StencilOp<float> &, const float *, Epi,
                                    const EpiArgs<float> &);
}  // namespace tofe
