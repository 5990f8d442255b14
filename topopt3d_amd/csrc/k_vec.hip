// Vector kernels of the multigrid-preconditioned CG solve: deterministic two-stage
// reductions, CG updates driven by device-side scalars (no host round trip for alpha
// and beta), the zero-guess Chebyshev step, trilinear restriction/prolongation
// (Interpolation.cpp:123-169 row-scaled, applied matrix-free) and the Jacobi-CG steps
// whose coefficients estimate lambda_max(D^-1 A) for the Chebyshev bounds.
#include "k_common.hpp"

namespace tofe {
namespace {

constexpr int NT = 256;

template <typename T>
__global__ __launch_bounds__(NT) void k_dot_partial(const T *__restrict__ a,
                                                    const T *__restrict__ b, int64_t n,
                                                    double *partial)
{
  __shared__ double sred[NT / 64];
  const int64_t vo = (int64_t)blockIdx.y * n;  // vector blockIdx.y
  a += vo;
  b += vo;
  partial += (int64_t)blockIdx.y * gridDim.x;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
    s += (double)a[i] * (double)b[i];
  const double t = block_sum<NT>(s, sred);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// mirror (may be null): device address of host memory that receives the vector's SL_COUNT
// scalars once they are final, so the host reads them without a copy launch.
__device__ __forceinline__ void reduce_final_block(const double *__restrict__ partial,
                                                   int nparts, double *out, int post,
                                                   double *scal, double *mirror = nullptr)
{
  __shared__ double sred[NT / 64];
  // vector blockIdx.y: its partials, and its scalars (out included) SL_COUNT further on
  partial += (int64_t)blockIdx.y * nparts;
  if (out) out += (int64_t)blockIdx.y * SL_COUNT;
  if (scal) scal += (int64_t)blockIdx.y * SL_COUNT;
  double s = 0;
  for (int i = threadIdx.x; i < nparts; i += NT) s += partial[i];
  const double v = block_sum<NT>(s, sred);
  if (threadIdx.x == 0) {
    if (out) *out = v;
    // a frozen right-hand side (converged in a block solve) takes no further steps
    const bool frozen = scal && scal[SL_FROZEN] != 0.0;
    switch (post) {
      case POST_ALPHA_CG: scal[SL_ALPHA] = frozen ? 0.0 : scal[SL_RZ] / v; break;
      case POST_BETA_CG:
        scal[SL_BETA] = frozen ? 0.0 : v / scal[SL_RZ];
        scal[SL_RZ] = v;
        break;
      case POST_ALPHA_FCG: scal[SL_ALPHA] = frozen ? 0.0 : scal[SL_PR] / v; break;
      case POST_BETA_FCG: scal[SL_BETA] = frozen ? 0.0 : -v / scal[SL_PQ]; break;
      case POST_CONV_FCG:
      case POST_CONV_CG:
        if (!frozen) {
          scal[SL_ITS] += 1.0;
          const double pq = scal[SL_PQ];
          double rs = 0.0;
          if (post == POST_CONV_CG && scal[SL_RZ] < 0) rs = TOFE_DIVERGED_INDEFINITE_PC;
          else if (!(pq > 0)) rs = (pq != pq) ? TOFE_DIVERGED_NANORINF : TOFE_DIVERGED_INDEFINITE_MAT;
          else if (!(v == v) || v > 1.7976931348623157e308) rs = TOFE_DIVERGED_NANORINF;
          else if (sqrt(v) <= scal[SL_TTOL]) rs = scal[SL_CONVCODE];
          if (rs != 0.0) {
            scal[SL_REASON] = rs;
            scal[SL_FROZEN] = 1.0;
          }
        }
        break;
      default: break;
    }
    if (mirror && scal) {
      double *mv = mirror + (int64_t)blockIdx.y * SL_COUNT;
      for (int i = 0; i < SL_COUNT; ++i) mv[i] = scal[i];
    }
  }
}

__global__ __launch_bounds__(NT) void k_reduce_final(const double *__restrict__ partial,
                                                     int nparts, double *out, int post,
                                                     double *scal, double *mirror)
{
  reduce_final_block(partial, nparts, out, post, scal, mirror);
}

__global__ __launch_bounds__(NT) void k_reduce_final_batch(const ReduceLaunch *__restrict__ tab,
                                                           const int *__restrict__ b0, int n)
{
  const ReduceLaunch L = tab[batch_entry(b0, n, blockIdx.x)];
  reduce_final_block(L.partial, L.nparts, L.out, L.post, L.scal);
}

__global__ void k_set_conv_state(double *scal, int nvec, ConvState cs, unsigned long long mask)
{
  const int v = threadIdx.x;
  if (v >= nvec) return;
  double *sv = scal + (int64_t)v * SL_COUNT;
  sv[SL_TTOL] = cs.ttol[v];
  sv[SL_CONVCODE] = (double)cs.conv[v];
  sv[SL_ITS] = 0.0;
  sv[SL_REASON] = 0.0;
  sv[SL_FROZEN] = ((mask >> v) & 1ull) ? 1.0 : 0.0;
}

// scal[v * SL_COUNT + SL_FROZEN] = bit v of mask, v < nvec (the right-hand sides of a block
// solve that have converged and take no further steps)
__global__ void k_set_frozen(double *scal, int nvec, unsigned long long mask)
{
  const int v = threadIdx.x;
  if (v < nvec) scal[(int64_t)v * SL_COUNT + SL_FROZEN] = ((mask >> v) & 1ull) ? 1.0 : 0.0;
}

template <typename T>
__global__ void k_cheb_zero(const T *__restrict__ b, const T *__restrict__ dinv, T c2, T *d,
                            T *x, int64_t n)
{
  const int64_t vo = (int64_t)blockIdx.y * n;  // vector blockIdx.y; dinv is shared
  b += vo;
  d += vo;
  x += vo;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const T v = c2 * (dinv[i] * b[i]);
    d[i] = v;
    x[i] = v;
  }
}

__global__ void k_cheb_zero_from64(const double *__restrict__ b64, float *__restrict__ b32,
                                   const float *__restrict__ dinv, float c2, float *d, float *x,
                                   int64_t n)
{
  const int64_t vo = (int64_t)blockIdx.y * n;  // vector blockIdx.y; dinv is shared
  b64 += vo;
  b32 += vo;
  d += vo;
  x += vo;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const float b = (float)b64[i];
    b32[i] = b;
    const float v = c2 * (dinv[i] * b);
    d[i] = v;
    x[i] = v;
  }
}

// Both kernels below visit their (up to 27 or 8) partners unconditionally: an index outside
// the grid is clamped to a valid node and its weight is zero, so no branch separates the
// loads; the sums run in the same ascending order as a skip-based loop and the zero terms
// leave them unchanged.

// b_c(I) = sum over fine nodes i within one fine step of 2I of w(i,I) r_f(i)  (P^T r)
template <typename T>
__device__ __forceinline__ void restrict_block(const T *__restrict__ rf, int Nfx, int Nfy, int Nfz,
                                               T *__restrict__ bc, int Ncx, int Ncy, int Ncz,
                                               FirstStep<T> fs, int blk)
{
  const int64_t ncn = (int64_t)Ncx * Ncy * Ncz;
  {  // vector blockIdx.y of fine residuals and coarse right-hand sides; dinv is shared
    const int64_t v = blockIdx.y;
    rf += v * 3 * ((int64_t)Nfx * Nfy * Nfz);
    bc += v * 3 * ncn;
    if (fs.d) {
      fs.d += v * 3 * ncn;
      fs.x += v * 3 * ncn;
    }
  }
  const int64_t I = (int64_t)blk * NT + threadIdx.x;
  if (I >= ncn) return;
  int Ix, Iy, Iz;
  node_ijk(I, Ncx, Ncy, Ix, Iy, Iz);
  int ix[3], iy[3], iz[3];
  double wx[3], wy[3], wz[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int fx = 2 * Ix - 1 + d, fy = 2 * Iy - 1 + d, fz = 2 * Iz - 1 + d;
    const bool okx = fx >= 0 && fx < Nfx, oky = fy >= 0 && fy < Nfy, okz = fz >= 0 && fz < Nfz;
    ix[d] = okx ? fx : 2 * Ix;
    iy[d] = oky ? fy : 2 * Iy;
    iz[d] = okz ? fz : 2 * Iz;
    wx[d] = okx ? w1d(fx, Ix, Nfx) : 0.0;
    wy[d] = oky ? w1d(fy, Iy, Nfy) : 0.0;
    wz[d] = okz ? w1d(fz, Iz, Nfz) : 0.0;
  }
  double acc[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double w = wx[a] * wy[b] * wz[c];
        const int64_t f = ix[a] + (int64_t)Nfx * (iy[b] + (int64_t)Nfy * iz[c]);
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q] += w * (double)rf[3 * f + q];
      }
#pragma unroll
  for (int q = 0; q < 3; ++q) bc[3 * I + q] = (T)acc[q];
  if (fs.d) {  // the coarse level's first Chebyshev step from x = 0 (as k_cheb_zero)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T v = fs.c2 * (fs.dinv[3 * I + q] * (T)acc[q]);
      fs.d[3 * I + q] = v;
      fs.x[3 * I + q] = v;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void k_restrict(const T *__restrict__ rf, int Nfx, int Nfy,
                                                 int Nfz, T *__restrict__ bc, int Ncx, int Ncy,
                                                 int Ncz, FirstStep<T> fs)
{
  restrict_block<T>(rf, Nfx, Nfy, Nfz, bc, Ncx, Ncy, Ncz, fs, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_restrict_batch(const RestrictLaunch<T> *__restrict__ tab,
                                                       const int *__restrict__ b0, int n)
{
  const int e = batch_entry(b0, n, blockIdx.x);
  const RestrictLaunch<T> L = tab[e];
  restrict_block<T>(L.rf, L.Nfx, L.Nfy, L.Nfz, L.bc, L.Ncx, L.Ncy, L.Ncz, L.fs, blockIdx.x - b0[e]);
}

// x_f(i) += sum over coarse parents J of w(i,J) x_c(J)  (P x_c)
template <typename T>
__device__ __forceinline__ void prolong_block(const T *__restrict__ xc, int Ncx, int Ncy, int Ncz,
                                              T *__restrict__ xf, int Nfx, int Nfy, int Nfz, int blk)
{
  const int64_t nfn = (int64_t)Nfx * Nfy * Nfz;
  xc += (int64_t)blockIdx.y * 3 * ((int64_t)Ncx * Ncy * Ncz);  // vector blockIdx.y
  xf += (int64_t)blockIdx.y * 3 * nfn;
  const int64_t f = (int64_t)blk * NT + threadIdx.x;
  if (f >= nfn) return;
  int fx, fy, fz;
  node_ijk(f, Nfx, Nfy, fx, fy, fz);
  // parents (i - 1) >> 1 and (i + 1) >> 1 (the same one for even i), weights w1d
  int Jx[2], Jy[2], Jz[2];
  double wx[2], wy[2], wz[2];
  auto parents = [](int i, int Nf, int Nc, int (&J)[2], double (&w)[2]) {
    const int j0 = (i - 1) >> 1, j1 = (i + 1) >> 1;
    const bool ok0 = j0 >= 0 && j0 < Nc, ok1 = j1 != j0 && j1 >= 0 && j1 < Nc;
    J[0] = ok0 ? j0 : j1;
    J[1] = ok1 ? j1 : J[0];
    w[0] = ok0 ? w1d(i, j0, Nf) : 0.0;
    w[1] = ok1 ? w1d(i, j1, Nf) : 0.0;
  };
  parents(fx, Nfx, Ncx, Jx, wx);
  parents(fy, Nfy, Ncy, Jy, wy);
  parents(fz, Nfz, Ncz, Jz, wz);
  double acc[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const double w = wx[a] * wy[b] * wz[c];
        const int64_t J = Jx[a] + (int64_t)Ncx * (Jy[b] + (int64_t)Ncy * Jz[c]);
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q] += w * (double)xc[3 * J + q];
      }
#pragma unroll
  for (int q = 0; q < 3; ++q) xf[3 * f + q] = (T)((double)xf[3 * f + q] + acc[q]);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_prolong_add(const T *__restrict__ xc, int Ncx, int Ncy,
                                                    int Ncz, T *__restrict__ xf, int Nfx, int Nfy,
                                                    int Nfz)
{
  prolong_block<T>(xc, Ncx, Ncy, Ncz, xf, Nfx, Nfy, Nfz, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(NT) void k_prolong_add_batch(const ProlongLaunch<T> *__restrict__ tab,
                                                          const int *__restrict__ b0, int n)
{
  const int e = batch_entry(b0, n, blockIdx.x);
  const ProlongLaunch<T> L = tab[e];
  prolong_block<T>(L.xc, L.Ncx, L.Ncy, L.Ncz, L.xf, L.Nfx, L.Nfy, L.Nfz, blockIdx.x - b0[e]);
}

template <typename Ta, typename Tb>
__global__ void k_convert(const Ta *__restrict__ a, Tb *__restrict__ b, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
    b[i] = (Tb)a[i];
}

template <typename T>
__global__ void k_fill(T *a, T v, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
    a[i] = v;
}

// r -= alpha q; partial ||r||^2 (the iterate's update x += alpha p follows in k_cg_xp); with
// fc.d, also the V-cycle's first fine-level Chebyshev step d = c2 D^-1 r (as k_cheb_zero).
__global__ __launch_bounds__(NT) void k_cg_r(double *__restrict__ r, const double *__restrict__ q,
                                             int64_t n, const double *__restrict__ scal, int alpha_slot,
                                             double *partial, FirstCheb64 fc)
{
  __shared__ double sred[NT / 64];
  const int64_t vo = (int64_t)blockIdx.y * n;  // vector blockIdx.y
  r += vo;
  q += vo;
  if (fc.d) fc.d += vo;
  scal += (int64_t)blockIdx.y * SL_COUNT;
  partial += (int64_t)blockIdx.y * gridDim.x;
  const double alpha = scal[alpha_slot];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += ri * ri;
    if (fc.d) fc.d[i] = fc.c2 * (fc.dinv[i] * ri);
  }
  const double t = block_sum<NT>(s, sred);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// x += alpha p (the iterate's update of the step that formed alpha), then p = z + beta p
__global__ void k_cg_xp(double *__restrict__ p, const double *__restrict__ z, double *__restrict__ x,
                        int64_t n, const double *__restrict__ scal, int beta_slot, int alpha_slot)
{
  const int64_t vo = (int64_t)blockIdx.y * n;  // vector blockIdx.y
  p += vo;
  z += vo;
  x += vo;
  scal += (int64_t)blockIdx.y * SL_COUNT;
  const double beta = scal[beta_slot], alpha = scal[alpha_slot];
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const double pi = p[i];
    // the iterate is touched once per iteration (non-temporal, as in fcg_p_block)
    __builtin_nontemporal_store(__builtin_nontemporal_load(x + i) + alpha * pi, x + i);
    p[i] = z[i] + beta * pi;
  }
}

// v_i = 0.5 + frac(i * 2654435761 / 2^32): deterministic start vector of the estimate
template <typename T>
__global__ void k_power_start(T *v, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const uint32_t h = (uint32_t)((uint64_t)i * 2654435761ull);
    v[i] = (T)(0.5 + (double)h / 4294967296.0);
  }
}

// z = dinv r, partial r.z (Jacobi-preconditioned CG of the eigenvalue estimate)
template <typename T>
__global__ __launch_bounds__(NT) void k_jacobi_dot(const T *__restrict__ r, const T *__restrict__ dinv,
                                                   T *z, int64_t n, double *partial)
{
  __shared__ double sred[NT / 64];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const T zi = dinv[i] * r[i];
    z[i] = zi;
    s += (double)r[i] * (double)zi;
  }
  const double t = block_sum<NT>(s, sred);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void k_record(const double *scal, int slot, double *dst)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = scal[slot];
}

// Sum of nparts per-block partial sums, formed by every block of the consuming kernel
// in the order k_reduce_final uses (thread-strided, wave butterfly, waves in order), so
// the value is the same bit for bit without a separate reduction launch.  Valid in all
// threads.
__device__ inline double all_sum(const double *__restrict__ partial, int nparts, double *smem)
{
  double s = 0;
  for (int i = threadIdx.x; i < nparts; i += NT) s += partial[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) smem[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) t += smem[i];
  __syncthreads();
  return t;
}

// Flexible-CG step without separate reduction launches, first half: alpha = p.r / p.Ap
// from the operator's two sets of partials (block 0 stores p.r, p.Ap and alpha in scal);
// r -= alpha q, partials of r.r (the iterate's update x += alpha p follows in k_fcg_p).
__device__ __forceinline__ void fcg_xr_block(double *__restrict__ r, const double *__restrict__ q,
                                             int64_t n, const double *__restrict__ pq_part,
                                             const double *__restrict__ pr_part, int nparts,
                                             double *scal, double *partial, FirstCheb fc, int blk,
                                             int nb, bool pre = false)
{
  __shared__ double sred[NT / 64];
  {  // vector blockIdx.y: its vectors, partials and scalars; dinv is shared
    const int64_t v = blockIdx.y, vo = v * n;
    r += vo;
    q += vo;
    pq_part += v * nparts;
    pr_part += v * nparts;
    scal += v * SL_COUNT;
    partial += v * nb;
    if (fc.b32) {
      fc.b32 += vo;
      fc.d += vo;
      if (fc.x) fc.x += vo;
    }
  }
  double alpha;
  if (pre) {
    alpha = scal[SL_ALPHA];
  } else {
    const double pr = all_sum(pr_part, nparts, sred);
    const double pq = all_sum(pq_part, nparts, sred);
    // a frozen right-hand side (converged in a block solve) takes no further steps
    alpha = scal[SL_FROZEN] != 0.0 ? 0.0 : pr / pq;
    if (blk == 0 && threadIdx.x == 0) {
      scal[SL_PR] = pr;
      scal[SL_PQ] = pq;
      scal[SL_ALPHA] = alpha;
    }
  }
  double s = 0;
  for (int64_t i = (int64_t)blk * NT + threadIdx.x; i < n; i += (int64_t)nb * NT) {
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    s += ri * ri;
    if (fc.b32) {  // the V-cycle's first step on the new residual (as k_cheb_zero_from64)
      const float b = (float)ri;
      fc.b32[i] = b;
      const float v = fc.c2 * (fc.dinv[i] * b);
      fc.d[i] = v;
      if (fc.x) fc.x[i] = v;
    }
  }
  const double t = block_sum<NT>(s, sred);
  if (threadIdx.x == 0) partial[blk] = t;
}

__global__ __launch_bounds__(NT) void k_fcg_xr(double *__restrict__ r, const double *__restrict__ q,
                                               int64_t n,
                                               const double *__restrict__ pq_part,
                                               const double *__restrict__ pr_part, int nparts,
                                               double *scal, double *partial, FirstCheb fc)
{
  fcg_xr_block(r, q, n, pq_part, pr_part, nparts, scal, partial, fc, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(NT) void k_fcg_xr_batch(const FcgXrLaunch *__restrict__ tab,
                                                     const int *__restrict__ b0, int n)
{
  const int e = batch_entry(b0, n, blockIdx.x);
  const FcgXrLaunch L = tab[e];
  fcg_xr_block(L.r, L.q, L.n, L.pq_part, L.pr_part, L.nparts, L.scal, L.partial, L.fc,
               blockIdx.x - b0[e], L.nblocks, L.pre != 0);
}

// Second half: beta = -z.Ap / p.Ap from the z.q partials (block 0 stores z.Ap and beta);
// x += alpha p with the first half's alpha, then p = z + beta p.
// z comes in FP32 (z32, the FP32 V-cycle's last step, whose values are FP32 numbers) or FP64.
__device__ __forceinline__ void fcg_p_block(double *__restrict__ p, const double *__restrict__ z,
                                            const float *__restrict__ z32,
                                            double *__restrict__ x, int64_t n,
                                            const double *__restrict__ zq_part, int nparts,
                                            double *scal, int blk, int nb, bool pre = false)
{
  __shared__ double sred[NT / 64];
  {  // vector blockIdx.y
    const int64_t v = blockIdx.y, vo = v * n;
    p += vo;
    if (z32) z32 += vo;
    else z += vo;
    x += vo;
    zq_part += v * nparts;
    scal += v * SL_COUNT;
  }
  double beta;
  if (pre) {
    beta = scal[SL_BETA];
  } else {
    const double zq = all_sum(zq_part, nparts, sred);
    beta = scal[SL_FROZEN] != 0.0 ? 0.0 : -zq / scal[SL_PQ];
    if (blk == 0 && threadIdx.x == 0) {
      scal[SL_ZQ] = zq;
      scal[SL_BETA] = beta;
    }
  }
  const double alpha = scal[SL_ALPHA];
  for (int64_t i = (int64_t)blk * NT + threadIdx.x; i < n; i += (int64_t)nb * NT) {
    const double pi = p[i];
    // The iterate is read and written once per iteration: non-temporal, so that it does not
    // displace from the memory-side cache (256 MB) the FP32 vectors and level-1 stencil that
    // the V-cycle reuses (measured: mixed solve of the 192x96x96 cantilever 6.24 -> 6.15 ms).
    __builtin_nontemporal_store(__builtin_nontemporal_load(x + i) + alpha * pi, x + i);
    const double zi = z32 ? (double)z32[i] : z[i];
    p[i] = zi + beta * pi;
  }
}

__global__ __launch_bounds__(NT) void k_fcg_p(double *__restrict__ p, const double *__restrict__ z,
                                              const float *__restrict__ z32,
                                              double *__restrict__ x, int64_t n,
                                              const double *__restrict__ zq_part, int nparts,
                                              double *scal)
{
  fcg_p_block(p, z, z32, x, n, zq_part, nparts, scal, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(NT) void k_fcg_p_batch(const FcgPLaunch *__restrict__ tab,
                                                    const int *__restrict__ b0, int n)
{
  const int e = batch_entry(b0, n, blockIdx.x);
  const FcgPLaunch L = tab[e];
  fcg_p_block(L.p, L.z, L.z32, L.x, L.n, L.zq_part, L.nparts, L.scal, blockIdx.x - b0[e],
              L.nblocks, L.pre != 0);
}

// x += alpha p with alpha = scal[SL_ALPHA]: the iterate's update when no k_fcg_p / k_cg_xp follows.
__global__ __launch_bounds__(NT) void k_fcg_x(double *__restrict__ x, const double *__restrict__ p,
                                              int64_t n, const double *scal)
{
  x += (int64_t)blockIdx.y * n;  // vector blockIdx.y
  p += (int64_t)blockIdx.y * n;
  scal += (int64_t)blockIdx.y * SL_COUNT;
  const double alpha = scal[SL_ALPHA];
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
    x[i] += alpha * p[i];
}

// Jacobi-CG step of the eigenvalue estimate, first half: alpha = rz_old / p.Ap from the
// operator's partials; r -= alpha q; partial r.(D^-1 r).  Block 0 records alpha.
template <typename T>
__global__ __launch_bounds__(NT) void k_lz_r(T *__restrict__ r, const T *__restrict__ q,
                                             const T *__restrict__ dinv, int64_t n,
                                             const double *__restrict__ pq_part, int npq,
                                             const double *__restrict__ scal, int rz_slot,
                                             double *rec_alpha, double *partial)
{
  __shared__ double sred[NT / 64];
  const double alpha = scal[rz_slot] / all_sum(pq_part, npq, sred);
  if (blockIdx.x == 0 && threadIdx.x == 0) *rec_alpha = alpha;
  const T a = (T)alpha;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const T ri = r[i] - a * q[i];
    r[i] = ri;
    const T zi = dinv[i] * ri;
    s += (double)ri * (double)zi;
  }
  const double t = block_sum<NT>(s, sred);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Second half: rz = sum of the partials, beta = rz / rz_old, p = D^-1 r + beta p.  Block 0
// records beta and stores rz in the other slot for the next step.
template <typename T>
__global__ __launch_bounds__(NT) void k_lz_p(T *__restrict__ p, const T *__restrict__ r,
                                             const T *__restrict__ dinv, int64_t n,
                                             const double *__restrict__ rz_part, int nrz,
                                             double *scal, int rz_old, int rz_new, double *rec_beta)
{
  __shared__ double sred[NT / 64];
  const double rz = all_sum(rz_part, nrz, sred);
  const double beta = rz / scal[rz_old];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *rec_beta = beta;
    scal[rz_new] = rz;
  }
  const T b = (T)beta;
  for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
    const T zi = dinv[i] * r[i];
    p[i] = zi + b * p[i];
  }
}

inline unsigned stride_blocks(int64_t n)
{
  const int64_t b = (n + NT - 1) / NT;
  return (unsigned)(b < kRedBlocks ? (b > 0 ? b : 1) : kRedBlocks);
}

}  // namespace

template <typename T>
void dot_partial(hipStream_t s, const T *a, const T *b, int64_t n, double *partial, int nvec)
{
  k_dot_partial<T><<<dim3(kRedBlocks, nvec), dim3(NT), 0, s>>>(a, b, n, partial);
  TOFE_LAUNCH_CHECK();
}

void reduce_final(hipStream_t s, const double *partial, int nparts, double *out, int post,
                  double *scal, int nvec, double *mirror)
{
  k_reduce_final<<<dim3(1, nvec), dim3(NT), 0, s>>>(partial, nparts, out, post, scal, mirror);
  TOFE_LAUNCH_CHECK();
}

void set_conv_state(hipStream_t s, double *scal, int nvec, const ConvState &cs,
                    unsigned long long frozen_mask)
{
  if (nvec < 1 || nvec > 16) throw Error(TOFE_ERR_ARG, "set_conv_state: 1..16 vectors");
  k_set_conv_state<<<dim3(1), dim3(64), 0, s>>>(scal, nvec, cs, frozen_mask);
  TOFE_LAUNCH_CHECK();
}

void set_frozen(hipStream_t s, double *scal, int nvec, unsigned long long mask)
{
  if (nvec < 1 || nvec > 64) throw Error(TOFE_ERR_ARG, "set_frozen: 1..64 vectors");
  k_set_frozen<<<dim3(1), dim3(64), 0, s>>>(scal, nvec, mask);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void cheb_zero(hipStream_t s, const T *b, const T *dinv, T c2, T *d, T *x, int64_t n, int nvec)
{
  k_cheb_zero<T><<<dim3(stride_blocks(n), nvec), dim3(NT), 0, s>>>(b, dinv, c2, d, x, n);
  TOFE_LAUNCH_CHECK();
}

void cheb_zero_from64(hipStream_t s, const double *b64, float *b32, const float *dinv, float c2,
                      float *d, float *x, int64_t n, int nvec)
{
  k_cheb_zero_from64<<<dim3(stride_blocks(n), nvec), dim3(NT), 0, s>>>(b64, b32, dinv, c2, d, x, n);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void restrict_op(hipStream_t s, const T *rf, int Nfx, int Nfy, int Nfz, T *bc, int Ncx, int Ncy,
                 int Ncz, const FirstStep<T> &fs, int nvec)
{
  const int64_t ncn = (int64_t)Ncx * Ncy * Ncz;
  k_restrict<T><<<dim3((unsigned)((ncn + NT - 1) / NT), nvec), dim3(NT), 0, s>>>(rf, Nfx, Nfy, Nfz, bc,
                                                                           Ncx, Ncy, Ncz, fs);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void prolong_add(hipStream_t s, const T *xc, int Ncx, int Ncy, int Ncz, T *xf, int Nfx, int Nfy,
                 int Nfz, int nvec)
{
  const int64_t nfn = (int64_t)Nfx * Nfy * Nfz;
  k_prolong_add<T><<<dim3((unsigned)((nfn + NT - 1) / NT), nvec), dim3(NT), 0, s>>>(xc, Ncx, Ncy, Ncz,
                                                                              xf, Nfx, Nfy, Nfz);
  TOFE_LAUNCH_CHECK();
}

template <typename Ta, typename Tb>
void convert(hipStream_t s, const Ta *a, Tb *b, int64_t n)
{
  k_convert<Ta, Tb><<<dim3(stride_blocks(n)), dim3(NT), 0, s>>>(a, b, n);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void fill(hipStream_t s, T *a, T v, int64_t n)
{
  k_fill<T><<<dim3(stride_blocks(n)), dim3(NT), 0, s>>>(a, v, n);
  TOFE_LAUNCH_CHECK();
}
template void fill<float>(hipStream_t, float *, float, int64_t);
template void fill<double>(hipStream_t, double *, double, int64_t);

void fill_d(hipStream_t s, double *a, double v, int64_t n) { fill<double>(s, a, v, n); }

void cg_update_r(hipStream_t s, double *r, const double *q, int64_t n, const double *scal,
                 int alpha_slot, double *partial, int nvec, const FirstCheb64 &fc)
{
  k_cg_r<<<dim3(kRedBlocks, nvec), dim3(NT), 0, s>>>(r, q, n, scal, alpha_slot, partial, fc);
  TOFE_LAUNCH_CHECK();
}

void cg_update_xp(hipStream_t s, double *p, const double *z, double *x, int64_t n, const double *scal,
                  int beta_slot, int alpha_slot, int nvec)
{
  k_cg_xp<<<dim3(stride_blocks(n), nvec), dim3(NT), 0, s>>>(p, z, x, n, scal, beta_slot, alpha_slot);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void power_start(hipStream_t s, T *v, int64_t n)
{
  k_power_start<T><<<dim3(stride_blocks(n)), dim3(NT), 0, s>>>(v, n);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void jacobi_dot(hipStream_t s, const T *r, const T *dinv, T *z, int64_t n, double *partial)
{
  k_jacobi_dot<T><<<dim3(kRedBlocks), dim3(NT), 0, s>>>(r, dinv, z, n, partial);
  TOFE_LAUNCH_CHECK();
}

void record_scalar(hipStream_t s, const double *scal, int slot, double *dst)
{
  k_record<<<dim3(1), dim3(64), 0, s>>>(scal, slot, dst);
  TOFE_LAUNCH_CHECK();
}

void fcg_update_r(hipStream_t s, double *r, const double *q, int64_t n, const double *pq_part,
                  const double *pr_part, int nparts, double *scal, double *partial, const FirstCheb &fc,
                  int nvec)
{
  k_fcg_xr<<<dim3(kRedBlocks, nvec), dim3(NT), 0, s>>>(r, q, n, pq_part, pr_part, nparts, scal, partial, fc);
  TOFE_LAUNCH_CHECK();
}

void fcg_update_xp(hipStream_t s, double *p, const double *z, const float *z32, double *x,
                   int64_t n, const double *zq_part, int nparts, double *scal, int nvec)
{
  k_fcg_p<<<dim3(stride_blocks(n), nvec), dim3(NT), 0, s>>>(p, z, z32, x, n, zq_part, nparts, scal);
  TOFE_LAUNCH_CHECK();
}

int fcg_update_r_blocks() { return kRedBlocks; }
int fcg_update_xp_blocks(int64_t n) { return (int)stride_blocks(n); }

void reduce_final_batch(hipStream_t s, const ReduceLaunch *tab, const int *b0, int n, int nvec)
{
  k_reduce_final_batch<<<dim3(n, nvec), dim3(NT), 0, s>>>(tab, b0, n);
  TOFE_LAUNCH_CHECK();
}

void fcg_update_r_batch(hipStream_t s, const FcgXrLaunch *tab, const int *b0, int n,
                        int nblocks_total, int nvec)
{
  k_fcg_xr_batch<<<dim3(nblocks_total, nvec), dim3(NT), 0, s>>>(tab, b0, n);
  TOFE_LAUNCH_CHECK();
}

void fcg_update_xp_batch(hipStream_t s, const FcgPLaunch *tab, const int *b0, int n,
                         int nblocks_total, int nvec)
{
  k_fcg_p_batch<<<dim3(nblocks_total, nvec), dim3(NT), 0, s>>>(tab, b0, n);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void restrict_batch(hipStream_t s, const RestrictLaunch<T> *tab, const int *b0, int n,
                    int nblocks_total, int nvec)
{
  k_restrict_batch<T><<<dim3(nblocks_total, nvec), dim3(NT), 0, s>>>(tab, b0, n);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void prolong_batch(hipStream_t s, const ProlongLaunch<T> *tab, const int *b0, int n,
                   int nblocks_total, int nvec)
{
  k_prolong_add_batch<T><<<dim3(nblocks_total, nvec), dim3(NT), 0, s>>>(tab, b0, n);
  TOFE_LAUNCH_CHECK();
}

void fcg_update_x(hipStream_t s, double *x, const double *p, int64_t n, const double *scal, int nvec)
{
  k_fcg_x<<<dim3(stride_blocks(n), nvec), dim3(NT), 0, s>>>(x, p, n, scal);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void lanczos_r(hipStream_t s, T *r, const T *q, const T *dinv, int64_t n, const double *pq_part,
               int npq, const double *scal, int rz_slot, double *rec_alpha, double *partial)
{
  k_lz_r<T><<<dim3(kRedBlocks), dim3(NT), 0, s>>>(r, q, dinv, n, pq_part, npq, scal, rz_slot,
                                               rec_alpha, partial);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void lanczos_p(hipStream_t s, T *p, const T *r, const T *dinv, int64_t n, const double *rz_part,
               double *scal, int rz_old, int rz_new, double *rec_beta)
{
  k_lz_p<T><<<dim3(stride_blocks(n)), dim3(NT), 0, s>>>(p, r, dinv, n, rz_part, kRedBlocks, scal,
                                                     rz_old, rz_new, rec_beta);
  TOFE_LAUNCH_CHECK();
}

#define TOFE_INST_T(T)                                                                      \
  template void dot_partial<T>(hipStream_t, const T *, const T *, int64_t, double *, int);  \
  template void cheb_zero<T>(hipStream_t, const T *, const T *, T, T *, T *, int64_t, int); \
  template void restrict_op<T>(hipStream_t, const T *, int, int, int, T *, int, int, int,        \
                               const FirstStep<T> &, int);                                    \
  template void prolong_add<T>(hipStream_t, const T *, int, int, int, T *, int, int, int, int); \
  template void restrict_batch<T>(hipStream_t, const RestrictLaunch<T> *, const int *, int, int, int); \
  template void prolong_batch<T>(hipStream_t, const ProlongLaunch<T> *, const int *, int, int, int); \

TOFE_INST_T(double)
TOFE_INST_T(float)
#define TOFE_INST_LZ(T)                                                                      \
  template void power_start<T>(hipStream_t, T *, int64_t);                                  \
  template void jacobi_dot<T>(hipStream_t, const T *, const T *, T *, int64_t, double *);   \
  template void lanczos_r<T>(hipStream_t, T *, const T *, const T *, int64_t, const double *, \
                             int, const double *, int, double *, double *);                 \
  template void lanczos_p<T>(hipStream_t, T *, const T *, const T *, int64_t, const double *, \
                             double *, int, int, double *);
TOFE_INST_LZ(double)
TOFE_INST_LZ(float)
template void convert<double, float>(hipStream_t, const double *, float *, int64_t);
template void convert<float, double>(hipStream_t, const float *, double *, int64_t);

}  // namespace tofe
