// Device helpers shared by the kernel translation units.
#pragma once

#include "tofe_internal.hpp"

namespace tofe {

// Element-local node order (0,0,0),(1,0,0),(1,1,0),(0,1,0), +z layer (RecMesh.cpp:486-512).
__host__ __device__ constexpr int lidx(int lx, int ly, int lz)
{
  return lz * 4 + (ly ? (lx ? 2 : 3) : (lx ? 1 : 0));
}
__host__ __device__ constexpr int lnx(int l) { return (l == 1 || l == 2 || l == 5 || l == 6) ? 1 : 0; }
__host__ __device__ constexpr int lny(int l) { return (l == 2 || l == 3 || l == 6 || l == 7) ? 1 : 0; }
__host__ __device__ constexpr int lnz(int l) { return l >= 4 ? 1 : 0; }

// Blocks are dispatched round-robin over the 8 XCDs of an MI355X; give each XCD a
// contiguous range of logical tiles so neighbouring tiles share its L2.
__device__ inline int xcd_remap(int bid, int nb)
{
  const int xcd = bid & 7, local = bid >> 3;
  const int per = nb >> 3, rem = nb & 7;
  return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// (i, j, k) of node n of an Nx x Ny grid (x fastest), in 32-bit unsigned arithmetic: node
// counts stay below 2^31 (checked when a context is created), and a 64-bit division costs
// an order of magnitude more instructions than the rest of a light kernel's index work.
__device__ inline void node_ijk(int64_t n, int Nx, int Ny, int &i, int &j, int &k)
{
  const unsigned u = (unsigned)n, t = u / (unsigned)Nx;
  i = (int)(u - t * (unsigned)Nx);
  k = (int)(t / (unsigned)Ny);
  j = (int)(t - (unsigned)k * (unsigned)Ny);
}

// q = t / d, r = t - q d for 0 <= t < 2^32
__device__ inline void udivmod(int64_t t, int64_t d, int &q, int64_t &r)
{
  const unsigned qq = (unsigned)t / (unsigned)d;
  q = (int)qq;
  r = (int64_t)((unsigned)t - qq * (unsigned)d);
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Deterministic block sum; the result is valid in thread 0.  smem holds NT/64 doubles.
template <int NT>
__device__ inline double block_sum(double v, double *smem)
{
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) smem[w] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) t += smem[i];
  }
  return t;
}

// Row-scaled 1-D trilinear transfer weight between fine node i and coarse node J of a
// grid with Nf fine nodes (Interpolation.cpp:135-163, 331-347).  The 3-D weight is the
// product of three of these; all values are exact dyadic numbers.
__host__ __device__ inline double w1d(int i, int J, int Nf)
{
  const int t = i - 2 * J;
  if (t < -1 || t > 1) return 0.0;
  if (t == 0) return 1.0;
  if ((Nf & 1) == 0 && i == Nf - 1) return 1.0;  // last fine node of an even grid: one parent
  return 0.5;
}

// Launches over several instances: one grid over the blocks of all of them, entry e owning
// blocks [b0[e], b0[e + 1]).  The entry of block b (uniform in the block).
__device__ inline int batch_entry(const int *__restrict__ b0, int n, int b)
{
  int lo = 0, hi = n - 1;  // the largest e with b0[e] <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Epilogue arguments and input of vector blockIdx.y of a launch over ep.nvec vectors stored
// back to back (ndof values each); the per-block partial sums of vector v start at
// v * nblocks (the launch's blocks per vector).  Vector 0 keeps the pointers as given.
template <typename T>
__device__ inline void epi_offset(EpiArgs<T> &ep, int64_t v, int64_t ndof, int nblocks)
{
  if (v == 0) return;
  const int64_t vo = v * ndof, po = v * (int64_t)nblocks;
  if (ep.y) ep.y += vo;
  if (ep.b) ep.b += vo;
  if (ep.d) ep.d += vo;
  if (ep.dout) ep.dout += vo;
  if (ep.xout) ep.xout += vo;
  if (ep.xout64) ep.xout64 += vo;
  if (ep.q64) ep.q64 += vo;
  if (ep.partial) ep.partial += po;
  if (ep.partial2) ep.partial2 += po;
}

// Vector blockIdx.y of a launch over several vectors (epi_offset and the operand u).
template <typename T>
__device__ inline void epi_select_vector(EpiArgs<T> &ep, const T *&u, int64_t ndof, int nblocks)
{
  const int64_t v = blockIdx.y;
  epi_offset(ep, v, ndof, nblocks);
  u += v * ndof;
}

// b0 u0 + b1 u1 + b2 u2 rounded as fma(b2, u2, fma(b1, u1, b0 u0)): the one operation order of
// a stencil block row times a node's three values, so that every kernel forming it (one vector
// or several per thread) gets the same bits.
template <typename T>
__device__ __forceinline__ T row3(T b0, T b1, T b2, T u0, T u1, T u2)
{
  if constexpr (sizeof(T) == 4)
    return __builtin_fmaf(b2, u2, __builtin_fmaf(b1, u1, b0 * u0));
  else
    return __builtin_fma(b2, u2, __builtin_fma(b1, u1, b0 * u0));
}

template <typename T, int EPI>
__device__ inline void apply_epilogue(const EpiArgs<T> &ep, int64_t dof, T Au, T uval,
                                      double &dsum, double &dsum2)
{
  if constexpr (EPI == EPI_APPLY) {
    ep.y[dof] = Au;
  } else if constexpr (EPI == EPI_APPLY_DOT) {
    ep.y[dof] = Au;
    dsum += (double)uval * (double)Au;
  } else if constexpr (EPI == EPI_APPLY_DOT2) {
    ep.y[dof] = Au;
    dsum += (double)uval * (double)Au;
    dsum2 += (double)uval * (double)ep.b[dof];
  } else if constexpr (EPI == EPI_RESID) {
    ep.y[dof] = ep.b[dof] - Au;
  } else if constexpr (EPI == EPI_CHEB) {
    const T r = ep.dinv[dof] * (ep.b[dof] - Au);
    const T dn = (ep.c1 == T(0)) ? ep.c2 * r : ep.c1 * ep.d[dof] + ep.c2 * r;
    (ep.dout ? ep.dout : ep.d)[dof] = dn;
    ep.xout[dof] = uval + dn;
  } else if constexpr (EPI == EPI_CHEB_OUT) {
    const T r = ep.dinv[dof] * (ep.b[dof] - Au);
    const T dn = (ep.c1 == T(0)) ? ep.c2 * r : ep.c1 * ep.d[dof] + ep.c2 * r;
    const T xt = uval + dn;
    const double x = (double)xt;
    if (ep.xout64) ep.xout64[dof] = x;
    else ep.xout[dof] = xt;  // the same value, kept in the working precision
    dsum += x * ep.q64[dof];
  }
}

// Epilogue of a node's three DOFs in two halves, so the inputs can be loaded a step ahead
// of the arithmetic that produces A u: epi_load issues every load the epilogue needs
// (u, the fixed-DOF mask, and per EPI b, D^-1, d, q), epi_store forms the outputs exactly as
// apply_epilogue does for each component.
template <typename T, int EPI>
struct EpiIn {
  T u[3], b[3], dinv[3], d[3];
  double q[3];
  unsigned fm;
};

template <typename T, int EPI>
__device__ inline void epi_load(const EpiArgs<T> &ep, const T *__restrict__ u, const uint8_t *__restrict__ fix,
                                int64_t n, EpiIn<T, EPI> &in)
{
  const int64_t d0 = 3 * n;
  in.fm = fix ? fix[n] : 0u;
#pragma unroll
  for (int c = 0; c < 3; ++c) in.u[c] = u[d0 + c];
  if constexpr (EPI == EPI_APPLY_DOT2 || EPI == EPI_RESID || EPI == EPI_CHEB || EPI == EPI_CHEB_OUT) {
    const T *__restrict__ b = ep.b;
#pragma unroll
    for (int c = 0; c < 3; ++c) in.b[c] = b[d0 + c];
  }
  if constexpr (EPI == EPI_CHEB || EPI == EPI_CHEB_OUT) {
    const T *__restrict__ dinv = ep.dinv;
#pragma unroll
    for (int c = 0; c < 3; ++c) in.dinv[c] = dinv[d0 + c];
    if (ep.c1 != T(0)) {
      const T *__restrict__ d = ep.d;
#pragma unroll
      for (int c = 0; c < 3; ++c) in.d[c] = d[d0 + c];
    }
  }
  if constexpr (EPI == EPI_CHEB_OUT) {
    const double *__restrict__ q = ep.q64;
#pragma unroll
    for (int c = 0; c < 3; ++c) in.q[c] = q[d0 + c];
  }
}

template <typename T, int EPI>
__device__ inline void epi_store(const EpiArgs<T> &ep, int64_t n, const EpiIn<T, EPI> &in, const T (&Ku)[3],
                                 double &dsum, double &dsum2)
{
  const int64_t d0 = 3 * n;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t dof = d0 + c;
    const T uval = in.u[c];
    const T Au = ((in.fm >> c) & 1) ? uval : Ku[c];
    if constexpr (EPI == EPI_APPLY) {
      ep.y[dof] = Au;
    } else if constexpr (EPI == EPI_APPLY_DOT) {
      ep.y[dof] = Au;
      dsum += (double)uval * (double)Au;
    } else if constexpr (EPI == EPI_APPLY_DOT2) {
      ep.y[dof] = Au;
      dsum += (double)uval * (double)Au;
      dsum2 += (double)uval * (double)in.b[c];
    } else if constexpr (EPI == EPI_RESID) {
      ep.y[dof] = in.b[c] - Au;
    } else if constexpr (EPI == EPI_CHEB) {
      const T r = in.dinv[c] * (in.b[c] - Au);
      const T dn = (ep.c1 == T(0)) ? ep.c2 * r : ep.c1 * in.d[c] + ep.c2 * r;
      (ep.dout ? ep.dout : ep.d)[dof] = dn;
      ep.xout[dof] = uval + dn;
    } else if constexpr (EPI == EPI_CHEB_OUT) {
      const T r = in.dinv[c] * (in.b[c] - Au);
      const T dn = (ep.c1 == T(0)) ? ep.c2 * r : ep.c1 * in.d[c] + ep.c2 * r;
      const T xt = uval + dn;
      const double x = (double)xt;
      if (ep.xout64) ep.xout64[dof] = x;
      else ep.xout[dof] = xt;  // the same value, kept in the working precision
      dsum += x * in.q[c];
    }
  }
}

}  // namespace tofe
