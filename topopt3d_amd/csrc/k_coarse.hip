// Coarse-level operators of the geometric multigrid hierarchy.
//
// Levels >= 1 hold the Galerkin operator A_{l+1} = P_l^T A_l P_l (PCMGSetGalerkin
// PC_MG_GALERKIN_BOTH, FEAnalysis.cpp:412) as a 27-point 3x3-block stencil on the
// coarse node grid.  With the row-scaled trilinear transfers of Interpolation.cpp the
// coarse operator couples each coarse node only with its 26 neighbours, so the
// stencil is exact, including the Dirichlet identity rows carried up from the fine
// level and the degenerate last cell of grids with an odd element count.
#include <cstdlib>

#include "k_common.hpp"

namespace tofe {
namespace {

constexpr int NT = 256;

// SPLIT (small grids): one thread per DOF instead of per node, component-major over the
// blocks (a = blockIdx / node blocks), so a few thousand nodes still fill many waves; each
// DOF's sum runs in the same order either way.
// Blocks are remapped so each XCD works on a contiguous range of nodes, whose
// neighbours' stencil blocks (read again as the transposed lower half) then sit in the
// same L2.
template <bool SPLIT>
__device__ inline void split_index(int64_t nn, int blk, int nb, int64_t &n, int &a0)
{
  const int b = xcd_remap(blk, nb);
  if constexpr (SPLIT) {
    const int nbn = (int)((nn + NT - 1) / NT);
    a0 = b / nbn;
    n = (int64_t)(b - a0 * nbn) * NT + threadIdx.x;
  } else {
    a0 = 0;
    n = (int64_t)b * NT + threadIdx.x;
  }
}

template <typename T, int EPI, bool SPLIT>
__global__ __launch_bounds__(NT) void k_stencil_op(StencilOp<T> op, const T *u0,
                                                   EpiArgs<T> ep, int64_t nn)
{
  __shared__ double sred[NT / 64];
  epi_select_vector(ep, u0, 3 * nn, gridDim.x);  // vector blockIdx.y
  const T *__restrict__ u = u0;
  int64_t n;
  int a0;
  split_index<SPLIT>(nn, blockIdx.x, gridDim.x, n, a0);
  const T *__restrict__ S = op.S;
  double dsum = 0, dsum2 = 0;
  if (n < nn) {
    const int Nx = op.Nx, Ny = op.Ny, Nz = op.Nz;
    int i, j, k;
    node_ijk(n, Nx, Ny, i, j, k);
    constexpr int NA = SPLIT ? 1 : 3;  // components summed by this thread
    T acc[NA];
#pragma unroll
    for (int c = 0; c < NA; ++c) acc[c] = 0;
#pragma unroll
    for (int o = 0; o < 27; ++o) {
      const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
      const int mi = i + dx, mj = j + dy, mk = k + dz;
      // Loads are unconditional (a missing neighbour reads node n) and the contribution is
      // selected afterwards, so no branch separates the 27 neighbours' loads.
      const bool in = mi >= 0 && mi < Nx && mj >= 0 && mj < Ny && mk >= 0 && mk < Nz;
      const int64_t m = in ? n + dx + (int64_t)Nx * (dy + (int64_t)Ny * dz) : n;
      const T u0 = u[3 * m], u1 = u[3 * m + 1], u2 = u[3 * m + 2];
#pragma unroll
      for (int c = 0; c < NA; ++c) {
        const int a = SPLIT ? a0 : c;
        T t = acc[c];
        t += S[(int64_t)(o * 9 + a * 3 + 0) * nn + n] * u0;
        t += S[(int64_t)(o * 9 + a * 3 + 1) * nn + n] * u1;
        t += S[(int64_t)(o * 9 + a * 3 + 2) * nn + n] * u2;
        acc[c] = in ? t : acc[c];
      }
    }
#pragma unroll
    for (int c = 0; c < NA; ++c) {
      const int64_t dof = 3 * n + (SPLIT ? a0 : c);
      apply_epilogue<T, EPI>(ep, dof, acc[c], u[dof], dsum, dsum2);
    }
  }
  if constexpr (EPI == EPI_APPLY_DOT || EPI == EPI_APPLY_DOT2 || EPI == EPI_CHEB_OUT) {
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
  // grid: x over the output nodes, y over the 243 (offset, component pair) values
  const int64_t node = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (node >= nout) return;
  const int oq = blockIdx.y;
  const int64_t t = (int64_t)oq * nout + node;
  const int o = oq / 9, q = oq % 9;
  int c[3];
  node_ijk(node, m[0], m[1], c[0], c[1], c[2]);
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
  // The nine (i, d) terms are loaded unconditionally (an i outside the grid reads node 2I,
  // a missing term has weight 0), so no branch separates the loads; the sum runs in the
  // same (i, d) order as with the missing terms skipped, which they leave unchanged.
  double acc = 0.0;
#pragma unroll
  for (int di = -1; di <= 1; ++di) {
    const int i = 2 * I + di;
    const bool iok = i >= 0 && i < Nf;
    const int ic = iok ? i : 2 * I;
    const double wi = iok ? w1d(i, I, Nf) : 0.0;
#pragma unroll
    for (int d = -1; d <= 1; ++d) {
      const int j = i + d;
      const double wj = (iok && j >= 0 && j < Nf) ? w1d(j, I + D, Nf) : 0.0;
      const int oin = o_base + (d + 1) * so;
      acc += wi * wj * in[(int64_t)(oin * 9 + q) * nin + base_node + (int64_t)ic * sa];
    }
  }
  out[t] = acc;
  if (upper && o != 13) {
    const int64_t J = (c[0] + off[0]) + (int64_t)m[0] * ((c[1] + off[1]) + (int64_t)m[1] * (c[2] + off[2]));
    const int a = q / 3, b = q % 3;
    out[(int64_t)((26 - o) * 9 + b * 3 + a) * nout + J] = acc;
  }
}

// The same 1-D Galerkin product for an intermediate step (no mirroring), three outputs per
// thread: for a fixed output node, (b, c)-offset and component pair, the outputs for the
// three offsets D along the axis read the same nine input values (input nodes 2I - 1 .. 2I + 1
// along the axis, input offsets -1 .. 1 along it), so each is loaded once instead of three
// times.  Each output is summed in the same (i, d) order as in k_galerkin_axis.
__global__ __launch_bounds__(NT) void k_galerkin_axis3(const double *__restrict__ in, int n0, int n1,
                                                       int n2, int axis, double *__restrict__ out)
{
  int m[3] = {n0, n1, n2};
  const int Nf = m[axis];
  const int Nc = (Nf + 1) / 2;
  m[axis] = Nc;
  const int64_t nin = (int64_t)n0 * n1 * n2, nout = (int64_t)m[0] * m[1] * m[2];
  // grid: x over the output nodes, y over the 81 (other-axes offset, component pair) values
  const int64_t node = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (node >= nout) return;
  const int ob = blockIdx.y / 9, q = blockIdx.y % 9;
  int c[3];
  node_ijk(node, m[0], m[1], c[0], c[1], c[2]);
  // offsets along the other two axes (increasing axis order) from ob
  const int ax1 = axis == 0 ? 1 : 0, ax2 = axis == 2 ? 1 : 2;
  int off[3];
  off[ax1] = ob % 3 - 1;
  off[ax2] = ob / 3 - 1;
  off[axis] = 0;
  bool other_in = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k == axis) continue;
    const int cc = c[k] + off[k];
    if (cc < 0 || cc >= m[k]) other_in = false;
  }
  const int I = c[axis];
  const int sa = axis == 0 ? 1 : (axis == 1 ? n0 : n0 * n1);  // input node stride along axis
  const int so = axis == 0 ? 1 : (axis == 1 ? 3 : 9);         // offset-index stride of axis
  const int o_mid = (off[0] + 1) + 3 * (off[1] + 1) + 9 * (off[2] + 1);  // axis offset 0
  int cin[3] = {c[0], c[1], c[2]};
  cin[axis] = 0;
  const int64_t base_node = cin[0] + (int64_t)n0 * (cin[1] + (int64_t)n1 * cin[2]);
  double v[3][3];  // v[di + 1][d + 1] = in(node 2I + di, axis offset d)
#pragma unroll
  for (int di = -1; di <= 1; ++di) {
    const int i = 2 * I + di;
    const int ic = (i >= 0 && i < Nf) ? i : 2 * I;
#pragma unroll
    for (int d = -1; d <= 1; ++d)
      v[di + 1][d + 1] = in[(int64_t)((o_mid + d * so) * 9 + q) * nin + base_node + (int64_t)ic * sa];
  }
#pragma unroll
  for (int D = -1; D <= 1; ++D) {
    const int o = o_mid + D * so;
    const int64_t t = (int64_t)(o * 9 + q) * nout + node;
    if (!other_in || I + D < 0 || I + D >= Nc) {
      out[t] = 0.0;  // the coupled output node does not exist
      continue;
    }
    double acc = 0.0;
#pragma unroll
    for (int di = -1; di <= 1; ++di) {
      const int i = 2 * I + di;
      const bool iok = i >= 0 && i < Nf;
      const double wi = iok ? w1d(i, I, Nf) : 0.0;
#pragma unroll
      for (int d = -1; d <= 1; ++d) {
        const int j = i + d;
        const double wj = (iok && j >= 0 && j < Nf) ? w1d(j, I + D, Nf) : 0.0;
        acc += wi * wj * v[di + 1][d + 1];
      }
    }
    out[t] = acc;
  }
}

// The same operator from the upper half of an exactly symmetric stencil: offsets
// o >= 13 are stored at the node (value (o-13)*9 + a*3 + b, 126 values per node) and
// the block of a lower offset o < 13 is the transpose of the neighbour's block 26 - o.  Each
// stored block is read by its node and by one neighbour, so the HBM traffic is about half of
// the full stencil's.
// Node-blocked layout of the upper-half stencil (FP32 and FP64): value v (0..125) of node n at
// (n / 64) * 126 * 64 + v * 64 + n % 64, so the 126 values of a wavefront's 64 nodes form one
// contiguous range (each value still one coalesced read of 64 consecutive values).
__host__ __device__ inline int64_t su_index(int v, int64_t n)
{
  return (n >> 6) * (126 * 64) + (int64_t)v * 64 + (n & 63);
}

// Neighbours per iteration of the rolled neighbour loop (3, 9 or 27).  27 for FP32 selects
// the fully unrolled form; FP64 keeps one dz plane (9) per iteration, 220 VGPRs and no spills.
// Level-1 Chebyshev step on the 192x96x96 cantilever: FP32 31 us (27) / 34 us (9 or 3); FP64
// 62 us (9) / 75 us (3); the fully unrolled FP64 instance spills.  The FP32 one-thread-per-DOF
// instances take the rolled form's branch-free body with all 27 neighbours unrolled: 5.3 us
// against 5.8 us per Chebyshev step on level 3 (4,225 nodes); per node that body is slower
// (33.6 against 31.2 us on level 1).
#ifndef TOFE_SYM32_G
#define TOFE_SYM32_G 27
#endif
#ifndef TOFE_SYM64_G
#define TOFE_SYM64_G 9
#endif
// One block (blk of nb per vector) of the upper-half stencil operator.
template <typename T, int EPI, bool SPLIT>
__device__ __forceinline__ void stencil_sym_block(const StencilOp<T> &op, const T *u0,
                                                  EpiArgs<T> ep, int64_t nn, int blk, int nb)
{
  __shared__ double sred[NT / 64];
  epi_select_vector(ep, u0, 3 * nn, nb);  // vector blockIdx.y
  const T *__restrict__ u = u0;
  int64_t n;
  int a0;
  split_index<SPLIT>(nn, blk, nb, n, a0);
  const T *__restrict__ S = op.S;
  double dsum = 0, dsum2 = 0;
  if (n < nn) {
    const int Nx = op.Nx, Ny = op.Ny, Nz = op.Nz;
    int i, j, k;
    node_ijk(n, Nx, Ny, i, j, k);
    constexpr int NA = SPLIT ? 1 : 3;
    T acc[NA];
#pragma unroll
    for (int c = 0; c < NA; ++c) acc[c] = 0;
    if constexpr (sizeof(T) == 4 && TOFE_SYM32_G == 27 && !SPLIT) {
      // FP32: fully unrolled; the compiler issues all 27 neighbours' loads up front (323
      // registers, one wave per SIMD), which measured faster than the rolled form below
#pragma unroll
      for (int o = 0; o < 27; ++o) {
        const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
        const int mi = i + dx, mj = j + dy, mk = k + dz;
        // unconditional loads, selected contribution (as in k_stencil_op)
        const bool in = mi >= 0 && mi < Nx && mj >= 0 && mj < Ny && mk >= 0 && mk < Nz;
        const int64_t m = in ? n + dx + (int64_t)Nx * (dy + (int64_t)Ny * dz) : n;
        const T u0 = u[3 * m], u1 = u[3 * m + 1], u2 = u[3 * m + 2];
#pragma unroll
        for (int c = 0; c < NA; ++c) {
          const int a = SPLIT ? a0 : c;
          T t;
          if (o >= 13) {
            const T *B = S + su_index((o - 13) * 9, n);
            t = row3(B[(a * 3 + 0) * 64], B[(a * 3 + 1) * 64], B[(a * 3 + 2) * 64], u0, u1, u2);
          } else {
            const T *B = S + su_index((13 - o) * 9, m);  // block (m, 26 - o), transposed
            t = row3(B[(0 * 3 + a) * 64], B[(1 * 3 + a) * 64], B[(2 * 3 + a) * 64], u0, u1, u2);
          }
          acc[c] = in ? acc[c] + t : acc[c];
        }
      }
    } else {
      // GN neighbours per iteration, the loop kept rolled so that one group's loads are in
      // flight per thread
      constexpr int GN = sizeof(T) == 4 ? (SPLIT ? 27 : TOFE_SYM32_G) : TOFE_SYM64_G;
#pragma unroll 1
      for (int g = 0; g < 27 / GN; ++g) {
#pragma unroll
        for (int ix = 0; ix < GN; ++ix) {
          const int o = GN * g + ix;
          const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
          const int mi = i + dx, mj = j + dy, mk = k + dz;
          // Unconditional loads (a missing neighbour reads node n) weighted by w = 0 or 1, with
          // no branch: fma(w, t, acc) is acc + t or acc exactly.  (A select of acc + t, or a
          // short-circuit bounds test, lets the compiler put each neighbour's loads behind a
          // branch of their own.)
          const bool in = ((unsigned)mi < (unsigned)Nx) & ((unsigned)mj < (unsigned)Ny) &
                          ((unsigned)mk < (unsigned)Nz);
          const T w = in ? T(1) : T(0);
          const int64_t m = in ? n + dx + (int64_t)Nx * (dy + (int64_t)Ny * dz) : n;
          const T u0 = u[3 * m], u1 = u[3 * m + 1], u2 = u[3 * m + 2];
          // o >= 13: stored block (n, o); o < 13: stored block (m, 26 - o), transposed
          const bool up = o >= 13;
          const T *B = S + (up ? su_index((o - 13) * 9, n) : su_index((13 - o) * 9, m));
          const int sr = up ? 3 : 1, sc = up ? 1 : 3;  // strides of (row a, column b)
#pragma unroll
          for (int c = 0; c < NA; ++c) {
            const int a = SPLIT ? a0 : c;
            const T t = row3(B[(a * sr + 0 * sc) * 64], B[(a * sr + 1 * sc) * 64],
                             B[(a * sr + 2 * sc) * 64], u0, u1, u2);
            acc[c] = __builtin_fma(w, t, acc[c]);
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NA; ++c) {
      const int64_t dof = 3 * n + (SPLIT ? a0 : c);
      apply_epilogue<T, EPI>(ep, dof, acc[c], u[dof], dsum, dsum2);
    }
  }
  if constexpr (EPI == EPI_APPLY_DOT || EPI == EPI_APPLY_DOT2 || EPI == EPI_CHEB_OUT) {
    const double t = block_sum<NT>(dsum, sred);
    if (threadIdx.x == 0) ep.partial[blk] = t;
    if constexpr (EPI == EPI_APPLY_DOT2) {
      __syncthreads();
      const double t2 = block_sum<NT>(dsum2, sred);
      if (threadIdx.x == 0) ep.partial2[blk] = t2;
    }
  }
}

template <typename T, int EPI, bool SPLIT>
__global__ __launch_bounds__(NT) void k_stencil_op_sym(StencilOp<T> op, const T *u0,
                                                       EpiArgs<T> ep, int64_t nn)
{
  stencil_sym_block<T, EPI, SPLIT>(op, u0, ep, nn, blockIdx.x, gridDim.x);
}

// The same operator for V vectors per thread (blockIdx.y numbers groups of V vectors): each
// stored stencil value is read once for the group instead of once per vector.  Per vector,
// the neighbour terms (row3), their order and the epilogue are those of stencil_sym_block, so
// every vector's result equals its one-vector launch bit for bit.
template <typename T, int EPI, bool SPLIT, int V>
__device__ __forceinline__ void stencil_sym_block_mv(const StencilOp<T> &op, const T *u0,
                                                     const EpiArgs<T> &ep0, int64_t nn, int blk,
                                                     int nb)
{
  __shared__ double sred[NT / 64];
  const int64_t ndof = 3 * nn;
  const int v0 = blockIdx.y * V;
  int64_t n;
  int a0;
  split_index<SPLIT>(nn, blk, nb, n, a0);
  const T *__restrict__ S = op.S;
  constexpr int NA = SPLIT ? 1 : 3;
  T acc[V][NA];
#pragma unroll
  for (int j = 0; j < V; ++j)
#pragma unroll
    for (int c = 0; c < NA; ++c) acc[j][c] = 0;
  double dsum[V], dsum2[V];
#pragma unroll
  for (int j = 0; j < V; ++j) dsum[j] = dsum2[j] = 0;
  if (n < nn) {
    const int Nx = op.Nx, Ny = op.Ny, Nz = op.Nz;
    int i, j, k;
    node_ijk(n, Nx, Ny, i, j, k);
#pragma unroll 3
    for (int o = 0; o < 27; ++o) {
      const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
      const int mi = i + dx, mj = j + dy, mk = k + dz;
      // unconditional loads (a missing neighbour reads node n), selected contribution
      const bool in = ((unsigned)mi < (unsigned)Nx) & ((unsigned)mj < (unsigned)Ny) &
                      ((unsigned)mk < (unsigned)Nz);
      const int64_t m = in ? n + dx + (int64_t)Nx * (dy + (int64_t)Ny * dz) : n;
      const bool up = o >= 13;
      const T *B = S + (up ? su_index((o - 13) * 9, n) : su_index((13 - o) * 9, m));
      const int sr = up ? 3 : 1, sc = up ? 1 : 3;
      T bv[NA][3];
#pragma unroll
      for (int c = 0; c < NA; ++c) {
        const int a = SPLIT ? a0 : c;
#pragma unroll
        for (int q = 0; q < 3; ++q) bv[c][q] = B[(a * sr + q * sc) * 64];
      }
#pragma unroll
      for (int jv = 0; jv < V; ++jv) {
        const T *__restrict__ u = u0 + (int64_t)(v0 + jv) * ndof;
        const T x0 = u[3 * m], x1 = u[3 * m + 1], x2 = u[3 * m + 2];
#pragma unroll
        for (int c = 0; c < NA; ++c) {
          const T t = row3(bv[c][0], bv[c][1], bv[c][2], x0, x1, x2);
          acc[jv][c] = in ? acc[jv][c] + t : acc[jv][c];
        }
      }
    }
#pragma unroll
    for (int jv = 0; jv < V; ++jv) {
      EpiArgs<T> ep = ep0;
      epi_offset(ep, v0 + jv, ndof, nb);
      const T *__restrict__ u = u0 + (int64_t)(v0 + jv) * ndof;
#pragma unroll
      for (int c = 0; c < NA; ++c) {
        const int64_t dof = 3 * n + (SPLIT ? a0 : c);
        apply_epilogue<T, EPI>(ep, dof, acc[jv][c], u[dof], dsum[jv], dsum2[jv]);
      }
    }
  }
  if constexpr (EPI == EPI_APPLY_DOT || EPI == EPI_APPLY_DOT2 || EPI == EPI_CHEB_OUT) {
#pragma unroll
    for (int jv = 0; jv < V; ++jv) {
      EpiArgs<T> ep = ep0;
      epi_offset(ep, v0 + jv, ndof, nb);
      const double t = block_sum<NT>(dsum[jv], sred);
      if (threadIdx.x == 0) ep.partial[blk] = t;
      __syncthreads();
      if constexpr (EPI == EPI_APPLY_DOT2) {
        const double t2 = block_sum<NT>(dsum2[jv], sred);
        if (threadIdx.x == 0) ep.partial2[blk] = t2;
        __syncthreads();
      }
    }
  }
}

template <typename T, int EPI, bool SPLIT, int V>
__global__ __launch_bounds__(NT) void k_stencil_op_sym_mv(StencilOp<T> op, const T *u0,
                                                          EpiArgs<T> ep, int64_t nn)
{
  stencil_sym_block_mv<T, EPI, SPLIT, V>(op, u0, ep, nn, blockIdx.x, gridDim.x);
}

template <typename T, int EPI, int V>
__global__ __launch_bounds__(NT) void k_stencil_op_sym_mv_batch(const StencilLaunch<T> *__restrict__ tab,
                                                                const int *__restrict__ b0, int n)
{
  const int e = batch_entry(b0, n, blockIdx.x);
  const StencilLaunch<T> L = tab[e];
  const int blk = blockIdx.x - b0[e];
  if (L.nn < L.op.split_nodes)
    stencil_sym_block_mv<T, EPI, true, V>(L.op, L.u, L.ep, L.nn, blk, L.nblocks);
  else
    stencil_sym_block_mv<T, EPI, false, V>(L.op, L.u, L.ep, L.nn, blk, L.nblocks);
}

// The operator of several instances' levels in one grid: entry e owns blocks
// [b0[e], b0[e + 1]); grids of either form (one thread per node or per DOF) share the launch.
template <typename T, int EPI>
__global__ __launch_bounds__(NT) void k_stencil_op_sym_batch(const StencilLaunch<T> *__restrict__ tab,
                                                             const int *__restrict__ b0, int n)
{
  const int e = batch_entry(b0, n, blockIdx.x);
  const StencilLaunch<T> L = tab[e];
  const int blk = blockIdx.x - b0[e];
  if (L.nn < L.op.split_nodes)
    stencil_sym_block<T, EPI, true>(L.op, L.u, L.ep, L.nn, blk, L.nblocks);
  else
    stencil_sym_block<T, EPI, false>(L.op, L.u, L.ep, L.nn, blk, L.nblocks);
}

// FP64 full stencil -> upper half (offsets 13..26) in the node-blocked layout, as To;
// blockIdx.y is the stored value v (0..125), so no thread divides by nn.
template <typename To>
__global__ __launch_bounds__(NT) void k_stencil_upper(const double *__restrict__ S, int64_t nn,
                                                      To *__restrict__ Su)
{
  // One thread per node, all 126 values: a block's four 64-node groups of the node-blocked
  // layout (32 KB each) are written in order and fill whole lines in the cache, where one value
  // per block left 256-byte pieces 32 KB apart for other blocks to fill much later; the reads
  // stay contiguous runs of NT values per stencil plane.
  const int64_t n = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (n >= nn) return;
  const double *__restrict__ src = S + (int64_t)(13 * 9) * nn + n;
  To *__restrict__ dst = Su + su_index(0, n);
#pragma unroll 9
  for (int v = 0; v < 126; ++v) dst[(int64_t)v * 64] = (To)src[(int64_t)v * nn];
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

// inv (may be null): the dense matrix's index of each natural DOF 3 node + c
__global__ void k_stencil_dense(const double *__restrict__ S, int Nx, int Ny, int Nz,
                                double *__restrict__ A, const int *__restrict__ inv)
{
  const int64_t nn = (int64_t)Nx * Ny * Nz, n = 3 * nn;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= 27 * nn) return;
  int o;
  int64_t I;
  udivmod(t, nn, o, I);
  const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
  int i, j, k;
  node_ijk(I, Nx, Ny, i, j, k);
  if (i + dx < 0 || i + dx >= Nx || j + dy < 0 || j + dy >= Ny || k + dz < 0 || k + dz >= Nz) return;
  const int64_t J = I + dx + (int64_t)Nx * (dy + (int64_t)Ny * dz);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int64_t r = inv ? inv[3 * I + a] : 3 * I + a, c = inv ? inv[3 * J + b] : 3 * J + b;
      A[r * n + c] = S[(int64_t)(o * 9 + a * 3 + b) * nn + I];
    }
}

// X = L^-1 (column-major, lower triangular; entries above the diagonal are not stored)
// -> row-major copies of X and of X^T in T, rows padded to ld with zeros outside the
// triangle, for the two triangular GEMVs of the coarse solve.
template <typename T>
__global__ void k_tri_pack(const double *__restrict__ X, int n, int ld, T *__restrict__ Xr,
                           T *__restrict__ XTr)
{
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= (int64_t)n * ld) return;
  const int row = (int)(t / ld), col = (int)(t - (int64_t)row * ld);
  // Xr(row, col) = X(row, col) for col <= row; XTr(row, col) = X(col, row) for col >= row
  Xr[t] = col <= row ? (T)X[row + (int64_t)col * n] : T(0);
  XTr[t] = (col >= row && col < n) ? (T)X[col + (int64_t)row * n] : T(0);
}

// y = A x for a row-major triangular A (padded rows of ld, zeros outside the triangle):
// LOWER rows use columns [0, row], upper rows [row, n).  Four rows per block (one per
// wavefront), x staged in LDS, 16-byte loads along the row, products summed in FP64 in a
// fixed order.  perm (may be null) maps the triangles' order to the vectors' natural order:
// the LOWER product gathers x[perm[i]], the upper one writes y[perm[row]].
template <typename T, bool LOWER>
__device__ __forceinline__ void trigemv_block(const T *__restrict__ A, int n, int ld,
                                              const T *__restrict__ x, T *__restrict__ y, int blk,
                                              const int *__restrict__ perm)
{
  x += (int64_t)blockIdx.y * n;  // vector blockIdx.y of a block of right-hand sides
  y += (int64_t)blockIdx.y * n;
  extern __shared__ unsigned char xs_raw[];
  T *xs = reinterpret_cast<T *>(xs_raw);
  const int row = blk * (NT / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const bool live = row < n;
  constexpr int V = 16 / (int)sizeof(T);  // values per 16-byte load
  typedef T V16 __attribute__((ext_vector_type(V)));
  const int c_lo = LOWER ? 0 : (live ? row / V : 0), c_hi = LOWER ? (live ? row / V + 1 : 0) : ld / V;
  const V16 *a4 = reinterpret_cast<const V16 *>(A + (int64_t)(live ? row : 0) * ld);
  // The row's first U loads per lane are issued before x is staged, so their latency overlaps
  // the staging and its barrier instead of following it one dependent load after another.
  constexpr int U = 8;
  V16 va[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int c = c_lo + lane + 64 * u;
    if (live && c < c_hi) va[u] = a4[c];
  }
  for (int i = threadIdx.x; i < ld; i += NT)
    xs[i] = i < n ? x[(LOWER && perm) ? perm[i] : i] : T(0);
  __syncthreads();
  if (!live) return;
  const V16 *x4 = reinterpret_cast<const V16 *>(xs);
  double s = 0;
  // the products are summed in ascending c per lane, as one load at a time would
  for (int c0 = c_lo + lane; c0 < c_hi; c0 += 64 * U) {
    if (c0 != c_lo + lane) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int c = c0 + 64 * u;
        if (c < c_hi) va[u] = a4[c];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int c = c0 + 64 * u;
      if (c < c_hi) {
        const V16 v = va[u], w = x4[c];
        if constexpr (sizeof(T) == 4)
          s += (double)(v.x * w.x) + (double)(v.y * w.y) + (double)(v.z * w.z) + (double)(v.w * w.w);
        else
          s += v.x * w.x + v.y * w.y;
      }
    }
  }
  s = wave_sum(s);
  if (lane == 0) y[(!LOWER && perm) ? perm[row] : row] = (T)s;
}

template <typename T, bool LOWER>
__global__ __launch_bounds__(NT) void k_trigemv(const T *__restrict__ A, int n, int ld,
                                                const T *__restrict__ x, T *__restrict__ y,
                                                const int *__restrict__ perm)
{
  trigemv_block<T, LOWER>(A, n, ld, x, y, blockIdx.x, perm);
}

// Several instances' products in one grid (the dynamic LDS holds the largest ld).
template <typename T, bool LOWER>
__global__ __launch_bounds__(NT) void k_trigemv_batch(const TrigemvLaunch<T> *__restrict__ tab,
                                                      const int *__restrict__ b0, int n)
{
  const int e = batch_entry(b0, n, blockIdx.x);
  const TrigemvLaunch<T> L = tab[e];
  trigemv_block<T, LOWER>(L.A, L.n, L.ld, L.x, L.y, blockIdx.x - b0[e], L.perm);
}

// ---- dense coarse solve: banded blocked Cholesky, X = L^-1, A^-1 = X^T X ------------
// The coarse matrix is n x n, column-major, SPD, and couples DOFs at most bw apart, so
// the Cholesky factor L stays inside that band.  Tiles are 64 x 64 (kCholNB).

constexpr int CB = kCholNB;

// Value of a double held by lane `lane` (uniform) of the wavefront.
__device__ inline double lane_value(double v, int lane)
{
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Ordering point for LDS traffic inside one wavefront: stores by any lane issued before
// it are seen by loads issued after it (a wavefront's LDS operations complete in order).
__device__ inline void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Cholesky factor of the kb x kb diagonal block at (k0, k0) of A (leading dimension lda),
// padded to 64 with the identity, in one wavefront: lane i keeps row i in registers, each
// finished column j of L is stored once to Ls (column-major) and the other lanes read the
// entries they need from there as broadcasts.  dinv[j] = 1 / L(j,j).  Returns the global
// pivot index + 1 of the first non-positive pivot (0 if none); that pivot is replaced by 1
// so the factorisation runs to the end.
template <bool FULL>
__device__ inline int chol_diag(const double *A, int64_t lda, int k0, int kb, double *Ls,
                                double *dinv)
{
  const int i = threadIdx.x & 63;
  double r[CB];
  if constexpr (FULL) {
    // The dense coarse matrix is stored full (symmetric), so the whole block is loaded
    // unconditionally, all 64 loads in flight together.  Entries above the diagonal then
    // go through the same arithmetic as the rest but are never read as L values: lane i
    // writes column j's entry only for j <= i, later columns read Ls only below the
    // diagonal, and the result is written out with zeros above it.
#pragma unroll
    for (int j = 0; j < CB; ++j) r[j] = A[(k0 + i) + (int64_t)(k0 + j) * lda];
  } else {
    // last, partial block: padded with the identity (clamped loads, selected afterwards)
    const int ic = i < kb ? i : 0;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      const double a = A[(k0 + ic) + (int64_t)(k0 + (j < kb ? j : 0)) * lda];
      r[j] = (i < kb && j < kb) ? (i >= j ? a : 0.0) : (i == j ? 1.0 : 0.0);
    }
  }
  // Software-pipelined: the pivot, square root and reciprocal of column j+1 are formed
  // right after entry j+1 is updated, ahead of the rest of column j's update, so their
  // latency overlaps that work.  A non-positive pivot is replaced by 1 (and reported).
  int bad = 0;
  double d, rd;
  // rd = 1/sqrt(p) from the hardware estimate and two Newton steps, d = sqrt(p) from rd with
  // one correction: about a dozen dependent operations per pivot instead of the twenty-five
  // of a correctly rounded square root and division (both results within about 1 ulp).
  auto pivot = [&](double p, int j) {
    const bool ok = p > 0;
    bad = (bad == 0 && !ok) ? k0 + j + 1 : bad;
    const double q = ok ? p : 1.0;
    double y = __builtin_amdgcn_rsq(q);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const double e = fma(-q * y, y, 1.0);
      y = fma(0.5 * y, e, y);
    }
    const double s0 = q * y;
    d = fma(fma(-s0, s0, q), 0.5 * y, s0);
    rd = y;
  };
  pivot(lane_value(r[0], 0), 0);
#pragma unroll
  for (int j = 0; j < CB; ++j) {
    int ii = i;
    asm volatile("" : "+v"(ii));  // keeps the lane test here instead of 64 hoisted masks
    const double lij = (ii == j) ? d : r[j] * rd;
    Ls[j * CB + i] = lij;
    dinv[j] = rd;  // same value from every lane
    wave_lds_sync();
    // trailing update of row i.  The empty asms pin the column (so its LDS reads are
    // issued together rather than one read and wait at a time) and each updated value
    // (left to itself the compiler defers the updates to each entry's first use and keeps
    // every broadcast column live, which spills).
    double lc[CB];
#pragma unroll
    for (int l = j + 1; l < CB; ++l) lc[l] = Ls[j * CB + l];
#pragma unroll
    for (int l = j + 1; l < CB; ++l) asm volatile("" : "+v"(lc[l]));
    if (j + 1 < CB) {
      r[j + 1] -= lij * lc[j + 1];
      asm volatile("" : "+v"(r[j + 1]));
      pivot(lane_value(r[j + 1], j + 1), j + 1);
    }
#pragma unroll
    for (int l = j + 2; l < CB; ++l) r[l] -= lij * lc[l];
#pragma unroll
    for (int l = j + 2; l < CB; ++l) asm volatile("" : "+v"(r[l]));
  }
  return bad;
}

// Lower tile (I, J), I >= J, of a triangle of tiles from its linear index.
__device__ inline void tile_ij(int b, int &I, int &J)
{
  int i = 0;
  while (b > i) {
    b -= i + 1;
    ++i;
  }
  I = i;
  J = b;
}

// ---- one launch of the pipelined coarse factorisation and inversion -------------------
// Launch k (0 <= k <= nbk + 1) runs three independent parts in one grid:
//  * step k of the banded right-looking Cholesky (k < nbk), one workgroup per lower tile
//    of the trailing window;
//  * row block k-1 of X = L^-1 (1 <= k <= nbk), one workgroup per quarter column block;
//  * the contributions X(k-2, I)^T X(k-2, J) to the lower tiles (I, J) of A^-1 (k >= 2).
// Each part reads only what earlier launches completed (row block k-1 of L is complete
// after launch k-1; X row block k-2 after launch k-1), so the three chains run as one
// pipeline and X and A^-1 are finished two launches after the last factorisation step.
// A^-1 accumulates in place of A: its tile (I, J) lies in column blocks that the
// factorisation has finished with, and the upper triangle is written at the last
// contribution, after the last factorisation step.
constexpr int XCPW = 4, XCW = 4 * XCPW, XPARTS = CB / XCW;
constexpr int kCoarseLds = CB * CB + CB + 3 * CB * CB;  // doubles

// Factorisation step: diagonal block (every workgroup, redundantly, so none waits for
// another), the panel rows P = A21 L11^-T of its row blocks I and J of the window, and
// C(I,J) -= P_I P_J^T.  L goes to Lb (the diagonal block from workgroup 0, the panel rows
// from the diagonal tiles); the panel of A is only read.
template <bool FULL>
__device__ void chol_part(double *A, double *Lb, int n, int k0, int kb, int m, int *info, int b,
                          double *lds)
{
  double *Ls = lds, *dinv = lds + CB * CB, *Pt0 = lds + CB * CB + CB, *Pt1 = Pt0 + CB * CB;
  double *Ct = Pt1 + CB * CB;  // tile C(I,J) of the trailing window, column-major
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  int I, J;
  tile_ij(b, I, J);
  // While wavefront 0 factors the diagonal block, wavefronts 1 and 2 load the panel rows of
  // row blocks I and J and wavefront 3 stages tile C(I,J), so no global load waits on the
  // steps after the factorisation.
  const int w0 = k0 + kb;  // first row and column of the trailing window
  const bool panel_wave = m > 0 && (w == 1 || (w == 2 && J != I));
  const int prow = 64 * (w == 2 ? J : I) + lane;  // this lane's row of the window
  const bool plive = panel_wave && prow < m;
  double x[CB];
  if (panel_wave) {
    const int rowc = plive ? prow : 0;
#pragma unroll
    for (int q = 0; q < CB; ++q) {
      const double a = A[(w0 + rowc) + (int64_t)(k0 + q) * n];
      x[q] = plive ? a : 0.0;
    }
  }
  if (w == 3 && m > 0) {
    const int r = 64 * I + lane, rc = r < m ? r : 0;
    for (int c = 0; c < CB; ++c) {
      const int col = 64 * J + c, cc = col < m ? col : 0;
      Ct[c * CB + lane] = A[(w0 + rc) + (int64_t)(w0 + cc) * n];
    }
  }
  if (w == 0) {
    const int bad = chol_diag<FULL>(A, n, k0, kb, Ls, dinv);
    if (b == 0 && lane == 0 && bad != 0 && info[0] == 0) info[0] = bad;
  }
  __syncthreads();
  if (b == 0) {
    for (int e = t; e < CB * CB; e += 256) {
      const int r = e & 63, cc = e >> 6;
      if (r < kb && cc < kb) Lb[(k0 + r) + (int64_t)(k0 + cc) * n] = r >= cc ? Ls[cc * CB + r] : 0.0;
    }
  }
  if (m <= 0) return;
  // m > 0 means a full diagonal block (kb == 64) with rows below it
  if (panel_wave) {
    double *Pt = (w == 2) ? Pt1 : Pt0;
    const int row = prow;
    const bool live = plive;
    // x L11^T = a: column by column, each finished x[q] updates the later entries
#pragma unroll
    for (int q = 0; q < CB; ++q) {
      x[q] *= dinv[q];
      double lc[CB];  // pinned as in chol_diag
#pragma unroll
      for (int l = q + 1; l < CB; ++l) lc[l] = Ls[q * CB + l];
#pragma unroll
      for (int l = q + 1; l < CB; ++l) asm volatile("" : "+v"(lc[l]));
#pragma unroll
      for (int l = q + 1; l < CB; ++l) x[l] -= x[q] * lc[l];
#pragma unroll
      for (int l = q + 1; l < CB; ++l) asm volatile("" : "+v"(x[l]));
    }
#pragma unroll
    for (int q = 0; q < CB; ++q) Pt[q * CB + lane] = x[q];
    if (I == J && live) {
#pragma unroll
      for (int q = 0; q < CB; ++q) Lb[(w0 + row) + (int64_t)(k0 + q) * n] = x[q];
    }
  }
  __syncthreads();
  // C(I,J) -= P_I P_J^T on the FP64 matrix cores: wavefront w forms rows 16w..16w+15 of the
  // tile as four 16x16 blocks, 16 k-steps of v_mfma_f64_16x16x4 each.  Operand layouts (lane
  // l): A(16x4) element (l % 16, l / 16), B(4x16) element (l / 16, l % 16); output item j of
  // lane l is D(l / 16 + 4 j, l % 16).
  const double *PI = Pt0, *PJ = (I == J) ? Pt0 : Pt1;
  typedef double v4d __attribute__((ext_vector_type(4)));
  const int mr = lane & 15, kq = lane >> 4;
  v4d acc[4];
#pragma unroll
  for (int cbk = 0; cbk < 4; ++cbk) acc[cbk] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int q0 = 0; q0 < CB; q0 += 4) {
    const double a = PI[(q0 + kq) * CB + 16 * w + mr];
#pragma unroll
    for (int cbk = 0; cbk < 4; ++cbk) {
      const double bv = PJ[(q0 + kq) * CB + 16 * cbk + mr];
      acc[cbk] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[cbk], 0, 0, 0);
    }
  }
#pragma unroll
  for (int cbk = 0; cbk < 4; ++cbk)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int row = 64 * I + 16 * w + kq + 4 * jj, col = 64 * J + 16 * cbk + mr;
      if (row < m && col < m && (I != J || row >= col))
        A[(w0 + row) + (int64_t)(w0 + col) * n] = Ct[(16 * cbk + mr) * CB + 16 * w + kq + 4 * jj] - acc[cbk][jj];
    }
}

// Row block ix of X = L^-1, columns [CB kc + XCW part, + XCW): T = E - sum_j L(ix,j) X(j,kc)
// over the row blocks kc <= j < ix inside the band (X(j,kc) staged through LDS), then
// X(ix,kc) = L(ix,ix)^-1 T by forward substitution with lane = row, XCPW columns per wave.
__device__ void x_part(const double *Lb, double *X, int n, int jb, int ix, int kc, int part,
                       double *lds)
{
  double *Ld = lds, *Xs = lds + CB * CB;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int cw0 = CB * kc + XCW * part;  // first column of the workgroup
  const int c0 = cw0 + XCPW * w;         // first column of this wave
  const int64_t grow = (int64_t)CB * ix + lane;
  const bool rok = grow < n;
  const int64_t growc = rok ? grow : 0;
  double T[XCPW];
#pragma unroll
  for (int v = 0; v < XCPW; ++v) T[v] = (grow == c0 + v) ? 1.0 : 0.0;
  const int jlo = (ix - jb > kc) ? ix - jb : kc;
  for (int j = jlo; j < ix; ++j) {
    __syncthreads();  // previous readers of Xs are done
#pragma unroll
    for (int e0 = 0; e0 < CB * XCW; e0 += 256) {
      const int e = e0 + t, r = e & 63, cc = e >> 6;
      const int64_t gc = cw0 + cc, gr = (int64_t)CB * j + r;  // gr < n: j < ix
      const bool ok = gc < n;
      const double v = X[gr + (ok ? gc : 0) * (int64_t)n];
      Xs[r * XCW + cc] = ok ? v : 0.0;
    }
    __syncthreads();
    const double *Lrow = Lb + growc + (int64_t)CB * j * n;
    double lv[CB];  // row of L(ix,j): all 64 loads in flight at once
#pragma unroll
    for (int p = 0; p < CB; ++p) lv[p] = Lrow[(int64_t)p * n];
#pragma unroll
    for (int p0 = 0; p0 < CB; p0 += 16) {
      double xv[16][XCPW];  // 16 rows of this wave's columns, read together (pinned)
#pragma unroll
      for (int p = 0; p < 16; ++p)
#pragma unroll
        for (int v = 0; v < XCPW; ++v) xv[p][v] = Xs[(p0 + p) * XCW + XCPW * w + v];
#pragma unroll
      for (int p = 0; p < 16; ++p)
#pragma unroll
        for (int v = 0; v < XCPW; ++v) asm volatile("" : "+v"(xv[p][v]));
#pragma unroll
      for (int p = 0; p < 16; ++p) {
        const double l = rok ? lv[p0 + p] : 0.0;
#pragma unroll
        for (int v = 0; v < XCPW; ++v) T[v] -= l * xv[p][v];
      }
    }
  }
  __syncthreads();  // Xs readers done before Ld overwrites the start of the buffer
#pragma unroll
  for (int e0 = 0; e0 < CB * CB; e0 += 256) {
    const int e = e0 + t, r = e & 63, cc = e >> 6;
    const int64_t gr = (int64_t)CB * ix + r, gc = (int64_t)CB * ix + cc;
    const bool ok = gr < n && gc < n;
    const double v = Lb[(ok ? gr : 0) + (ok ? gc : 0) * n];
    Ld[e] = ok ? v : (r == cc ? 1.0 : 0.0);
  }
  __syncthreads();
  // Forward substitution with L(ix,ix) scaled to a unit diagonal row by row (lane r
  // divides its row and right-hand side by L(r,r)), so no step divides: at step p lane p
  // holds x_p, and the lanes below subtract L(r,p)/L(r,r) x_p.
  const double rown = 1.0 / Ld[lane * CB + lane];
  double lr[CB];
#pragma unroll
  for (int p = 0; p < CB; ++p) lr[p] = Ld[p * CB + lane];
#pragma unroll
  for (int p = 0; p < CB; ++p) asm volatile("" : "+v"(lr[p]));
#pragma unroll
  for (int v = 0; v < XCPW; ++v) T[v] *= rown;
#pragma unroll
  for (int p = 0; p < CB; ++p) {
    int ii = lane;
    asm volatile("" : "+v"(ii));  // keeps the lane test here instead of hoisted masks
    const double lp = lr[p] * rown;
#pragma unroll
    for (int v = 0; v < XCPW; ++v) {
      const double xp = lane_value(T[v], p);
      const double u = T[v] - lp * xp;
      T[v] = (ii > p) ? u : T[v];
    }
  }
#pragma unroll
  for (int v = 0; v < XCPW; ++v)
    if (rok && c0 + v < n) X[grow + (int64_t)(c0 + v) * n] = T[v];
}

template <bool FULL>
__global__ __launch_bounds__(256) void k_coarse_launch(double *A, double *Lb, double *X, int n,
                                                       int nbk, int jb, int k, int kb, int m,
                                                       int nchol, int nx, int *info)
{
  __shared__ double lds[kCoarseLds];
  const int b = blockIdx.x;
  if (b < nchol) {
    chol_part<FULL>(A, Lb, n, CB * k, kb, m, info, b, lds);
  } else {
    const int xb = b - nchol;
    x_part(Lb, X, n, jb, k - 1, xb / XPARTS, xb % XPARTS, lds);
  }
}

}  // namespace

void coarse_inverse_factor(hipStream_t s, double *A, double *Lb, double *X, int n, int bw, int *info)
{
  const int nbk = (n + CB - 1) / CB;
  const int jb = std::min((bw + CB - 1) / CB, nbk);
  TOFE_HIP(hipMemsetAsync(Lb, 0, sizeof(double) * (size_t)n * n, s));
  for (int k = 0; k <= nbk; ++k) {
    int kb = 0, m = 0, nchol = 0;
    if (k < nbk) {
      kb = std::min(CB, n - CB * k);
      m = std::min(n - CB * k - kb, bw);
      const int nI = (m + CB - 1) / CB;
      nchol = m > 0 ? nI * (nI + 1) / 2 : 1;
    }
    const int nx = k >= 1 ? k * XPARTS : 0;
    const dim3 grid((unsigned)(nchol + nx)), block(256);
    if (kb == CB || k >= nbk)
      k_coarse_launch<true><<<grid, block, 0, s>>>(A, Lb, X, n, nbk, jb, k, kb, m, nchol, nx, info);
    else
      k_coarse_launch<false><<<grid, block, 0, s>>>(A, Lb, X, n, nbk, jb, k, kb, m, nchol, nx, info);
    TOFE_LAUNCH_CHECK();
  }
}

template <typename T>
int stencil_op_blocks(const StencilOp<T> &op)
{
  const int64_t nn = (int64_t)op.Nx * op.Ny * op.Nz;
  const int nb = (int)((nn + NT - 1) / NT);
  return nn < op.split_nodes ? 3 * nb : nb;
}

#define TOFE_STENCIL_SWITCH(KERNEL)                                                         \
  do {                                                                                      \
    const int64_t nn = (int64_t)op.Nx * op.Ny * op.Nz;                                      \
    const bool split = nn < op.split_nodes;                                                 \
    dim3 grid((unsigned)stencil_op_blocks<T>(op), ep.nvec), block(NT);                      \
    switch (epi) {                                                                          \
      case EPI_APPLY:                                                                       \
        if (split) KERNEL<T, EPI_APPLY, true><<<grid, block, 0, s>>>(op, u, ep, nn);        \
        else KERNEL<T, EPI_APPLY, false><<<grid, block, 0, s>>>(op, u, ep, nn);             \
        break;                                                                              \
      case EPI_APPLY_DOT:                                                                   \
        if (split) KERNEL<T, EPI_APPLY_DOT, true><<<grid, block, 0, s>>>(op, u, ep, nn);    \
        else KERNEL<T, EPI_APPLY_DOT, false><<<grid, block, 0, s>>>(op, u, ep, nn);         \
        break;                                                                              \
      case EPI_APPLY_DOT2:                                                                  \
        if (split) KERNEL<T, EPI_APPLY_DOT2, true><<<grid, block, 0, s>>>(op, u, ep, nn);   \
        else KERNEL<T, EPI_APPLY_DOT2, false><<<grid, block, 0, s>>>(op, u, ep, nn);        \
        break;                                                                              \
      case EPI_RESID:                                                                       \
        if (split) KERNEL<T, EPI_RESID, true><<<grid, block, 0, s>>>(op, u, ep, nn);        \
        else KERNEL<T, EPI_RESID, false><<<grid, block, 0, s>>>(op, u, ep, nn);             \
        break;                                                                              \
      case EPI_CHEB:                                                                        \
        if (split) KERNEL<T, EPI_CHEB, true><<<grid, block, 0, s>>>(op, u, ep, nn);         \
        else KERNEL<T, EPI_CHEB, false><<<grid, block, 0, s>>>(op, u, ep, nn);              \
        break;                                                                              \
      default:                                                                              \
        throw Error(TOFE_ERR_ARG, "epilogue not available for stencil levels");             \
    }                                                                                       \
    TOFE_LAUNCH_CHECK();                                                                    \
  } while (0)

template <typename T>
void stencil_op(hipStream_t s, const StencilOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep)
{
  TOFE_STENCIL_SWITCH(k_stencil_op);
}

// Vectors per thread of the multi-vector stencil kernels for a launch over nvec vectors of
// nblocks blocks each (1: the one-vector kernel).  Only the V-cycle's epilogues (Chebyshev
// step, residual) have them, and only launches that still fill the GPU with V vectors per
// thread (kMvMinBlocks blocks) take them: on a small grid the vectors' blocks running side by
// side finish sooner than one block doing V vectors one after the other.  TOFE_STENCIL_MV=0
// keeps every launch on the one-vector kernel, TOFE_STENCIL_MV=2 puts every eligible launch on
// the multi-vector kernels whatever its grid.
constexpr int64_t kMvMinBlocks = 1024;
int stencil_mv_width(int nvec, Epi epi, int64_t nblocks)
{
  const char *e = std::getenv("TOFE_STENCIL_MV");
  const int mode = e ? std::atoi(e) : 1;
  if (mode == 0 || (epi != EPI_CHEB && epi != EPI_RESID)) return 1;
  for (int V : {4, 3, 2})
    if (nvec % V == 0 && (mode == 2 || nblocks * (nvec / V) >= kMvMinBlocks)) return V;
  return 1;
}

template <typename T, int EPI, int V>
void launch_sym_mv(hipStream_t s, const StencilOp<T> &op, const T *u, const EpiArgs<T> &ep)
{
  const int64_t nn = (int64_t)op.Nx * op.Ny * op.Nz;
  const dim3 grid((unsigned)stencil_op_blocks<T>(op), ep.nvec / V), block(NT);
  if (nn < op.split_nodes) k_stencil_op_sym_mv<T, EPI, true, V><<<grid, block, 0, s>>>(op, u, ep, nn);
  else k_stencil_op_sym_mv<T, EPI, false, V><<<grid, block, 0, s>>>(op, u, ep, nn);
}

template <typename T, int EPI>
void launch_sym_mv(hipStream_t s, const StencilOp<T> &op, const T *u, const EpiArgs<T> &ep, int V)
{
  if (V == 4) launch_sym_mv<T, EPI, 4>(s, op, u, ep);
  else if (V == 3) launch_sym_mv<T, EPI, 3>(s, op, u, ep);
  else launch_sym_mv<T, EPI, 2>(s, op, u, ep);
}

template <typename T>
void stencil_op_sym(hipStream_t s, const StencilOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep)
{
  const int V = stencil_mv_width(ep.nvec, epi, stencil_op_blocks<T>(op));
  if (V > 1) {
    if (epi == EPI_CHEB) launch_sym_mv<T, EPI_CHEB>(s, op, u, ep, V);
    else launch_sym_mv<T, EPI_RESID>(s, op, u, ep, V);
    TOFE_LAUNCH_CHECK();
    return;
  }
  TOFE_STENCIL_SWITCH(k_stencil_op_sym);
}

template <typename T, int EPI>
void launch_sym_batch(hipStream_t s, const StencilLaunch<T> *tab, const int *b0, int n,
                      int nblocks_total, int nvec)
{
  k_stencil_op_sym_batch<T, EPI><<<dim3(nblocks_total, nvec), dim3(NT), 0, s>>>(tab, b0, n);
}

template <typename T, int EPI>
void launch_sym_mv_batch(hipStream_t s, const StencilLaunch<T> *tab, const int *b0, int n,
                         int nblocks_total, int nvec, int V)
{
  const dim3 grid(nblocks_total, nvec / V), block(NT);
  if (V == 4) k_stencil_op_sym_mv_batch<T, EPI, 4><<<grid, block, 0, s>>>(tab, b0, n);
  else if (V == 3) k_stencil_op_sym_mv_batch<T, EPI, 3><<<grid, block, 0, s>>>(tab, b0, n);
  else k_stencil_op_sym_mv_batch<T, EPI, 2><<<grid, block, 0, s>>>(tab, b0, n);
}

template <typename T>
void stencil_sym_batch(hipStream_t s, Epi epi, const StencilLaunch<T> *tab, const int *b0, int n,
                       int nblocks_total, int nvec)
{
  const int V = stencil_mv_width(nvec, epi, nblocks_total);
  if (V > 1) {
    if (epi == EPI_CHEB) launch_sym_mv_batch<T, EPI_CHEB>(s, tab, b0, n, nblocks_total, nvec, V);
    else launch_sym_mv_batch<T, EPI_RESID>(s, tab, b0, n, nblocks_total, nvec, V);
    TOFE_LAUNCH_CHECK();
    return;
  }
  switch (epi) {
    case EPI_APPLY: launch_sym_batch<T, EPI_APPLY>(s, tab, b0, n, nblocks_total, nvec); break;
    case EPI_APPLY_DOT: launch_sym_batch<T, EPI_APPLY_DOT>(s, tab, b0, n, nblocks_total, nvec); break;
    case EPI_APPLY_DOT2: launch_sym_batch<T, EPI_APPLY_DOT2>(s, tab, b0, n, nblocks_total, nvec); break;
    case EPI_RESID: launch_sym_batch<T, EPI_RESID>(s, tab, b0, n, nblocks_total, nvec); break;
    case EPI_CHEB: launch_sym_batch<T, EPI_CHEB>(s, tab, b0, n, nblocks_total, nvec); break;
    case EPI_CHEB_OUT: launch_sym_batch<T, EPI_CHEB_OUT>(s, tab, b0, n, nblocks_total, nvec); break;
  }
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void trigemv_batch(hipStream_t s, bool lower, const TrigemvLaunch<T> *tab, const int *b0, int n,
                   int nblocks_total, int nvec, int ld_max)
{
  const dim3 grid(nblocks_total, nvec), block(NT);
  const size_t lds = sizeof(T) * (size_t)ld_max;
  if (lower) k_trigemv_batch<T, true><<<grid, block, lds, s>>>(tab, b0, n);
  else k_trigemv_batch<T, false><<<grid, block, lds, s>>>(tab, b0, n);
  TOFE_LAUNCH_CHECK();
}

template <typename To>
void stencil_upper(hipStream_t s, const double *S, int64_t nn, To *Su)
{
  if (nn <= 0) return;
  k_stencil_upper<To><<<dim3((unsigned)((nn + NT - 1) / NT)), dim3(NT), 0, s>>>(S, nn, Su);
  TOFE_LAUNCH_CHECK();
}

void galerkin(hipStream_t s, const double *Sf, int Nfx, int Nfy, int Nfz, double *Sc, double *T1,
              double *T2)
{
  // x: (Nfx, Nfy, Nfz) -> (Ncx, Nfy, Nfz); y: -> (Ncx, Ncy, Nfz); z: -> (Ncx, Ncy, Ncz)
  const int Ncx = (Nfx + 1) / 2, Ncy = (Nfy + 1) / 2, Ncz = (Nfz + 1) / 2;
  const int64_t n1 = (int64_t)Ncx * Nfy * Nfz, n2 = (int64_t)Ncx * Ncy * Nfz,
                n3 = (int64_t)Ncx * Ncy * Ncz;
  k_galerkin_axis3<<<dim3((unsigned)((n1 + NT - 1) / NT), 81), dim3(NT), 0, s>>>(Sf, Nfx, Nfy, Nfz, 0, T1);
  TOFE_LAUNCH_CHECK();
  k_galerkin_axis3<<<dim3((unsigned)((n2 + NT - 1) / NT), 81), dim3(NT), 0, s>>>(T1, Ncx, Nfy, Nfz, 1, T2);
  TOFE_LAUNCH_CHECK();
  k_galerkin_axis<<<dim3((unsigned)((n3 + NT - 1) / NT), 243), dim3(NT), 0, s>>>(T2, Ncx, Ncy, Nfz, 2,
                                                                                Sc, 1);
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

void stencil_to_dense(hipStream_t s, const double *S, int Nx, int Ny, int Nz, double *A,
                      const int *inv)
{
  const int64_t nn = (int64_t)Nx * Ny * Nz;
  TOFE_HIP(hipMemsetAsync(A, 0, sizeof(double) * 9 * nn * nn, s));
  k_stencil_dense<<<dim3((unsigned)((27 * nn + NT - 1) / NT)), dim3(NT), 0, s>>>(S, Nx, Ny, Nz, A,
                                                                                 inv);
  TOFE_LAUNCH_CHECK();
}


template <typename T>
void tri_pack(hipStream_t s, const double *X, int n, int ld, T *Xr, T *XTr)
{
  k_tri_pack<T><<<dim3((unsigned)(((int64_t)n * ld + NT - 1) / NT)), dim3(NT), 0, s>>>(X, n, ld, Xr, XTr);
  TOFE_LAUNCH_CHECK();
}

template <typename T>
void coarse_solve_tri(hipStream_t s, const T *Xr, const T *XTr, int n, int ld, const T *b, T *tmp, T *x,
                      int nvec, const int *perm)
{
  if (ld % (16 / (int)sizeof(T)) != 0 || ld < n)
    throw Error(TOFE_ERR_ARG, "coarse solve: bad padded leading dimension");
  const dim3 grid((unsigned)((n + 3) / 4), nvec), block(NT);
  const size_t lds = sizeof(T) * (size_t)ld;
  k_trigemv<T, true><<<grid, block, lds, s>>>(Xr, n, ld, b, tmp, perm);
  TOFE_LAUNCH_CHECK();
  k_trigemv<T, false><<<grid, block, lds, s>>>(XTr, n, ld, tmp, x, perm);
  TOFE_LAUNCH_CHECK();
}

template int stencil_op_blocks<double>(const StencilOp<double> &);
template int stencil_op_blocks<float>(const StencilOp<float> &);
template void stencil_op<double>(hipStream_t, const StencilOp<double> &, const double *, Epi,
                                 const EpiArgs<double> &);
template void stencil_op<float>(hipStream_t, const StencilOp<float> &, const float *, Epi,
                                const EpiArgs<float> &);
template void tri_pack<double>(hipStream_t, const double *, int, int, double *, double *);
template void tri_pack<float>(hipStream_t, const double *, int, int, float *, float *);
template void coarse_solve_tri<double>(hipStream_t, const double *, const double *, int, int, const double *,
                                       double *, double *, int, const int *);
template void coarse_solve_tri<float>(hipStream_t, const float *, const float *, int, int, const float *,
                                      float *, float *, int, const int *);

template void stencil_sym_batch<float>(hipStream_t, Epi, const StencilLaunch<float> *, const int *,
                                       int, int, int);
template void stencil_sym_batch<double>(hipStream_t, Epi, const StencilLaunch<double> *, const int *,
                                        int, int, int);
template void trigemv_batch<float>(hipStream_t, bool, const TrigemvLaunch<float> *, const int *, int,
                                   int, int, int);
template void trigemv_batch<double>(hipStream_t, bool, const TrigemvLaunch<double> *, const int *,
                                    int, int, int, int);
template void stencil_upper<float>(hipStream_t, const double *, int64_t, float *);
template void stencil_upper<double>(hipStream_t, const double *, int64_t, double *);
template void stencil_op_sym<double>(hipStream_t, const StencilOp<double> &, const double *, Epi,
                                     const EpiArgs<double> &);
template void stencil_op_sym<float>(hipStream_t, const StencilOp<float> &, const float *, Epi,
                                    const EpiArgs<float> &);
}  // namespace tofe
