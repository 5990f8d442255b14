// Matrix-free SIMP stiffness operator on the matrix cores.
//
// y = K(E) u with K = sum_e E_e KE (FEAnalysis.cpp:448-464) and the Dirichlet rows and
// columns of MatZeroRowsColumns(K, fixed, 1.0) (FEAnalysis.cpp:473).
//
// The element contraction Y_e = KE (E_e u_e) is a GEMM with M = K = 24: per 16 elements,
// Y (24 x 16) = KE (24 x 24) V (24 x 16), V(:,e) = E_e u_e, computed with
// v_mfma_{f32,f64}_16x16x4 as 2 row tiles (rows 16..31 padded with zeros) x 6 k-steps.
// KE lives in the A-operand registers for the whole kernel.
//
// A 512-thread block (8 wavefronts) owns a 31 x 15 tile of node columns and marches up a
// chunk of z.  Per element layer it stages one node plane of u (masked at Dirichlet DOFs)
// and the layer's 32 x 16 element scales in LDS, then processes the layer's elements in
// the four parity classes (ex mod 2, ey mod 2): elements of one class share no node, so
// each class scatters its 24 outputs per element into LDS node accumulators without
// conflicts or atomics.  After the four classes the lower node plane is complete and is
// written through the epilogue.  Summation order is fixed, so results are deterministic.
#include "k_common.hpp"

namespace tofe {
namespace {

constexpr int ETX = 32, ETY = 16, NTH = 512;  // element tile per layer, threads
constexpr int NTX = ETX + 1, NTY = ETY + 1;   // node tile incl. the one-node halo
constexpr int OWNX = ETX - 1, OWNY = ETY - 1; // nodes whose four in-plane elements are all in the tile
constexpr int PL = NTX * NTY * 3;             // values per staged node plane
constexpr int ZL = 12;                        // node planes per z chunk

// Operand layouts (lane l of 64): A(16x4) element (l % 16, l / 16), B(4x16) element
// (l / 16, l % 16).  Output item j of lane l is D(row(l, j), l % 16) with
// row = 4 (l / 16) + j for the FP32 instruction and row = l / 16 + 4 j for the FP64 one.
template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mac(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int kk, int j) { return 4 * kk + j; }
};
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mac(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int kk, int j) { return kk + 4 * j; }
};

struct MfmaGrid {
  int tiles_x, tiles_y, zchunks, nblocks;
};

template <typename T>
MfmaGrid mfma_grid(const FineOp<T> &op)
{
  MfmaGrid g;
  g.tiles_x = (op.nx + 1 + OWNX - 1) / OWNX;
  g.tiles_y = (op.ny + 1 + OWNY - 1) / OWNY;
  g.zchunks = (op.nz + 1 + ZL - 1) / ZL;
  g.nblocks = g.tiles_x * g.tiles_y * g.zchunks;
  return g;
}

template <typename T, int EPI>
__global__ __launch_bounds__(NTH) void k_fine_mfma(FineOp<T> op, const T *u0,
                                                   EpiArgs<T> ep, int tiles_x, int tiles_y,
                                                   int nblocks)
{
  typedef typename Mfma<T>::v4 v4;
  __shared__ T su[2][PL];
  __shared__ T sacc[2][PL];
  __shared__ T sE[ETX * ETY];
  __shared__ double sred[NTH / 64];

  const int nx = op.nx, ny = op.ny, nz = op.nz;
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  epi_select_vector(ep, u0, 3 * (int64_t)Nx * Ny * Nz, gridDim.x);  // vector blockIdx.y
  const T *__restrict__ u = u0;
  const int bid = xcd_remap(blockIdx.x, nblocks);
  const int txy = tiles_x * tiles_y;
  const int bz = bid / txy;
  const int rem = bid - bz * txy;
  const int by = rem / tiles_x, bx = rem - by * tiles_x;
  const int X0 = bx * OWNX, Y0 = by * OWNY;  // first owned node; tile node/element origin is X0-1
  const int k0 = bz * ZL, k1 = min(k0 + ZL, Nz);
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int m = lane & 15, kk = lane >> 4;
  const T *__restrict__ E = op.E;
  const uint8_t *__restrict__ fix = op.fix;

  // A operand: KE rows 16t + m, columns 4s + kk (rows >= 24 are zero padding)
  T afr[2][6];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int r = 16 * t + m;
      afr[t][s] = r < 24 ? op.KE[r * 24 + 4 * s + kk] : T(0);
    }
  // B gather: element DOF q = 4s + kk -> (local node q/3, component q%3); offset inside a
  // node plane relative to the element's (0,0) node and which plane (0 = lower, 1 = upper)
  int goff[6], gup[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int q = 4 * s + kk, L = q / 3, c = q - 3 * (q / 3);
    const int lx = (L == 1 || L == 2 || L == 5 || L == 6) ? 1 : 0;
    const int ly = (L == 2 || L == 3 || L == 6 || L == 7) ? 1 : 0;
    goff[s] = (ly * NTX + lx) * 3 + c;
    gup[s] = L >= 4 ? 1 : 0;
  }
  // scatter targets of the D rows this lane holds: r = 16t + row(kk, j); rows 24..31 are
  // padding (sval = false) and are never scattered
  int soff[2][4], sup[2][4];
  bool sval[2][4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int r = 16 * t + Mfma<T>::row(kk, j);
      sval[t][j] = r < 24;
      if (r > 23) r = 23;
      const int L = r / 3, c = r - 3 * (r / 3);
      const int lx = (L == 1 || L == 2 || L == 5 || L == 6) ? 1 : 0;
      const int ly = (L == 2 || L == 3 || L == 6 || L == 7) ? 1 : 0;
      soff[t][j] = (ly * NTX + lx) * 3 + c;
      sup[t][j] = L >= 4 ? 1 : 0;
    }

  auto stage_u = [&](int buf, int gz) {
    const bool zin = gz >= 0 && gz < Nz;
    for (int v = tid; v < PL; v += NTH) {
      const int nl = v / 3, c = v - 3 * nl;
      const int ly = nl / NTX, lx = nl - ly * NTX;
      const int gx = X0 - 1 + lx, gy = Y0 - 1 + ly;
      T val = 0;
      if (zin && gx >= 0 && gx < Nx && gy >= 0 && gy < Ny) {
        const int64_t n = gx + (int64_t)Nx * (gy + (int64_t)Ny * gz);
        val = u[3 * n + c];
        if (fix && ((fix[n] >> c) & 1)) val = 0;
      }
      su[buf][v] = val;
    }
  };

  for (int v = tid; v < PL; v += NTH) {
    sacc[0][v] = 0;
    sacc[1][v] = 0;
  }
  stage_u(0, k0 - 1);
  double dsum = 0, dsum2 = 0;

  for (int ez = k0 - 1, st = 0; ez < k1; ++ez, ++st) {
    const int lo = st & 1, hi = lo ^ 1;
    stage_u(hi, ez + 1);
    {
      const int ex = X0 - 1 + (tid & (ETX - 1)), ey = Y0 - 1 + (tid / ETX);
      T val = 0;
      if (ez >= 0 && ez < nz && ex >= 0 && ex < nx && ey >= 0 && ey < ny)
        val = E[ex + (int64_t)nx * (ey + (int64_t)ny * ez)];
      sE[tid] = val;
    }
    __syncthreads();
#pragma unroll
    for (int cls = 0; cls < 4; ++cls) {
      const int ex_l = (cls & 1) + 2 * m, ey_l = (cls >> 1) + 2 * w;
      const T Ee = sE[ey_l * ETX + ex_l];
      const int base = (ey_l * NTX + ex_l) * 3;
      v4 d0 = {0, 0, 0, 0}, d1 = {0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const T bval = Ee * su[gup[s] ? hi : lo][base + goff[s]];
        d0 = Mfma<T>::mac(afr[0][s], bval, d0);
        d1 = Mfma<T>::mac(afr[1][s], bval, d1);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) sacc[sup[0][j] ? hi : lo][base + soff[0][j]] += d0[j];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (sval[1][j]) sacc[sup[1][j] ? hi : lo][base + soff[1][j]] += d1[j];
      __syncthreads();
    }
    if (ez >= k0) {
      for (int v = tid; v < OWNX * OWNY; v += NTH) {
        const int ox = v % OWNX, oy = v / OWNX;
        const int gx = X0 + ox, gy = Y0 + oy;
        if (gx >= Nx || gy >= Ny) continue;
        const int64_t n = gx + (int64_t)Nx * (gy + (int64_t)Ny * ez);
        const uint8_t fm = fix ? fix[n] : 0;
        const int lb = ((oy + 1) * NTX + ox + 1) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int64_t dof = 3 * n + c;
          const T uval = u[dof];
          const T Ku = sacc[lo][lb + c];
          const T Au = ((fm >> c) & 1) ? uval : (op.spk ? Ku + op.spk[dof] * uval : Ku);
          apply_epilogue<T, EPI>(ep, dof, Au, uval, dsum, dsum2);
        }
      }
    }
    __syncthreads();
    for (int v = tid; v < PL; v += NTH) sacc[lo][v] = 0;
  }
  if constexpr (EPI == EPI_APPLY_DOT || EPI == EPI_APPLY_DOT2 || EPI == EPI_CHEB_OUT) {
    __syncthreads();
    const double t = block_sum<NTH>(dsum, sred);
    if (tid == 0) ep.partial[blockIdx.x] = t;
    if constexpr (EPI == EPI_APPLY_DOT2) {
      __syncthreads();
      const double t2 = block_sum<NTH>(dsum2, sred);
      if (tid == 0) ep.partial2[blockIdx.x] = t2;
    }
  }
}

}  // namespace

template <typename T>
int fine_mfma_blocks(const FineOp<T> &op)
{
  return mfma_grid(op).nblocks;
}

template <typename T>
void fine_mfma(hipStream_t s, const FineOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep)
{
  const MfmaGrid g = mfma_grid(op);
  dim3 grid(g.nblocks, ep.nvec), block(NTH);
  switch (epi) {
    case EPI_APPLY:
      k_fine_mfma<T, EPI_APPLY><<<grid, block, 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y, g.nblocks);
      break;
    case EPI_APPLY_DOT:
      k_fine_mfma<T, EPI_APPLY_DOT><<<grid, block, 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y, g.nblocks);
      break;
    case EPI_APPLY_DOT2:
      k_fine_mfma<T, EPI_APPLY_DOT2><<<grid, block, 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y, g.nblocks);
      break;
    case EPI_RESID:
      k_fine_mfma<T, EPI_RESID><<<grid, block, 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y, g.nblocks);
      break;
    case EPI_CHEB:
      k_fine_mfma<T, EPI_CHEB><<<grid, block, 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y, g.nblocks);
      break;
    case EPI_CHEB_OUT:
      k_fine_mfma<T, EPI_CHEB_OUT><<<grid, block, 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y, g.nblocks);
      break;
  }
  TOFE_LAUNCH_CHECK();
}

template int fine_mfma_blocks<double>(const FineOp<double> &);
template int fine_mfma_blocks<float>(const FineOp<float> &);
template void fine_mfma<double>(hipStream_t, const FineOp<double> &, const double *, Epi,
                                const EpiArgs<double> &);
template void fine_mfma<float>(hipStream_t, const FineOp<float> &, const float *, Epi,
                               const EpiArgs<float> &);

}  // namespace tofe
