// Level-1 Galerkin operator P^T K P computed element by element from E and KE.
//
// Every fine element e interpolates from the 2x2x2 window of coarse nodes starting at
// (ex/2, ey/2, ez/2); per direction the 2x2 interpolation block q takes one of three
// forms (even element, odd element, last element of a grid with an even node count whose
// last node has a single parent, Interpolation.cpp:331-347).  So
//   P^T K P = sum_e E_e Q_e^T KE Q_e,   Q_e = q_x (x) q_y (x) q_z (x) I_3,
// with 27 precomputed matrices KEC_f = Q_f^T KE Q_f.  Elements touching a Dirichlet DOF
// use their own masked matrix E_e Q^T (I-M_e) KE (I-M_e) Q, and the unit diagonal of the
// fixed rows (MatZeroRowsColumns, FEAnalysis.cpp:473) contributes P^T M P.  The fine
// stencil is never formed.
#include "k_common.hpp"

namespace tofe {
namespace {

constexpr int NT = 256;

// KCm[m] = E_e Q_f^T (I - M_e) KE (I - M_e) Q_f for masked element m (one block each).
__global__ __launch_bounds__(NT) void k_masked_kc(const MaskedElems me, const double *__restrict__ E,
                                                  const double *__restrict__ KE,
                                                  const double *__restrict__ Qforms,
                                                  double *__restrict__ KCm)
{
  __shared__ double sK[576], sX[576], sQ[64];
  const int m = blockIdx.x;
  const int64_t e = me.elem[m];
  const int f = me.form[m];
  const uint32_t mask = me.mask[m];
  for (int v = threadIdx.x; v < 576; v += NT) {
    const int r = v / 24, c = v - 24 * r;
    sK[v] = (((mask >> r) & 1) || ((mask >> c) & 1)) ? 0.0 : KE[v];
  }
  for (int v = threadIdx.x; v < 64; v += NT) sQ[v] = Qforms[f * 64 + v];
  __syncthreads();
  // X = KEm Qd, Qd[(3l+a)][(3L+b)] = Q[l][L] delta_ab
  for (int v = threadIdx.x; v < 576; v += NT) {
    const int r = v / 24, c = v - 24 * r;
    const int L = c / 3, b = c - 3 * L;
    double s = 0;
#pragma unroll
    for (int l = 0; l < 8; ++l) s += sK[r * 24 + 3 * l + b] * sQ[l * 8 + L];
    sX[v] = s;
  }
  __syncthreads();
  const double Ee = E[e];
  for (int v = threadIdx.x; v < 576; v += NT) {
    const int r = v / 24, c = v - 24 * r;
    const int L = r / 3, a = r - 3 * L;
    double s = 0;
#pragma unroll
    for (int l = 0; l < 8; ++l) s += sQ[l * 8 + L] * sX[(3 * l + a) * 24 + c];
    KCm[(int64_t)m * 576 + v] = Ee * s;
  }
}

__device__ inline int elem_form(int ex, int nel, int Nf)
{
  if (ex & 1) return 1;
  if ((Nf & 1) == 0 && ex == Nf - 2) return 2;
  return 0;
}

// Block (I, I + D) of the level-1 operator for the compile-time offset O: the sum over
// the fine elements whose coarse window holds both nodes, in ascending element order.
// Per direction the candidate elements are ex = 2 I - 2 + d, d = 0..3 (window start
// I - 1 + d / 2); both nodes lie in the window for d = 0..3 (D = 0), 2..3 (D = +1) or
// 0..1 (D = -1), so positions, local node indices and the interpolation form of a regular
// element are compile-time constants and its Q^T KE Q block comes through uniform
// (scalar) loads.  Elements with a Dirichlet DOF (their own masked matrix) and the last
// element of a grid with an even node count (single-parent form) add their own blocks.
template <int O>
__device__ inline void galerkin_block(const double *__restrict__ E, const int *__restrict__ emap,
                                      const double *__restrict__ KEC,
                                      const double *__restrict__ KCm, const int (&Ic)[3],
                                      const int (&nel)[3], const int (&Nf)[3], double (&acc)[9])
{
  constexpr int DX = O % 3 - 1, DY = (O / 3) % 3 - 1, DZ = O / 9 - 1;
  constexpr int X0 = DX > 0 ? 2 : 0, X1 = DX < 0 ? 2 : 4;
  constexpr int Y0 = DY > 0 ? 2 : 0, Y1 = DY < 0 ? 2 : 4;
  constexpr int Z0 = DZ > 0 ? 2 : 0, Z1 = DZ < 0 ? 2 : 4;
  // Every element of the window is visited without a branch, so its loads are issued with
  // the others: an element outside the grid (clamped load, weight 0) or of a special form adds
  // w = 0 times its regular block, which leaves the sum unchanged, so the sum runs over the same
  // elements in the same order as with those skipped.  A special element then adds its own
  // block right away (a branch taken only near Dirichlet faces and at the last element of an
  // axis with an even node count).
#pragma unroll
  for (int dz = Z0; dz < Z1; ++dz) {
    const int ez = 2 * Ic[2] - 2 + dz;
    const bool zin = ez >= 0 && ez < nel[2];
#pragma unroll
    for (int dy = Y0; dy < Y1; ++dy) {
      const int ey = 2 * Ic[1] - 2 + dy;
      const bool yin = ey >= 0 && ey < nel[1];
#pragma unroll
      for (int dx = X0; dx < X1; ++dx) {
        const int ex = 2 * Ic[0] - 2 + dx;
        const bool in = zin && yin && ex >= 0 && ex < nel[0];
        const int li = lidx(1 - (dx >> 1), 1 - (dy >> 1), 1 - (dz >> 1));
        const int lj = lidx(1 - (dx >> 1) + DX, 1 - (dy >> 1) + DY, 1 - (dz >> 1) + DZ);
        const int64_t e = in ? ex + (int64_t)nel[0] * (ey + (int64_t)nel[1] * ez) : 0;
        const int m = in ? emap[e] : -1;
        const double Ee = E[e];
        // single-parent form: even position (even ex) that is the last element of a grid
        // with an even node count
        const bool sx = !(dx & 1) && (Nf[0] & 1) == 0 && ex == Nf[0] - 2;
        const bool sy = !(dy & 1) && (Nf[1] & 1) == 0 && ey == Nf[1] - 2;
        const bool sz = !(dz & 1) && (Nf[2] & 1) == 0 && ez == Nf[2] - 2;
        const bool special = in && (m >= 0 || sx || sy || sz);
        const double w = (in && !special) ? Ee : 0.0;
        const int F = (dx & 1) + 3 * (dy & 1) + 9 * (dz & 1);
        const double *K = KEC + F * 576 + 3 * li * 24 + 3 * lj;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b) acc[a * 3 + b] = __builtin_fma(w, K[a * 24 + b], acc[a * 3 + b]);
        if (special) {
          if (m < 0) {
            const int Fs = (sx ? 2 : (dx & 1)) + 3 * (sy ? 2 : (dy & 1)) + 9 * (sz ? 2 : (dz & 1));
            const double *Ks = KEC + (int64_t)Fs * 576 + 3 * li * 24 + 3 * lj;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
              for (int b = 0; b < 3; ++b) acc[a * 3 + b] = __builtin_fma(Ee, Ks[a * 24 + b], acc[a * 3 + b]);
          } else {
            const double *Km = KCm + (int64_t)m * 576 + 3 * li * 24 + 3 * lj;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
              for (int b = 0; b < 3; ++b) acc[a * 3 + b] += Km[a * 24 + b];
          }
        }
      }
    }
  }
}

// One thread per (coarse node I, coarse offset o >= 13); the transposed block goes to
// (J, 26 - o) so the level-1 operator is exactly symmetric.
// The couplings of level-1 node I to a neighbour I + D outside the grid are zero; no thread
// of k_galerkin_fine writes them (it returns for such offsets), so they are cleared here, the
// only entries of Sc that k_galerkin_fine leaves alone.
__global__ __launch_bounds__(NT) void k_galerkin_zero_outside(int nx, int ny, int nz, double *__restrict__ Sc)
{
  const int Nc[3] = {(nx + 2) / 2, (ny + 2) / 2, (nz + 2) / 2};
  const int64_t ncn = (int64_t)Nc[0] * Nc[1] * Nc[2];
  const int64_t I = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (I >= ncn) return;
  const int o = (int)blockIdx.y;
  const int D[3] = {o % 3 - 1, (o / 3) % 3 - 1, o / 9 - 1};
  int Ic[3];
  node_ijk(I, Nc[0], Nc[1], Ic[0], Ic[1], Ic[2]);
  bool out = false;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (Ic[d] + D[d] < 0 || Ic[d] + D[d] >= Nc[d]) out = true;
  if (!out) return;
#pragma unroll
  for (int q = 0; q < 9; ++q) Sc[(int64_t)(o * 9 + q) * ncn + I] = 0.0;
}

__global__ __launch_bounds__(NT) void k_galerkin_fine(const double *__restrict__ E,
                                                      const int *__restrict__ emap,
                                                      const double *__restrict__ KEC,
                                                      const double *__restrict__ KCm,
                                                      const uint8_t *__restrict__ fix, int nx,
                                                      int ny, int nz, double *__restrict__ Sc)
{
  const int Nf[3] = {nx + 1, ny + 1, nz + 1};
  const int nel[3] = {nx, ny, nz};
  const int Nc[3] = {(Nf[0] + 1) / 2, (Nf[1] + 1) / 2, (Nf[2] + 1) / 2};
  const int64_t ncn = (int64_t)Nc[0] * Nc[1] * Nc[2];
  // grid: x over the coarse nodes, y over the 14 upper offsets
  const int64_t I = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (I >= ncn) return;
  const int o = 13 + (int)blockIdx.y;
  const int D[3] = {o % 3 - 1, (o / 3) % 3 - 1, o / 9 - 1};
  int Ic[3];
  node_ijk(I, Nc[0], Nc[1], Ic[0], Ic[1], Ic[2]);
  int Jc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    Jc[d] = Ic[d] + D[d];
    if (Jc[d] < 0 || Jc[d] >= Nc[d]) return;
  }
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  switch (o) {
    case 13: galerkin_block<13>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 14: galerkin_block<14>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 15: galerkin_block<15>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 16: galerkin_block<16>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 17: galerkin_block<17>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 18: galerkin_block<18>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 19: galerkin_block<19>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 20: galerkin_block<20>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 21: galerkin_block<21>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 22: galerkin_block<22>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 23: galerkin_block<23>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 24: galerkin_block<24>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    case 25: galerkin_block<25>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
    default: galerkin_block<26>(E, emap, KEC, KCm, Ic, nel, Nf, acc); break;
  }
  // P^T M P: unit diagonal of the fixed fine DOFs
  if (fix) {
    int il[3][3], cn[3];
    double wl[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      int c = 0;
      for (int i = 2 * Ic[d] - 1; i <= 2 * Ic[d] + 1; ++i) {
        if (i < 0 || i >= Nf[d]) continue;
        const double w = w1d(i, Ic[d], Nf[d]) * w1d(i, Jc[d], Nf[d]);
        if (w == 0.0) continue;
        il[d][c] = i;
        wl[d][c] = w;
        ++c;
      }
      cn[d] = c;
    }
    for (int cz = 0; cz < cn[2]; ++cz)
      for (int cy = 0; cy < cn[1]; ++cy)
        for (int cx = 0; cx < cn[0]; ++cx) {
          const int64_t n = il[0][cx] + (int64_t)Nf[0] * (il[1][cy] + (int64_t)Nf[1] * il[2][cz]);
          const uint8_t fm = fix[n];
          if (!fm) continue;
          const double w = wl[0][cx] * wl[1][cy] * wl[2][cz];
#pragma unroll
          for (int a = 0; a < 3; ++a)
            if ((fm >> a) & 1) acc[a * 4] += w;
        }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) Sc[(int64_t)(o * 9 + q) * ncn + I] = acc[q];
  if (o != 13) {
    const int64_t J = Jc[0] + (int64_t)Nc[0] * (Jc[1] + (int64_t)Nc[1] * Jc[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) Sc[(int64_t)((26 - o) * 9 + b * 3 + a) * ncn + J] = acc[a * 3 + b];
  }
}

// Jacobi diagonal of the fine K with Dirichlet rows: sum over the node's elements in
// ascending order of the rounded products E_e KE_ii (the assembly order of
// FEAnalysis.cpp:448-464), 1 on fixed DOFs, and diag <= 0 -> 1 (476-484).
__global__ void k_fine_diag(const double *__restrict__ E, const double *__restrict__ KE,
                            const uint8_t *__restrict__ fix, int nx, int ny, int nz,
                            double *diag, double *dinv, const double *__restrict__ spk)
{
#pragma clang fp contract(off)
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  const int64_t nn = (int64_t)Nx * Ny * Nz;
  const int64_t t = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (t >= 3 * nn) return;
  const int64_t n = t / 3;
  const int a = (int)(t - 3 * n);
  int i, j, k;
  node_ijk(n, Nx, Ny, i, j, k);
  double s = 0.0;
  for (int ez = k - 1; ez <= k; ++ez) {
    if (ez < 0 || ez >= nz) continue;
    for (int ey = j - 1; ey <= j; ++ey) {
      if (ey < 0 || ey >= ny) continue;
      for (int ex = i - 1; ex <= i; ++ex) {
        if (ex < 0 || ex >= nx) continue;
        const int l = lidx(i - ex, j - ey, k - ez);
        s = s + E[ex + (int64_t)nx * (ey + (int64_t)ny * ez)] * KE[(3 * l + a) * 25];
      }
    }
  }
  if (spk) s = s + spk[t];  // MatDiagonalSet(K, spKVec, ADD_VALUES), FEAnalysis.cpp:471
  if (fix && ((fix[n] >> a) & 1)) s = 1.0;
  if (!(s > 0)) s = 1.0;
  diag[t] = s;
  dinv[t] = 1.0 / s;
}

// Sc(I, I + o) += sum over fine nodes f shared by the supports of I and I + o of
// w(f, I) w(f, I + o) spk(f) on the diagonal of the 3x3 block (one thread per coarse node and
// offset o = blockIdx.y); identity: coarse == fine, offset 13 only.
__global__ __launch_bounds__(NT) void k_galerkin_springs(const double *__restrict__ spk, int Nfx,
                                                         int Nfy, int Nfz, int identity,
                                                         double *__restrict__ Sc)
{
  const int Ncx = identity ? Nfx : (Nfx + 1) / 2, Ncy = identity ? Nfy : (Nfy + 1) / 2;
  const int Ncz = identity ? Nfz : (Nfz + 1) / 2;
  const int64_t ncn = (int64_t)Ncx * Ncy * Ncz;
  const int64_t I = (int64_t)blockIdx.x * NT + threadIdx.x;
  const int o = identity ? 13 : (int)blockIdx.y;
  if (I >= ncn) return;
  int Ix, Iy, Iz;
  node_ijk(I, Ncx, Ncy, Ix, Iy, Iz);
  const int Jx = Ix + o % 3 - 1, Jy = Iy + (o / 3) % 3 - 1, Jz = Iz + o / 9 - 1;
  if (Jx < 0 || Jx >= Ncx || Jy < 0 || Jy >= Ncy || Jz < 0 || Jz >= Ncz) return;
  double acc[3] = {0.0, 0.0, 0.0};
  if (identity) {
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] = spk[3 * I + a];
  } else {
    for (int fz = 2 * Iz - 1; fz <= 2 * Iz + 1; ++fz) {
      if (fz < 0 || fz >= Nfz) continue;
      const double wz = w1d(fz, Iz, Nfz) * w1d(fz, Jz, Nfz);
      if (wz == 0.0) continue;
      for (int fy = 2 * Iy - 1; fy <= 2 * Iy + 1; ++fy) {
        if (fy < 0 || fy >= Nfy) continue;
        const double wy = w1d(fy, Iy, Nfy) * w1d(fy, Jy, Nfy);
        if (wy == 0.0) continue;
        for (int fx = 2 * Ix - 1; fx <= 2 * Ix + 1; ++fx) {
          if (fx < 0 || fx >= Nfx) continue;
          const double w = w1d(fx, Ix, Nfx) * w1d(fx, Jx, Nfx) * wy * wz;
          if (w == 0.0) continue;
          const int64_t f = fx + (int64_t)Nfx * (fy + (int64_t)Nfy * fz);
#pragma unroll
          for (int a = 0; a < 3; ++a) acc[a] += w * spk[3 * f + a];
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) Sc[(int64_t)(o * 9 + a * 3 + a) * ncn + I] += acc[a];
}

}  // namespace

void galerkin_springs(hipStream_t s, const double *spk, int Nfx, int Nfy, int Nfz, bool identity,
                      double *Sc)
{
  const int64_t ncn = identity ? (int64_t)Nfx * Nfy * Nfz
                               : (int64_t)((Nfx + 1) / 2) * ((Nfy + 1) / 2) * ((Nfz + 1) / 2);
  k_galerkin_springs<<<dim3((unsigned)((ncn + NT - 1) / NT), identity ? 1 : 27), dim3(NT), 0, s>>>(
      spk, Nfx, Nfy, Nfz, identity ? 1 : 0, Sc);
  TOFE_LAUNCH_CHECK();
}

void galerkin_fine(hipStream_t s, const double *E, const int *emap, const double *KEC,
                   const MaskedElems &me, const double *KE, const double *Qforms, double *KCm,
                   const uint8_t *fix, int nx, int ny, int nz, double *Sc)
{
  if (me.count > 0) {
    k_masked_kc<<<dim3((unsigned)me.count), dim3(NT), 0, s>>>(me, E, KE, Qforms, KCm);
    TOFE_LAUNCH_CHECK();
  }
  const int64_t ncn = (int64_t)((nx + 2) / 2) * ((ny + 2) / 2) * ((nz + 2) / 2);
  k_galerkin_zero_outside<<<dim3((unsigned)((ncn + NT - 1) / NT), 27), dim3(NT), 0, s>>>(nx, ny, nz, Sc);
  TOFE_LAUNCH_CHECK();
  k_galerkin_fine<<<dim3((unsigned)((ncn + NT - 1) / NT), 14), dim3(NT), 0, s>>>(E, emap, KEC, KCm,
                                                                                 fix, nx, ny, nz, Sc);
  TOFE_LAUNCH_CHECK();
}

void fine_diag(hipStream_t s, const double *E, const double *KE, const uint8_t *fix, int nx, int ny,
               int nz, double *diag, double *dinv, const double *spk)
{
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  k_fine_diag<<<dim3((unsigned)((3 * nn + NT - 1) / NT)), dim3(NT), 0, s>>>(E, KE, fix, nx, ny, nz,
                                                                          diag, dinv, spk);
  TOFE_LAUNCH_CHECK();
}

}  // namespace tofe
