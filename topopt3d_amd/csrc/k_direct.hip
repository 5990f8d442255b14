// Coarsest multigrid operator straight from the fine level.
//
// The Galerkin hierarchy (PCMGSetGalerkin, FEAnalysis.cpp:412) forms
// A_{L-1} = P_{L-2}^T ... P_0^T K P_0 ... P_{L-2}.  When every element count is a multiple of
// r = 2^(L-1), the row-scaled trilinear transfers of Interpolation.cpp:123-169 compose to
// plain trilinear interpolation from the coarsest grid (nodes every r fine nodes), so each
// fine element lies inside one coarsest cell and
//   A_{L-1} = sum over cells c, positions p in the cell of E(c,p) M_p
//           - sum over elements e touching a Dirichlet DOF of E_e W_e^T (KE - KE_masked,e) W_e
//           + sum over fine nodes n of W_n^T diag(fixed ? 1 : spring stiffness) W_n,
// with M_p = W_p^T KE W_p (24 x 24) the element matrix at position p seen from the cell's 8
// nodes (W_p: trilinear weights of the element's nodes on the cell's nodes).  The first sum is
// one GEMM over all cells, [cells x r^3] (the element scales) times [r^3 x 576] (the table of
// M_p).  The Dirichlet set is fixed per context, so the second sum is E_e times a per-context
// table row per element (one row per distinct position and node masks), and the third is a
// constant per cell (recomputed on the host when the springs change).  The coarse
// factorisation can then start from E alone, beside the Galerkin chain of the smoothed levels,
// instead of after it.  Summation orders are fixed (no atomics): results are reproducible.
#include <algorithm>
#include <map>
#include <utility>

#include "k_common.hpp"

namespace tofe {
namespace {

constexpr int NT = 256;
constexpr int DC = 576;  // a cell's 24 x 24 matrix, row-major over (3 J + a, 3 K + b)

// weight of a node at cell-local offset t (0..r) along one axis on the cell's node j (0 or 1);
// exact in binary for r a power of two
__host__ __device__ inline double wlin(int t, int j, int r)
{
  return j ? (double)t / r : 1.0 - (double)t / r;
}

__host__ __device__ inline double wnode(int tx, int ty, int tz, int J, int r)
{
  return wlin(tx, J & 1, r) * wlin(ty, (J >> 1) & 1, r) * wlin(tz, J >> 2, r);
}

// part[chunk][cell][e] = sum over the chunk's kc positions p of E(cell, p) M[p][e], on the FP64
// matrix cores: a workgroup takes 16 cells and all 576 entries (wavefront w: entries
// 144 w .. 144 w + 143 as nine 16 x 16 tiles), 4 positions per v_mfma_f64_16x16x4.  Operand
// layouts (lane l): A(16x4) element (l % 16, l / 16), B(4x16) element (l / 16, l % 16); output
// item j of lane l is D(l / 16 + 4 j, l % 16).
__global__ __launch_bounds__(NT) void k_direct_gemm(const double *__restrict__ E, int nx, int ny,
                                                    int r, int ncx, int ncy, int ncell,
                                                    const double *__restrict__ M, int kc,
                                                    double *__restrict__ part)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c0 = blockIdx.x * 16, chunk = blockIdx.y;
  const int kq = lane >> 4, mr = lane & 15;
  const int cA = c0 + mr;
  const bool cok = cA < ncell;
  const int cc = cok ? cA : 0;
  const int cx = cc % ncx, cy = (cc / ncx) % ncy, cz = cc / (ncx * ncy);
  const int64_t ebase = (int64_t)cx * r + (int64_t)nx * ((int64_t)cy * r + (int64_t)ny * ((int64_t)cz * r));
  typedef double v4d __attribute__((ext_vector_type(4)));
  v4d acc[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
  const int pb = chunk * kc;
#pragma unroll 4
  for (int p0 = pb; p0 < pb + kc; p0 += 4) {
    const int p = p0 + kq;
    const int px = p % r, py = (p / r) % r, pz = p / (r * r);
    const double a = cok ? E[ebase + px + (int64_t)nx * (py + (int64_t)ny * pz)] : 0.0;
    const double *Mr = M + (int64_t)p * DC + 144 * w + mr;
    double bv[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) bv[t] = Mr[16 * t];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv[t], acc[t], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int cell = c0 + kq + 4 * j;
      if (cell < ncell) part[((int64_t)chunk * ncell + cell) * DC + 144 * w + 16 * t + mr] = acc[t][j];
    }
}

// Dirichlet part of the elements of one cell that touch a fixed DOF, E_e C_t(e), summed in
// element order over slice blockIdx.y of kMaskSplit equal slices of the cell's list (blockIdx.x
// numbers the cells with such elements): corr[slot][slice][e].  C_t = W_p^T (KE - KE_masked) W_p
// is the table row of the element's position and node masks (fixed per context).  Scales and
// row numbers are staged through LDS; eight rows' loads are issued together.
constexpr int kMaskSplit = kDirectMaskSplit;
__global__ __launch_bounds__(NT) void k_direct_masked(const int *__restrict__ range,
                                                      const int64_t *__restrict__ elem,
                                                      const int *__restrict__ tidx,
                                                      const double *__restrict__ table,
                                                      const double *__restrict__ E,
                                                      double *__restrict__ corr)
{
  __shared__ double es[NT];
  __shared__ int ts[NT];
  const int slot = blockIdx.x, slice = blockIdx.y, t = threadIdx.x;
  const int r0 = range[slot], len = range[slot + 1] - r0;
  const int q0 = r0 + (int)((int64_t)len * slice / kMaskSplit);
  const int q1 = r0 + (int)((int64_t)len * (slice + 1) / kMaskSplit);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int qb = q0; qb < q1; qb += NT) {
    const int nq = min(NT, q1 - qb);
    __syncthreads();  // previous chunk read
    if (t < nq) {
      es[t] = E[elem[qb + t]];
      ts[t] = tidx[qb + t];
    }
    for (int q = nq; q < ((nq + 7) & ~7); ++q)
      if (t == q) {  // padding of the last group of eight: zero scale, row 0
        es[t] = 0.0;
        ts[t] = 0;
      }
    __syncthreads();
    for (int q = 0; q < nq; q += 8) {
      double v[8][3];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double *row = table + (int64_t)ts[q + u] * DC;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int e = t + NT * i;
          v[u][i] = e < DC ? row[e] : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[i] += es[q + u] * v[u][i];
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int e = t + NT * i;
    if (e < DC) corr[((int64_t)slot * kMaskSplit + slice) * DC + e] = acc[i];
  }
}

// Cell matrices: the GEMM chunks summed in order, plus the cell's constant part (the fixed DOFs'
// unit diagonal and the springs of the nodes it owns), minus its Dirichlet element part.
__global__ __launch_bounds__(NT) void k_direct_cells(const double *__restrict__ part, int nchunk,
                                                     int ncell, const double *__restrict__ cconst,
                                                     const int *__restrict__ slot_of,
                                                     const double *__restrict__ corr,
                                                     double *__restrict__ Cc)
{
  const int cell = blockIdx.x;
  const int slot = slot_of[cell];
  for (int e = threadIdx.x; e < DC; e += NT) {
    double s = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) s += part[((int64_t)ch * ncell + cell) * DC + e];
    s += cconst[(int64_t)cell * DC + e];
    if (slot >= 0)
      for (int sl = 0; sl < kMaskSplit; ++sl) s -= corr[((int64_t)slot * kMaskSplit + sl) * DC + e];
    Cc[(int64_t)cell * DC + e] = s;
  }
}

// Dense coarsest matrix (column-major, n = 3 Ncx Ncy Ncz): entry (i, j) sums the cell matrices
// of the cells holding both nodes, in cell order.  perm (may be null): the natural DOF of each
// row and column of the dense order.
__global__ __launch_bounds__(NT) void k_direct_dense(const double *__restrict__ Cc, int Ncx, int Ncy,
                                                     int Ncz, int n, double *__restrict__ A,
                                                     const int *__restrict__ perm)
{
  const int64_t idx = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int p = (int)(idx % n), q = (int)(idx / n);
  const int i = perm ? perm[p] : p, j = perm ? perm[q] : q;
  const int I = i / 3, a = i % 3, J = j / 3, b = j % 3;
  const int Ix = I % Ncx, Iy = (I / Ncx) % Ncy, Iz = I / (Ncx * Ncy);
  const int Jx = J % Ncx, Jy = (J / Ncx) % Ncy, Jz = J / (Ncx * Ncy);
  const int ncx = Ncx - 1, ncy = Ncy - 1, ncz = Ncz - 1;
  double s = 0.0;
  if (abs(Ix - Jx) <= 1 && abs(Iy - Jy) <= 1 && abs(Iz - Jz) <= 1) {
    for (int cz = max(max(Iz, Jz) - 1, 0); cz <= min(min(Iz, Jz), ncz - 1); ++cz)
      for (int cy = max(max(Iy, Jy) - 1, 0); cy <= min(min(Iy, Jy), ncy - 1); ++cy)
        for (int cx = max(max(Ix, Jx) - 1, 0); cx <= min(min(Ix, Jx), ncx - 1); ++cx) {
          const int lI = (Ix - cx) + 2 * (Iy - cy) + 4 * (Iz - cz);
          const int lJ = (Jx - cx) + 2 * (Jy - cy) + 4 * (Jz - cz);
          const int cell = cx + ncx * (cy + ncy * cz);
          s += Cc[(int64_t)cell * DC + (3 * lI + a) * 24 + 3 * lJ + b];
        }
  }
  A[idx] = s;
}

}  // namespace

bool direct_coarse_ok(int nx, int ny, int nz, int levels)
{
  if (levels < 2) return false;
  const int r = 1 << (levels - 1);
  return r <= kDirectMaxR && nx % r == 0 && ny % r == 0 && nz % r == 0;
}

void direct_coarse_table(const double KE[576], int r, double *M)
{
  const int r3 = r * r * r;
  for (int p = 0; p < r3; ++p) {
    const int px = p % r, py = (p / r) % r, pz = p / (r * r);
    double W[8][8];
    for (int k = 0; k < 8; ++k)
      for (int J = 0; J < 8; ++J) W[k][J] = wnode(px + lnx(k), py + lny(k), pz + lnz(k), J, r);
    double T[24][24];  // T[(k,a)][(K,b)] = sum_l KE[(k,a)][(l,b)] W[l][K]
    for (int ka = 0; ka < 24; ++ka)
      for (int K = 0; K < 8; ++K)
        for (int b = 0; b < 3; ++b) {
          double s = 0.0;
          for (int l = 0; l < 8; ++l) s += KE[ka * 24 + 3 * l + b] * W[l][K];
          T[ka][3 * K + b] = s;
        }
    double *Mp = M + (size_t)p * DC;
    for (int J = 0; J < 8; ++J)
      for (int a = 0; a < 3; ++a)
        for (int Kb = 0; Kb <= 3 * J + a; ++Kb) {
          double s = 0.0;
          for (int k = 0; k < 8; ++k) s += W[k][J] * T[3 * k + a][Kb];
          Mp[(3 * J + a) * 24 + Kb] = s;
          Mp[Kb * 24 + 3 * J + a] = s;  // exactly symmetric
        }
  }
}

int direct_coarse_chunks(int ncell, int r)
{
  const int r3 = r * r * r, nb = (ncell + 15) / 16;
  int nchunk = 1;
  while (nchunk * 2 <= r3 / 16 && nb * nchunk * 2 <= 1024) nchunk *= 2;
  return nchunk;
}

void direct_coarse(hipStream_t s, const DirectCoarse &d, double *A)
{
  const int r = d.r, ncx = d.nx / r, ncy = d.ny / r, ncz = d.nz / r;
  const int ncell = ncx * ncy * ncz, r3 = r * r * r;
  const int nchunk = direct_coarse_chunks(ncell, r), kc = r3 / nchunk;
  if (kc % 4 != 0 || r > kDirectMaxR) throw Error(TOFE_ERR_ARG, "direct coarse operator: bad shape");
  k_direct_gemm<<<dim3((unsigned)((ncell + 15) / 16), (unsigned)nchunk), dim3(NT), 0, s>>>(
      d.E, d.nx, d.ny, r, ncx, ncy, ncell, d.M, kc, d.part);
  TOFE_LAUNCH_CHECK();
  if (d.nmasked > 0) {
    k_direct_masked<<<dim3((unsigned)d.nmasked, kMaskSplit), dim3(NT), 0, s>>>(d.range, d.elem, d.tidx,
                                                                              d.table, d.E, d.corr);
    TOFE_LAUNCH_CHECK();
  }
  k_direct_cells<<<dim3((unsigned)ncell), dim3(NT), 0, s>>>(d.part, nchunk, ncell, d.cconst, d.slot_of,
                                                           d.corr, d.cells);
  TOFE_LAUNCH_CHECK();
  const int Ncx = ncx + 1, Ncy = ncy + 1, Ncz = ncz + 1, n = 3 * Ncx * Ncy * Ncz;
  const int64_t nn = (int64_t)n * n;
  k_direct_dense<<<dim3((unsigned)((nn + NT - 1) / NT)), dim3(NT), 0, s>>>(d.cells, Ncx, Ncy, Ncz, n, A,
                                                                          d.perm);
  TOFE_LAUNCH_CHECK();
}

void direct_coarse_static(const double KE[576], int nx, int ny, int nz, int r, const uint8_t *fixmask,
                          DirectStatic &st)
{
  const int ncx = nx / r, ncy = ny / r, ncz = nz / r, ncell = ncx * ncy * ncz;
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  st = DirectStatic();
  st.cconst.assign((size_t)ncell * DC, 0.0);
  st.slot_of.assign(ncell, -1);
  st.range.push_back(0);
  if (!fixmask) return;
  auto node = [&](int fx, int fy, int fz) { return (int64_t)fx + (int64_t)Nx * (fy + (int64_t)Ny * fz); };
  std::map<std::pair<int, uint32_t>, int> key;  // (position, node masks) -> table row
  for (int c = 0; c < ncell; ++c) {
    const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
    const size_t first = st.elem.size();
    for (int pz = 0; pz < r; ++pz)
      for (int py = 0; py < r; ++py)
        for (int px = 0; px < r; ++px) {
          const int ex = cx * r + px, ey = cy * r + py, ez = cz * r + pz;
          uint32_t m = 0;
          for (int k = 0; k < 8; ++k)
            m |= (uint32_t)(fixmask[node(ex + lnx(k), ey + lny(k), ez + lnz(k))] & 7u) << (3 * k);
          if (!m) continue;
          const int p = px + r * (py + r * pz);
          auto it = key.find({p, m});
          if (it == key.end()) {
            const int row = (int)key.size();
            it = key.emplace(std::make_pair(p, m), row).first;
            st.table.resize((size_t)(row + 1) * DC, 0.0);
            double W[8][8];
            for (int k = 0; k < 8; ++k)
              for (int J = 0; J < 8; ++J) W[k][J] = wnode(px + lnx(k), py + lny(k), pz + lnz(k), J, r);
            double *C = &st.table[(size_t)row * DC];
            for (int R = 0; R < 24; ++R)
              for (int Q = 0; Q < 24; ++Q) {
                const int J = R / 3, a = R % 3, K = Q / 3, b = Q % 3;
                double v = 0.0;
                for (int k = 0; k < 8; ++k)
                  for (int l = 0; l < 8; ++l)
                    if (((m >> (3 * k + a)) & 1u) || ((m >> (3 * l + b)) & 1u))
                      v += W[k][J] * KE[(3 * k + a) * 24 + 3 * l + b] * W[l][K];
                C[R * 24 + Q] = v;
              }
          }
          st.elem.push_back(ex + (int64_t)nx * (ey + (int64_t)ny * ez));
          st.tidx.push_back(it->second);
        }
    if (st.elem.size() > first) {
      st.slot_of[c] = (int)st.cells.size();
      st.cells.push_back(c);
      st.range.push_back((int)st.elem.size());
    }
  }
  // the fixed DOFs' unit diagonal, by the cell that owns each node
  for (int fz = 0; fz < Nz; ++fz)
    for (int fy = 0; fy < Ny; ++fy)
      for (int fx = 0; fx < Nx; ++fx) {
        const unsigned m = fixmask[node(fx, fy, fz)] & 7u;
        if (!m) continue;
        const int cx = std::min(fx / r, ncx - 1), cy = std::min(fy / r, ncy - 1), cz = std::min(fz / r, ncz - 1);
        const int tx = fx - cx * r, ty = fy - cy * r, tz = fz - cz * r;
        double *C = &st.cconst[(size_t)(cx + ncx * (cy + ncy * cz)) * DC];
        for (int a = 0; a < 3; ++a) {
          if (!((m >> a) & 1u)) continue;
          for (int J = 0; J < 8; ++J)
            for (int K = 0; K < 8; ++K)
              C[(3 * J + a) * 24 + 3 * K + a] += wnode(tx, ty, tz, J, r) * wnode(tx, ty, tz, K, r);
        }
      }
}

void direct_coarse_add_springs(int nx, int ny, int nz, int r, const double *spk, double *cconst)
{
  const int ncx = nx / r, ncy = ny / r, ncz = nz / r;
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  for (int fz = 0; fz < Nz; ++fz)
    for (int fy = 0; fy < Ny; ++fy)
      for (int fx = 0; fx < Nx; ++fx) {
        const int64_t n = (int64_t)fx + (int64_t)Nx * (fy + (int64_t)Ny * fz);
        const int cx = std::min(fx / r, ncx - 1), cy = std::min(fy / r, ncy - 1), cz = std::min(fz / r, ncz - 1);
        const int tx = fx - cx * r, ty = fy - cy * r, tz = fz - cz * r;
        double *C = &cconst[(size_t)(cx + ncx * (cy + ncy * cz)) * DC];
        for (int a = 0; a < 3; ++a) {
          const double k = spk[3 * n + a];
          if (k == 0.0) continue;
          for (int J = 0; J < 8; ++J)
            for (int K = 0; K < 8; ++K)
              C[(3 * J + a) * 24 + 3 * K + a] += wnode(tx, ty, tz, J, r) * k * wnode(tx, ty, tz, K, r);
        }
      }
}

}  // namespace tofe
