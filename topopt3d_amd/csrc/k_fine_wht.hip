// Matrix-free SIMP stiffness operator, element-centric in the Walsh-Hadamard basis.
//
// Same operator as k_fine_op (y = K(E) u with K = sum_e E_e KE, FEAnalysis.cpp:448-464, and
// the MatZeroRowsColumns rows and columns of FEAnalysis.cpp:473): fixed entries of u are
// zeroed on load and fixed rows return u itself.  Each element product is formed as
// KE u_e = H (M (H u_e)) (two 8-point Walsh-Hadamard transforms per displacement component
// and the 45 terms of the block-diagonal M, about 190 operations instead of 576
// multiply-adds), valid whenever KE has the reflection symmetry of the box element, which
// the host checks before it selects this kernel.
//
// A 256-thread block is a 16 x 16 tile of element columns marching up a chunk of z; thread
// (tx, ty) forms element (x0 - 1 + tx, y0 - 1 + ty) of each layer and owns the node at its
// (1, 1) corner, so the block owns 15 x 15 node columns.  A node's four elements of the
// layer are its own, its right neighbour's (lane + 1 in the same row of 16 lanes: DPP) and
// the two of the row above, whose partial sums come through LDS.  Node planes of u are
// staged in a ring of three LDS planes, loaded into registers one step ahead.
#include "k_common.hpp"
#include "ke_wht.hpp"

namespace tofe {
namespace {

constexpr int WT = 16;          // element columns per tile edge, one thread each
constexpr int WN = WT - 1;      // owned node columns per tile edge
constexpr int NW = WT * WT;     // threads per block
constexpr int WP = WT + 1;      // staged node window edge
constexpr int WPL = WP * WP;    // nodes per staged plane

// coordinate index m = x + 2y + 4z of element-local node l
constexpr int mof(int l) { return lnx(l) + 2 * lny(l) + 4 * lnz(l); }

// unnormalised 8-point Walsh-Hadamard transform over the coordinate index
template <typename T>
__device__ inline void wht8(T (&v)[8])
{
#pragma unroll
  for (int b = 1; b < 8; b <<= 1)
#pragma unroll
    for (int m = 0; m < 8; ++m)
      if (!(m & b)) {
        const T a = v[m], c = v[m | b];
        v[m] = a + c;
        v[m | b] = a - c;
      }
}

// value held by lane + 1 of the same row of 16 lanes (0 for the last lane of a row)
template <typename T>
__device__ inline T from_next_lane(T x)
{
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x101, 0xf, 0xf, true));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x101, 0xf, 0xf, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x101, 0xf, 0xf, true);
    return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
  }
}

__device__ inline __amdgpu_buffer_rsrc_t wplane_rsrc(const void *base, int bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, bytes, 0x00020000);
}
template <typename T>
__device__ inline T wplane_load(__amdgpu_buffer_rsrc_t r, unsigned off)
{
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// stored value of block entry B_t[i][j] (cubic elements: its group value)
template <bool CUBE>
constexpr int wht_slot(int t, int i, int j)
{
  return CUBE ? kWhtCubeGroup[wht_index(t, i, j)] : wht_index(t, i, j);
}

// y = E * KE u for the element whose 24 values are v[c][m], in place.
template <typename T, bool CUBE, int NV>
__device__ inline void element_product(T (&v)[3][8], const T (&kb)[NV], T Ee)
{
  wht8(v[0]);
  wht8(v[1]);
  wht8(v[2]);
  // block t couples the entries (j, t ^ e_j), j = 0..2, and writes the same three
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    T x[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) x[j] = v[j][t ^ (1 << j)];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      T acc = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j)
        if (wht_index(t, i, j) >= 0) acc += kb[wht_slot<CUBE>(t, i, j)] * x[j];
      v[i][t ^ (1 << i)] = Ee * acc;
    }
  }
  wht8(v[0]);
  wht8(v[1]);
  wht8(v[2]);
}

// One block of the operator: block `bx` of the launch's `nblocks` (per vector).
template <typename T, int EPI, bool CUBE>
__device__ __forceinline__ void fine_wht_block(const FineOp<T> &op, const T *u0, EpiArgs<T> ep,
                                               int tiles_x, int tiles_y, int nblocks, int zlen,
                                               int blk)
{
  // partial sums handed to the row below (lower, upper node): double-buffered in FP32; one
  // buffer and a second barrier per step in FP64, which keeps the block at 33 KB of LDS (four
  // resident blocks per CU instead of three)
  constexpr int NPB = sizeof(T) == 8 ? 1 : 2;
  __shared__ T su[3][3][WPL];     // ring of three node planes, one array per component
  __shared__ T sp[NPB][6][NW];
  __shared__ double sred[NW / 64];

  const int nx = op.nx, ny = op.ny, nz = op.nz;
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  epi_select_vector(ep, u0, 3 * (int64_t)Nx * Ny * Nz, nblocks);  // vector blockIdx.y
  const T *__restrict__ u = u0;
  const int bid = xcd_remap(blk, nblocks);
  const int txy = tiles_x * tiles_y;
  const int bz = bid / txy;
  const int rem = bid - bz * txy;
  const int by = rem / tiles_x, bx = rem - by * tiles_x;
  const int x0 = bx * WN, y0 = by * WN;  // first owned node column
  const int k0 = bz * zlen, k1 = min(k0 + zlen, Nz);
  const int t = threadIdx.x;
  const int tx = t % WT, ty = t / WT;
  const int ex = x0 - 1 + tx, ey = y0 - 1 + ty;  // element column of this thread
  const int i = x0 + tx, j = y0 + ty;            // owned node column
  const bool own = tx < WN && ty < WN && i < Nx && j < Ny;
  constexpr unsigned kNone = 0xffffffffu;
  const unsigned eidx = (ex >= 0 && ex < nx && ey >= 0 && ey < ny) ? unsigned(ex + nx * ey) : kNone;
  const T *__restrict__ E = op.E;
  const uint8_t *__restrict__ fix = op.fix;
  const int NxNy = Nx * Ny, nxny = nx * ny;

  constexpr int NV = CUBE ? kWhtCubeVals : kWhtVals;
  T kb[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) kb[k] = op.KE[(CUBE ? kWhtCubeOffset : kWhtOffset) + k];

  // staging map: node slot q of this thread is node t + q * NW of the staged window
  // (x0 - 1 .. x0 + 15) x (y0 - 1 .. y0 + 15), its in-plane index or kNone
  constexpr int NU = (WPL + NW - 1) / NW;
  unsigned u_node[NU];
#pragma unroll
  for (int q = 0; q < NU; ++q) {
    const int nl = t + q * NW;
    const int ly = nl / WP, lx = nl - ly * WP;
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
    u_node[q] = (nl < WPL && gx >= 0 && gx < Nx && gy >= 0 && gy < Ny) ? unsigned(gx + Nx * gy) : kNone;
  }
  auto load_u = [&](T (&pu)[NU][3], unsigned (&pf)[NU], int gz) {
    const bool zin = gz >= 0 && gz < Nz;
    const int64_t z = zin ? gz : 0;
    const __amdgpu_buffer_rsrc_t ur = wplane_rsrc(u + 3 * NxNy * z, 3 * NxNy * (int)sizeof(T));
    const __amdgpu_buffer_rsrc_t fr = wplane_rsrc(fix ? fix + NxNy * z : fix, NxNy);
#pragma unroll
    for (int q = 0; q < NU; ++q) {
      pu[q][0] = pu[q][1] = pu[q][2] = 0;
      pf[q] = 0;
      if (zin && u_node[q] != kNone) {
        const unsigned off = u_node[q] * (3 * sizeof(T));
#pragma unroll
        for (int c = 0; c < 3; ++c) pu[q][c] = wplane_load<T>(ur, off + c * sizeof(T));
        if (fix) pf[q] = __builtin_amdgcn_raw_buffer_load_b8(fr, u_node[q], 0, 0);
      }
    }
  };
  // the fixed-DOF mask is applied at the store, after the arithmetic that hides the loads
  auto store_u = [&](const T (&pu)[NU][3], const unsigned (&pf)[NU], int slot) {
#pragma unroll
    for (int q = 0; q < NU; ++q) {
      const int nl = t + q * NW;
      if (nl < WPL) {
#pragma unroll
        for (int c = 0; c < 3; ++c) su[slot][c][nl] = ((pf[q] >> c) & 1u) ? T(0) : pu[q][c];
      }
    }
  };
  auto load_E = [&](int ez) -> T {
    return (ez >= 0 && ez < nz && eidx != kNone) ? E[eidx + (int64_t)nxny * ez] : T(0);
  };

  T alo[3], ahi[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) ahi[c] = 0;
  double dsum = 0, dsum2 = 0;

  T Ecur;
  {
    T pu[NU][3];
    unsigned pf[NU];
    load_u(pu, pf, k0 - 1);
    store_u(pu, pf, 0);
    load_u(pu, pf, k0);
    store_u(pu, pf, 1);
    Ecur = load_E(k0 - 1);
  }
  __syncthreads();
  // Step st (element layer ez = k0 - 1 + st) reads node planes ez and ez + 1 from ring slots
  // st % 3 and (st + 1) % 3, writes partial sums to sp[st % 2] and node plane ez + 2 to slot
  // (st + 2) % 3, then passes the step's barrier.  Those two were last read before the
  // barrier of step st - 1 (FP32); with the single sp buffer of FP64 a second barrier ahead
  // of the sp writes orders them after the previous step's reads.
  int st = 0;
  for (int ez = k0 - 1; ez < k1; ++ez, ++st) {
    const int slo = st % 3, shi = (st + 1) % 3, pb = NPB == 2 ? (st & 1) : 0;
    const bool more = ez + 1 < k1;
    T pu[NU][3];
    unsigned pf[NU];
    T Enext = 0;
    if (more) {
      load_u(pu, pf, ez + 2);
      Enext = load_E(ez + 1);
    }
    // this step's epilogue inputs, loaded ahead of the element arithmetic
    const bool emit = own && ez >= k0;
    const int64_t n = i + (int64_t)Nx * (j + (int64_t)Ny * ez);
    EpiIn<T, EPI> ein;
    if (emit) epi_load<T, EPI>(ep, u, fix, n, ein);
    T v[3][8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      const int nl = (ty + lny(l)) * WP + tx + lnx(l);
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c][mof(l)] = su[lnz(l) ? shi : slo][c][nl];
    }
    element_product<T, CUBE, NV>(v, kb, Ecur);
    // node (i, j): own element's corner (1, 1) [m = 3 lower, 7 upper] + right element's
    // corner (0, 1) [m = 2, 6]; the row above contributes corners (1, 0) and (0, 0)
    // [m = 1, 5 and 0, 4] of its two elements, summed there and passed through LDS
    T plo[3], phi[3];
    if constexpr (NPB == 1) __syncthreads();  // the previous step's readers of sp are done
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      plo[c] = v[c][3] + from_next_lane(v[c][2]);
      phi[c] = v[c][7] + from_next_lane(v[c][6]);
      sp[pb][c][t] = v[c][1] + from_next_lane(v[c][0]);
      sp[pb][3 + c][t] = v[c][5] + from_next_lane(v[c][4]);
    }
    if (more) store_u(pu, pf, (st + 2) % 3);
    __syncthreads();
    const int tb = min(t + WT, NW - 1);  // thread (tx, ty + 1); ty = 15 owns no node
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      alo[c] = ahi[c] + (plo[c] + sp[pb][c][tb]);
      ahi[c] = phi[c] + sp[pb][3 + c][tb];
    }
    if (emit) {
      if (op.spk) {  // spring supports: K + diag(spk) on free DOFs (fixed rows return u below)
#pragma unroll
        for (int c = 0; c < 3; ++c) alo[c] += op.spk[3 * n + c] * ein.u[c];
      }
      epi_store<T, EPI>(ep, n, ein, alo, dsum, dsum2);
    }
    Ecur = Enext;
  }
  if constexpr (EPI == EPI_APPLY_DOT || EPI == EPI_APPLY_DOT2 || EPI == EPI_CHEB_OUT) {
    __syncthreads();
    const double s1 = block_sum<NW>(dsum, sred);
    if (t == 0) ep.partial[blk] = s1;
    if constexpr (EPI == EPI_APPLY_DOT2) {
      __syncthreads();
      const double s2 = block_sum<NW>(dsum2, sred);
      if (t == 0) ep.partial2[blk] = s2;
    }
  }
}

template <typename T, int EPI, bool CUBE>
__global__ __launch_bounds__(NW) void k_fine_wht(FineOp<T> op, const T *u0, EpiArgs<T> ep,
                                                 int tiles_x, int tiles_y, int nblocks, int zlen)
{
  fine_wht_block<T, EPI, CUBE>(op, u0, ep, tiles_x, tiles_y, nblocks, zlen, blockIdx.x);
}

// The operator of several instances in one grid: entry e of the table owns blocks
// [b0[e], b0[e + 1]).  (Cubic and general elements take separate launches: one kernel holding
// both bodies would hold both bodies' LDS arrays.)
template <typename T, int EPI, bool CUBE>
__global__ __launch_bounds__(NW) void k_fine_wht_batch(const FineLaunch<T> *__restrict__ tab,
                                                       const int *__restrict__ b0, int n)
{
  const int e = batch_entry(b0, n, blockIdx.x);
  const FineLaunch<T> L = tab[e];
  fine_wht_block<T, EPI, CUBE>(L.op, L.u, L.ep, L.tiles_x, L.tiles_y, L.nblocks, L.zlen,
                               blockIdx.x - b0[e]);
}

struct WhtGrid {
  int tiles_x, tiles_y, nblocks;
};

template <typename T>
WhtGrid wht_grid(const FineOp<T> &op)
{
  if (op.zlen < 1) throw Error(TOFE_ERR_ARG, "fine operator: invalid z-chunk length");
  if ((int64_t)3 * (op.nx + 1) * (op.ny + 1) * (int64_t)sizeof(T) >= (int64_t(1) << 31))
    throw Error(TOFE_ERR_ARG, "fine operator: node plane of 2 GiB or more");
  WhtGrid g;
  g.tiles_x = (op.nx + 1 + WN - 1) / WN;
  g.tiles_y = (op.ny + 1 + WN - 1) / WN;
  g.nblocks = g.tiles_x * g.tiles_y * ((op.nz + 1 + op.zlen - 1) / op.zlen);
  return g;
}

}  // namespace

int fine_wht_zlen(int nx, int ny, int nz, int ncu, int hide)
{
  // every step forms a full element layer, so a chunk of zlen node planes costs zlen + 1
  // steps; the launch lasts about ceil(blocks / CUs) block-times once `hide` blocks per CU
  // are resident (the larger chunk wins ties)
  const int tiles = ((nx + 1 + WN - 1) / WN) * ((ny + 1 + WN - 1) / WN);
  int best = 8;
  double best_cost = 1e300;
  ncu = std::max(ncu, 1);
  for (int zl = 4; zl <= 24; ++zl) {
    const int nb = tiles * ((nz + 1 + zl - 1) / zl);
    const double cost = (zl + 1.0) * std::max((nb + ncu - 1) / ncu, hide);
    if (cost < best_cost - 1e-9 || (cost < best_cost + 1e-9 && zl > best)) {
      best_cost = cost;
      best = zl;
    }
  }
  return best;
}

template <typename T>
int fine_wht_blocks(const FineOp<T> &op)
{
  return wht_grid(op).nblocks;
}

template <typename T, int EPI>
void launch_wht(hipStream_t s, const FineOp<T> &op, const T *u, const EpiArgs<T> &ep, const WhtGrid &g)
{
  dim3 grid(g.nblocks, ep.nvec), block(NW);
  if (op.wht_cube)
    k_fine_wht<T, EPI, true><<<grid, block, 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y, g.nblocks, op.zlen);
  else
    k_fine_wht<T, EPI, false><<<grid, block, 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y, g.nblocks, op.zlen);
}

template <typename T>
void fine_wht(hipStream_t s, const FineOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep)
{
  const WhtGrid g = wht_grid(op);
  switch (epi) {
    case EPI_APPLY: launch_wht<T, EPI_APPLY>(s, op, u, ep, g); break;
    case EPI_APPLY_DOT: launch_wht<T, EPI_APPLY_DOT>(s, op, u, ep, g); break;
    case EPI_APPLY_DOT2: launch_wht<T, EPI_APPLY_DOT2>(s, op, u, ep, g); break;
    case EPI_RESID: launch_wht<T, EPI_RESID>(s, op, u, ep, g); break;
    case EPI_CHEB: launch_wht<T, EPI_CHEB>(s, op, u, ep, g); break;
    case EPI_CHEB_OUT: launch_wht<T, EPI_CHEB_OUT>(s, op, u, ep, g); break;
  }
  TOFE_LAUNCH_CHECK();
}

template <typename T>
int fine_wht_grid(const FineOp<T> &op, int *tiles_x, int *tiles_y)
{
  const WhtGrid g = wht_grid(op);
  *tiles_x = g.tiles_x;
  *tiles_y = g.tiles_y;
  return g.nblocks;
}

template <typename T, int EPI>
void launch_wht_batch(hipStream_t s, bool cube, const FineLaunch<T> *tab, const int *b0, int n,
                      int nblocks_total, int nvec)
{
  dim3 grid(nblocks_total, nvec), block(NW);
  if (cube) k_fine_wht_batch<T, EPI, true><<<grid, block, 0, s>>>(tab, b0, n);
  else k_fine_wht_batch<T, EPI, false><<<grid, block, 0, s>>>(tab, b0, n);
}

template <typename T>
void fine_wht_batch(hipStream_t s, Epi epi, bool cube, const FineLaunch<T> *tab, const int *b0,
                    int n, int nblocks_total, int nvec)
{
  switch (epi) {
    case EPI_APPLY: launch_wht_batch<T, EPI_APPLY>(s, cube, tab, b0, n, nblocks_total, nvec); break;
    case EPI_APPLY_DOT: launch_wht_batch<T, EPI_APPLY_DOT>(s, cube, tab, b0, n, nblocks_total, nvec); break;
    case EPI_APPLY_DOT2: launch_wht_batch<T, EPI_APPLY_DOT2>(s, cube, tab, b0, n, nblocks_total, nvec); break;
    case EPI_RESID: launch_wht_batch<T, EPI_RESID>(s, cube, tab, b0, n, nblocks_total, nvec); break;
    case EPI_CHEB: launch_wht_batch<T, EPI_CHEB>(s, cube, tab, b0, n, nblocks_total, nvec); break;
    case EPI_CHEB_OUT: launch_wht_batch<T, EPI_CHEB_OUT>(s, cube, tab, b0, n, nblocks_total, nvec); break;
  }
  TOFE_LAUNCH_CHECK();
}

template int fine_wht_grid<double>(const FineOp<double> &, int *, int *);
template int fine_wht_grid<float>(const FineOp<float> &, int *, int *);
template void fine_wht_batch<double>(hipStream_t, Epi, bool, const FineLaunch<double> *, const int *,
                                     int, int, int);
template void fine_wht_batch<float>(hipStream_t, Epi, bool, const FineLaunch<float> *, const int *,
                                    int, int, int);
template int fine_wht_blocks<double>(const FineOp<double> &);
template int fine_wht_blocks<float>(const FineOp<float> &);
template void fine_wht<double>(hipStream_t, const FineOp<double> &, const double *, Epi, const EpiArgs<double> &);
template void fine_wht<float>(hipStream_t, const FineOp<float> &, const float *, Epi, const EpiArgs<float> &);

}  // namespace tofe
