// Matrix-free SIMP stiffness operator on the structured Q1 hexahedral mesh.
//
// y = K(E) u with K = sum_e E_e KE (FEAnalysis.cpp:448-464) and the Dirichlet rows and
// columns of MatZeroRowsColumns(K, fixed, 1.0) (FEAnalysis.cpp:473): fixed entries of u
// are zeroed on load and fixed rows return u itself.
//
// Node-centric, deterministic, no atomics: a 256-thread block owns a TX x TY column tile of
// nodes (one of a few compiled shapes, chosen per mesh so the tiles waste few lanes on the
// N = 16k + 1 node counts of the usual meshes) and marches up a chunk of z.  Each step stages one node plane of u (with a
// one-node halo) and one element layer of E in LDS; every thread keeps the 3x3 node
// neighbourhood of the two planes around the current element layer in registers and
// adds the layer's four elements to its lower node (which is then complete and written
// through the epilogue) and to its upper node.  Per node: 8 elements x 3 rows x 24 =
// 576 FMAs.  KE is indexed with compile-time constants only, so it is read through
// scalar loads, uniform across the wavefront.
#include "k_common.hpp"
#include "ke_orbits.hpp"

#include <utility>

namespace tofe {
namespace {

constexpr int NT = 256;  // threads per block (4 wavefronts)

// Compiled column-tile shapes (TX x TY nodes, at most 256); the index is FineOp::tile.
constexpr int kTileShape[kFineTiles][2] = {{32, 8}, {28, 9}, {25, 10}, {22, 11}};

typedef double v8d __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v8f __attribute__((ext_vector_type(8)));

// One 24-entry KE row, read with scalar (SMEM) loads into SGPRs.  The loads are issued
// explicitly so the compiler can neither hoist all 576 KE values out of the z loop nor
// batch them up front; at most one row (48 SGPRs in FP64) is live at a time and every
// FMA takes its KE operand straight from an SGPR.
template <typename T, int OFF>
__device__ inline void ke_row(const T *base, T (&k)[24])
{
  if constexpr (sizeof(T) == 8) {
    v8d a, b, c;
    asm volatile(
        "s_load_dwordx16 %0, %3, %4\n\t"
        "s_load_dwordx16 %1, %3, %5\n\t"
        "s_load_dwordx16 %2, %3, %6\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&s"(a), "=&s"(b), "=&s"(c)
        : "s"(base), "i"(OFF * 8), "i"(OFF * 8 + 64), "i"(OFF * 8 + 128));
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      k[q] = a[q];
      k[8 + q] = b[q];
      k[16 + q] = c[q];
    }
  } else {
    v16f a;
    v8f b;
    asm volatile(
        "s_load_dwordx16 %0, %2, %3\n\t"
        "s_load_dwordx8 %1, %2, %4\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&s"(a), "=&s"(b)
        : "s"(base), "i"(OFF * 4), "i"(OFF * 4 + 64));
#pragma unroll
    for (int q = 0; q < 16; ++q) k[q] = a[q];
#pragma unroll
    for (int q = 0; q < 8; ++q) k[16 + q] = b[q];
  }
}

// Rows 3*L+r of KE applied to the element vector ue, for compile-time L and r.
template <typename T, int L, int R>
__device__ inline T ke_row_dot(const T *base, const T (&ue)[24])
{
  T k[24];
  ke_row<T, (3 * L + R) * 24>(base, k);
  T t = 0;
#pragma unroll
  for (int q = 0; q < 24; ++q) t += k[q] * ue[q];
  // keep this row's FMAs ahead of the next row's loads
  asm volatile("" : "+v"(t));
  __builtin_amdgcn_sched_barrier(0);
  return t;
}

// Cubic elements: KE(x, y) = +-v[orbit(x, y)] with the 12 orbit values v held in scalar
// registers for the whole kernel (loaded once), so no KE row is fetched inside the z loop.
template <int ROW, int Q>
struct KeTerm {
  static constexpr int idx = kKeCube.idx[ROW][Q];
  static constexpr bool neg = kKeCube.neg[ROW][Q];
};
template <typename T, int ROW, int... Q>
__device__ inline T ke_row_dot_sym(const T (&kv)[kKeOrbitsMax], const T (&ue)[24],
                                   std::integer_sequence<int, Q...>)
{
  T t = 0;
  ((t += (KeTerm<ROW, Q>::neg ? -kv[KeTerm<ROW, Q>::idx] : kv[KeTerm<ROW, Q>::idx]) * ue[Q]), ...);
  return t;
}

// Row 3L + R of KE applied to ue, from scalar loads (SYM false) or the orbit values.
template <typename T, bool SYM, int L, int R>
__device__ inline T row_dot(const T *base, const T (&kv)[kKeOrbitsMax], const T (&ue)[24])
{
  if constexpr (SYM)
    return ke_row_dot_sym<T, 3 * L + R>(kv, ue, std::make_integer_sequence<int, 24>());
  else
    return ke_row_dot<T, L, R>(base, ue);
}

// The rows of element (A, B) of the layer for the thread's lower node (want_lo) and upper
// node (want_hi); the flags are uniform across the block.
template <typename T, bool SYM, int A, int B>
__device__ inline void element_rows(const T *base, const T (&kv)[kKeOrbitsMax], const T (&ue)[24], T Ee,
                                    T (&alo)[3], T (&ahi)[3], bool want_lo, bool want_hi)
{
  constexpr int Llo = lidx(1 - A, 1 - B, 0), Lhi = lidx(1 - A, 1 - B, 1);
  if (want_lo) {
    alo[0] += Ee * row_dot<T, SYM, Llo, 0>(base, kv, ue);
    alo[1] += Ee * row_dot<T, SYM, Llo, 1>(base, kv, ue);
    alo[2] += Ee * row_dot<T, SYM, Llo, 2>(base, kv, ue);
  }
  if (want_hi) {
    ahi[0] += Ee * row_dot<T, SYM, Lhi, 0>(base, kv, ue);
    ahi[1] += Ee * row_dot<T, SYM, Lhi, 1>(base, kv, ue);
    ahi[2] += Ee * row_dot<T, SYM, Lhi, 2>(base, kv, ue);
  }
}

// Buffer resource over one staged plane (byte size below 2^31, checked on the host): the
// staging loads then take one 32-bit offset register each instead of a 64-bit address.
__device__ inline __amdgpu_buffer_rsrc_t plane_rsrc(const void *base, int bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, bytes, 0x00020000);
}
template <typename T>
__device__ inline T plane_load(__amdgpu_buffer_rsrc_t r, unsigned off)
{
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

struct FineGrid {
  int tiles_x, tiles_y, zchunks, nblocks;
};

FineGrid fine_grid_of(int nx, int ny, int nz, int tile, int zlen)
{
  FineGrid g;
  const int TX = kTileShape[tile][0], TY = kTileShape[tile][1];
  g.tiles_x = (nx + 1 + TX - 1) / TX;
  g.tiles_y = (ny + 1 + TY - 1) / TY;
  g.zchunks = (nz + 1 + zlen - 1) / zlen;
  g.nblocks = g.tiles_x * g.tiles_y * g.zchunks;
  return g;
}

template <typename T>
FineGrid fine_grid(const FineOp<T> &op)
{
  if (op.tile < 0 || op.tile >= kFineTiles || op.zlen < 1)
    throw Error(TOFE_ERR_ARG, "fine operator: invalid tile shape or z-chunk length");
  if ((int64_t)3 * (op.nx + 1) * (op.ny + 1) * (int64_t)sizeof(T) >= (int64_t(1) << 31))
    throw Error(TOFE_ERR_ARG, "fine operator: node plane of 2 GiB or more");
  return fine_grid_of(op.nx, op.ny, op.nz, op.tile, op.zlen);
}

template <typename T, int EPI, int TX, int TY, bool SYM>
__global__ __launch_bounds__(NT) void k_fine_op(FineOp<T> op, const T *u0,
                                                EpiArgs<T> ep, int tiles_x, int tiles_y,
                                                int nblocks, int zlen)
{
  // threads beyond TX * TY (at most one partial row) mirror row TY - 1 and own no node
  static_assert(TX * TY <= NT && NT - TX * TY < TX, "tile shape");
  constexpr int UW = TX + 2, UH = TY + 2, PLN = UW * UH, UPL = 3 * PLN;  // staged node plane
  constexpr int EW = TX + 1, EH = TY + 1, EPL = EW * EH;                 // staged element layer
  __shared__ T su[3][3][PLN];   // ring of three node planes, one array per component
  __shared__ T sE[2][EPL];
  __shared__ double sred[NT / 64];

  const int nx = op.nx, ny = op.ny, nz = op.nz;
  const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  epi_select_vector(ep, u0, 3 * (int64_t)Nx * Ny * Nz, gridDim.x);  // vector blockIdx.y
  const T *__restrict__ u = u0;
  const int bid = xcd_remap(blockIdx.x, nblocks);
  const int txy = tiles_x * tiles_y;
  const int bz = bid / txy;
  const int rem = bid - bz * txy;
  const int by = rem / tiles_x, bx = rem - by * tiles_x;
  const int x0 = bx * TX, y0 = by * TY;
  const int k0 = bz * zlen, k1 = min(k0 + zlen, Nz);
  const int trow = threadIdx.x / TX;
  const int tx = threadIdx.x - trow * TX, ty = min(trow, TY - 1);
  const int i = x0 + tx, j = y0 + ty;
  const bool own = (trow < TY) && (i < Nx) && (j < Ny);
  const T *KEbase = op.KE;
  T kv[kKeOrbitsMax];
  if constexpr (SYM) {
#pragma unroll
    for (int k = 0; k < kKeOrbitsMax; ++k) kv[k] = op.KE[576 + k];
  }
  const T *__restrict__ E = op.E;
  const uint8_t *__restrict__ fix = op.fix;

  // Staging map, fixed for the block: node slot q of this thread is node nl = t + q * NT of
  // the staged (TX + 2) x (TY + 2) window (its three DOFs are loaded together; lanes read
  // consecutive nodes, so each load covers one contiguous span), element slot q is entry
  // t + q * NT of the staged (TX + 1) x (TY + 1) element window.  Stored: the in-plane node
  // / element index, or kNone outside the mesh or the window.
  constexpr int NU = (PLN + NT - 1) / NT, NE = (EPL + NT - 1) / NT;
  constexpr unsigned kNone = 0xffffffffu;
  const int NxNy = Nx * Ny, nxny = nx * ny;
  unsigned u_node[NU], e_idx[NE];
#pragma unroll
  for (int q = 0; q < NU; ++q) {
    const int nl = threadIdx.x + q * NT;
    const int ly = nl / UW, lx = nl - ly * UW;
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
    u_node[q] = (nl < PLN && gx >= 0 && gx < Nx && gy >= 0 && gy < Ny) ? unsigned(gx + Nx * gy) : kNone;
  }
#pragma unroll
  for (int q = 0; q < NE; ++q) {
    const int v = threadIdx.x + q * NT;
    const int ly = v / EW, lx = v - ly * EW;
    const int ex = x0 - 1 + lx, ey = y0 - 1 + ly;
    e_idx[q] = (v < EPL && ex >= 0 && ex < nx && ey >= 0 && ey < ny) ? unsigned(ex + nx * ey) : kNone;
  }
  // Node plane gz of u (Dirichlet entries zeroed, zero outside the mesh) into registers, and
  // from registers into ring slot `slot`; likewise element layer ez of E.  The loads of the
  // next step's plane and layer are issued before this step's FMAs and stored after them,
  // so their latency hides behind the arithmetic; the fixed-DOF mask is applied at the
  // store (a select at the load would wait for the data there).
  auto load_u = [&](T (&pu)[NU][3], unsigned (&pf)[NU], int gz) {
    const bool zin = gz >= 0 && gz < Nz;
    const int64_t z = zin ? gz : 0;
    const __amdgpu_buffer_rsrc_t ur = plane_rsrc(u + 3 * NxNy * z, 3 * NxNy * (int)sizeof(T));
    const __amdgpu_buffer_rsrc_t fr = plane_rsrc(fix ? fix + NxNy * z : fix, NxNy);
#pragma unroll
    for (int q = 0; q < NU; ++q) {
      pu[q][0] = pu[q][1] = pu[q][2] = 0;
      pf[q] = 0;
      if (zin && u_node[q] != kNone) {
        const unsigned off = u_node[q] * (3 * sizeof(T));
#pragma unroll
        for (int c = 0; c < 3; ++c) pu[q][c] = plane_load<T>(ur, off + c * sizeof(T));
        if (fix) pf[q] = __builtin_amdgcn_raw_buffer_load_b8(fr, u_node[q], 0, 0);
      }
    }
  };
  auto store_u = [&](const T (&pu)[NU][3], const unsigned (&pf)[NU], int slot) {
#pragma unroll
    for (int q = 0; q < NU; ++q) {
      const int nl = threadIdx.x + q * NT;
      if (nl < PLN) {
#pragma unroll
        for (int c = 0; c < 3; ++c) su[slot][c][nl] = ((pf[q] >> c) & 1u) ? T(0) : pu[q][c];
      }
    }
  };
  auto load_E = [&](T (&pe)[NE], int ez) {
    const bool zin = ez >= 0 && ez < nz;
    const __amdgpu_buffer_rsrc_t er = plane_rsrc(E + (int64_t)nxny * (zin ? ez : 0), nxny * (int)sizeof(T));
#pragma unroll
    for (int q = 0; q < NE; ++q)
      pe[q] = (zin && e_idx[q] != kNone) ? plane_load<T>(er, e_idx[q] * (unsigned)sizeof(T)) : T(0);
  };
  auto store_E = [&](const T (&pe)[NE], int buf) {
#pragma unroll
    for (int q = 0; q < NE; ++q)
      if (threadIdx.x + q * NT < EPL) sE[buf][threadIdx.x + q * NT] = pe[q];
  };

  T alo[3], ahi[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) ahi[c] = 0;
  double dsum = 0, dsum2 = 0;

  {
    T pu[NU][3], pe[NE];
    unsigned pf[NU];
    load_u(pu, pf, k0 - 1);
    store_u(pu, pf, 0);
    load_u(pu, pf, k0);
    store_u(pu, pf, 1);
    load_E(pe, k0 - 1);
    store_E(pe, 0);
  }
  __syncthreads();
  // Step st (element layer ez = k0 - 1 + st) reads node planes ez (slot st % 3) and ez + 1
  // (slot (st + 1) % 3) and layer ez (buffer st % 2), and stages plane ez + 2 into slot
  // (st + 2) % 3 and layer ez + 1 into buffer (st + 1) % 2: both were last read in step
  // st - 1, which every wavefront has left at the barrier that ends it.
  int st = 0;
  for (int ez = k0 - 1; ez < k1; ++ez, ++st) {
    const int slo = st % 3, shi = (st + 1) % 3, eb = st & 1;
    const bool more = ez + 1 < k1;
    T pu[NU][3], pe[NE];
    unsigned pf[NU];
    if (more) {
      load_u(pu, pf, ez + 2);
      load_E(pe, ez + 1);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      alo[c] = ahi[c];
      ahi[c] = 0;
    }
    // Lower-node rows are needed once node plane ez is emitted (not in the chunk's first
    // step), upper-node rows only while plane ez + 1 belongs to the chunk (not in its last
    // step); layers outside the mesh contribute nothing.
    const bool layer_in = ez >= 0 && ez < nz;
    const bool want_lo = layer_in && ez >= k0, want_hi = layer_in && ez + 1 < k1;
    // the four elements of the layer around the node column, each gathered from LDS
    // just before use so only one element vector is live
    auto element = [&](auto A_, auto B_) {
      constexpr int A = decltype(A_)::value, B = decltype(B_)::value;
      asm volatile("" ::: "memory");
      T ue[24];
#pragma unroll
      for (int l = 0; l < 8; ++l) {
        const int nl = (ty + B + lny(l)) * UW + tx + A + lnx(l);
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[3 * l + c] = su[lnz(l) ? shi : slo][c][nl];
      }
      element_rows<T, SYM, A, B>(KEbase, kv, ue, sE[eb][(ty + B) * EW + tx + A], alo, ahi, want_lo,
                                 want_hi);
    };
    element(std::integral_constant<int, 0>(), std::integral_constant<int, 0>());
    element(std::integral_constant<int, 1>(), std::integral_constant<int, 0>());
    element(std::integral_constant<int, 0>(), std::integral_constant<int, 1>());
    element(std::integral_constant<int, 1>(), std::integral_constant<int, 1>());
    if (own && ez >= k0) {
      const int64_t n = i + (int64_t)Nx * (j + (int64_t)Ny * ez);
      const uint8_t fm = fix ? fix[n] : 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int64_t dof = 3 * n + c;
        const T uval = u[dof];
        const T Au = ((fm >> c) & 1) ? uval : (op.spk ? alo[c] + op.spk[dof] * uval : alo[c]);
        apply_epilogue<T, EPI>(ep, dof, Au, uval, dsum, dsum2);
      }
    }
    if (more) {
      store_u(pu, pf, (st + 2) % 3);
      store_E(pe, (st + 1) & 1);
      __syncthreads();
    }
  }
  if constexpr (EPI == EPI_APPLY_DOT || EPI == EPI_APPLY_DOT2 || EPI == EPI_CHEB_OUT) {
    __syncthreads();
    const double t = block_sum<NT>(dsum, sred);
    if (threadIdx.x == 0) ep.partial[blockIdx.x] = t;
    if constexpr (EPI == EPI_APPLY_DOT2) {
      __syncthreads();
      const double t2 = block_sum<NT>(dsum2, sred);
      if (threadIdx.x == 0) ep.partial2[blockIdx.x] = t2;
    }
  }
}

// Assembled fine stencil: for node n and neighbour m = n + off(o), block(a,b) =
// sum over shared elements (ascending element order) of E_e * KE, each product
// rounded before it is added exactly as FEAssemble's ke = E*ke[0] then ADD_VALUES
// (FEAnalysis.cpp:448-464); then the MatZeroRowsColumns rows/columns (473) and the
// diag <= 0 -> 1 rule (476-484).  Used to form the level-1 Galerkin operator and the
// fine Jacobi diagonal.
__global__ __launch_bounds__(256) void k_fine_stencil(FineOp<double> op, double *S, int64_t nn)
{
#pragma clang fp contract(off)
  __shared__ double sKE[576];
  for (int v = threadIdx.x; v < 576; v += 256) sKE[v] = op.KE[v];
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= 27 * nn) return;
  int o;
  int64_t n;
  udivmod(t, nn, o, n);
  const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
  const int Nx = op.nx + 1, Ny = op.ny + 1, Nz = op.nz + 1;
  int i, j, k;
  node_ijk(n, Nx, Ny, i, j, k);
  const int mi = i + dx, mj = j + dy, mk = k + dz;
  double blk[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) blk[q] = 0.0;
  const bool inside = mi >= 0 && mi < Nx && mj >= 0 && mj < Ny && mk >= 0 && mk < Nz;
  if (inside) {
    for (int ez = max(k, mk) - 1; ez <= min(k, mk); ++ez) {
      if (ez < 0 || ez >= op.nz) continue;
      for (int ey = max(j, mj) - 1; ey <= min(j, mj); ++ey) {
        if (ey < 0 || ey >= op.ny) continue;
        for (int ex = max(i, mi) - 1; ex <= min(i, mi); ++ex) {
          if (ex < 0 || ex >= op.nx) continue;
          const double Ee = op.E[ex + (int64_t)op.nx * (ey + (int64_t)op.ny * ez)];
          const int ln = lidx(i - ex, j - ey, k - ez), lm = lidx(mi - ex, mj - ey, mk - ez);
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
              blk[a * 3 + b] = blk[a * 3 + b] + Ee * sKE[(3 * ln + a) * 24 + 3 * lm + b];
        }
      }
    }
    if (op.fix) {
      const int64_t m = mi + (int64_t)Nx * (mj + (int64_t)Ny * mk);
      const uint8_t fn = op.fix[n], fmm = op.fix[m];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
          if (((fn >> a) & 1) || ((fmm >> b) & 1)) blk[a * 3 + b] = (o == 13 && a == b) ? 1.0 : 0.0;
    }
    if (o == 13) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (!(blk[a * 4] > 0)) blk[a * 4] = 1.0;
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) S[(int64_t)(o * 9 + q) * nn + n] = blk[q];
}

}  // namespace

template <typename T, int EPI, int TX, int TY>
void launch_tile(hipStream_t s, const FineOp<T> &op, const T *u, const EpiArgs<T> &ep, const FineGrid &g)
{
  if (op.sym)
    k_fine_op<T, EPI, TX, TY, true><<<dim3(g.nblocks, ep.nvec), dim3(NT), 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y,
                                                                         g.nblocks, op.zlen);
  else
    k_fine_op<T, EPI, TX, TY, false><<<dim3(g.nblocks, ep.nvec), dim3(NT), 0, s>>>(op, u, ep, g.tiles_x, g.tiles_y,
                                                                          g.nblocks, op.zlen);
}

template <typename T, int EPI>
void launch_fine(hipStream_t s, const FineOp<T> &op, const T *u, const EpiArgs<T> &ep, const FineGrid &g)
{
  switch (op.tile) {
    case 0: launch_tile<T, EPI, 32, 8>(s, op, u, ep, g); break;
    case 1: launch_tile<T, EPI, 28, 9>(s, op, u, ep, g); break;
    case 2: launch_tile<T, EPI, 25, 10>(s, op, u, ep, g); break;
    case 3: launch_tile<T, EPI, 22, 11>(s, op, u, ep, g); break;
  }
}

FineTile fine_tile_choose(int nx, int ny, int nz, int ncu, int hide)
{
  // Every block runs (zlen + 1) z-steps, the first and last with half the rows; blocks are
  // spread evenly over the CUs and run together, so a launch lasts about ceil(blocks / CUs)
  // block-times, and a CU needs about `hide` resident blocks (wavefronts per SIMD) before
  // its latency is hidden.  Among equal costs the larger z-chunk (less restaging) wins.
  FineTile best{0, 8};
  double best_cost = 1e300;
  ncu = std::max(ncu, 1);
  for (int t = 0; t < kFineTiles; ++t)
    for (int zl = 4; zl <= 16; ++zl) {
      const FineGrid g = fine_grid_of(nx, ny, nz, t, zl);
      const int per_cu = (g.nblocks + ncu - 1) / ncu;
      const double cost = (zl + 0.5) * std::max(per_cu, hide);
      if (cost < best_cost - 1e-9 || (cost < best_cost + 1e-9 && zl > best.zlen)) {
        best_cost = cost;
        best = FineTile{t, zl};
      }
    }
  return best;
}

template <typename T>
int fine_op_blocks(const FineOp<T> &op)
{
  return fine_grid(op).nblocks;
}

template <typename T>
void fine_op(hipStream_t s, const FineOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep)
{
  const FineGrid g = fine_grid(op);
  switch (epi) {
    case EPI_APPLY: launch_fine<T, EPI_APPLY>(s, op, u, ep, g); break;
    case EPI_APPLY_DOT: launch_fine<T, EPI_APPLY_DOT>(s, op, u, ep, g); break;
    case EPI_APPLY_DOT2: launch_fine<T, EPI_APPLY_DOT2>(s, op, u, ep, g); break;
    case EPI_RESID: launch_fine<T, EPI_RESID>(s, op, u, ep, g); break;
    case EPI_CHEB: launch_fine<T, EPI_CHEB>(s, op, u, ep, g); break;
    case EPI_CHEB_OUT: launch_fine<T, EPI_CHEB_OUT>(s, op, u, ep, g); break;
  }
  TOFE_LAUNCH_CHECK();
}

void fine_stencil_build(hipStream_t s, const FineOp<double> &op, double *S)
{
  const int64_t nn = (int64_t)(op.nx + 1) * (op.ny + 1) * (op.nz + 1);
  const int64_t nthreads = 27 * nn;
  if (nthreads >= (int64_t(1) << 32))
    throw Error(TOFE_ERR_ARG, "single-level fine stencil too large");
  dim3 grid((unsigned)((nthreads + 255) / 256)), block(256);
  k_fine_stencil<<<grid, block, 0, s>>>(op, S, nn);
  TOFE_LAUNCH_CHECK();
}

template int fine_op_blocks<double>(const FineOp<double> &);
template int fine_op_blocks<float>(const FineOp<float> &);
template void fine_op<double>(hipStream_t, const FineOp<double> &, const double *, Epi,
                              const EpiArgs<double> &);
template void fine_op<float>(hipStream_t, const FineOp<float> &, const float *, Epi,
                             const EpiArgs<float> &);

}  // namespace tofe
