// Internal declarations shared by the HIP translation units of libtofe.
//
// Data layout in HBM (one context, one GPU):
//   vectors      3*nnode values, DOF = 3*node + component, nodes x-fastest;
//   element data nelem values, elements x-fastest;
//   fixed DOFs   one byte per node, bit c set when component c is a Dirichlet DOF;
//   stencils     level >= 1 operators as 27-point 3x3-block stencils, structure of
//                arrays: S[(o*9 + a*3 + b) * nnode + n] couples DOF (n,a) with DOF
//                (n + off(o), b), o = (dx+1) + 3(dy+1) + 9(dz+1).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tofe.h"

namespace tofe {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define TOFE_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::tofe::Error(TOFE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define TOFE_LAUNCH_CHECK() TOFE_HIP(hipGetLastError())

// Operator epilogues: what a K-apply kernel does with (Au)_i for DOF i.
enum Epi : int {
  EPI_APPLY = 0,      // y = Au
  EPI_APPLY_DOT = 1,  // y = Au, partial += u.Au
  EPI_RESID = 2,      // y = b - Au
  EPI_CHEB = 3,       // r = dinv (b - Au); d = c1 d + c2 r; xout = u + d
  EPI_APPLY_DOT2 = 4, // y = Au, partial += u.Au, partial2 += u.b
  EPI_CHEB_OUT = 5    // last Chebyshev step: x = u + (c1 d + c2 dinv (b - Au)) written in
                      // FP64 to xout64 (or, with xout64 null, to xout), partial += x.q64 (d is
                      // not updated)
};

template <typename T>
struct EpiArgs {
  T *y = nullptr;
  const T *b = nullptr;
  const T *dinv = nullptr;
  T *d = nullptr;
  T *dout = nullptr;  // Chebyshev steps: where the new d goes (null: over d, in place)
  T *xout = nullptr;
  double *partial = nullptr;
  double *partial2 = nullptr;
  double *xout64 = nullptr;
  const double *q64 = nullptr;
  T c1 = 0, c2 = 0;
  // Vectors per launch: vector v (blockIdx.y) of u, y, b, d, xout, xout64 and q64 starts v
  // level-DOF counts after the first and its partial sums v gridDim.x entries after the
  // first; dinv and the operator are shared by all vectors.
  int nvec = 1;
};

// Matrix-free fine operator y = K(E) u with MatZeroRowsColumns Dirichlet rows/cols.
template <typename T>
struct FineOp {
  const T *E = nullptr;        // nelem element stiffness scales
  const uint8_t *fix = nullptr;// nnode fixed-DOF bitmasks (may be null)
  const T *KE = nullptr;       // 576, row-major
  int nx = 0, ny = 0, nz = 0;  // elements
  int tile = 0, zlen = 8;      // column-tile shape index and z-chunk length of k_fine_op
  bool sym = false;            // cubic element: KE[576 + k] holds the cube-orbit values of KE
  bool wht = false;            // KE[592 + k] holds its Walsh-Hadamard blocks: use k_fine_wht
  bool wht_cube = false;       // cubic element: KE[625 + k] holds the 9 cube block values
  const T *spk = nullptr;      // spring stiffness per DOF added to free diagonals (may be null)
};

// Launch shape of the fine operator: one of kFineTiles compiled column tiles and the number of
// node planes per block, chosen from the mesh size, the number of CUs and the number of
// resident blocks per CU that hide the kernel's latency (measured on MI355X at 5.4M DOF:
// 3 for FP64, 6 for FP32).
constexpr int kFineTiles = 4;
struct FineTile {
  int tile, zlen;
};
FineTile fine_tile_choose(int nx, int ny, int nz, int ncu, int hide);
// Element-centric fine operator in the Walsh-Hadamard basis of KE (k_fine_wht).
int fine_wht_zlen(int nx, int ny, int nz, int ncu, int hide);
template <typename T>
int fine_wht_blocks(const FineOp<T> &op);
template <typename T>
void fine_wht(hipStream_t s, const FineOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep);

// Stored 27-point block stencil operator of a coarse level.
// Stencil grids with fewer nodes than this run one thread per DOF instead of one per node
// (level 2 of the 192x96x96 cantilever, 30k nodes, is faster per node; level 1, 233k, per DOF
// 44 us against 31 us).  TOFE_STENCIL_SPLIT_NODES overrides it per context.
constexpr int64_t kStencilSplitNodes = 16384;

template <typename T>
struct StencilOp {
  const T *S = nullptr;        // 243 * nnode
  int Nx = 0, Ny = 0, Nz = 0;  // nodes
  int64_t split_nodes = kStencilSplitNodes;  // grids below this run one thread per DOF
};

constexpr int kRedBlocks = 1024;   // grid of the grid-stride reduction kernels


// ---- launches over several instances (batches) ----------------------------
// Parameters of one instance's part of a batched launch; a table of them (device memory) and
// the first block of each entry (n + 1 offsets) drive one grid over every instance's blocks.
template <typename T>
struct FineLaunch {
  FineOp<T> op;
  const T *u = nullptr;
  EpiArgs<T> ep;
  int tiles_x = 0, tiles_y = 0, nblocks = 0, zlen = 0;
};
template <typename T>
struct StencilLaunch {
  StencilOp<T> op;
  const T *u = nullptr;
  EpiArgs<T> ep;
  int64_t nn = 0;
  int nblocks = 0;
};
template <typename T>
struct TrigemvLaunch {
  const T *A = nullptr;
  int n = 0, ld = 0;
  const T *x = nullptr;
  T *y = nullptr;
  int nblocks = 0;
  const int *perm = nullptr;  // triangles' order -> natural order of x and y (may be null)
};
// Batched launches (one grid, nblocks_total blocks, nvec rows of vectors): table `tab` and
// block offsets `b0` (n + 1 entries) in device memory.
template <typename T>
void fine_wht_batch(hipStream_t s, Epi epi, bool cube, const FineLaunch<T> *tab, const int *b0,
                    int n, int nblocks_total, int nvec);
template <typename T>
int fine_wht_grid(const FineOp<T> &op, int *tiles_x, int *tiles_y);
template <typename T>
void stencil_sym_batch(hipStream_t s, Epi epi, const StencilLaunch<T> *tab, const int *b0, int n,
                       int nblocks_total, int nvec);
// coarse solve halves (lower: tmp = X b, upper: x = X^T tmp); the LDS holds ld_max values
template <typename T>
void trigemv_batch(hipStream_t s, bool lower, const TrigemvLaunch<T> *tab, const int *b0, int n,
                   int nblocks_total, int nvec, int ld_max);

// ---- fine operator (k_fine.hip) -------------------------------------------
template <typename T>
int fine_op_blocks(const FineOp<T> &op);
template <typename T>
void fine_op(hipStream_t s, const FineOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep);
void fine_stencil_build(hipStream_t s, const FineOp<double> &op, double *S);
// ---- fine operator on the matrix cores (k_fine_mfma.hip) -------------------
template <typename T>
int fine_mfma_blocks(const FineOp<T> &op);
template <typename T>
void fine_mfma(hipStream_t s, const FineOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep);

// ---- coarse operators and Galerkin products (k_coarse.hip) ---------------
template <typename T>
int stencil_op_blocks(const StencilOp<T> &op);
template <typename T>
void stencil_op(hipStream_t s, const StencilOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep);
// Coarse stencil of P^T A P, restricted axis by axis through the scratch fields T1 and T2
// (sizes from galerkin_scratch: values in T1, and in T2 through n2).
// Stencil operator from the upper half of a symmetric stencil (126 values per node) and
// the FP64 full -> FP32 upper-half conversion.
template <typename T>
void stencil_op_sym(hipStream_t s, const StencilOp<T> &op, const T *u, Epi epi, const EpiArgs<T> &ep);
// FP64 full stencil -> its upper half (offsets 13..26) in k_stencil_op_sym's layout
template <typename To>
void stencil_upper(hipStream_t s, const double *S, int64_t nn, To *Su);
void galerkin(hipStream_t s, const double *Sf, int Nfx, int Nfy, int Nfz, double *Sc, double *T1,
              double *T2);
int64_t galerkin_scratch(int Nfx, int Nfy, int Nfz, int64_t *n2);
void stencil_diag(hipStream_t s, const double *S, int64_t nn, double *diag, double *dinv);
// inv (may be null): the dense matrix's index of each natural DOF
void stencil_to_dense(hipStream_t s, const double *S, int Nx, int Ny, int Nz, double *A,
                      const int *inv = nullptr);
// ---- coarsest operator straight from the fine level (k_direct.hip) ---------
// Valid when levels >= 2 and every element count is a multiple of r = 2^(levels-1) <= 16.
constexpr int kDirectMaxR = 16;
constexpr int kDirectMaskSplit = 8;  // slices of a cell's Dirichlet elements (corr rows per cell)
bool direct_coarse_ok(int nx, int ny, int nz, int levels);
// M (host, r^3 x 576): the element matrix at each cell position seen from the cell's 8 nodes
void direct_coarse_table(const double KE[576], int r, double *M);
struct DirectCoarse {
  const double *E = nullptr;       // FP64 element scales
  const double *M = nullptr;       // device table of M_p
  // Dirichlet elements (per context): nmasked cells, element ranges, element indices, table rows
  int nmasked = 0;
  const int *range = nullptr;
  const int64_t *elem = nullptr;
  const int *tidx = nullptr;
  const double *table = nullptr;
  const int *slot_of = nullptr;    // cell -> its row of corr, -1 without Dirichlet elements
  const double *cconst = nullptr;  // cells x 576: fixed DOFs' unit diagonal and springs
  double *part = nullptr, *cells = nullptr, *corr = nullptr;  // scratch (direct_coarse_chunks)
  int nx = 0, ny = 0, nz = 0, r = 0;
  const int *perm = nullptr;  // dense order -> natural DOF of the coarsest grid (may be null)
};
// Host tables of the Dirichlet part (fixmask: per node, bit a = component a fixed; may be null).
struct DirectStatic {
  std::vector<int> cells, range, tidx, slot_of;
  std::vector<int64_t> elem;
  std::vector<double> table, cconst;
};
void direct_coarse_static(const double KE[576], int nx, int ny, int nz, int r, const uint8_t *fixmask,
                          DirectStatic &st);
// cconst += the springs (per DOF, zero on fixed DOFs) of every cell's owned nodes
void direct_coarse_add_springs(int nx, int ny, int nz, int r, const double *spk, double *cconst);
int direct_coarse_chunks(int ncell, int r);
// Dense coarsest matrix (column-major) into A, as stencil_to_dense of the Galerkin chain's last
// level gives it (to rounding).
void direct_coarse(hipStream_t s, const DirectCoarse &d, double *A);
constexpr int kCholNB = 64;  // tile of the coarse Cholesky factorisation
// A (n x n SPD, column-major, half-bandwidth bw): banded blocked Cholesky A = L L^T into
// Lb and X = L^-1 into X (column-major, lower triangle), so A^-1 = X^T X.  info[0] = first
// non-positive pivot index + 1 (0 if none).  Lb and X are n x n scratch.
void coarse_inverse_factor(hipStream_t s, double *A, double *Lb, double *X, int n, int bw, int *info);
// Row-major copies of X and X^T in T, rows padded to ld (a multiple of 16 bytes / sizeof(T)),
// zero outside the triangle.
template <typename T>
void tri_pack(hipStream_t s, const double *X, int n, int ld, T *Xr, T *XTr);
// x = A^-1 b = X^T (X b) with the packed triangles for nvec right-hand sides; b, tmp and x
// hold nvec vectors of n values one after another.
template <typename T>
void coarse_solve_tri(hipStream_t s, const T *Xr, const T *XTr, int n, int ld, const T *b, T *tmp, T *x,
                      int nvec = 1, const int *perm = nullptr);

// ---- level-1 Galerkin operator from E and KE (k_galerkin.hip) -------------
struct MaskedElems {
  const int64_t *elem = nullptr;   // element index
  const int *form = nullptr;       // interpolation form fx + 3 fy + 9 fz
  const uint32_t *mask = nullptr;  // bit r set: local DOF r is a Dirichlet DOF
  int count = 0;
};
void galerkin_fine(hipStream_t s, const double *E, const int *emap, const double *KEC,
                   const MaskedElems &me, const double *KE, const double *Qforms, double *KCm,
                   const uint8_t *fix, int nx, int ny, int nz, double *Sc);
void fine_diag(hipStream_t s, const double *E, const double *KE, const uint8_t *fix, int nx, int ny,
               int nz, double *diag, double *dinv, const double *spk = nullptr);
// Spring supports in the Galerkin hierarchy: Sc += P^T diag(spk) P on the coarse grid of a fine
// grid of Nfx x Nfy x Nfz nodes (spk per fine DOF, zero at Dirichlet DOFs), or, with identity
// transfer (coarse == fine), the diagonal itself.
void galerkin_springs(hipStream_t s, const double *spk, int Nfx, int Nfy, int Nfz, bool identity,
                      double *Sc);

// ---- vector kernels (k_vec.hip) -------------------------------------------
// nvec > 1: nvec vectors of n values stored one after another (vector v at v * n), each
// with its own per-block partial sums (v * grid blocks on) and scalars (v * SL_COUNT on);
// dinv, the operator and the transfer weights are shared.
template <typename T>
void dot_partial(hipStream_t s, const T *a, const T *b, int64_t n, double *partial, int nvec = 1);
// mirror (may be null): device address of mapped host memory that also receives each vector's
// SL_COUNT scalars after the reduction
void reduce_final(hipStream_t s, const double *partial, int nparts, double *out, int post,
                  double *scal, int nvec = 1, double *mirror = nullptr);
// Right-hand sides of a block solve that have converged (bit v of mask): their alpha and
// beta are forced to zero from the next reduction on.
void set_frozen(hipStream_t s, double *scal, int nvec, unsigned long long mask);
template <typename T>
void cheb_zero(hipStream_t s, const T *b, const T *dinv, T c2, T *d, T *x, int64_t n, int nvec = 1);
// first Chebyshev step from an FP64 right-hand side: b32 = (float) b64, d = x = c2 dinv b32
void cheb_zero_from64(hipStream_t s, const double *b64, float *b32, const float *dinv, float c2,
                      float *d, float *x, int64_t n, int nvec = 1);
// Optional output of restrict_op: the coarse level's first Chebyshev step from x = 0,
// d = x = c2 dinv b (k_cheb_zero), when d is set.
template <typename T>
struct FirstStep {
  const T *dinv = nullptr;
  T c2 = 0;
  T *d = nullptr, *x = nullptr;
};
template <typename T>
void restrict_op(hipStream_t s, const T *rf, int Nfx, int Nfy, int Nfz, T *bc, int Ncx, int Ncy,
                 int Ncz, const FirstStep<T> &fs = FirstStep<T>(), int nvec = 1);
template <typename T>
void prolong_add(hipStream_t s, const T *xc, int Ncx, int Ncy, int Ncz, T *xf, int Nfx, int Nfy,
                 int Nfz, int nvec = 1);
template <typename Ta, typename Tb>
void convert(hipStream_t s, const Ta *a, Tb *b, int64_t n);
void fill_d(hipStream_t s, double *a, double v, int64_t n);
// a fill kernel, not hipMemsetAsync: it may be captured into a graph (no memset nodes there)
template <typename T>
void fill(hipStream_t s, T *a, T v, int64_t n);
struct FirstCheb64;
void cg_update_r(hipStream_t s, double *r, const double *q, int64_t n, const double *scal,
                 int alpha_slot, double *partial, int nvec, const FirstCheb64 &fc);
// x += scal[alpha_slot] p (the iterate's update, deferred to where p is read), p = z + beta p
void cg_update_xp(hipStream_t s, double *p, const double *z, double *x, int64_t n, const double *scal,
                  int beta_slot, int alpha_slot, int nvec = 1);
template <typename T>
void power_start(hipStream_t s, T *v, int64_t n);
template <typename T>
void jacobi_dot(hipStream_t s, const T *r, const T *dinv, T *z, int64_t n, double *partial);
void record_scalar(hipStream_t s, const double *scal, int slot, double *dst);
// Flexible-CG updates that sum the producers' partials themselves: fcg_update_r forms
// alpha = sum(pr_part) / sum(pq_part) (stored with both sums in scal), r -= alpha q, partials
// of r.r; fcg_update_xp forms beta = -sum(zq_part) / scal[PQ], applies the iterate's update
// x += scal[ALPHA] p of the same step (deferred to where p is read anyway) and p = z + beta p;
// fcg_update_x applies x += scal[ALPHA] p alone (a step that ends the solve).
// Optional output of fcg_update_r: the mixed V-cycle's first Chebyshev step on the new
// residual (b32 = (float) r, d = x = c2 dinv b32), when b32 is set.
struct FirstCheb {
  float *b32 = nullptr, *d = nullptr;
  float *x = nullptr;  // null: the iterate is not written (it equals d after the first step)
  const float *dinv = nullptr;
  float c2 = 0;
};
// The FP64 V-cycle's first fine-level Chebyshev step formed by the CG residual update:
// d = c2 D^-1 r (d == nullptr: none).  The iterate is not written (it equals d).
struct FirstCheb64 {
  double *d = nullptr;
  const double *dinv = nullptr;
  double c2 = 0;
};
void fcg_update_r(hipStream_t s, double *r, const double *q, int64_t n, const double *pq_part,
                  const double *pr_part, int nparts, double *scal, double *partial,
                  const FirstCheb &fc = FirstCheb(), int nvec = 1);
void fcg_update_xp(hipStream_t s, double *p, const double *z, const float *z32, double *x,
                   int64_t n, const double *zq_part, int nparts, double *scal, int nvec = 1);
void fcg_update_x(hipStream_t s, double *x, const double *p, int64_t n, const double *scal,
                  int nvec = 1);
// Batched forms of the V-cycle's transfers and of the flexible-CG updates: one entry per
// instance, with the blocks per vector of the single launch (restriction: coarse nodes / 256,
// prolongation: fine nodes / 256, fcg_update_r_blocks, fcg_update_xp_blocks, reduction: 1).
template <typename T>
struct RestrictLaunch {
  const T *rf = nullptr;
  int Nfx = 0, Nfy = 0, Nfz = 0;
  T *bc = nullptr;
  int Ncx = 0, Ncy = 0, Ncz = 0;
  FirstStep<T> fs;
  int nblocks = 0;
};
template <typename T>
struct ProlongLaunch {
  const T *xc = nullptr;
  int Ncx = 0, Ncy = 0, Ncz = 0;
  T *xf = nullptr;
  int Nfx = 0, Nfy = 0, Nfz = 0;
  int nblocks = 0;
};
// pre != 0: alpha (FcgXr) or beta (FcgP) already stands in the scalars, formed by reduction
// launches (POST_ALPHA_FCG, POST_BETA_FCG) with the sum every block would otherwise form
struct FcgXrLaunch {
  double *r = nullptr;
  const double *q = nullptr;
  int64_t n = 0;
  const double *pq_part = nullptr, *pr_part = nullptr;
  int nparts = 0;
  double *scal = nullptr, *partial = nullptr;
  FirstCheb fc;
  int nblocks = 0;
  int pre = 0;
};
struct FcgPLaunch {
  double *p = nullptr;
  const double *z = nullptr;
  const float *z32 = nullptr;  // z held in FP32 (the V-cycle's last step), or null: z
  double *x = nullptr;
  int64_t n = 0;
  const double *zq_part = nullptr;
  int nparts = 0;
  double *scal = nullptr;
  int nblocks = 0;
  int pre = 0;
};
struct ReduceLaunch {
  const double *partial = nullptr;
  int nparts = 0;
  double *out = nullptr;
  int post = 0;
  double *scal = nullptr;
  int nblocks = 1;
};
int fcg_update_r_blocks();
int fcg_update_xp_blocks(int64_t n);
void reduce_final_batch(hipStream_t s, const ReduceLaunch *tab, const int *b0, int n, int nvec);
void fcg_update_r_batch(hipStream_t s, const FcgXrLaunch *tab, const int *b0, int n,
                        int nblocks_total, int nvec);
void fcg_update_xp_batch(hipStream_t s, const FcgPLaunch *tab, const int *b0, int n,
                         int nblocks_total, int nvec);
template <typename T>
void restrict_batch(hipStream_t s, const RestrictLaunch<T> *tab, const int *b0, int n,
                    int nblocks_total, int nvec);
template <typename T>
void prolong_batch(hipStream_t s, const ProlongLaunch<T> *tab, const int *b0, int n,
                   int nblocks_total, int nvec);
// Fused Jacobi-CG steps of the eigenvalue estimate (no separate reduction launches):
// lanczos_r: alpha = scal[rz_slot] / sum(pq_part[0..npq)), recorded; r -= alpha q; partials
// of r.(D^-1 r).  lanczos_p: beta = sum(rz_part) / scal[rz_old], recorded; scal[rz_new] =
// the sum; p = D^-1 r + beta p.
template <typename T>
void lanczos_r(hipStream_t s, T *r, const T *q, const T *dinv, int64_t n, const double *pq_part,
               int npq, const double *scal, int rz_slot, double *rec_alpha, double *partial);
template <typename T>
void lanczos_p(hipStream_t s, T *p, const T *r, const T *dinv, int64_t n, const double *rz_part,
               double *scal, int rz_old, int rz_new, double *rec_beta);

// device scalar slots (ctx->scal)
enum Slot : int {
  SL_RZ = 0, SL_PQ = 1, SL_RR = 2, SL_ALPHA = 3, SL_BETA = 4, SL_BB = 5, SL_PR = 6, SL_ZQ = 7,
  SL_RZA = 8, SL_RZB = 9, SL_CONVCODE = 10, SL_TMP = 11, SL_FROZEN = 12, SL_TTOL = 13,
  SL_ITS = 14, SL_REASON = 15, SL_COUNT = 16
};
// reduce_final post-processing
enum Post : int {
  POST_NONE = 0,
  POST_ALPHA_CG = 1,   // scal[ALPHA] = scal[RZ] / v
  POST_BETA_CG = 2,    // scal[BETA] = v / scal[RZ]; scal[RZ] = v
  POST_ALPHA_FCG = 3,  // scal[ALPHA] = scal[PR] / v   (v = p.Ap)
  POST_BETA_FCG = 4,   // scal[BETA] = -v / scal[PQ]  (v = z.Ap)
  // v = r.r after a step of a solve whose test runs on the device: unless frozen, count the
  // step in scal[ITS] and freeze with scal[REASON] on convergence (sqrt(v) <= scal[TTOL]:
  // reason scal[CONVCODE]), on p.Ap <= 0 or NaN, on a non-finite residual, or (CG only) on
  // r.z < 0
  POST_CONV_FCG = 5,
  POST_CONV_CG = 6
};
// Per-column state of a solve tested on the device: tolerances, convergence reason codes,
// iteration counts and reasons reset, frozen flags as in the mask.
struct ConvState {
  double ttol[16];
  int conv[16];
};
void set_conv_state(hipStream_t s, double *scal, int nvec, const ConvState &cs,
                    unsigned long long frozen_mask);

// ---- design-side kernels (k_design.hip) -----------------------------------
struct FilterTemplate {
  const int *off = nullptr;    // 3 per entry (di, dj, dk)
  const double *w = nullptr;   // 1 - dist/R per entry
  int ntpl = 0;
  int nx = 0, ny = 0, nz = 0;
};
void filter_rowsum_inv(hipStream_t s, const FilterTemplate &t, double *inv);
void filter_apply(hipStream_t s, const FilterTemplate &t, const double *inv, const double *x,
                  double *rho);
void filter_apply_T(hipStream_t s, const FilterTemplate &t, const double *inv, const double *g,
                    double *gx);
void filter_count(hipStream_t s, const FilterTemplate &t, int64_t *cnt);
void filter_fill(hipStream_t s, const FilterTemplate &t, const double *inv, const int64_t *rowptr,
                 int32_t *col, double *w);
struct SimpParams {
  double penal = 3, eps = 1e-10, vd_penal = 1, vd_min = -1;
  int interp = TOFE_INTERP_SIMP;
  double p0 = 0, p1 = 0;
};
void void_pre(hipStream_t s, double *rho, int64_t ne, double q, double *zt);
void rfilter_madd(hipStream_t s, const FilterTemplate &t, const double *REdge, const double *zt,
                  double inv_vdmin, double *y);
void rfilter_T(hipStream_t s, const FilterTemplate &t, const double *g, double *gx);
void simp(hipStream_t s, const double *rho, const double *y, int64_t ne, const SimpParams &sp,
          double *yout, double *rhoq, double *E, double *dEdz, double *Es, double *dEsdz, float *E32);
void chain_mid(hipStream_t s, const double *dfdE, const double *y, const double *rhoq,
               const double *RtT, double *out, int64_t ne);
void mul(hipStream_t s, const double *a, const double *b, double *c, int64_t n);
// Element geometry for the stress stiffness: D (6x6), dN_n/dx_a at the Gauss points, det J.
struct ShapeData {
  double D[36];
  double dN[8][3][8];
  double detJ[8];
};
void stress_h(hipStream_t s, const double *U, int nx, int ny, int nz, const ShapeData &sd,
              double *H);
// y = Ks x for nvec vectors stored one after another (H read once for up to four of them)
void apply_ks(hipStream_t s, const double *H, const double *Es, const uint8_t *fix, int nx,
              int ny, int nz, const double *x, double *y, int nvec = 1);
void ks_quad(hipStream_t s, const double *H, const double *phi, int nx, int ny, int nz,
             int64_t ndof, int nvec, double *q);
void stab_rhs(hipStream_t s, const double *phi, const double *Es, const uint8_t *fix, int nx,
              int ny, int nz, const ShapeData &sd, double *r);
void elem_bilinear(hipStream_t s, const double *A, const double *B, const double *KE, int nx,
                   int ny, int nz, int64_t ndof, int nvec, double *out);
void mass_diag(hipStream_t s, const double *V, int nx, int ny, int nz, double m0,
               const double *mlump, const uint8_t *fix, double *M);
void diag_mul(hipStream_t s, const double *d, const double *x, double *y, int64_t n, int nrhs);
// kq = phi_e^T KE phi_e; mq = sum of phi_e^2 (lumped mass) or phi_e^T (Q (x) I_3) phi_e
// (consistent mass over m0), per element and vector
void elem_quad(hipStream_t s, const double *phi, const double *KE, int nx, int ny, int nz,
               int64_t ndof, int nvec, double *kq, double *mq, bool consistent = false);
// y = M x for nvec vectors with the consistent mass sum_e V_e m0 Q (x) I_3 + diag(mlump) and
// the Dirichlet rows / columns of MatZeroRowsColumns(M, fixed, 1e-8)
void apply_mass_consistent(hipStream_t s, const double *V, double m0, const double *mlump,
                           const uint8_t *fix, int nx, int ny, int nz, const double *x, double *y,
                           int nvec);
void sensitivity(hipStream_t s, const double *U, const double *dEdz, const double *KE, int nx,
                 int ny, int nz, double *g);

// ---- context-free entry point (k_noalloc.hip) -------------------------------
// message of the calling thread's last failing tofe_compliance_noalloc
const char *noalloc_last_error();

}  // namespace tofe
