// C ABI of libtofe: context life cycle, design update, solver set-up and the
// multigrid-preconditioned CG solve on one GPU.
//
// FP64 path: CG (KSPCG semantics, FEAnalysis.cpp:387-395) preconditioned by a Galerkin
// V-cycle with Chebyshev-Jacobi smoothing and an FP64 dense coarse solve.
// Mixed path: the same V-cycle in FP32 (FP32 operator, stencils, vectors) inside an
// FP64 flexible CG (truncated FCG(1)); the outer residual, the dot products and the
// coarse solve stay FP64.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "host_setup.hpp"
#include "ke_orbits.hpp"
#include "ke_wht.hpp"
#include "tofe_internal.hpp"

using namespace tofe;

namespace {

// roctx range around one C-ABI call (or one phase of it): rocprofv3 --marker-trace shows
// which kernels each entry point launched
struct Range {
  explicit Range(const char *name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range &) = delete;
  Range &operator=(const Range &) = delete;
};

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void alloc(size_t count)
  {
    if (count == n && p) return;
    release();
    if (count) {
      TOFE_HIP(hipMalloc(&p, count * sizeof(T)));
      n = count;
    }
  }
  void zero(hipStream_t s)
  {
    if (p) TOFE_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
  }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct Grid {
  int Nx = 0, Ny = 0, Nz = 0;
  int64_t nn() const { return (int64_t)Nx * Ny * Nz; }
  int64_t ndof() const { return 3 * nn(); }
};

template <typename T>
struct Level {
  Grid g;
  DBuf<T> S;     // 27-point block stencil (levels >= 1)
  DBuf<T> Su;    // FP64 solve: upper half of S in k_stencil_op_sym's layout, valid while su_ok
  bool su_ok = false;
  DBuf<T> dinv;  // inverse Jacobi diagonal
  DBuf<T> b, x, x2, r, d;
};

constexpr int kMaxEig = 64;
// Right-hand sides per block pass: solves, V-cycles and operator applies with more vectors
// run in chunks of this many (each chunk one launch per kernel).
constexpr int kMaxBlockVec = 16;
constexpr int kCoarseCUs = 32;  // compute units left to the coarse factorisation during set-up

double now_s()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The CU-masked streams of one device, shared by every context of the process.  A CU mask is
// a property of a hardware queue, so each CU-masked stream holds a queue of its own; once a
// process holds more queues than the hardware scheduler maps at once, the scheduler
// time-slices all of them, and busy kernels then stall for whole quanta (measured: 21 ms
// stalls every 31.5 ms of busy time with 16 contexts of two masked streams each).  Created on
// first use and kept until the process ends.
struct MaskedStreams {
  hipStream_t a = nullptr, b = nullptr;  // off the first first_cu compute units
  hipStream_t full = nullptr;            // every compute unit, a hardware queue of its own
};
MaskedStreams masked_streams(int device, int ncu, int first_cu)
{
  static std::mutex mu;
  static std::map<int, MaskedStreams> per_device;
  std::lock_guard<std::mutex> lock(mu);
  auto it = per_device.find(device);
  if (it != per_device.end()) return it->second;
  std::vector<uint32_t> mask((ncu + 31) / 32, 0u);
  for (int cu = first_cu; cu < ncu; ++cu) mask[cu / 32] |= 1u << (cu % 32);
  MaskedStreams m;
  TOFE_HIP(hipExtStreamCreateWithCUMask(&m.a, (uint32_t)mask.size(), mask.data()));
  TOFE_HIP(hipExtStreamCreateWithCUMask(&m.b, (uint32_t)mask.size(), mask.data()));
  std::vector<uint32_t> all((ncu + 31) / 32, 0u);
  for (int cu = 0; cu < ncu; ++cu) all[cu / 32] |= 1u << (cu % 32);
  TOFE_HIP(hipExtStreamCreateWithCUMask(&m.full, (uint32_t)all.size(), all.data()));
  per_device[device] = m;
  return m;
}

// Launches of one instance's solve iteration, recorded instead of enqueued, so that a batch
// of instances can run each step as one launch over all of them.  An op of a kind with a
// batched kernel keeps its launch parameters (desc, blocks per vector); any other op is kept
// as a closure that enqueues it on a given stream.
enum BatchKind : int {
  BK_CLOSURE = 0,
  BK_FINE = 1,      // fine operator, element-centric Walsh-Hadamard form
  BK_STENCIL = 2,   // upper-half stencil operator of a coarse level
  BK_RESTRICT = 3,
  BK_PROLONG = 4,
  BK_TRI_LO = 5,    // coarse solve, tmp = X b
  BK_TRI_UP = 6,    // coarse solve, x = X^T tmp
  BK_FCG_XR = 7,
  BK_FCG_P = 8,
  BK_REDUCE = 9,
};
// kind, precision, element form (fine operator) and epilogue in one key: ops with equal keys
// share a launch
constexpr int batch_key(int kind, bool fp32, int epi = 0, bool cube = false)
{
  return (kind << 8) | ((fp32 ? 1 : 0) << 7) | ((cube ? 1 : 0) << 6) | epi;
}
struct RecOp {
  int key = BK_CLOSURE;
  int nblocks = 0;  // blocks per vector
  int ld = 0;       // coarse solve: padded row length (dynamic LDS)
  std::vector<unsigned char> desc;
  std::function<void(hipStream_t)> run;
};
struct Recorder {
  std::vector<RecOp> ops;
};

}  // namespace

struct tofe_ctx {
  tofe_desc desc{};
  int device = 0;
  int L = 1;
  std::vector<Grid> grids;
  int64_t ne = 0, nn = 0, ndof = 0;
  hipStream_t stream = nullptr;
  Recorder *rec = nullptr;  // set while a batch records this context's iteration
  // second stream: the coarsest-level factorisation runs beside the Galerkin chain and the
  // Lanczos estimates
  hipStream_t stream2 = nullptr;
  // third stream: the Lanczos estimates of the coarser smoothed levels run beside the
  // fine level's
  hipStream_t stream3 = nullptr;
  // fourth stream: the fine level's Lanczos estimate.  The third and fourth streams are
  // kept off the first kCoarseCUs compute units, which the latency-bound coarse
  // factorisation on the second stream then has to itself.  The second to fourth streams are
  // CU-masked streams shared by every context of the device (masked_streams), not destroyed
  // with the context: a masked stream owns its hardware queue, where ordinary streams share
  // the process's few queues round the pool, so that two of a context's streams could land on
  // one queue and run in turn (measured: set-up of the 192x96x96 cantilever 2.5 or 3.5 ms
  // depending on the streams other contexts had created before).
  hipStream_t stream4 = nullptr;
  bool shared_helpers = false;
  hipEvent_t ev_dense = nullptr, ev_inv = nullptr, ev_ops = nullptr, ev_lz = nullptr;
  hipEvent_t ev_scal = nullptr, ev_lz0 = nullptr, ev_diag = nullptr;
  double *hscal = nullptr;      // pinned, mapped host copy of the device scalars
  double *hscal_dev = nullptr;  // its device address (the solve's reductions write it directly)
  double KE[576];
  std::vector<int> tpl_off;
  std::vector<double> tpl_w;

  DBuf<double> dKE;
  DBuf<float> dKE32;
  DBuf<uint8_t> dfix;
  bool has_fix = false;
  DBuf<int> dtpl_off;
  DBuf<double> dtpl_w, filt_inv;
  DBuf<double> x, rho, E, dEdz, Es, dEsdz;
  DBuf<float> E32;
  bool design_set = false;
  // maximum length scale: unit-weight filter R, REdge, y, rhoq (Material_Interpolation.cpp:37-60)
  SimpParams sp;
  bool has_vd = false;
  std::vector<int> rtpl_off;
  DBuf<int> drtpl_off;
  DBuf<double> drtpl_w, REdge, y, rhoq, ebuf3;
  // spring stiffness per DOF (spKVec, FEAnalysis.cpp:360-368), zero at Dirichlet DOFs, FP64 and
  // FP32 (the mixed V-cycle's fine operator)
  DBuf<double> spk;
  DBuf<float> spk32;
  bool has_springs = false;
  // mass matrix: the lumped diagonal (Frequency::DiagMassFnc), or the consistent element mass
  // V_e m0 Q (x) I_3 applied matrix-free from a snapshot of V and the point masses
  DBuf<double> Mdiag, mV, mlump;
  int mass_form = TOFE_MASS_LUMPED;
  double mass_m0 = 0;
  bool has_mass = false, has_mlump = false;
  // stress state of the buckling analysis: H_e per element (Stability::StressFnc)
  ShapeData shape{};
  DBuf<double> Hs;
  bool has_stress = false;

  std::vector<std::unique_ptr<Level<double>>> lv64;
  std::vector<std::unique_ptr<Level<float>>> lv32;
  DBuf<double> fineS, diag0;
  // level-1 Galerkin from E: interpolation forms, masked (Dirichlet) elements
  DBuf<double> dQforms, dKEC, dKCm;
  DBuf<int> demap, dme_form;
  DBuf<int64_t> dme_elem;
  DBuf<uint32_t> dme_mask;
  int n_masked = 0;
  DBuf<double> gal_t1, gal_t2;  // axis-by-axis Galerkin scratch
  DBuf<double> Ainv;
  // Numbering of the coarsest grid's DOFs in its dense matrix: natural, or (coarse_perm_on)
  // with the longest axis slowest, which narrows the factor's band (coarse_order).
  std::vector<int> hperm, hinv;
  DBuf<int> cperm, cinv;       // dense order -> natural DOF, natural DOF -> dense order
  bool cperm_ready = false, cperm_on = false, coarse_perm_on = false;
  int cbw_nat = 0, cbw_perm = 0, coarse_bw = 0;
  // coarsest operator formed from E directly (k_direct.hip) when the mesh allows it
  bool direct = false;
  int direct_nmasked = 0;
  DBuf<double> direct_M, direct_part, direct_cells, direct_corr, direct_table, direct_cconst;
  DBuf<int> direct_range, direct_tidx, direct_slot;
  DBuf<int64_t> direct_elem;
  std::vector<double> direct_cfix;  // host: the cells' fixed-DOF part of cconst
  // X = L^-1 of the coarse operator as padded row-major triangles X and X^T, in the V-cycle's
  // precision, and an n-vector for the solve's intermediate X b
  DBuf<double> X64r, XT64r, ctmp64;
  DBuf<float> X32r, XT32r, ctmp32;
  int ldc = 0;  // padded row length of the packed coarse triangles
  DBuf<int> dinfo;
  DBuf<double> chol_l, chol_x;
  int64_t nc = 0;
  // vectors per block pass the work buffers hold (level vectors, CG vectors, partial sums,
  // scalars); grows on demand up to kMaxBlockVec
  int nvcap = 1;
  bool frozen_dirty = false;  // some SL_FROZEN flag of scal may be set
  DBuf<double> partial, partial2, partial3, scal, eig_ab;
  DBuf<double> partial_b, partial2_b, scal_b;  // reductions of the third stream
  DBuf<double> partial_c, partial2_c, scal_c;  // set-up reductions of the main stream
  int npart = 0;
  std::vector<double> lam;
  std::vector<double> c1, c2;  // Chebyshev coefficients, per level * kmax

  DBuf<double> vr, vz, vp, vq, vF, vU;
  DBuf<float> vr32, vz32;
  DBuf<double> ebuf, ebuf2;

  bool assembled = false;
  // The coarsest level's factorisation (second stream) after tofe_assemble: coarse_unjoined
  // while the main stream has not waited for it, coarse_unchecked while its pivot report has
  // not been read (coarse_join, coarse_settle).
  bool coarse_unjoined = false, coarse_unchecked = false;
  bool fine_mfma = false;  // TOFE_FINE_MFMA=1: element-centric MFMA fine operator
  int64_t split_nodes = kStencilSplitNodes;  // TOFE_STENCIL_SPLIT_NODES: one thread per DOF below
  FineTile fine_tile{0, 8}, fine_tile32{0, 8};  // k_fine_op launch shapes, FP64 and FP32
                                                // (TOFE_FINE_TILE="<shape>:<zlen>" sets both)
  bool ke_sym = false;  // KE matches the cube-orbit table: fine operator keeps KE in registers
  bool ke_wht = false;  // KE is block diagonal in the Walsh-Hadamard basis: k_fine_wht
  // tofe_time_kernels: an event pair around each operator launch on the main stream, with
  // its kind (level, precision, epilogue)
  bool kt_on = false;
  size_t kt_n = 0;
  std::vector<hipEvent_t> kt_ev;
  std::vector<int> kt_kind;
  bool ke_wht_cube = false;  // ... with the 9-value block structure of a cubic element
  int wht_zlen = 8, wht_zlen32 = 8;  // its z-chunk lengths, FP64 and FP32
  int precision = TOFE_PREC_FP64;
  int built_precision = -1;
  tofe_stats stats{};
  std::string err;

  int kmax() const { return std::max(std::max(desc.nu_pre, desc.nu_post), 1); }

  tofe_ctx() = default;
  tofe_ctx(const tofe_ctx &) = delete;
  tofe_ctx &operator=(const tofe_ctx &) = delete;
  ~tofe_ctx()
  {
    if (stream) (void)hipStreamSynchronize(stream);
    if (stream2) (void)hipStreamSynchronize(stream2);
    if (stream3) (void)hipStreamSynchronize(stream3);
    if (stream4) (void)hipStreamSynchronize(stream4);
    for (hipEvent_t e : {ev_dense, ev_inv, ev_ops, ev_lz, ev_scal, ev_lz0, ev_diag})
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : kt_ev)
      if (e) (void)hipEventDestroy(e);
    if (hscal) (void)hipHostFree(hscal);
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2 && !shared_helpers) (void)hipStreamDestroy(stream2);
    if (stream3 && !shared_helpers) (void)hipStreamDestroy(stream3);
    if (stream4 && !shared_helpers) (void)hipStreamDestroy(stream4);
  }
};

namespace {

FilterTemplate rfilter_tpl(const tofe_ctx *c)
{
  FilterTemplate t;
  t.off = c->drtpl_off.p;
  t.w = c->drtpl_w.p;
  t.ntpl = (int)(c->rtpl_off.size() / 3);
  t.nx = c->desc.nx;
  t.ny = c->desc.ny;
  t.nz = c->desc.nz;
  return t;
}

FilterTemplate filter_tpl(const tofe_ctx *c)
{
  FilterTemplate t;
  t.off = c->dtpl_off.p;
  t.w = c->dtpl_w.p;
  t.ntpl = (int)c->tpl_w.size();
  t.nx = c->desc.nx;
  t.ny = c->desc.ny;
  t.nz = c->desc.nz;
  return t;
}

template <typename T>
std::vector<std::unique_ptr<Level<T>>> &levels(tofe_ctx *c);
template <>
std::vector<std::unique_ptr<Level<double>>> &levels<double>(tofe_ctx *c) { return c->lv64; }
template <>
std::vector<std::unique_ptr<Level<float>>> &levels<float>(tofe_ctx *c) { return c->lv32; }

template <typename T>
FineOp<T> fine_of(tofe_ctx *c);
template <>
FineOp<double> fine_of<double>(tofe_ctx *c)
{
  FineOp<double> f;
  f.E = c->E.p;
  f.fix = c->has_fix ? c->dfix.p : nullptr;
  f.KE = c->dKE.p;
  f.nx = c->desc.nx; f.ny = c->desc.ny; f.nz = c->desc.nz;
  f.tile = c->fine_tile.tile; f.zlen = c->fine_tile.zlen;
  f.sym = c->ke_sym;
  f.wht = c->ke_wht;
  f.wht_cube = c->ke_wht_cube;
  if (f.wht) f.zlen = c->wht_zlen;
  f.spk = c->has_springs ? c->spk.p : nullptr;
  return f;
}
template <>
FineOp<float> fine_of<float>(tofe_ctx *c)
{
  FineOp<float> f;
  f.E = c->E32.p;
  f.fix = c->has_fix ? c->dfix.p : nullptr;
  f.KE = c->dKE32.p;
  f.nx = c->desc.nx; f.ny = c->desc.ny; f.nz = c->desc.nz;
  f.tile = c->fine_tile32.tile; f.zlen = c->fine_tile32.zlen;
  f.sym = c->ke_sym;
  f.wht = c->ke_wht;
  f.wht_cube = c->ke_wht_cube;
  if (f.wht) f.zlen = c->wht_zlen32;
  f.spk = c->has_springs ? c->spk32.p : nullptr;
  return f;
}

// tofe_time_kernels: record the start event of an operator launch on the main stream
bool kernel_timer_begin(tofe_ctx *c, hipStream_t s, int level, bool fp32, Epi epi)
{
  if (!c->kt_on || s != c->stream || 2 * c->kt_n + 1 >= c->kt_ev.size()) return false;
  TOFE_HIP(hipEventRecord(c->kt_ev[2 * c->kt_n], s));
  c->kt_kind[c->kt_n] = (level * 2 + (fp32 ? 1 : 0)) * 8 + (int)epi;
  return true;
}

void kernel_timer_end(tofe_ctx *c, hipStream_t s)
{
  TOFE_HIP(hipEventRecord(c->kt_ev[2 * c->kt_n + 1], s));
  ++c->kt_n;
}

template <typename D>
void rec_push(tofe_ctx *c, int key, const D &d, int nblocks, int ld = 0)
{
  static_assert(std::is_trivially_copyable<D>::value, "launch parameters are copied as bytes");
  RecOp op;
  op.key = key;
  op.nblocks = nblocks;
  op.ld = ld;
  op.desc.resize(sizeof(D));
  std::memcpy(op.desc.data(), &d, sizeof(D));
  c->rec->ops.push_back(std::move(op));
}

// Enqueue f(stream) on the context's stream, or record it while a batch records the context.
void issue(tofe_ctx *c, std::function<void(hipStream_t)> f)
{
  if (c->rec) {
    RecOp op;
    op.run = std::move(f);
    c->rec->ops.push_back(std::move(op));
  } else {
    f(c->stream);
  }
}

template <typename T>
void level_apply_now(tofe_ctx *c, int l, const T *u, Epi epi, const EpiArgs<T> &ep, hipStream_t s);

// Records a level's operator launch (recording mode of level_apply).
template <typename T>
void level_record(tofe_ctx *c, int l, const T *u, Epi epi, const EpiArgs<T> &ep)
{
  constexpr bool f32 = sizeof(T) == 4;
  if (l == 0) {
    const FineOp<T> f = fine_of<T>(c);
    if (f.wht && !c->fine_mfma) {
      FineLaunch<T> d;
      d.op = f;
      d.u = u;
      d.ep = ep;
      d.nblocks = fine_wht_grid<T>(f, &d.tiles_x, &d.tiles_y);
      d.zlen = f.zlen;
      rec_push(c, batch_key(BK_FINE, f32, epi, f.wht_cube), d, d.nblocks);
      return;
    }
  } else {
    auto &lv = *levels<T>(c)[l];
    if (f32 || lv.su_ok) {
      StencilLaunch<T> d;
      d.op.S = f32 ? lv.S.p : lv.Su.p;
      d.op.Nx = lv.g.Nx; d.op.Ny = lv.g.Ny; d.op.Nz = lv.g.Nz;
      d.op.split_nodes = c->split_nodes;
      d.u = u;
      d.ep = ep;
      d.nn = lv.g.nn();
      d.nblocks = stencil_op_blocks<T>(d.op);
      rec_push(c, batch_key(BK_STENCIL, f32, epi), d, d.nblocks);
      return;
    }
  }
  issue(c, [=](hipStream_t s) { level_apply_now<T>(c, l, u, epi, ep, s); });
}

template <typename T>
void level_apply(tofe_ctx *c, int l, const T *u, Epi epi, const EpiArgs<T> &ep,
                 hipStream_t s = nullptr)
{
  if (c->rec && (!s || s == c->stream)) {
    level_record<T>(c, l, u, epi, ep);
    return;
  }
  level_apply_now<T>(c, l, u, epi, ep, s ? s : c->stream);
}

template <typename T>
void level_apply_now(tofe_ctx *c, int l, const T *u, Epi epi, const EpiArgs<T> &ep, hipStream_t s)
{
  if (l == 0) {
    const FineOp<T> f = fine_of<T>(c);
    const bool timed = kernel_timer_begin(c, s, l, sizeof(T) == 4, epi);
    if (c->fine_mfma) fine_mfma<T>(s, f, u, epi, ep);
    else if (f.wht) fine_wht<T>(s, f, u, epi, ep);
    else fine_op<T>(s, f, u, epi, ep);
    if (timed) kernel_timer_end(c, s);
  } else {
    StencilOp<T> op;
    auto &lv = *levels<T>(c)[l];
    op.S = lv.S.p;
    op.Nx = lv.g.Nx; op.Ny = lv.g.Ny; op.Nz = lv.g.Nz;
    op.split_nodes = c->split_nodes;
    const bool timed = kernel_timer_begin(c, s, l, sizeof(T) == 4, epi);
    // FP32 levels keep the upper half of the (exactly symmetric) Galerkin stencil; FP64 levels
    // read their upper-half copy once the FP64 set-up has formed it (half the stencil traffic)
    if constexpr (sizeof(T) == 4) {
      stencil_op_sym<T>(s, op, u, epi, ep);
    } else if (lv.su_ok) {
      op.S = lv.Su.p;
      stencil_op_sym<T>(s, op, u, epi, ep);
    } else {
      stencil_op<T>(s, op, u, epi, ep);
    }
    if (timed) kernel_timer_end(c, s);
  }
}

template <typename T>
int level_blocks(tofe_ctx *c, int l)
{
  if (l == 0) {
    const FineOp<T> f = fine_of<T>(c);
    return c->fine_mfma ? fine_mfma_blocks<T>(f) : f.wht ? fine_wht_blocks<T>(f) : fine_op_blocks<T>(f);
  }
  StencilOp<T> op;
  op.Nx = c->grids[l].Nx; op.Ny = c->grids[l].Ny; op.Nz = c->grids[l].Nz;
  op.split_nodes = c->split_nodes;
  return stencil_op_blocks<T>(op);
}

void alloc_level(tofe_ctx *c, Level<double> &lv, int l)
{
  lv.g = c->grids[l];
  const int64_t nd = lv.g.ndof(), ndv = nd * c->nvcap;
  if (l > 0) lv.S.alloc((size_t)243 * lv.g.nn());
  lv.dinv.alloc(nd);
  if (l > 0) lv.b.alloc(ndv);
  lv.x.alloc(ndv);
  lv.x2.alloc(ndv);
  lv.r.alloc(ndv);
  lv.d.alloc(ndv);
  lv.d.zero(c->stream);
}

void alloc_level32(tofe_ctx *c, Level<float> &lv, int l)
{
  lv.g = c->grids[l];
  const int64_t nd = lv.g.ndof(), ndv = nd * c->nvcap;
  // upper half of the symmetric stencil, node-blocked by 64 (k_stencil_op_sym)
  if (l > 0) lv.S.alloc((size_t)126 * ((lv.g.nn() + 63) / 64 * 64));
  lv.dinv.alloc(nd);
  lv.b.alloc(ndv);
  lv.x.alloc(ndv);
  lv.x2.alloc(ndv);
  lv.r.alloc(ndv);
  lv.d.alloc(ndv);
  lv.d.zero(c->stream);
}

// Grow the per-vector work buffers to hold nvec vectors (block solves, V-cycles and applies).
// Every stream is idle on return; nothing a later call reads survives in them.
void ensure_nvec(tofe_ctx *c, int nvec)
{
  if (nvec < 1 || nvec > kMaxBlockVec) throw Error(TOFE_ERR_ARG, "vectors per block pass out of range");
  if (nvec <= c->nvcap) return;
  for (hipStream_t st : {c->stream, c->stream2, c->stream3, c->stream4})
    if (st) TOFE_HIP(hipStreamSynchronize(st));
  c->nvcap = nvec;
  for (size_t l = 0; l < c->lv64.size(); ++l) alloc_level(c, *c->lv64[l], (int)l);
  for (size_t l = 0; l < c->lv32.size(); ++l) alloc_level32(c, *c->lv32[l], (int)l);
  const size_t nd = (size_t)c->ndof * nvec;
  for (auto *b : {&c->vr, &c->vz, &c->vp, &c->vq, &c->vF, &c->vU}) b->alloc(nd);
  if (c->vr32.p) {
    c->vr32.alloc(nd);
    c->vz32.alloc(nd);
  }
  if (c->ctmp64.p) c->ctmp64.alloc((size_t)c->nc * nvec);
  if (c->ctmp32.p) c->ctmp32.alloc((size_t)c->nc * nvec);
  for (auto *b : {&c->partial, &c->partial2, &c->partial3}) b->alloc((size_t)c->npart * nvec);
  c->scal.alloc((size_t)SL_COUNT * nvec);
  c->scal.zero(c->stream);
  c->frozen_dirty = false;
  if (c->hscal) (void)hipHostFree(c->hscal);
  c->hscal = c->hscal_dev = nullptr;
  TOFE_HIP(hipHostMalloc((void **)&c->hscal, (size_t)SL_COUNT * nvec * sizeof(double),
                         hipHostMallocMapped));
  TOFE_HIP(hipHostGetDevicePointer((void **)&c->hscal_dev, c->hscal, 0));
  TOFE_HIP(hipStreamSynchronize(c->stream));
}

// tofe_assemble leaves the coarsest level's factorisation running on the second stream (PETSc
// factors inside the solve's PCSetUp): the main stream waits for it just before its first use,
// so the solve's work up to the first coarse solve overlaps the factorisation's end.
void coarse_join(tofe_ctx *c)
{
  if (!c->coarse_unjoined) return;
  TOFE_HIP(hipStreamWaitEvent(c->stream, c->ev_inv, 0));
  c->coarse_unjoined = false;
}

// Exact solve with the coarsest operator, x = A^-1 b = X^T (X b) with X = L^-1, in the
// V-cycle's precision.
template <typename T>
void coarse_apply(tofe_ctx *c, const T *b, T *x, int nvec = 1)
{
  if (c->rec) {
    TrigemvLaunch<T> lo, up;
    lo.n = up.n = (int)c->nc;
    lo.ld = up.ld = c->ldc;
    lo.nblocks = up.nblocks = (lo.n + 3) / 4;
    T *tmp;
    if constexpr (sizeof(T) == 4) {
      lo.A = c->X32r.p; up.A = c->XT32r.p; tmp = c->ctmp32.p;
    } else {
      lo.A = c->X64r.p; up.A = c->XT64r.p; tmp = c->ctmp64.p;
    }
    lo.x = b; lo.y = tmp;
    up.x = tmp; up.y = x;
    lo.perm = up.perm = c->coarse_perm_on ? c->cperm.p : nullptr;
    rec_push(c, batch_key(BK_TRI_LO, sizeof(T) == 4), lo, lo.nblocks, lo.ld);
    rec_push(c, batch_key(BK_TRI_UP, sizeof(T) == 4), up, up.nblocks, up.ld);
    return;
  }
  coarse_join(c);
  const int *perm = c->coarse_perm_on ? c->cperm.p : nullptr;
  if constexpr (sizeof(T) == 4)
    coarse_solve_tri<float>(c->stream, c->X32r.p, c->XT32r.p, (int)c->nc, c->ldc, b, c->ctmp32.p, x,
                            nvec, perm);
  else
    coarse_solve_tri<double>(c->stream, c->X64r.p, c->XT64r.p, (int)c->nc, c->ldc, b, c->ctmp64.p, x,
                             nvec, perm);
}

// FP64 ends of the mixed-precision V-cycle, fused into its first and last fine-level
// kernels: b64 is converted while the first Chebyshev step is formed, and the last step
// writes x in FP64 to x64 with the partial sums of x.q64 (one per fine-operator block).
struct VcycleIO {
  const double *b64 = nullptr;
  double *x64 = nullptr;
  const double *q64 = nullptr;
  double *partial = nullptr;
  bool first_done = false;  // b0, d and the first iterate already hold the first step
  bool fused_out = false;   // set when the last step wrote x64 and the partials
  // want_z32: the caller reads the result in FP32 (its values are FP32 numbers): the last step
  // then writes it to the level's free FP32 buffer, returned in z32, instead of x64
  bool want_z32 = false;
  const float *z32 = nullptr;
};

// Where the mixed V-cycle's first fine-level Chebyshev step goes (b0, d, the iterate it
// starts from), for kernels that form that step themselves; x.b32 == nullptr when the
// V-cycle has no such step.
FirstCheb vcycle_first_step(tofe_ctx *c)
{
  FirstCheb f;
  const int nu1 = c->desc.nu_pre, nu2 = c->desc.nu_post;
  if (c->built_precision != TOFE_PREC_MIXED || c->L == 1 || nu1 == 0) return f;
  Level<float> &lv = *c->lv32[0];
  const int flips = std::max(nu1 - 1, 0) + nu2;
  f.b32 = c->vr32.p;
  f.d = lv.d.p;
  // With a second pre-smoothing step, that step reads the iterate from d (the two are equal
  // after the first step from zero), so the first step does not write it.
  f.x = nu1 >= 2 ? nullptr : (flips & 1) ? lv.x2.p : c->vz32.p;
  f.dinv = lv.dinv.p;
  f.c2 = (float)c->c2[0];
  return f;
}

// The FP64 V-cycle's first fine-level Chebyshev step for the CG residual update to form
// (with a second pre-smoothing step, which then reads its operand from d); d == nullptr: none.
FirstCheb64 vcycle_first_step64(tofe_ctx *c)
{
  FirstCheb64 f;
  if (c->built_precision != TOFE_PREC_FP64 || c->L == 1 || c->desc.nu_pre < 2) return f;
  f.d = c->lv64[0]->d.p;
  f.dinv = c->lv64[0]->dinv.p;
  f.c2 = c->c2[0];
  return f;
}

// Multigrid V-cycle: x0 = M^-1 b0 (PCMG V-cycle, Chebyshev-Jacobi smoothers,
// Galerkin coarse operators, direct coarse solve).
template <typename T>
void vcycle(tofe_ctx *c, const T *b0, T *x0, VcycleIO *io = nullptr, int nvec = 1)
{
  auto &LV = levels<T>(c);
  const int L = c->L;
  const hipStream_t s = c->stream;
  const bool mixed_io = io && sizeof(T) == 4;
  if (L == 1) {
    if (mixed_io && io->b64) {
      const double *b64 = io->b64;
      const int64_t nd = c->ndof * nvec;
      issue(c, [=](hipStream_t st) { convert<double, T>(st, b64, const_cast<T *>(b0), nd); });
    }
    coarse_apply<T>(c, b0, x0, nvec);
    return;
  }
  const int nu1 = c->desc.nu_pre, nu2 = c->desc.nu_post, km = c->kmax();
  std::vector<T *> cur(L, nullptr), oth(L, nullptr), dcur(L, nullptr);  // dcur: where d lives
  for (int l = 0; l < L - 1; ++l) {
    Level<T> &lv = *LV[l];
    const T *b = (l == 0) ? b0 : lv.b.p;
    T *xa = (l == 0) ? x0 : lv.x.p, *xb = lv.x2.p;
    const int flips = std::max(nu1 - 1, 0) + nu2;
    T *xc = (flips & 1) ? xb : xa, *xo = (xc == xa) ? xb : xa;
    const int64_t nd = lv.g.ndof();
    const double *c1 = &c->c1[(size_t)l * km], *c2 = &c->c2[(size_t)l * km];
    const bool from64 = l == 0 && mixed_io && io->b64;
    const double *b64 = from64 ? io->b64 : nullptr;
    if (from64 && nu1 == 0)
      issue(c, [=](hipStream_t st) { convert<double, T>(st, b64, const_cast<T *>(b0), nd * nvec); });
    if (nu1 > 0) {
      // the first step from x = 0 on levels l >= 1 was formed by the restriction into l
      const T *dinv = lv.dinv.p;
      T *d = lv.d.p;
      const T c20 = (T)c2[0];
      if constexpr (sizeof(T) == 4) {
        if (from64 && io->first_done) {
          // formed by the kernel that produced b64
        } else if (from64)
          issue(c, [=](hipStream_t st) {
            cheb_zero_from64(st, b64, const_cast<float *>(b0), dinv, c20, d, xc, nd, nvec);
          });
        else if (l == 0)
          issue(c, [=](hipStream_t st) { cheb_zero<T>(st, b, dinv, c20, d, xc, nd, nvec); });
      } else {
        if (l == 0 && !(io && io->first_done))  // else formed by the CG residual update
          issue(c, [=](hipStream_t st) { cheb_zero<T>(st, b, dinv, c20, d, xc, nd, nvec); });
      }
      // After a first step formed without the iterate (vcycle_first_step), the iterate equals
      // d: the second step reads its operand from d and writes the new d to the iterate's unused
      // buffer (not over d, which other blocks read as operand), and d lives there from then on.
      bool x_in_d = l == 0 && io && io->first_done && nu1 >= 2 && (sizeof(T) == 8 || from64);
      T *dc = lv.d.p;
      for (int k = 1; k < nu1; ++k) {
        EpiArgs<T> ep;
        ep.b = b; ep.dinv = lv.dinv.p; ep.d = dc; ep.xout = xo;
        ep.c1 = (T)c1[k]; ep.c2 = (T)c2[k];
        ep.nvec = nvec;
        if (x_in_d) {
          ep.dout = xc;
          level_apply<T>(c, l, dc, EPI_CHEB, ep);
          T *const freed = dc;
          dc = xc;
          xc = xo;
          xo = freed;
          x_in_d = false;
        } else {
          level_apply<T>(c, l, xc, EPI_CHEB, ep);
          std::swap(xc, xo);
        }
      }
      dcur[l] = dc;
    } else {
      issue(c, [=](hipStream_t st) { fill<T>(st, xc, T(0), nd * nvec); });
    }
    EpiArgs<T> er;
    er.y = lv.r.p; er.b = b;
    er.nvec = nvec;
    level_apply<T>(c, l, xc, EPI_RESID, er);
    const Grid &gf = lv.g, &gc = LV[l + 1]->g;
    FirstStep<T> fs;
    if (nu1 > 0 && l + 1 < L - 1) {  // level l + 1 is smoothed: its first step here
      Level<T> &ln = *LV[l + 1];
      fs.dinv = ln.dinv.p;
      fs.c2 = (T)c->c2[(size_t)(l + 1) * km];
      fs.d = ln.d.p;
      fs.x = (flips & 1) ? ln.x2.p : ln.x.p;  // the level's xc (same flip count on every level)
    }
    if (c->rec) {
      RestrictLaunch<T> d;
      d.rf = lv.r.p; d.Nfx = gf.Nx; d.Nfy = gf.Ny; d.Nfz = gf.Nz;
      d.bc = LV[l + 1]->b.p; d.Ncx = gc.Nx; d.Ncy = gc.Ny; d.Ncz = gc.Nz;
      d.fs = fs;
      d.nblocks = (int)((gc.nn() + 255) / 256);
      rec_push(c, batch_key(BK_RESTRICT, sizeof(T) == 4), d, d.nblocks);
    } else {
      restrict_op<T>(s, lv.r.p, gf.Nx, gf.Ny, gf.Nz, LV[l + 1]->b.p, gc.Nx, gc.Ny, gc.Nz, fs, nvec);
    }
    cur[l] = xc;
    oth[l] = xo;
  }
  Level<T> &lc = *LV[L - 1];
  coarse_apply<T>(c, lc.b.p, lc.x.p, nvec);
  cur[L - 1] = lc.x.p;
  for (int l = L - 2; l >= 0; --l) {
    Level<T> &lv = *LV[l];
    const T *b = (l == 0) ? b0 : lv.b.p;
    T *xc = cur[l], *xo = oth[l];
    const Grid &gf = lv.g, &gc = LV[l + 1]->g;
    if (c->rec) {
      ProlongLaunch<T> d;
      d.xc = cur[l + 1]; d.Ncx = gc.Nx; d.Ncy = gc.Ny; d.Ncz = gc.Nz;
      d.xf = xc; d.Nfx = gf.Nx; d.Nfy = gf.Ny; d.Nfz = gf.Nz;
      d.nblocks = (int)((gf.nn() + 255) / 256);
      rec_push(c, batch_key(BK_PROLONG, sizeof(T) == 4), d, d.nblocks);
    } else {
      prolong_add<T>(s, cur[l + 1], gc.Nx, gc.Ny, gc.Nz, xc, gf.Nx, gf.Ny, gf.Nz, nvec);
    }
    const double *c1 = &c->c1[(size_t)l * km], *c2 = &c->c2[(size_t)l * km];
    for (int k = 0; k < nu2; ++k) {
      EpiArgs<T> ep;
      ep.b = b; ep.dinv = lv.dinv.p; ep.d = dcur[l] ? dcur[l] : lv.d.p; ep.xout = xo;
      ep.c1 = (T)(k == 0 ? 0.0 : c1[k]); ep.c2 = (T)c2[k];
      ep.nvec = nvec;
      if (l == 0 && k == nu2 - 1 && mixed_io && io->x64) {
        ep.xout64 = io->want_z32 ? nullptr : io->x64;
        ep.q64 = io->q64; ep.partial = io->partial;
        level_apply<T>(c, l, xc, EPI_CHEB_OUT, ep);
        io->fused_out = true;
        if constexpr (sizeof(T) == 4)
          if (io->want_z32) io->z32 = xo;
        return;
      }
      level_apply<T>(c, l, xc, EPI_CHEB, ep);
      std::swap(xc, xo);
    }
    cur[l] = xc;
  }
  if (cur[0] != x0) {
    const T *src = cur[0];
    const size_t bytes = c->ndof * nvec * sizeof(T);
    issue(c, [=](hipStream_t st) {
      TOFE_HIP(hipMemcpyAsync(x0, src, bytes, hipMemcpyDeviceToDevice, st));
    });
  }
}

// z = M^-1 r.  With q given (flexible CG), also the partial sums of z.q into c->partial;
// returns how many there are.  With z32 given, a mixed V-cycle whose last step forms z.q may
// leave z in FP32 instead (*z32 then points to it, else *z32 = null and z holds it).
int precondition(tofe_ctx *c, const double *r, double *z, const double *q = nullptr,
                 bool first_done = false, int nvec = 1, const float **z32 = nullptr)
{
  const hipStream_t s = c->stream;
  if (z32) *z32 = nullptr;
  if (c->built_precision == TOFE_PREC_MIXED) {
    VcycleIO io;
    io.b64 = r;
    io.first_done = first_done;
    if (q && c->L > 1 && !c->fine_mfma) {
      io.x64 = z; io.q64 = q; io.partial = c->partial.p;
      io.want_z32 = z32 != nullptr;
    }
    vcycle<float>(c, c->vr32.p, c->vz32.p, &io, nvec);
    if (io.fused_out) {
      if (z32) *z32 = io.z32;
      return level_blocks<float>(c, 0);
    }
    const float *z32 = c->vz32.p;
    const int64_t nd = c->ndof * nvec;
    issue(c, [=](hipStream_t st) { convert<float, double>(st, z32, z, nd); });
  } else {
    VcycleIO io;
    io.first_done = first_done;  // the CG residual update formed the first fine-level step
    vcycle<double>(c, r, z, first_done ? &io : nullptr, nvec);
  }
  if (!q) return 0;
  double *part = c->partial.p;
  const int64_t nd = c->ndof;
  issue(c, [=](hipStream_t st) { dot_partial<double>(st, z, q, nd, part, nvec); });
  return kRedBlocks;
}

// Final stage of a reduction (k_reduce_final), enqueued or recorded.
void reduce_into(tofe_ctx *c, const double *partial, int nparts, double *out, int post, double *scal,
                 int nvec)
{
  if (c->rec) {
    ReduceLaunch d;
    d.partial = partial; d.nparts = nparts; d.out = out; d.post = post; d.scal = scal;
    rec_push(c, batch_key(BK_REDUCE, false), d, 1);
  } else {
    reduce_final(c->stream, partial, nparts, out, post, scal, nvec);
  }
}

// scal[v * SL_COUNT + slot] = a_v . b_v for the nvec vectors of length n
void dot_into(tofe_ctx *c, const double *a, const double *b, int64_t n, int slot, int post = POST_NONE,
              int nvec = 1)
{
  double *part = c->partial.p;
  issue(c, [=](hipStream_t st) { dot_partial<double>(st, a, b, n, part, nvec); });
  reduce_into(c, c->partial.p, kRedBlocks, c->scal.p + slot, post, c->scal.p, nvec);
}

// Device scalars (of nvec vectors) to the host through the pinned buffer: queue the copy and
// an event ...
void scal_to_host_async(tofe_ctx *c, int nvec = 1)
{
  TOFE_HIP(hipMemcpyAsync(c->hscal, c->scal.p, (size_t)SL_COUNT * nvec * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
  TOFE_HIP(hipEventRecord(c->ev_scal, c->stream));
}

// ... and wait for them.
void scal_wait(tofe_ctx *c, double *h, int nvec = 1)
{
  TOFE_HIP(hipEventSynchronize(c->ev_scal));
  std::memcpy(h, c->hscal, (size_t)SL_COUNT * nvec * sizeof(double));
}

void read_scal(tofe_ctx *c, double *h)
{
  scal_to_host_async(c);
  scal_wait(c, h);
}

// CG / flexible CG on the fine level with the KSPConvergedDefault test
// ||r|| <= max(rtol ||b||, 1e-50) (FEAnalysis.cpp:390-393), for nvec right-hand sides at
// once: every launch serves all of them (vector v of F, U and the work vectors at v * ndof,
// its scalars at v * SL_COUNT), each with its own recurrence, test and reason.  A right-hand
// side that has finished is frozen: its alpha and beta are zero from then on, so it takes
// no further steps while the others go on.  its_out, reason_out, rn_out and bn_out hold nvec
// values.
void solve_block(tofe_ctx *c, const double *F, double *U, int nvec, double rtol, int maxit,
                 int *its_out, int *reason_out, double *rn_out, double *bn_out)
{
  const hipStream_t s = c->stream;
  const int64_t n = c->ndof;
  double *r = c->vr.p, *z = c->vz.p, *p = c->vp.p, *q = c->vq.p;
  const bool flex = c->built_precision == TOFE_PREC_MIXED;
  const int nfb = level_blocks<double>(c, 0);
  std::vector<double> hv((size_t)SL_COUNT * nvec);
  double *h = hv.data();
  auto H = [&](int v, int slot) { return h[(size_t)v * SL_COUNT + slot]; };

  if (c->frozen_dirty) {
    set_frozen(s, c->scal.p, nvec, 0ull);
    c->frozen_dirty = false;
  }
  EpiArgs<double> er;
  er.y = r; er.b = F; er.nvec = nvec;
  level_apply<double>(c, 0, U, EPI_RESID, er);
  dot_into(c, F, F, n, SL_BB, POST_NONE, nvec);
  dot_into(c, r, r, n, SL_RR, POST_NONE, nvec);
  scal_to_host_async(c, nvec);
  scal_wait(c, h, nvec);
  std::vector<double> rn(nvec), rnorm0(nvec), ttol(nvec), rn_prev(nvec), rate(nvec, 0.5);
  std::vector<int> conv(nvec), reason(nvec, 0), its(nvec, 0);
  int live = 0;
  unsigned long long done_mask = 0;
  for (int v = 0; v < nvec; ++v) {
    const double bnorm = std::sqrt(H(v, SL_BB));
    rn[v] = std::sqrt(H(v, SL_RR));
    rnorm0[v] = bnorm > 0 ? bnorm : rn[v];
    ttol[v] = std::max(rtol * rnorm0[v], 1e-50);
    conv[v] = (rtol * rnorm0[v] > 1e-50) ? TOFE_CONVERGED_RTOL : TOFE_CONVERGED_ATOL;
    bn_out[v] = bnorm;
    if (std::isnan(rn[v]) || std::isinf(rn[v])) reason[v] = TOFE_DIVERGED_NANORINF;
    else if (rn[v] <= ttol[v]) reason[v] = conv[v];
    if (reason[v]) done_mask |= 1ull << v;
    else ++live;
    rn_prev[v] = rn[v];
  }
  int k = 0;
  if (live) {
    if (done_mask) {  // right-hand sides already solved take no steps
      set_frozen(s, c->scal.p, nvec, done_mask);
      c->frozen_dirty = true;
    }
    precondition(c, r, z, nullptr, false, nvec);
    dot_into(c, r, z, n, SL_RZ, POST_NONE, nvec);
    TOFE_HIP(hipMemcpyAsync(p, z, n * nvec * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  // The next preconditioner application and search direction depend on the device
  // scalars only, so they are queued before the host reads this step's residual norms
  // whenever no right-hand side is predicted (from its last contraction rate) to converge:
  // the device then never waits for the host's test.  A converged step leaves them unused.
  // mixed: the FCG update also forms the next V-cycle's first fine-level step
  const FirstCheb first = flex ? vcycle_first_step(c) : FirstCheb();
  // CG: the residual update forms the next V-cycle's first fine-level step
  const FirstCheb64 first64 = flex ? FirstCheb64() : vcycle_first_step64(c);
  // CG and flexible CG apply the iterate's update x += alpha p in the next search-direction
  // kernel, which reads p anyway; x_pending marks a step whose update is still owed
  bool x_pending = false;
  auto next_direction = [&] {
    if (flex) {
      const float *z32 = nullptr;
      const int nparts = precondition(c, r, z, q, first.b32 != nullptr, nvec, &z32);
      fcg_update_xp(s, p, z, z32, U, n, c->partial.p, nparts, c->scal.p, nvec);
    } else {
      precondition(c, r, z, nullptr, first64.d != nullptr, nvec);
      dot_into(c, r, z, n, SL_TMP, POST_BETA_CG, nvec);
      cg_update_xp(s, p, z, U, n, c->scal.p, SL_BETA, SL_ALPHA, nvec);
    }
    x_pending = false;
  };
  while (live) {
    if (k >= maxit) {
      for (int v = 0; v < nvec; ++v)
        if (!reason[v]) {
          reason[v] = TOFE_DIVERGED_ITS;
          its[v] = k;
        }
      break;
    }
    EpiArgs<double> ep;
    ep.y = q; ep.partial = c->partial.p; ep.nvec = nvec;
    if (flex) {
      ep.b = r;
      ep.partial2 = c->partial2.p;
      level_apply<double>(c, 0, p, EPI_APPLY_DOT2, ep);
      fcg_update_r(s, r, q, n, c->partial.p, c->partial2.p, nfb, c->scal.p, c->partial3.p, first,
                   nvec);
      x_pending = true;
      reduce_final(s, c->partial3.p, kRedBlocks, c->scal.p + SL_RR, POST_NONE, c->scal.p, nvec,
                   c->hscal_dev);
    } else {
      level_apply<double>(c, 0, p, EPI_APPLY_DOT, ep);
      reduce_final(s, c->partial.p, nfb, c->scal.p + SL_PQ, POST_ALPHA_CG, c->scal.p, nvec);
      cg_update_r(s, r, q, n, c->scal.p, SL_ALPHA, c->partial.p, nvec, first64);
      x_pending = true;
      reduce_final(s, c->partial.p, kRedBlocks, c->scal.p + SL_RR, POST_NONE, c->scal.p, nvec,
                   c->hscal_dev);
    }
    // the reduction has written this iteration's scalars to the mapped host copy: no copy
    // launch between it and the next direction's kernels
    TOFE_HIP(hipEventRecord(c->ev_scal, s));
    bool ahead = k + 1 < maxit;
    for (int v = 0; v < nvec && ahead; ++v)
      if (!reason[v] && !(rn_prev[v] * std::min(rate[v], 1.0) > 3 * ttol[v])) ahead = false;
    if (ahead) next_direction();
    scal_wait(c, h, nvec);
    ++k;
    unsigned long long newly = 0;
    for (int v = 0; v < nvec; ++v) {
      if (reason[v]) continue;
      int rs = 0;
      const double rnv = std::sqrt(H(v, SL_RR));
      if (!flex && H(v, SL_RZ) < 0) rs = TOFE_DIVERGED_INDEFINITE_PC;
      else if (!(H(v, SL_PQ) > 0))
        rs = std::isnan(H(v, SL_PQ)) ? TOFE_DIVERGED_NANORINF : TOFE_DIVERGED_INDEFINITE_MAT;
      else {
        rn[v] = rnv;
        if (std::isnan(rnv) || std::isinf(rnv)) rs = TOFE_DIVERGED_NANORINF;
        else if (rnv <= ttol[v]) rs = conv[v];
        else if (rnv >= 1e5 * rnorm0[v]) rs = TOFE_DIVERGED_DTOL;
      }
      if (rs) {
        reason[v] = rs;
        its[v] = k;
        newly |= 1ull << v;
        --live;
      } else {
        if (rn_prev[v] > 0) rate[v] = rnv / rn_prev[v];
        rn_prev[v] = rnv;
      }
    }
    if (!live) break;
    if (newly) {
      // frozen from here on; a finished right-hand side's owed update of this step is still
      // applied by the next search-direction kernel (its alpha is already formed)
      done_mask |= newly;
      set_frozen(s, c->scal.p, nvec, done_mask);
      c->frozen_dirty = true;
    }
    if (!ahead) next_direction();
  }
  // the last step's owed update (right-hand sides frozen earlier have alpha = 0)
  if (x_pending) fcg_update_x(s, U, p, n, c->scal.p, nvec);
  for (int v = 0; v < nvec; ++v) {
    its_out[v] = its[v];
    reason_out[v] = reason[v];
    rn_out[v] = rn[v];
  }
}

// lambda_max(D^-1 A) of level l from eig_its steps of Jacobi-preconditioned CG: the CG
// coefficients alpha_k, beta_k (recorded on the device, read back once per assemble)
// form the Lanczos tridiagonal whose top eigenvalue is the estimate; PETSc's Chebyshev
// smoother estimates the same quantity (KSPChebyshevEstEigSet 0, 0.1, 0, 1.1).
template <typename T>
void lanczos_coefficients(tofe_ctx *c, int l, hipStream_t s, T *r, T *p, T *q, double *part,
                          double *part2, double *scal)
{
  Level<T> &lv = *levels<T>(c)[l];
  const int64_t nd = lv.g.ndof();
  const int nb = level_blocks<T>(c, l);
  const int its = std::min(c->desc.eig_its, kMaxEig);
  double *alpha = c->eig_ab.p + (size_t)(2 * l) * kMaxEig;
  double *beta = c->eig_ab.p + (size_t)(2 * l + 1) * kMaxEig;
  // Three launches per step: the operator with its p.Ap partials, the r update with the
  // r.z partials, the p update; each consumer sums the producer's partials itself.  The
  // iterate x of the CG is never needed for the estimate and is not formed.
  power_start<T>(s, r, nd);
  jacobi_dot<T>(s, r, lv.dinv.p, p, nd, part);
  reduce_final(s, part, kRedBlocks, scal + SL_RZA, POST_NONE, scal);
  for (int k = 0; k < its; ++k) {
    const int rz_old = (k & 1) ? SL_RZB : SL_RZA, rz_new = (k & 1) ? SL_RZA : SL_RZB;
    EpiArgs<T> ep;
    ep.y = q;
    ep.partial = part;
    level_apply<T>(c, l, p, EPI_APPLY_DOT, ep, s);
    lanczos_r<T>(s, r, q, lv.dinv.p, nd, part, nb, scal, rz_old, alpha + k, part2);
    lanczos_p<T>(s, p, r, lv.dinv.p, nd, part2, scal, rz_old, rz_new, beta + k);
  }
}

// The smoothed levels' estimates are independent: level 1 runs on the third stream, the fine
// level on the fourth and levels 2.. in order on the main stream, each on its level's V-cycle
// work vectors and its stream's reduction buffers, once the operators are built on the main
// stream.  In the mixed-precision build they run on
// the FP32 operators of the V-cycle (scalars and recorded coefficients stay FP64), whose
// copies are made on the same streams first; the FP64 build first copies each level's
// stencil into its upper-half layout there.  ev_lz / ev_lz0 mark their ends.
// The fine level's estimate on the fourth stream, once the fine diagonal (main stream) is
// formed: it needs nothing of the Galerkin hierarchy.  ev_lz0 marks its end.
// after: an event of the main stream recorded once E and the fine diagonal are formed (null:
// record one now)
void lanczos_fine_enqueue(tofe_ctx *c, hipEvent_t after = nullptr)
{
  const int L = c->L;
  const hipStream_t s4 = c->stream4;
  c->eig_ab.alloc((size_t)2 * L * kMaxEig);
  if (!after) {
    TOFE_HIP(hipEventRecord(c->ev_ops, c->stream));
    after = c->ev_ops;
  }
  TOFE_HIP(hipStreamWaitEvent(s4, after, 0));
  if (c->precision == TOFE_PREC_MIXED) {
    if (c->lv32.empty()) {
      for (int l = 0; l < L; ++l) {
        c->lv32.emplace_back(new Level<float>());
        alloc_level32(c, *c->lv32.back(), l);
      }
      c->vr32.alloc((size_t)c->ndof * c->nvcap);
      c->vz32.alloc((size_t)c->ndof * c->nvcap);
      TOFE_HIP(hipStreamSynchronize(c->stream));
    }
    convert<double, float>(s4, c->lv64[0]->dinv.p, c->lv32[0]->dinv.p, c->ndof);
    convert<double, float>(s4, c->E.p, c->E32.p, c->ne);
    if (L > 1) {
      Level<float> &lv = *c->lv32[0];
      lanczos_coefficients<float>(c, 0, s4, lv.r.p, lv.x.p, lv.x2.p, c->partial.p,
                                  c->partial2.p, c->scal.p);
    }
  } else if (L > 1) {
    lanczos_coefficients<double>(c, 0, s4, c->vr.p, c->vp.p, c->vq.p, c->partial.p,
                                 c->partial2.p, c->scal.p);
  }
  TOFE_HIP(hipEventRecord(c->ev_lz0, s4));
}

void lanczos_enqueue(tofe_ctx *c, bool fine_done)
{
  const int L = c->L;
  // level 1 on the third stream, the fine level on the fourth (unless enqueued already),
  // levels 2.. on the main stream (idle until it joins the others)
  const hipStream_t s0 = c->stream, s3 = c->stream3, s4 = c->stream4;
  if (!fine_done) lanczos_fine_enqueue(c);
  TOFE_HIP(hipEventRecord(c->ev_ops, c->stream));
  TOFE_HIP(hipStreamWaitEvent(s3, c->ev_ops, 0));
  (void)s4;
  if (c->precision == TOFE_PREC_MIXED) {
    for (int l = 1; l < L; ++l) {
      const hipStream_t sl = l == 1 ? s3 : s0;
      convert<double, float>(sl, c->lv64[l]->dinv.p, c->lv32[l]->dinv.p, c->grids[l].ndof());
      stencil_upper<float>(sl, c->lv64[l]->S.p, c->grids[l].nn(), c->lv32[l]->S.p);
    }
    for (int l = 1; l + 1 < L; ++l) {
      Level<float> &lv = *c->lv32[l];
      if (l == 1)
        lanczos_coefficients<float>(c, l, s3, lv.r.p, lv.x.p, lv.x2.p, c->partial_b.p,
                                    c->partial2_b.p, c->scal_b.p);
      else
        lanczos_coefficients<float>(c, l, s0, lv.r.p, lv.x.p, lv.x2.p, c->partial_c.p,
                                    c->partial2_c.p, c->scal_c.p);
    }
  } else {
    // FP64: upper-half copies of the level stencils for the Lanczos steps and the solve
    for (int l = 1; l < L; ++l) {
      Level<double> &lv = *c->lv64[l];
      lv.Su.alloc((size_t)126 * ((lv.g.nn() + 63) / 64 * 64));
      stencil_upper<double>(l == 1 ? s3 : s0, lv.S.p, lv.g.nn(), lv.Su.p);
      lv.su_ok = true;
    }
    for (int l = 1; l + 1 < L; ++l) {
      Level<double> &lv = *c->lv64[l];
      if (l == 1)
        lanczos_coefficients<double>(c, l, s3, lv.r.p, lv.x.p, lv.x2.p, c->partial_b.p,
                                     c->partial2_b.p, c->scal_b.p);
      else
        lanczos_coefficients<double>(c, l, s0, lv.r.p, lv.x.p, lv.x2.p, c->partial_c.p,
                                     c->partial2_c.p, c->scal_c.p);
    }
  }
  TOFE_HIP(hipEventRecord(c->ev_lz, s3));
}

void lanczos_lambdas(tofe_ctx *c)
{
  const int L = c->L;
  const int its = std::min(c->desc.eig_its, kMaxEig);
  std::vector<double> ab((size_t)2 * L * kMaxEig);
  TOFE_HIP(hipMemcpyAsync(ab.data(), c->eig_ab.p, ab.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  TOFE_HIP(hipStreamSynchronize(c->stream));
  c->lam.assign(L, 0.0);
  for (int l = 0; l + 1 < L; ++l) {
    const double *alpha = &ab[(size_t)(2 * l) * kMaxEig], *beta = &ab[(size_t)(2 * l + 1) * kMaxEig];
    std::vector<double> ta, tb;
    for (int k = 0; k < its; ++k) {
      if (!(alpha[k] > 0) || !(beta[k] >= 0) || !std::isfinite(alpha[k]) || !std::isfinite(beta[k])) break;
      ta.push_back(1.0 / alpha[k] + (k > 0 ? beta[k - 1] / alpha[k - 1] : 0.0));
      tb.push_back(std::sqrt(beta[k]) / alpha[k]);
    }
    c->lam[l] = ta.empty() ? 1.0 : tridiag_max_eig((int)ta.size(), ta.data(), tb.data());
  }
}

void chebyshev_coefficients(tofe_ctx *c)
{
  const int km = c->kmax();
  c->c1.assign((size_t)c->L * km, 0.0);
  c->c2.assign((size_t)c->L * km, 0.0);
  for (int l = 0; l + 1 < c->L; ++l) {
    const double emax = c->desc.cheb_hi * c->lam[l], emin = c->desc.cheb_lo * c->lam[l];
    const double theta = 0.5 * (emax + emin), delta = 0.5 * (emax - emin);
    const double sigma = theta / delta;
    double rho_old = 1.0 / sigma;
    for (int k = 0; k < km; ++k) {
      double a = 0.0, b = 1.0 / theta;
      if (c->desc.smoother == TOFE_SMOOTHER_CHEBYSHEV && k > 0) {
        const double rho = 1.0 / (2 * sigma - rho_old);
        a = rho * rho_old;
        b = 2 * rho / delta;
        rho_old = rho;
      }
      c->c1[(size_t)l * km + k] = a;
      c->c2[(size_t)l * km + k] = b;
    }
  }
}

// A (dense SPD, column-major, in c->Ainv) -> X = L^-1 on the second stream: banded blocked
// Cholesky (one launch per 64-column step, each also forming the previous row block of X),
// then X and X^T packed as padded row-major triangles for the per-V-cycle solve
// X^T (X b).  The coarse operator couples DOFs at most bw apart in the dense order of
// coarse_order (27-point stencil) and the factor's fill stays inside that band.
// Tables of the direct coarsest operator (k_direct.hip) when TOFE_DIRECT_COARSE=1 and the mesh
// allows it: M_p, the Dirichlet elements' table rows, the cells' constant part.  Off by default:
// the set-up is bound by the bandwidth of the Galerkin chain and the Lanczos estimates, which the
// coarse factorisation overlaps either way (192x96x96 cantilever, set-up per call: mixed 2.53 ms
// direct / 2.54 ms chain, FP64 3.48 / 3.36 ms).
void direct_setup(tofe_ctx *c, const uint8_t *fixmask)
{
  const tofe_desc &d = c->desc;
  const char *env = std::getenv("TOFE_DIRECT_COARSE");
  c->direct = direct_coarse_ok(d.nx, d.ny, d.nz, c->L) && env && env[0] == '1';
  if (!c->direct) return;
  const int r = 1 << (c->L - 1);
  const int ncell = (d.nx / r) * (d.ny / r) * (d.nz / r);
  std::vector<double> M((size_t)r * r * r * 576);
  direct_coarse_table(c->KE, r, M.data());
  c->direct_M.alloc(M.size());
  TOFE_HIP(hipMemcpy(c->direct_M.p, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
  DirectStatic st;
  direct_coarse_static(c->KE, d.nx, d.ny, d.nz, r, fixmask, st);
  c->direct_nmasked = (int)st.cells.size();
  auto up = [](auto &buf, const auto &v) {
    buf.alloc(std::max<size_t>(v.size(), 1));
    if (!v.empty())
      TOFE_HIP(hipMemcpy(buf.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
  };
  up(c->direct_range, st.range);
  up(c->direct_tidx, st.tidx);
  up(c->direct_elem, st.elem);
  up(c->direct_table, st.table);
  up(c->direct_slot, st.slot_of);
  up(c->direct_cconst, st.cconst);
  c->direct_cfix = std::move(st.cconst);
  c->direct_part.alloc((size_t)direct_coarse_chunks(ncell, r) * ncell * 576);
  c->direct_cells.alloc((size_t)ncell * 576);
  c->direct_corr.alloc((size_t)std::max(c->direct_nmasked, 1) * 576 * kDirectMaskSplit);
}

// Numbering of the coarsest grid in its dense matrix.  With nodes x-fastest the factor's
// half-bandwidth is 3 (Nx Ny + Nx + 1) + 2 DOFs; numbering the two shortest axes fastest
// narrows it (13 x 7 x 7 nodes: 317 -> 173), which shortens every step of the factorisation
// (fewer trailing tiles, fewer band blocks per row block of X).  The coarse solve gathers b and
// scatters x through the permutation.  Computed once per context on the main stream.
void coarse_order(tofe_ctx *c)
{
  if (c->cperm_ready) return;
  const Grid &g = c->grids[c->L - 1];
  const int N[3] = {g.Nx, g.Ny, g.Nz};
  int ax[3] = {0, 1, 2};  // axes by node count, shortest first (stable)
  std::stable_sort(ax, ax + 3, [&](int a, int b) { return N[a] < N[b]; });
  c->cbw_nat = 3 * (g.Nx * g.Ny + g.Nx + 1) + 2;
  c->cbw_perm = 3 * (N[ax[0]] * N[ax[1]] + N[ax[0]] + 1) + 2;
  c->cperm_on = c->cbw_perm < c->cbw_nat;
  if (c->cperm_on) {
    const int64_t nn = g.nn();
    c->hperm.assign((size_t)(3 * nn), 0);
    c->hinv.assign((size_t)(3 * nn), 0);
    int64_t q = 0;  // dense-order node index
    int idx[3];
    for (idx[ax[2]] = 0; idx[ax[2]] < N[ax[2]]; ++idx[ax[2]])
      for (idx[ax[1]] = 0; idx[ax[1]] < N[ax[1]]; ++idx[ax[1]])
        for (idx[ax[0]] = 0; idx[ax[0]] < N[ax[0]]; ++idx[ax[0]], ++q) {
          const int64_t nat = idx[0] + (int64_t)g.Nx * (idx[1] + (int64_t)g.Ny * idx[2]);
          for (int comp = 0; comp < 3; ++comp) {
            c->hperm[(size_t)(3 * q + comp)] = (int)(3 * nat + comp);
            c->hinv[(size_t)(3 * nat + comp)] = (int)(3 * q + comp);
          }
        }
    c->cperm.alloc(c->hperm.size());
    c->cinv.alloc(c->hinv.size());
    TOFE_HIP(hipMemcpyAsync(c->cperm.p, c->hperm.data(), c->hperm.size() * sizeof(int),
                            hipMemcpyHostToDevice, c->stream));
    TOFE_HIP(hipMemcpyAsync(c->cinv.p, c->hinv.data(), c->hinv.size() * sizeof(int),
                            hipMemcpyHostToDevice, c->stream));
  }
  c->cperm_ready = true;
}

void coarse_inverse_enqueue(tofe_ctx *c)
{
  const hipStream_t s = c->stream2;
  const int n = (int)c->nc;
  c->chol_l.alloc((size_t)n * n);
  c->chol_x.alloc((size_t)n * n);
  TOFE_HIP(hipMemsetAsync(c->dinfo.p, 0, 2 * sizeof(int), s));
  const int bw = c->coarse_bw;
  // a failed factorisation is reported by coarse_inverse_finish; the inverse is then unused
  coarse_inverse_factor(s, c->Ainv.p, c->chol_l.p, c->chol_x.p, n, bw, c->dinfo.p);
  c->ldc = (n + 255) / 256 * 256;
  if (c->precision == TOFE_PREC_MIXED) {
    c->X32r.alloc((size_t)n * c->ldc);
    c->XT32r.alloc((size_t)n * c->ldc);
    c->ctmp32.alloc((size_t)n * c->nvcap);
    tri_pack<float>(s, c->chol_x.p, n, c->ldc, c->X32r.p, c->XT32r.p);
  } else {
    c->X64r.alloc((size_t)n * c->ldc);
    c->XT64r.alloc((size_t)n * c->ldc);
    c->ctmp64.alloc((size_t)n * c->nvcap);
    tri_pack<double>(s, c->chol_x.p, n, c->ldc, c->X64r.p, c->XT64r.p);
  }
}

void coarse_inverse_finish(tofe_ctx *c)
{
  int info[2] = {0, 0};
  TOFE_HIP(hipMemcpyAsync(info, c->dinfo.p, sizeof(int), hipMemcpyDeviceToHost, c->stream2));
  TOFE_HIP(hipStreamSynchronize(c->stream2));
  if (info[0] != 0)
    throw Error(TOFE_ERR_NUMERIC, "coarse operator is not positive definite (pivot " +
                                      std::to_string(info[0]) + ")");
}

// Joins the factorisation and reads its pivot report: at the end of a solve, and at the start
// of every other call on the context (its buffers and the second stream are then settled).  A
// failed factorisation leaves the context unassembled.
void coarse_settle(tofe_ctx *c)
{
  coarse_join(c);
  if (!c->coarse_unchecked) return;
  c->coarse_unchecked = false;
  try {
    coarse_inverse_finish(c);
  } catch (...) {
    c->assembled = false;
    throw;
  }
}

void assemble(tofe_ctx *c)
{
  Range range(c->precision == TOFE_PREC_MIXED ? "assemble:mixed" : "assemble:fp64");
  const double t0 = now_s();
  const hipStream_t s = c->stream;
  const int L = c->L;
  coarse_order(c);  // its first copies precede every event the second stream waits for
  // fine level: Jacobi diagonal of K; level 1 = P^T K P element by element from E
  c->diag0.alloc(c->ndof);
  for (auto &lv : c->lv64) lv->su_ok = false;  // the stencils are formed anew below
  const uint8_t *fix = c->has_fix ? c->dfix.p : nullptr;
  const double *spk = c->has_springs ? c->spk.p : nullptr;
  fine_diag(s, c->E.p, c->dKE.p, fix, c->desc.nx, c->desc.ny, c->desc.nz, c->diag0.p,
            c->lv64[0]->dinv.p, spk);
  if (spk && c->precision == TOFE_PREC_MIXED) {
    c->spk32.alloc(c->ndof);
    convert<double, float>(s, spk, c->spk32.p, c->ndof);
  }
  TOFE_HIP(hipEventRecord(c->ev_diag, s));
  const Grid &gL = c->grids[L - 1];
  c->nc = gL.ndof();
  c->Ainv.alloc((size_t)c->nc * c->nc);
  c->dinfo.alloc(2);
  if (c->direct) {
    // The coarsest operator straight from E on the second stream, and its factorisation right
    // after, beside the Galerkin chain of the smoothed levels on the main stream (enqueued
    // first: it is the longest chain of the set-up).
    TOFE_HIP(hipStreamWaitEvent(c->stream2, c->ev_diag, 0));
    DirectCoarse dc;
    dc.E = c->E.p;
    dc.M = c->direct_M.p;
    dc.nmasked = c->direct_nmasked;
    dc.range = c->direct_range.p;
    dc.elem = c->direct_elem.p;
    dc.tidx = c->direct_tidx.p;
    dc.table = c->direct_table.p;
    dc.slot_of = c->direct_slot.p;
    dc.cconst = c->direct_cconst.p;
    dc.part = c->direct_part.p;
    dc.cells = c->direct_cells.p;
    dc.corr = c->direct_corr.p;
    dc.nx = c->desc.nx; dc.ny = c->desc.ny; dc.nz = c->desc.nz;
    dc.r = 1 << (L - 1);
    c->coarse_perm_on = c->cperm_on;
    c->coarse_bw = c->cperm_on ? c->cbw_perm : c->cbw_nat;
    dc.perm = c->coarse_perm_on ? c->cperm.p : nullptr;
    direct_coarse(c->stream2, dc, c->Ainv.p);
    coarse_inverse_enqueue(c);
    TOFE_HIP(hipEventRecord(c->ev_inv, c->stream2));
    c->coarse_unjoined = c->coarse_unchecked = true;
  }
  if (L > 1) {
    MaskedElems me;
    me.elem = c->dme_elem.p;
    me.form = c->dme_form.p;
    me.mask = c->dme_mask.p;
    me.count = c->n_masked;
    galerkin_fine(s, c->E.p, c->demap.p, c->dKEC.p, me, c->dKE.p, c->dQforms.p, c->dKCm.p, fix,
                  c->desc.nx, c->desc.ny, c->desc.nz, c->lv64[1]->S.p);
    if (spk) galerkin_springs(s, spk, c->grids[0].Nx, c->grids[0].Ny, c->grids[0].Nz, false,
                              c->lv64[1]->S.p);
    stencil_diag(s, c->lv64[1]->S.p, c->grids[1].nn(), nullptr, c->lv64[1]->dinv.p);
  } else {
    c->fineS.alloc((size_t)243 * c->nn);
    fine_stencil_build(s, fine_of<double>(c), c->fineS.p);
    if (spk) galerkin_springs(s, spk, c->grids[0].Nx, c->grids[0].Ny, c->grids[0].Nz, true,
                              c->fineS.p);
  }
  for (int l = 1; l + 1 < L; ++l) {
    const Grid &gf = c->grids[l], &gc = c->grids[l + 1];
    (void)gc;
    int64_t n2 = 0;
    const int64_t n1 = galerkin_scratch(gf.Nx, gf.Ny, gf.Nz, &n2);
    if (c->gal_t1.n < (size_t)n1) c->gal_t1.alloc(n1);
    if (c->gal_t2.n < (size_t)n2) c->gal_t2.alloc(n2);
    galerkin(s, c->lv64[l]->S.p, gf.Nx, gf.Ny, gf.Nz, c->lv64[l + 1]->S.p, c->gal_t1.p, c->gal_t2.p);
    stencil_diag(s, c->lv64[l + 1]->S.p, gc.nn(), nullptr, c->lv64[l + 1]->dinv.p);
  }
  if (!c->direct) {
    // coarsest level from the chain's last stencil: dense Cholesky factor and explicit inverse
    // (FP64) on the second stream
    const double *SL = (L == 1) ? c->fineS.p : c->lv64[L - 1]->S.p;
    c->coarse_perm_on = c->cperm_on;
    c->coarse_bw = c->cperm_on ? c->cbw_perm : c->cbw_nat;
    stencil_to_dense(s, SL, gL.Nx, gL.Ny, gL.Nz, c->Ainv.p, c->coarse_perm_on ? c->cinv.p : nullptr);
    TOFE_HIP(hipEventRecord(c->ev_dense, s));
    TOFE_HIP(hipStreamWaitEvent(c->stream2, c->ev_dense, 0));
    coarse_inverse_enqueue(c);
    TOFE_HIP(hipEventRecord(c->ev_inv, c->stream2));
    c->coarse_unjoined = c->coarse_unchecked = true;
  }
  // The fine level's Lanczos estimate needs E and the fine diagonal only, but the fourth stream
  // starts it after the Galerkin chain: the chain feeds the other estimates, and the estimate's
  // bandwidth-bound kernels beside it slow it down (measured: 0.07 ms more set-up per call on the
  // 192x96x96 cantilever when started with the chain).
  lanczos_fine_enqueue(c);
  // Four streams from here: the coarsest level's factorisation on the second, the Lanczos
  // estimates of level 1 on the third, of the fine level on the fourth (enqueued above) and of
  // levels 2.. on the main stream, which then joins the other three.
  lanczos_enqueue(c, true);
  TOFE_HIP(hipStreamWaitEvent(s, c->ev_lz, 0));
  TOFE_HIP(hipStreamWaitEvent(s, c->ev_lz0, 0));
  lanczos_lambdas(c);
  chebyshev_coefficients(c);
  // the factorisation (flagged where it was enqueued, so that a call after a failure here
  // still waits for it) is joined by the first coarse solve and checked by coarse_settle
  TOFE_HIP(hipStreamSynchronize(s));
  c->built_precision = c->precision;
  c->assembled = true;
  c->stats.setup_seconds = now_s() - t0;
}

void check_ctx(const tofe_ctx *c)
{
  if (!c) throw Error(TOFE_ERR_ARG, "null context");
  TOFE_HIP(hipSetDevice(c->device));
}

// Runs an entry point: errors become return codes and the last error message.  The coarse
// factorisation of the last tofe_assemble is settled first, except for tofe_solve_dev
// (settle_first false), which settles it at its end.
template <typename F>
int guarded(tofe_ctx *c, const char *name, F &&f, bool settle_first = true)
{
  if (!c) return TOFE_ERR_ARG;
  Range range(name);
  try {
    check_ctx(c);
    if (settle_first) coarse_settle(c);
    f();
    return TOFE_OK;
  } catch (const Error &e) {
    c->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    c->err = "host allocation failed";
    return TOFE_ERR_NOMEM;
  } catch (const std::exception &e) {
    c->err = e.what();
    return TOFE_ERR_HIP;
  }
}

void upload_design(tofe_ctx *c, const double *dx)
{
  const hipStream_t s = c->stream;
  const FilterTemplate t = filter_tpl(c);
  if (dx != c->x.p)
    TOFE_HIP(hipMemcpyAsync(c->x.p, dx, c->ne * sizeof(double), hipMemcpyDeviceToDevice, s));
  filter_apply(s, t, c->filt_inv.p, c->x.p, c->rho.p);
  if (c->has_vd) {
    // rho <- min(rho, 1); y = min(1, (REdge + R (1-rho)^q) / vdMin)
    void_pre(s, c->rho.p, c->ne, c->sp.vd_penal, c->ebuf3.p);
    rfilter_madd(s, rfilter_tpl(c), c->REdge.p, c->ebuf3.p, 1 / c->sp.vd_min, c->y.p);
  }
  simp(s, c->rho.p, c->has_vd ? c->y.p : nullptr, c->ne, c->sp, c->y.p, c->rhoq.p, c->E.p,
       c->dEdz.p, c->Es.p, c->dEsdz.p, c->E32.p);
  TOFE_HIP(hipStreamSynchronize(s));
  c->design_set = true;
  c->assembled = false;
}

// Chain_Filter for dfdE, device to device (Material_Interpolation.cpp:164-183), with the
// element buffers t1, t2 as scratch of the maximum-length-scale terms.
void chain_dev(tofe_ctx *c, const double *g, double *out, double *t1, double *t2)
{
  const hipStream_t s = c->stream;
  if (c->has_vd) {
    mul(s, g, c->rho.p, t1, c->ne);
    rfilter_T(s, rfilter_tpl(c), t1, t2);
    chain_mid(s, g, c->y.p, c->rhoq.p, t2, t1, c->ne);
    filter_apply_T(s, filter_tpl(c), c->filt_inv.p, t1, out);
  } else {
    filter_apply_T(s, filter_tpl(c), c->filt_inv.p, g, out);
  }
}

void chain_dev(tofe_ctx *c, const double *g, double *out)
{
  chain_dev(c, g, out, c->ebuf3.p, c->ebuf2.p);
}

void need_assembled(tofe_ctx *c)
{
  if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set (call tofe_set_design or tofe_set_E)");
  if (!c->assembled || c->built_precision != c->precision)
    throw Error(TOFE_ERR_STATE, "operator not assembled for the current design (call tofe_assemble)");
}

void solve_dev(tofe_ctx *c, const double *dF, double *dU, int nrhs, double rtol, int maxit,
               int *its, int *reason)
{
  need_assembled(c);
  if (nrhs < 1 || !dF || !dU) throw Error(TOFE_ERR_ARG, "bad solve arguments");
  Range range(c->built_precision == TOFE_PREC_MIXED ? "solve:mixed" : "solve:fp64");
  const double t0 = now_s();
  ensure_nvec(c, std::min(nrhs, kMaxBlockVec));
  int kmax = 0, rs_out = 0;
  double rn_last = 0, bn_last = 0;
  for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec) {
    const int nv = std::min(kMaxBlockVec, nrhs - i0);
    int k[kMaxBlockVec], rs[kMaxBlockVec];
    double rn[kMaxBlockVec], bn[kMaxBlockVec];
    solve_block(c, dF + (size_t)i0 * c->ndof, dU + (size_t)i0 * c->ndof, nv, rtol, maxit, k, rs,
                rn, bn);
    for (int v = 0; v < nv; ++v) {
      kmax = std::max(kmax, k[v]);
      // the first failing right-hand side's reason, else the last one's
      if (rs_out >= 0) rs_out = rs[v];
      rn_last = rn[v];
      bn_last = bn[v];
    }
  }
  TOFE_HIP(hipStreamSynchronize(c->stream));
  coarse_settle(c);
  c->stats.solve_seconds = now_s() - t0;
  c->stats.its = kmax;
  c->stats.reason = rs_out;
  c->stats.rnorm = rn_last;
  c->stats.bnorm = bn_last;
  if (its) *its = kmax;
  if (reason) *reason = rs_out;
}

// Y = M X for nrhs vectors with the assembled mass form (chunks of up to 16 vectors per
// launch for the consistent form)
void apply_mass(tofe_ctx *c, const double *X, double *Y, int nrhs)
{
  if (c->mass_form == TOFE_MASS_CONSISTENT) {
    for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec)
      apply_mass_consistent(c->stream, c->mV.p, c->mass_m0, c->has_mlump ? c->mlump.p : nullptr,
                            c->has_fix ? c->dfix.p : nullptr, c->desc.nx, c->desc.ny, c->desc.nz,
                            X + (size_t)i0 * c->ndof, Y + (size_t)i0 * c->ndof,
                            std::min(kMaxBlockVec, nrhs - i0));
  } else {
    diag_mul(c->stream, c->Mdiag.p, X, Y, c->ndof, nrhs);
  }
}

// Average ms per launch of nrep back-to-back launches of one level operator kernel on the
// context's stream (HIP events), on a deterministic input with the level's work vectors.
double time_level_op(tofe_ctx *c, int level, bool f32, Epi epi, int nrep)
{
  const hipStream_t s = c->stream;
  const int64_t nd = c->grids[level].ndof();
  hipEvent_t e0, e1;
  TOFE_HIP(hipEventCreate(&e0));
  TOFE_HIP(hipEventCreate(&e1));
  auto run = [&](auto tag) {
    using T = decltype(tag);
    auto &lv = *levels<T>(c)[level];
    Level<double> &l64 = *c->lv64[level];
    power_start(s, l64.r.p, nd);
    if constexpr (sizeof(T) == 4) convert<double, float>(s, l64.r.p, lv.x.p, nd);
    else TOFE_HIP(hipMemcpyAsync(lv.x.p, l64.r.p, nd * sizeof(double), hipMemcpyDeviceToDevice, s));
    TOFE_HIP(hipMemsetAsync(lv.d.p, 0, nd * sizeof(T), s));
    EpiArgs<T> ep;
    ep.y = lv.r.p;
    ep.b = lv.x.p;
    ep.dinv = lv.dinv.p;
    ep.d = lv.d.p;
    ep.xout = lv.x2.p;
    ep.c1 = (T)0.5;
    ep.c2 = (T)0.5;
    level_apply<T>(c, level, lv.x.p, epi, ep);  // warm-up
    TOFE_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < nrep; ++k) level_apply<T>(c, level, lv.x.p, epi, ep);
    TOFE_HIP(hipEventRecord(e1, s));
  };
  if (f32) run(float{});
  else run(double{});
  TOFE_HIP(hipEventSynchronize(e1));
  float t = 0;
  TOFE_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)t / nrep;
}

// Launch shapes of the current precision's operators from measurement: the fine operator's
// z-chunk length (element-centric form) over candidate lengths, and the node-count
// threshold below which a stencil level runs one thread per DOF, chosen to minimise the
// summed Chebyshev-step time of the coarse levels.  The chosen shapes stay in the context.
void autotune(tofe_ctx *c, int nrep, int *zlen_out, int64_t *split_out)
{
  need_assembled(c);
  const bool f32 = c->built_precision == TOFE_PREC_MIXED;
  int best_z = f32 ? c->wht_zlen32 : c->wht_zlen;
  if (c->ke_wht && !c->fine_mfma) {
    double best = 1e300;
    const int keep = best_z;
    for (int zl : {4, 5, 6, 8, 10, 12, 16, 20, 24}) {
      if (zl > c->desc.nz + 1 && zl != keep) continue;
      (f32 ? c->wht_zlen32 : c->wht_zlen) = zl;
      const double t = time_level_op(c, 0, f32, EPI_CHEB, nrep);
      if (t < best * 0.995) {  // a longer chunk must win clearly (fewer blocks, less halo)
        best = t;
        best_z = zl;
      }
    }
    (f32 ? c->wht_zlen32 : c->wht_zlen) = best_z;
  }
  // per coarse level, the time of both stencil modes; thresholds between level sizes
  std::vector<double> t_node(c->L, 0.0), t_dof(c->L, 0.0);
  const int64_t keep_split = c->split_nodes;
  for (int l = 1; l < c->L; ++l) {
    c->split_nodes = 0;  // every grid one thread per node
    t_node[l] = time_level_op(c, l, f32, EPI_CHEB, nrep);
    c->split_nodes = INT64_MAX;  // every grid one thread per DOF
    t_dof[l] = time_level_op(c, l, f32, EPI_CHEB, nrep);
  }
  int64_t best_split = keep_split;
  double best_sum = 1e300;
  std::vector<int64_t> cand = {0};
  for (int l = 1; l < c->L; ++l) cand.push_back(c->grids[l].nn() + 1);
  for (int64_t T : cand) {
    double sum = 0;
    for (int l = 1; l < c->L; ++l) sum += c->grids[l].nn() < T ? t_dof[l] : t_node[l];
    if (sum < best_sum) {
      best_sum = sum;
      best_split = T;
    }
  }
  c->split_nodes = c->L > 1 ? best_split : keep_split;
  if (zlen_out) *zlen_out = best_z;
  if (split_out) *split_out = c->split_nodes;
}

// One iteration of the solve loop for nvec right-hand sides with its convergence test on the
// device (the reduction of r.r freezes converged columns): the same kernels as solve_block's
// step and search-direction update, with nothing read back, so it can be captured in a graph.
void iteration_on_device(tofe_ctx *c, double *U, int nvec)
{
  const hipStream_t s = c->stream;
  const int64_t n = c->ndof;
  double *r = c->vr.p, *z = c->vz.p, *p = c->vp.p, *q = c->vq.p;
  const int nfb = level_blocks<double>(c, 0);
  EpiArgs<double> ep;
  ep.y = q; ep.partial = c->partial.p; ep.nvec = nvec;
  if (c->built_precision == TOFE_PREC_MIXED) {
    const FirstCheb first = vcycle_first_step(c);
    ep.b = r;
    ep.partial2 = c->partial2.p;
    level_apply<double>(c, 0, p, EPI_APPLY_DOT2, ep);
    if (c->rec) {
      // p.r and p.Ap (then alpha) by two small reductions instead of in every block of the
      // update: in a batched launch the update's blocks, each re-summing the partials, would
      // cost more than the update itself
      reduce_into(c, c->partial2.p, nfb, c->scal.p + SL_PR, POST_NONE, c->scal.p, nvec);
      reduce_into(c, c->partial.p, nfb, c->scal.p + SL_PQ, POST_ALPHA_FCG, c->scal.p, nvec);
      FcgXrLaunch d;
      d.r = r; d.q = q; d.n = n; d.pq_part = c->partial.p; d.pr_part = c->partial2.p;
      d.nparts = nfb; d.scal = c->scal.p; d.partial = c->partial3.p; d.fc = first;
      d.nblocks = fcg_update_r_blocks();
      d.pre = 1;
      rec_push(c, batch_key(BK_FCG_XR, false), d, d.nblocks);
    } else {
      fcg_update_r(s, r, q, n, c->partial.p, c->partial2.p, nfb, c->scal.p, c->partial3.p, first,
                   nvec);
    }
    reduce_into(c, c->partial3.p, kRedBlocks, c->scal.p + SL_RR, POST_CONV_FCG, c->scal.p, nvec);
    const float *z32 = nullptr;
    const int nparts = precondition(c, r, z, q, first.b32 != nullptr, nvec, &z32);
    if (c->rec) {
      reduce_into(c, c->partial.p, nparts, c->scal.p + SL_ZQ, POST_BETA_FCG, c->scal.p, nvec);
      FcgPLaunch d;
      d.p = p; d.z = z; d.z32 = z32; d.x = U; d.n = n; d.zq_part = c->partial.p; d.nparts = nparts;
      d.scal = c->scal.p;
      d.nblocks = fcg_update_xp_blocks(n);
      d.pre = 1;
      rec_push(c, batch_key(BK_FCG_P, false), d, d.nblocks);
    } else {
      fcg_update_xp(s, p, z, z32, U, n, c->partial.p, nparts, c->scal.p, nvec);
    }
  } else {
    level_apply<double>(c, 0, p, EPI_APPLY_DOT, ep);
    reduce_into(c, c->partial.p, nfb, c->scal.p + SL_PQ, POST_ALPHA_CG, c->scal.p, nvec);
    double *scal = c->scal.p, *part = c->partial.p;
    const FirstCheb64 first64 = vcycle_first_step64(c);
    issue(c, [=](hipStream_t st) { cg_update_r(st, r, q, n, scal, SL_ALPHA, part, nvec, first64); });
    reduce_into(c, c->partial.p, kRedBlocks, c->scal.p + SL_RR, POST_CONV_CG, c->scal.p, nvec);
    precondition(c, r, z, nullptr, first64.d != nullptr, nvec);
    dot_into(c, r, z, n, SL_TMP, POST_BETA_CG, nvec);
    issue(c, [=](hipStream_t st) { cg_update_xp(st, p, z, U, n, scal, SL_BETA, SL_ALPHA, nvec); });
  }
}

// Start of a solve tested on the device, queued on the context's stream: r = F - K U, ||F||,
// ||r||, read back (the host forms each column's tolerance and start-up reason), then the
// device test's state, z = M^-1 r, r.z and p = z.  Returns the columns still to solve.
int start_on_device(tofe_ctx *c, const double *F, double *U, int nvec, double rtol, int *reason)
{
  const hipStream_t s = c->stream;
  const int64_t n = c->ndof;
  double *r = c->vr.p, *z = c->vz.p, *p = c->vp.p;
  EpiArgs<double> er;
  er.y = r; er.b = F; er.nvec = nvec;
  level_apply<double>(c, 0, U, EPI_RESID, er);
  dot_into(c, F, F, n, SL_BB, POST_NONE, nvec);
  dot_into(c, r, r, n, SL_RR, POST_NONE, nvec);
  std::vector<double> h((size_t)SL_COUNT * nvec);
  scal_to_host_async(c, nvec);
  scal_wait(c, h.data(), nvec);
  ConvState cs{};
  unsigned long long done = 0;
  int live = 0;
  for (int v = 0; v < nvec; ++v) {
    const double bnorm = std::sqrt(h[(size_t)v * SL_COUNT + SL_BB]);
    const double rn = std::sqrt(h[(size_t)v * SL_COUNT + SL_RR]);
    const double rnorm0 = bnorm > 0 ? bnorm : rn;
    cs.ttol[v] = std::max(rtol * rnorm0, 1e-50);
    cs.conv[v] = (rtol * rnorm0 > 1e-50) ? TOFE_CONVERGED_RTOL : TOFE_CONVERGED_ATOL;
    reason[v] = 0;
    if (std::isnan(rn) || std::isinf(rn)) reason[v] = TOFE_DIVERGED_NANORINF;
    else if (rn <= cs.ttol[v]) reason[v] = cs.conv[v];
    if (reason[v]) done |= 1ull << v;
    else ++live;
  }
  set_conv_state(s, c->scal.p, nvec, cs, done);
  c->frozen_dirty = true;
  if (live) {
    precondition(c, r, z, nullptr, false, nvec);
    dot_into(c, r, z, n, SL_RZ, POST_NONE, nvec);
    TOFE_HIP(hipMemcpyAsync(p, z, n * nvec * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  return live;
}

void h2d(tofe_ctx *c, double *d, const double *h, size_t n)
{
  TOFE_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
}
void d2h(tofe_ctx *c, double *h, const double *d, size_t n)
{
  TOFE_HIP(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TOFE_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace

extern "C" {

int tofe_create(const tofe_desc *desc, tofe_ctx **out)
{
  if (!desc || !out) return TOFE_ERR_ARG;
  Range range("tofe_create");
  *out = nullptr;
  std::unique_ptr<tofe_ctx> c(new tofe_ctx());
  c->desc = *desc;
  tofe_desc &d = c->desc;
  if (d.nx < 1 || d.ny < 1 || d.nz < 1 || !(d.hx > 0) || !(d.hy > 0) || !(d.hz > 0) ||
      !(d.E0 > 0) || !(d.nu > -1 && d.nu < 0.5) || d.Emin < 0 || !(d.rmin > 0)) {
    return TOFE_ERR_ARG;
  }
  if (d.nu_pre <= 0 && d.nu_post <= 0) d.nu_pre = d.nu_post = 2;
  if (d.nu_pre < 0) d.nu_pre = 0;
  if (d.nu_post < 0) d.nu_post = 0;
  if (!(d.cheb_hi > 0)) d.cheb_hi = 1.1;
  if (!(d.cheb_lo > 0)) d.cheb_lo = 0.1;
  if (d.eig_its <= 0) d.eig_its = 10;
  if (d.eig_its > kMaxEig || !(d.cheb_lo < d.cheb_hi)) return TOFE_ERR_ARG;
  if (d.penal == 0) d.penal = 3.0;
  if (d.vd_penal == 0) d.vd_penal = 1.0;
  if (d.interpolation < TOFE_INTERP_SIMP || d.interpolation > TOFE_INTERP_SIMP_LOGISTIC)
    return TOFE_ERR_ARG;
  if (d.vd_min > 0 && !(d.rmax > 0)) return TOFE_ERR_ARG;
  if (d.rmax < 0) return TOFE_ERR_ARG;
  c->sp.penal = d.penal;
  c->sp.eps = d.Emin;
  c->sp.vd_penal = d.vd_penal;
  c->sp.vd_min = d.vd_min > 0 ? d.vd_min : -1.0;
  c->sp.interp = d.interpolation;
  c->sp.p0 = d.interp_param[0];
  c->sp.p1 = d.interp_param[1];
  c->has_vd = d.vd_min > 0;
  d.fixed_dof = nullptr;
  c->precision = (desc->precision == TOFE_PREC_MIXED) ? TOFE_PREC_MIXED : TOFE_PREC_FP64;
  try {
    TOFE_HIP(hipGetDevice(&c->device));
    if (const char *fm = std::getenv("TOFE_FINE_MFMA")) c->fine_mfma = fm[0] == '1';
    if (const char *sn = std::getenv("TOFE_STENCIL_SPLIT_NODES")) c->split_nodes = std::atoll(sn);
    // level count: floor(log2(min nel)) unless given (RecMesh.cpp:552-554), truncated
    // once a coarse grid has <= 10 DOFs (Interpolation.cpp:70-71)
    int L = d.mg_levels;
    if (L <= 0) L = (int)std::floor(std::log2((double)std::min(d.nx, std::min(d.ny, d.nz))));
    if (L < 1) L = 1;
    if (L > 16) L = 16;
    Grid g;
    g.Nx = d.nx + 1; g.Ny = d.ny + 1; g.Nz = d.nz + 1;
    c->grids.push_back(g);
    for (int i = 0; i + 1 < L; ++i) {
      Grid gc;
      gc.Nx = (g.Nx + 1) / 2; gc.Ny = (g.Ny + 1) / 2; gc.Nz = (g.Nz + 1) / 2;
      c->grids.push_back(gc);
      g = gc;
      if (3 * gc.nn() <= 10) break;
    }
    c->L = (int)c->grids.size();
    d.mg_levels = c->L;
    c->ne = (int64_t)d.nx * d.ny * d.nz;
    c->nn = c->grids[0].nn();
    c->ndof = 3 * c->nn;
    // kernels index nodes in 32-bit arithmetic
    if (c->nn >= (int64_t(1) << 31))
      throw Error(TOFE_ERR_ARG, "mesh has 2^31 or more nodes");
    // dense direct solve on the coarsest level: n^2 doubles, half-bandwidth below 31 tiles
    const Grid &gc = c->grids[c->L - 1];
    if (gc.ndof() > 20000 || 3 * (gc.Nx * gc.Ny + gc.Nx + 1) + 2 > 31 * kCholNB - 63)
      throw Error(TOFE_ERR_ARG, "coarsest level has " + std::to_string(gc.ndof()) +
                                    " DOFs (" + std::to_string(gc.Nx) + " x " + std::to_string(gc.Ny) +
                                    " x " + std::to_string(gc.Nz) +
                                    " nodes); use more multigrid levels");

    TOFE_HIP(hipStreamCreate(&c->stream));
    {
      int ncu = 0;
      TOFE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
      c->fine_tile = fine_tile_choose(c->desc.nx, c->desc.ny, c->desc.nz, ncu, 3);
      c->fine_tile32 = fine_tile_choose(c->desc.nx, c->desc.ny, c->desc.nz, ncu, 6);
      c->wht_zlen = fine_wht_zlen(c->desc.nx, c->desc.ny, c->desc.nz, ncu, 3);
      c->wht_zlen32 = fine_wht_zlen(c->desc.nx, c->desc.ny, c->desc.nz, ncu, 5);
      if (const char *wz = std::getenv("TOFE_FINE_WHT_Z")) {
        int z64 = 0, z32 = 0;
        if (std::sscanf(wz, "%d:%d", &z64, &z32) != 2 || z64 < 1 || z64 > 64 || z32 < 1 || z32 > 64)
          throw Error(TOFE_ERR_ARG, std::string("TOFE_FINE_WHT_Z must be <zlen fp64>:<zlen fp32>, got ") + wz);
        c->wht_zlen = z64;
        c->wht_zlen32 = z32;
      }
      if (const char *ft = std::getenv("TOFE_FINE_TILE")) {
        int t = -1, z = 0;
        if (std::sscanf(ft, "%d:%d", &t, &z) != 2 || t < 0 || t >= kFineTiles || z < 1 || z > 64)
          throw Error(TOFE_ERR_ARG, std::string("TOFE_FINE_TILE must be <shape 0..") +
                                        std::to_string(kFineTiles - 1) + ">:<zlen 1..64>, got " + ft);
        c->fine_tile = c->fine_tile32 = FineTile{t, z};
      }
      if (ncu >= 4 * kCoarseCUs) {
        const MaskedStreams m = masked_streams(c->device, ncu, kCoarseCUs);
        c->stream2 = m.full;
        c->stream3 = m.a;
        c->stream4 = m.b;
        c->shared_helpers = true;
      } else {
        TOFE_HIP(hipStreamCreate(&c->stream2));
        TOFE_HIP(hipStreamCreate(&c->stream3));
        TOFE_HIP(hipStreamCreate(&c->stream4));
      }
    }
    for (hipEvent_t *e : {&c->ev_dense, &c->ev_inv, &c->ev_ops, &c->ev_lz, &c->ev_scal, &c->ev_lz0,
                          &c->ev_diag})
      TOFE_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    TOFE_HIP(hipHostMalloc((void **)&c->hscal, SL_COUNT * sizeof(double), hipHostMallocMapped));
    TOFE_HIP(hipHostGetDevicePointer((void **)&c->hscal_dev, c->hscal, 0));

    compute_ke(d.E0, d.nu, d.hx, d.hy, d.hz, c->KE);
    dmatrix(d.E0, d.nu, c->shape.D);
    shape_gradients(d.hx, d.hy, d.hz, c->shape.dN, c->shape.detJ);
    // KE, then (cubic elements) the values of its 12 orbits under the symmetries of the
    // cube, used by the fine operator when every entry matches its orbit value up to the
    // rounding of the quadrature.
    double ke64[kKeBufLen] = {};
    std::copy(c->KE, c->KE + 576, ke64);
    {
      double kmax = 0, dev = 0;
      for (int k = 0; k < kKeCube.n; ++k) ke64[576 + k] = c->KE[24 * kKeCube.rep[k][0] + kKeCube.rep[k][1]];
      for (int x = 0; x < 24; ++x)
        for (int y = 0; y < 24; ++y) {
          const double v = ke64[576 + kKeCube.idx[x][y]];
          kmax = std::max(kmax, std::fabs(c->KE[24 * x + y]));
          dev = std::max(dev, std::fabs(c->KE[24 * x + y] - (kKeCube.neg[x][y] ? -v : v)));
        }
      c->ke_sym = dev <= 1e-14 * kmax;
      if (const char *fs = std::getenv("TOFE_FINE_SYM")) c->ke_sym = c->ke_sym && fs[0] != '0';
    }
    // Walsh-Hadamard blocks B_t[i][j] = (H KE H)_((t ^ e_i, i), (t ^ e_j, j)) / 64 over the node
    // coordinate index m = x + 2y + 4z; k_fine_wht is used when every entry of H KE H / 64
    // outside the stored blocks vanishes and the stored ones are symmetric, to rounding.
    static_assert(kWhtOffset == 576 + kKeOrbitsMax, "KE buffer layout");
    {
      int mo[8];
      for (int l = 0; l < 8; ++l) mo[l] = ko_coord(l, 0) + 2 * ko_coord(l, 1) + 4 * ko_coord(l, 2);
      auto h = [](int s, int m) { return (__builtin_popcount(s & m) & 1) ? -1.0 : 1.0; };
      double M[24][24];
      double mmax = 0;
      for (int s1 = 0; s1 < 8; ++s1)
        for (int i = 0; i < 3; ++i)
          for (int s2 = 0; s2 < 8; ++s2)
            for (int j = 0; j < 3; ++j) {
              double acc = 0;
              for (int l1 = 0; l1 < 8; ++l1)
                for (int l2 = 0; l2 < 8; ++l2)
                  acc += h(s1, mo[l1]) * c->KE[24 * (3 * l1 + i) + 3 * l2 + j] * h(s2, mo[l2]);
              M[3 * s1 + i][3 * s2 + j] = acc / 64;
              mmax = std::max(mmax, std::fabs(acc / 64));
            }
      for (int t = 0; t < 8; ++t)
        for (int i = 0; i < 3; ++i)
          for (int j = i; j < 3; ++j)
            if (wht_index(t, i, j) >= 0)
              ke64[kWhtOffset + wht_index(t, i, j)] = M[3 * (t ^ (1 << i)) + i][3 * (t ^ (1 << j)) + j];
      double dev = 0;
      for (int s1 = 0; s1 < 8; ++s1)
        for (int i = 0; i < 3; ++i)
          for (int s2 = 0; s2 < 8; ++s2)
            for (int j = 0; j < 3; ++j) {
              const int t = s1 ^ (1 << i);
              const int k = (t == (s2 ^ (1 << j))) ? wht_index(t, i, j) : -1;
              const double want = k >= 0 ? ke64[kWhtOffset + k] : 0.0;
              dev = std::max(dev, std::fabs(M[3 * s1 + i][3 * s2 + j] - want));
            }
      c->ke_wht = dev <= 1e-13 * mmax;
      if (const char *fw = std::getenv("TOFE_FINE_WHT")) c->ke_wht = c->ke_wht && fw[0] != '0';
      // cubic element: the values of each group agree to rounding; the group value is the
      // first member's
      double gdev = 0;
      for (int k = 0; k < kWhtVals; ++k) {
        int first = 0;
        while (kWhtCubeGroup[first] != kWhtCubeGroup[k]) ++first;
        ke64[kWhtCubeOffset + kWhtCubeGroup[k]] = ke64[kWhtOffset + first];
        gdev = std::max(gdev, std::fabs(ke64[kWhtOffset + k] - ke64[kWhtOffset + first]));
      }
      c->ke_wht_cube = c->ke_wht && gdev <= 1e-13 * mmax;
      if (const char *fc = std::getenv("TOFE_FINE_WHT_CUBE")) c->ke_wht_cube = c->ke_wht_cube && fc[0] != '0';
    }
    constexpr int kKeBuf = kKeBufLen;
    float ke32[kKeBuf];
    for (int i = 0; i < kKeBuf; ++i) ke32[i] = (float)ke64[i];
    c->dKE.alloc(kKeBuf);
    c->dKE32.alloc(kKeBuf);
    h2d(c.get(), c->dKE.p, ke64, kKeBuf);
    TOFE_HIP(hipMemcpyAsync(c->dKE32.p, ke32, sizeof(ke32), hipMemcpyHostToDevice, c->stream));
    TOFE_HIP(hipStreamSynchronize(c->stream));  // the host copies are stack arrays

    if (filter_template(d.hx, d.hy, d.hz, d.rmin, c->tpl_off, c->tpl_w) < 1 || c->tpl_w.size() > 1024)
      throw Error(TOFE_ERR_ARG, "filter radius gives an unsupported neighbourhood");
    c->dtpl_off.alloc(c->tpl_off.size());
    c->dtpl_w.alloc(c->tpl_w.size());
    TOFE_HIP(hipMemcpyAsync(c->dtpl_off.p, c->tpl_off.data(), c->tpl_off.size() * sizeof(int),
                            hipMemcpyHostToDevice, c->stream));
    h2d(c.get(), c->dtpl_w.p, c->tpl_w.data(), c->tpl_w.size());
    c->filt_inv.alloc(c->ne);
    filter_rowsum_inv(c->stream, filter_tpl(c.get()), c->filt_inv.p);
    c->y.alloc(c->ne);
    c->rhoq.alloc(c->ne);
    if (d.rmax > 0) {
      // R: the same neighbourhood rule with radius Rmax and unit weights (RecMesh.cpp:544,
      // 706-718); REdge = max_i rowsum_i - rowsum (Material_Interpolation.cpp:37-41)
      std::vector<double> rw;
      if (filter_template(d.hx, d.hy, d.hz, d.rmax, c->rtpl_off, rw) < 1 || rw.size() > 1024)
        throw Error(TOFE_ERR_ARG, "maximum length scale radius gives an unsupported neighbourhood");
      std::fill(rw.begin(), rw.end(), 1.0);
      c->drtpl_off.alloc(c->rtpl_off.size());
      c->drtpl_w.alloc(rw.size());
      TOFE_HIP(hipMemcpyAsync(c->drtpl_off.p, c->rtpl_off.data(), c->rtpl_off.size() * sizeof(int),
                              hipMemcpyHostToDevice, c->stream));
      h2d(c.get(), c->drtpl_w.p, rw.data(), rw.size());
      DBuf<int64_t> cnt;
      cnt.alloc(c->ne);
      filter_count(c->stream, rfilter_tpl(c.get()), cnt.p);
      std::vector<int64_t> hc(c->ne);
      TOFE_HIP(hipMemcpyAsync(hc.data(), cnt.p, c->ne * sizeof(int64_t), hipMemcpyDeviceToHost,
                              c->stream));
      TOFE_HIP(hipStreamSynchronize(c->stream));
      int64_t mx = 0;
      for (int64_t v : hc) mx = std::max(mx, v);
      std::vector<double> re(c->ne);
      for (int64_t i = 0; i < c->ne; ++i) re[i] = -1 * (double)hc[i] + (double)mx;
      c->REdge.alloc(c->ne);
      h2d(c.get(), c->REdge.p, re.data(), c->ne);
      c->ebuf3.alloc(c->ne);
    }

    c->dfix.alloc(c->nn);
    if (desc->fixed_dof) {
      std::vector<uint8_t> m(c->nn, 0);
      for (int64_t n = 0; n < c->nn; ++n)
        for (int a = 0; a < 3; ++a)
          if (desc->fixed_dof[3 * n + a]) m[n] |= (uint8_t)(1u << a);
      TOFE_HIP(hipMemcpyAsync(c->dfix.p, m.data(), c->nn, hipMemcpyHostToDevice, c->stream));
      TOFE_HIP(hipStreamSynchronize(c->stream));
      c->has_fix = true;
      direct_setup(c.get(), m.data());
    } else {
      c->dfix.zero(c->stream);
      direct_setup(c.get(), nullptr);
    }

    {
      // interpolation forms and the elements that touch a Dirichlet DOF
      std::vector<double> Qf(27 * 64), KEC(27 * 576);
      galerkin_forms(c->KE, Qf.data(), KEC.data());
      c->dQforms.alloc(Qf.size());
      c->dKEC.alloc(KEC.size());
      h2d(c.get(), c->dQforms.p, Qf.data(), Qf.size());
      h2d(c.get(), c->dKEC.p, KEC.data(), KEC.size());
      std::vector<int> emap(c->ne, -1), mform;
      std::vector<int64_t> melem;
      std::vector<uint32_t> mmask;
      if (desc->fixed_dof) {
        const int nx = d.nx, ny = d.ny, nz = d.nz;
        const int64_t Nx = nx + 1, Ny = ny + 1;
        auto form = [](int ex, int nel) { return (ex & 1) ? 1 : (((nel + 1) % 2 == 0 && ex == nel - 1) ? 2 : 0); };
        static const int LN[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
        for (int ez = 0; ez < nz; ++ez)
          for (int ey = 0; ey < ny; ++ey)
            for (int ex = 0; ex < nx; ++ex) {
              uint32_t mask = 0;
              for (int l = 0; l < 8; ++l) {
                const int64_t n = (ex + LN[l][0]) + Nx * ((ey + LN[l][1]) + Ny * (ez + LN[l][2]));
                for (int a = 0; a < 3; ++a)
                  if (desc->fixed_dof[3 * n + a]) mask |= 1u << (3 * l + a);
              }
              if (!mask) continue;
              const int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
              emap[e] = (int)melem.size();
              melem.push_back(e);
              mform.push_back(form(ex, nx) + 3 * form(ey, ny) + 9 * form(ez, nz));
              mmask.push_back(mask);
            }
      }
      c->n_masked = (int)melem.size();
      c->demap.alloc(c->ne);
      TOFE_HIP(hipMemcpyAsync(c->demap.p, emap.data(), c->ne * sizeof(int), hipMemcpyHostToDevice, c->stream));
      if (c->n_masked) {
        c->dme_elem.alloc(melem.size());
        c->dme_form.alloc(mform.size());
        c->dme_mask.alloc(mmask.size());
        c->dKCm.alloc((size_t)576 * melem.size());
        TOFE_HIP(hipMemcpyAsync(c->dme_elem.p, melem.data(), melem.size() * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
        TOFE_HIP(hipMemcpyAsync(c->dme_form.p, mform.data(), mform.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        TOFE_HIP(hipMemcpyAsync(c->dme_mask.p, mmask.data(), mmask.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
      }
      TOFE_HIP(hipStreamSynchronize(c->stream));
    }
    for (auto *b : {&c->x, &c->rho, &c->E, &c->dEdz, &c->Es, &c->dEsdz, &c->ebuf, &c->ebuf2})
      b->alloc(c->ne);
    c->E32.alloc(c->ne);
    for (int l = 0; l < c->L; ++l) {
      c->lv64.emplace_back(new Level<double>());
      alloc_level(c.get(), *c->lv64.back(), l);
    }
    for (auto *b : {&c->vr, &c->vz, &c->vp, &c->vq, &c->vF, &c->vU}) b->alloc(c->ndof);
    int np = kRedBlocks;
    for (int l = 0; l < c->L; ++l) {
      np = std::max(np, level_blocks<double>(c.get(), l));
      np = std::max(np, level_blocks<float>(c.get(), l));
    }
    c->npart = np;
    c->partial.alloc(np);
    c->partial2.alloc(np);
    c->partial3.alloc(np);
    c->scal.alloc(SL_COUNT);
    c->scal.zero(c->stream);
    c->partial_b.alloc(np);
    c->partial2_b.alloc(np);
    c->scal_b.alloc(SL_COUNT);
    c->scal_b.zero(c->stream);
    c->partial_c.alloc(np);
    c->partial2_c.alloc(np);
    c->scal_c.alloc(SL_COUNT);
    c->scal_c.zero(c->stream);
    TOFE_HIP(hipStreamSynchronize(c->stream));
    c->stats.levels = c->L;
    c->stats.coarse_size = c->grids[c->L - 1].ndof();
  } catch (const Error &e) {
    return e.code;
  } catch (...) {
    return TOFE_ERR_HIP;
  }
  *out = c.release();
  return TOFE_OK;
}

int tofe_destroy(tofe_ctx *c)
{
  if (!c) return TOFE_OK;
  Range range("tofe_destroy");
  (void)hipSetDevice(c->device);
  delete c;  // ~tofe_ctx synchronises and releases streams, events and handles
  return TOFE_OK;
}

const char *tofe_last_error(const tofe_ctx *c)
{
  // without a context: the message of this thread's last failing context-free call
  if (!c) {
    const char *m = noalloc_last_error();
    return (m && *m) ? m : "null context";
  }
  return c->err.c_str();
}

int tofe_sizes(const tofe_ctx *c, int64_t *nelem, int64_t *nnode, int64_t *ndof, int *levels)
{
  if (!c) return TOFE_ERR_ARG;
  if (nelem) *nelem = c->ne;
  if (nnode) *nnode = c->nn;
  if (ndof) *ndof = c->ndof;
  if (levels) *levels = c->L;
  return TOFE_OK;
}

int tofe_get_ke(const tofe_ctx *c, double *ke)
{
  if (!c || !ke) return TOFE_ERR_ARG;
  std::memcpy(ke, c->KE, sizeof(c->KE));
  return TOFE_OK;
}

int tofe_set_design(tofe_ctx *c, const double *x)
{
  return guarded(c, "tofe_set_design", [&] {
    if (!x) throw Error(TOFE_ERR_ARG, "null design");
    h2d(c, c->x.p, x, c->ne);
    upload_design(c, c->x.p);
  });
}

int tofe_set_design_dev(tofe_ctx *c, const double *dx)
{
  return guarded(c, "tofe_set_design_dev", [&] {
    if (!dx) throw Error(TOFE_ERR_ARG, "null design");
    upload_design(c, dx);
  });
}

int tofe_set_E(tofe_ctx *c, const double *E)
{
  return guarded(c, "tofe_set_E", [&] {
    if (!E) throw Error(TOFE_ERR_ARG, "null E");
    h2d(c, c->E.p, E, c->ne);
    convert<double, float>(c->stream, c->E.p, c->E32.p, c->ne);
    TOFE_HIP(hipStreamSynchronize(c->stream));
    c->design_set = true;
    c->assembled = false;
  });
}

int tofe_get_field(const tofe_ctx *cc, int field, double *out)
{
  tofe_ctx *c = const_cast<tofe_ctx *>(cc);
  return guarded(c, "tofe_get_field", [&] {
    if (!out) throw Error(TOFE_ERR_ARG, "null output");
    const double *src = nullptr;
    int64_t n = c->ne;
    switch (field) {
      case TOFE_FIELD_X: src = c->x.p; break;
      case TOFE_FIELD_RHO: src = c->rho.p; break;
      case TOFE_FIELD_V: src = c->rho.p; break;
      case TOFE_FIELD_E: src = c->E.p; break;
      case TOFE_FIELD_DEDZ: src = c->dEdz.p; break;
      case TOFE_FIELD_ES: src = c->Es.p; break;
      case TOFE_FIELD_DESDZ: src = c->dEsdz.p; break;
      case TOFE_FIELD_Y: src = c->y.p; break;
      case TOFE_FIELD_MDIAG:
        if (!c->has_mass) throw Error(TOFE_ERR_STATE, "no mass matrix (call tofe_assemble_mass)");
        src = c->Mdiag.p;
        n = c->ndof;
        break;
      case TOFE_FIELD_RHOQ: src = c->rhoq.p; break;
      case TOFE_FIELD_REDGE:
        if (!c->REdge.p) throw Error(TOFE_ERR_STATE, "no maximum-length-scale filter (rmax <= 0)");
        src = c->REdge.p;
        break;
      case TOFE_FIELD_DIAG:
        if (!c->assembled) throw Error(TOFE_ERR_STATE, "diagonal needs tofe_assemble");
        src = c->diag0.p;
        n = c->ndof;
        break;
      default: throw Error(TOFE_ERR_ARG, "unknown field");
    }
    d2h(c, out, src, n);
  });
}

int tofe_assemble(tofe_ctx *c)
{
  return guarded(c, "tofe_assemble", [&] {
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    assemble(c);
  });
}

int tofe_solve_dev(tofe_ctx *c, const double *dF, double *dU, int nrhs, double rtol, int maxit,
                   int *its, int *reason)
{
  return guarded(
      c, "tofe_solve_dev", [&] { solve_dev(c, dF, dU, nrhs, rtol, maxit, its, reason); }, false);
}

int tofe_solve(tofe_ctx *c, const double *F, double *U, int nrhs, double rtol, int maxit,
               int *its, int *reason)
{
  return guarded(c, "tofe_solve", [&] {
    if (!F || !U || nrhs < 1) throw Error(TOFE_ERR_ARG, "bad solve arguments");
    need_assembled(c);
    ensure_nvec(c, std::min(nrhs, kMaxBlockVec));
    int kmax = 0, rs_out = 0;
    for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec) {
      const int nv = std::min(kMaxBlockVec, nrhs - i0);
      const size_t off = (size_t)i0 * c->ndof, len = (size_t)nv * c->ndof;
      h2d(c, c->vF.p, F + off, len);
      h2d(c, c->vU.p, U + off, len);
      int k = 0, rs = 0;
      solve_dev(c, c->vF.p, c->vU.p, nv, rtol, maxit, &k, &rs);
      d2h(c, U + off, c->vU.p, len);
      kmax = std::max(kmax, k);
      if (rs_out >= 0) rs_out = rs;
    }
    c->stats.its = kmax;
    c->stats.reason = rs_out;
    if (its) *its = kmax;
    if (reason) *reason = rs_out;
  });
}

int tofe_apply_K_dev(tofe_ctx *c, const double *dX, double *dY, int nrhs)
{
  return guarded(c, "tofe_apply_K_dev", [&] {
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    if (!dX || !dY || nrhs < 1) throw Error(TOFE_ERR_ARG, "bad apply arguments");
    for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec) {  // one launch per chunk of vectors
      EpiArgs<double> ep;
      ep.y = dY + (size_t)i0 * c->ndof;
      ep.nvec = std::min(kMaxBlockVec, nrhs - i0);
      level_apply<double>(c, 0, dX + (size_t)i0 * c->ndof, EPI_APPLY, ep);
    }
    TOFE_HIP(hipStreamSynchronize(c->stream));
  });
}

int tofe_apply_K(tofe_ctx *c, const double *X, double *Y, int nrhs)
{
  return guarded(c, "tofe_apply_K", [&] {
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    if (!X || !Y || nrhs < 1) throw Error(TOFE_ERR_ARG, "bad apply arguments");
    ensure_nvec(c, std::min(nrhs, kMaxBlockVec));
    for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec) {
      const int nv = std::min(kMaxBlockVec, nrhs - i0);
      const size_t off = (size_t)i0 * c->ndof, len = (size_t)nv * c->ndof;
      h2d(c, c->vp.p, X + off, len);
      EpiArgs<double> ep;
      ep.y = c->vq.p;
      ep.nvec = nv;
      level_apply<double>(c, 0, c->vp.p, EPI_APPLY, ep);
      d2h(c, Y + off, c->vq.p, len);
    }
  });
}

int tofe_assemble_mass(tofe_ctx *c, double density, const double *mlump)
{
  return guarded(c, "tofe_assemble_mass", [&] {
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    if (!(density > 0)) throw Error(TOFE_ERR_ARG, "density must be positive");
    // elemSize = 1 * dx * dy * dz (RecMesh.cpp:529-533); m0 = 1/2^3/3 * elemSize * density
    double elem = 1.0;
    elem *= c->desc.hx;
    elem *= c->desc.hy;
    elem *= c->desc.hz;
    const double m0 = 1.0 / std::pow(2.0, 3) / 3 * elem * density;
    c->mass_m0 = m0;
    if (c->mass_form == TOFE_MASS_CONSISTENT) {
      c->mV.alloc(c->ne);
      TOFE_HIP(hipMemcpyAsync(c->mV.p, c->rho.p, c->ne * sizeof(double), hipMemcpyDeviceToDevice,
                              c->stream));
      c->has_mlump = mlump != nullptr;
      if (mlump) {
        c->mlump.alloc(c->ndof);
        h2d(c, c->mlump.p, mlump, c->ndof);
      }
    } else {
      if (!c->Mdiag.p) c->Mdiag.alloc(c->ndof);
      const double *dl = nullptr;
      if (mlump) {
        h2d(c, c->vq.p, mlump, c->ndof);
        dl = c->vq.p;
      }
      mass_diag(c->stream, c->rho.p, c->desc.nx, c->desc.ny, c->desc.nz, m0, dl,
                c->has_fix ? c->dfix.p : nullptr, c->Mdiag.p);
    }
    TOFE_HIP(hipStreamSynchronize(c->stream));
    c->has_mass = true;
  });
}

int tofe_set_springs(tofe_ctx *c, const double *spk)
{
  return guarded(c, "tofe_set_springs", [&] {
    c->assembled = false;  // the diagonal and the Galerkin hierarchy change
    if (!spk) {
      c->has_springs = false;
      if (c->direct)
        TOFE_HIP(hipMemcpy(c->direct_cconst.p, c->direct_cfix.data(),
                           c->direct_cfix.size() * sizeof(double), hipMemcpyHostToDevice));
      return;
    }
    std::vector<double> h(spk, spk + c->ndof);
    if (c->has_fix) {  // MatZeroRowsColumns after MatDiagonalSet: nothing stays on fixed rows
      std::vector<uint8_t> fm(c->nn);
      TOFE_HIP(hipMemcpy(fm.data(), c->dfix.p, c->nn, hipMemcpyDeviceToHost));
      for (int64_t d = 0; d < c->ndof; ++d)
        if ((fm[d / 3] >> (d % 3)) & 1) h[d] = 0.0;
    }
    c->spk.alloc(c->ndof);
    h2d(c, c->spk.p, h.data(), c->ndof);
    TOFE_HIP(hipStreamSynchronize(c->stream));
    c->has_springs = true;
    if (c->direct) {  // the direct coarsest operator's constant part: fixed diagonal + springs
      std::vector<double> cc = c->direct_cfix;
      direct_coarse_add_springs(c->desc.nx, c->desc.ny, c->desc.nz, 1 << (c->L - 1), h.data(), cc.data());
      TOFE_HIP(hipMemcpy(c->direct_cconst.p, cc.data(), cc.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  });
}

int tofe_set_mass_form(tofe_ctx *c, int form)
{
  return guarded(c, "tofe_set_mass_form", [&] {
    if (form != TOFE_MASS_LUMPED && form != TOFE_MASS_CONSISTENT)
      throw Error(TOFE_ERR_ARG, "unknown mass form");
    if (form != c->mass_form) c->has_mass = false;  // assemble the mass again
    c->mass_form = form;
  });
}

int tofe_apply_M_dev(tofe_ctx *c, const double *dX, double *dY, int nrhs)
{
  return guarded(c, "tofe_apply_M_dev", [&] {
    if (!c->has_mass) throw Error(TOFE_ERR_STATE, "no mass matrix (call tofe_assemble_mass)");
    if (!dX || !dY || nrhs < 1) throw Error(TOFE_ERR_ARG, "bad apply arguments");
    apply_mass(c, dX, dY, nrhs);
    TOFE_HIP(hipStreamSynchronize(c->stream));
  });
}

int tofe_apply_M(tofe_ctx *c, const double *X, double *Y, int nrhs)
{
  return guarded(c, "tofe_apply_M", [&] {
    if (!c->has_mass) throw Error(TOFE_ERR_STATE, "no mass matrix (call tofe_assemble_mass)");
    if (!X || !Y || nrhs < 1) throw Error(TOFE_ERR_ARG, "bad apply arguments");
    ensure_nvec(c, std::min(nrhs, kMaxBlockVec));
    for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec) {
      const int nv = std::min(kMaxBlockVec, nrhs - i0);
      const size_t off = (size_t)i0 * c->ndof, len = (size_t)nv * c->ndof;
      h2d(c, c->vp.p, X + off, len);
      apply_mass(c, c->vp.p, c->vq.p, nv);
      d2h(c, Y + off, c->vq.p, len);
    }
  });
}

int tofe_element_quadratic(tofe_ctx *c, const double *phi, int nvec, double *kq, double *mq)
{
  return guarded(c, "tofe_element_quadratic", [&] {
    if (!phi || !kq || nvec < 1) throw Error(TOFE_ERR_ARG, "bad arguments");
    const size_t nd = (size_t)c->ndof * nvec, ne = (size_t)c->ne * nvec;
    DBuf<double> dphi, dk, dm;
    dphi.alloc(nd);
    dk.alloc(ne);
    if (mq) dm.alloc(ne);
    h2d(c, dphi.p, phi, nd);
    elem_quad(c->stream, dphi.p, c->dKE.p, c->desc.nx, c->desc.ny, c->desc.nz, c->ndof, nvec, dk.p,
              mq ? dm.p : nullptr, c->mass_form == TOFE_MASS_CONSISTENT);
    d2h(c, kq, dk.p, ne);
    if (mq) d2h(c, mq, dm.p, ne);
  });
}

int tofe_set_stress_dev(tofe_ctx *c, const double *dU)
{
  return guarded(c, "tofe_set_stress_dev", [&] {
    if (!dU) throw Error(TOFE_ERR_ARG, "null displacement");
    if (!c->Hs.p) c->Hs.alloc((size_t)64 * c->ne);
    stress_h(c->stream, dU, c->desc.nx, c->desc.ny, c->desc.nz, c->shape, c->Hs.p);
    TOFE_HIP(hipStreamSynchronize(c->stream));
    c->has_stress = true;
  });
}

int tofe_set_stress(tofe_ctx *c, const double *U)
{
  return guarded(c, "tofe_set_stress", [&] {
    if (!U) throw Error(TOFE_ERR_ARG, "null displacement");
    h2d(c, c->vp.p, U, c->ndof);
    if (!c->Hs.p) c->Hs.alloc((size_t)64 * c->ne);
    stress_h(c->stream, c->vp.p, c->desc.nx, c->desc.ny, c->desc.nz, c->shape, c->Hs.p);
    TOFE_HIP(hipStreamSynchronize(c->stream));
    c->has_stress = true;
  });
}

int tofe_apply_Ks_dev(tofe_ctx *c, const double *dX, double *dY, int nrhs)
{
  return guarded(c, "tofe_apply_Ks_dev", [&] {
    if (!c->has_stress) throw Error(TOFE_ERR_STATE, "no stress state (call tofe_set_stress)");
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    if (!dX || !dY || nrhs < 1) throw Error(TOFE_ERR_ARG, "bad apply arguments");
    for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec)
      apply_ks(c->stream, c->Hs.p, c->Es.p, c->has_fix ? c->dfix.p : nullptr, c->desc.nx,
               c->desc.ny, c->desc.nz, dX + (size_t)i0 * c->ndof, dY + (size_t)i0 * c->ndof,
               std::min(kMaxBlockVec, nrhs - i0));
    TOFE_HIP(hipStreamSynchronize(c->stream));
  });
}

int tofe_apply_Ks(tofe_ctx *c, const double *X, double *Y, int nrhs)
{
  return guarded(c, "tofe_apply_Ks", [&] {
    if (!c->has_stress) throw Error(TOFE_ERR_STATE, "no stress state (call tofe_set_stress)");
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    if (!X || !Y || nrhs < 1) throw Error(TOFE_ERR_ARG, "bad apply arguments");
    ensure_nvec(c, std::min(nrhs, kMaxBlockVec));
    for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec) {
      const int nv = std::min(kMaxBlockVec, nrhs - i0);
      const size_t off = (size_t)i0 * c->ndof, len = (size_t)nv * c->ndof;
      h2d(c, c->vp.p, X + off, len);
      apply_ks(c->stream, c->Hs.p, c->Es.p, c->has_fix ? c->dfix.p : nullptr, c->desc.nx,
               c->desc.ny, c->desc.nz, c->vp.p, c->vq.p, nv);
      d2h(c, Y + off, c->vq.p, len);
    }
  });
}

int tofe_stability_terms(tofe_ctx *c, const double *phi, int nvec, double *pksp, double *dksdu)
{
  return guarded(c, "tofe_stability_terms", [&] {
    if (!phi || nvec < 1) throw Error(TOFE_ERR_ARG, "bad arguments");
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    const size_t nd = (size_t)c->ndof * nvec, ne = (size_t)c->ne * nvec;
    DBuf<double> dphi;
    dphi.alloc(nd);
    h2d(c, dphi.p, phi, nd);
    if (pksp) {
      if (!c->has_stress) throw Error(TOFE_ERR_STATE, "no stress state (call tofe_set_stress)");
      DBuf<double> dq;
      dq.alloc(ne);
      ks_quad(c->stream, c->Hs.p, dphi.p, c->desc.nx, c->desc.ny, c->desc.nz, c->ndof, nvec, dq.p);
      d2h(c, pksp, dq.p, ne);
    }
    if (dksdu) {
      DBuf<double> dr;
      dr.alloc(nd);
      for (int j = 0; j < nvec; ++j)
        stab_rhs(c->stream, dphi.p + (size_t)j * c->ndof, c->Es.p, c->has_fix ? c->dfix.p : nullptr,
                 c->desc.nx, c->desc.ny, c->desc.nz, c->shape, dr.p + (size_t)j * c->ndof);
      d2h(c, dksdu, dr.p, nd);
    }
  });
}

int tofe_element_bilinear(tofe_ctx *c, const double *A, int nvec, const double *B, double *out)
{
  return guarded(c, "tofe_element_bilinear", [&] {
    if (!A || !B || !out || nvec < 1) throw Error(TOFE_ERR_ARG, "bad arguments");
    const size_t nd = (size_t)c->ndof * nvec, ne = (size_t)c->ne * nvec;
    DBuf<double> da, dout;
    da.alloc(nd);
    dout.alloc(ne);
    h2d(c, da.p, A, nd);
    h2d(c, c->vq.p, B, c->ndof);
    elem_bilinear(c->stream, da.p, c->vq.p, c->dKE.p, c->desc.nx, c->desc.ny, c->desc.nz, c->ndof,
                  nvec, dout.p);
    d2h(c, out, dout.p, ne);
  });
}

int tofe_pc_apply_dev(tofe_ctx *c, const double *dR, double *dZ, int nrhs)
{
  return guarded(c, "tofe_pc_apply_dev", [&] {
    need_assembled(c);
    if (!dR || !dZ || nrhs < 1) throw Error(TOFE_ERR_ARG, "bad pc arguments");
    ensure_nvec(c, std::min(nrhs, kMaxBlockVec));
    for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec)  // one V-cycle pass per chunk of vectors
      precondition(c, dR + (size_t)i0 * c->ndof, dZ + (size_t)i0 * c->ndof, nullptr, false,
                   std::min(kMaxBlockVec, nrhs - i0));
    TOFE_HIP(hipStreamSynchronize(c->stream));
  });
}

int tofe_pc_apply(tofe_ctx *c, const double *R, double *Z, int nrhs)
{
  return guarded(c, "tofe_pc_apply", [&] {
    need_assembled(c);
    if (!R || !Z || nrhs < 1) throw Error(TOFE_ERR_ARG, "bad pc arguments");
    ensure_nvec(c, std::min(nrhs, kMaxBlockVec));
    for (int i0 = 0; i0 < nrhs; i0 += kMaxBlockVec) {
      const int nv = std::min(kMaxBlockVec, nrhs - i0);
      const size_t off = (size_t)i0 * c->ndof, len = (size_t)nv * c->ndof;
      h2d(c, c->vp.p, R + off, len);
      precondition(c, c->vp.p, c->vq.p, nullptr, false, nv);
      d2h(c, Z + off, c->vq.p, len);
    }
  });
}

int tofe_compliance_dev(tofe_ctx *c, const double *dU, const double *dF, double *cval,
                        double *d_dcdx)
{
  return guarded(c, "tofe_compliance_dev", [&] {
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    if (!dU || !dF) throw Error(TOFE_ERR_ARG, "bad compliance arguments");
    if (cval) {
      dot_into(c, dU, dF, c->ndof, SL_TMP);
      double h[SL_COUNT];
      read_scal(c, h);
      *cval = h[SL_TMP];
    }
    if (d_dcdx) {
      sensitivity(c->stream, dU, c->dEdz.p, c->dKE.p, c->desc.nx, c->desc.ny, c->desc.nz, c->ebuf.p);
      chain_dev(c, c->ebuf.p, d_dcdx);
    }
    TOFE_HIP(hipStreamSynchronize(c->stream));
  });
}

int tofe_compliance(tofe_ctx *c, const double *U, const double *F, double *cval, double *dcdx)
{
  return guarded(c, "tofe_compliance", [&] {
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    if (!U || !F) throw Error(TOFE_ERR_ARG, "bad compliance arguments");
    h2d(c, c->vp.p, U, c->ndof);
    h2d(c, c->vq.p, F, c->ndof);
    if (cval) {
      dot_into(c, c->vp.p, c->vq.p, c->ndof, SL_TMP);
      double h[SL_COUNT];
      read_scal(c, h);
      *cval = h[SL_TMP];
    }
    if (dcdx) {
      sensitivity(c->stream, c->vp.p, c->dEdz.p, c->dKE.p, c->desc.nx, c->desc.ny, c->desc.nz, c->ebuf.p);
      chain_dev(c, c->ebuf.p, c->vr.p);
      d2h(c, dcdx, c->vr.p, c->ne);
    }
    TOFE_HIP(hipStreamSynchronize(c->stream));
  });
}

int tofe_sensitivity(tofe_ctx *c, const double *U, double *g)
{
  return guarded(c, "tofe_sensitivity", [&] {
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    if (!U || !g) throw Error(TOFE_ERR_ARG, "bad arguments");
    h2d(c, c->vp.p, U, c->ndof);
    sensitivity(c->stream, c->vp.p, c->dEdz.p, c->dKE.p, c->desc.nx, c->desc.ny, c->desc.nz, c->ebuf.p);
    d2h(c, g, c->ebuf.p, c->ne);
  });
}

int tofe_set_penalty(tofe_ctx *c, double penal, double vd_penal)
{
  return guarded(c, "tofe_set_penalty", [&] {
    if (!(penal > 0) || !(vd_penal > 0)) throw Error(TOFE_ERR_ARG, "penalties must be positive");
    c->desc.penal = c->sp.penal = penal;
    c->desc.vd_penal = c->sp.vd_penal = vd_penal;
  });
}

int tofe_chain_filter(tofe_ctx *c, double *dfdV, double *dfdE)
{
  return guarded(c, "tofe_chain_filter", [&] {
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    if (dfdV) {
      h2d(c, c->ebuf.p, dfdV, c->ne);
      filter_apply_T(c->stream, filter_tpl(c), c->filt_inv.p, c->ebuf.p, c->vr.p);
      d2h(c, dfdV, c->vr.p, c->ne);
    }
    if (dfdE) {
      h2d(c, c->ebuf.p, dfdE, c->ne);
      chain_dev(c, c->ebuf.p, c->vr.p);
      d2h(c, dfdE, c->vr.p, c->ne);
    }
  });
}

int tofe_filter(tofe_ctx *c, const double *x, double *rho)
{
  return guarded(c, "tofe_filter", [&] {
    if (!x || !rho) throw Error(TOFE_ERR_ARG, "bad arguments");
    h2d(c, c->ebuf.p, x, c->ne);
    filter_apply(c->stream, filter_tpl(c), c->filt_inv.p, c->ebuf.p, c->ebuf2.p);
    d2h(c, rho, c->ebuf2.p, c->ne);
  });
}

int tofe_filter_T(tofe_ctx *c, const double *g, double *gx)
{
  return guarded(c, "tofe_filter_T", [&] {
    if (!g || !gx) throw Error(TOFE_ERR_ARG, "bad arguments");
    h2d(c, c->ebuf.p, g, c->ne);
    filter_apply_T(c->stream, filter_tpl(c), c->filt_inv.p, c->ebuf.p, c->ebuf2.p);
    d2h(c, gx, c->ebuf2.p, c->ne);
  });
}

namespace {

// CSR of a template filter on the host: row counts, prefix sums, then the device fill
// (weights w_q * inv_e; inv == nullptr -> unit weights).
void template_pattern(tofe_ctx *c, const FilterTemplate &t, const double *inv, int64_t *rowptr,
                      int32_t *col, double *w)
{
  const int64_t ne = c->ne;
  DBuf<int64_t> cnt;
  cnt.alloc(ne + 1);
  filter_count(c->stream, t, cnt.p);
  std::vector<int64_t> h(ne + 1);
  TOFE_HIP(hipMemcpyAsync(h.data() + 1, cnt.p, ne * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  TOFE_HIP(hipStreamSynchronize(c->stream));
  h[0] = 0;
  for (int64_t e = 0; e < ne; ++e) h[e + 1] += h[e];
  std::memcpy(rowptr, h.data(), (ne + 1) * sizeof(int64_t));
  if (col && w) {
    const int64_t nnz = h[ne];
    DBuf<int32_t> dcol;
    DBuf<double> dw, ones;
    dcol.alloc(nnz);
    dw.alloc(nnz);
    if (!inv) {
      ones.alloc(ne);
      fill_d(c->stream, ones.p, 1.0, ne);
      inv = ones.p;
    }
    TOFE_HIP(hipMemcpyAsync(cnt.p, h.data(), (ne + 1) * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    filter_fill(c->stream, t, inv, cnt.p, dcol.p, dw.p);
    TOFE_HIP(hipMemcpyAsync(col, dcol.p, nnz * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    TOFE_HIP(hipMemcpyAsync(w, dw.p, nnz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TOFE_HIP(hipStreamSynchronize(c->stream));
  }
}

void write_template_matrix(tofe_ctx *c, const std::string &path, const FilterTemplate &t,
                           const double *inv, int index_bytes)
{
  std::vector<int64_t> rp(c->ne + 1);
  template_pattern(c, t, inv, rp.data(), nullptr, nullptr);
  std::vector<int32_t> col(rp[c->ne]);
  std::vector<double> w(rp[c->ne]);
  template_pattern(c, t, inv, rp.data(), col.data(), w.data());
  const int rc = tofe_petsc_aij_write(path.c_str(), c->ne, c->ne, rp.data(), col.data(), w.data(),
                                      index_bytes);
  if (rc) throw Error(rc, "cannot write " + path);
}

void write_vec(tofe_ctx *c, const std::string &path, const double *dsrc, int64_t n, int index_bytes)
{
  std::vector<double> h(n);
  d2h(c, h.data(), dsrc, n);
  const int rc = tofe_petsc_vec_write(path.c_str(), h.data(), n, index_bytes);
  if (rc) throw Error(rc, "cannot write " + path);
}

}  // namespace

int tofe_filter_pattern(tofe_ctx *c, int64_t *rowptr, int32_t *col, double *w)
{
  return guarded(c, "tofe_filter_pattern", [&] {
    if (!rowptr) throw Error(TOFE_ERR_ARG, "null rowptr");
    template_pattern(c, filter_tpl(c), c->filt_inv.p, rowptr, col, w);
  });
}

int tofe_mesh_out(tofe_ctx *c, const char *dir, const double origin[3], int index_bytes)
{
  return guarded(c, "tofe_mesh_out", [&] {
    if (!dir || (index_bytes != 4 && index_bytes != 8)) throw Error(TOFE_ERR_ARG, "bad arguments");
    const std::string d(dir);
    const double zero[3] = {0, 0, 0};
    const tofe_desc &q = c->desc;
    int rc = write_mesh_files(d, q.nx, q.ny, q.nz, q.hx, q.hy, q.hz, origin ? origin : zero,
                              index_bytes);
    if (rc) throw Error(rc, "cannot write the mesh files under " + d);
    write_template_matrix(c, d + "/Filter.bin", filter_tpl(c), c->filt_inv.p, index_bytes);
    if (c->REdge.p) {
      write_template_matrix(c, d + "/Max_Filter.bin", rfilter_tpl(c), nullptr, index_bytes);
      write_vec(c, d + "/Void_Edge_Volume.bin", c->REdge.p, c->ne, index_bytes);
    }
    for (int l = 0; l + 1 < c->L; ++l) {
      const Grid &g = c->grids[l];
      std::vector<int64_t> rp;
      std::vector<int32_t> col;
      std::vector<double> val;
      build_interpolation(g.Nx, g.Ny, g.Nz, rp, col, val);
      const std::string path = d + "/P" + std::to_string(l) + ".bin";
      rc = tofe_petsc_aij_write(path.c_str(), 3 * g.nn(), c->grids[l + 1].ndof(), rp.data(),
                                col.data(), val.data(), index_bytes);
      if (rc) throw Error(rc, "cannot write " + path);
    }
  });
}

int tofe_print_vals(tofe_ctx *c, const char *dir, const char *suffix, const double *U,
                    int index_bytes)
{
  return guarded(c, "tofe_print_vals", [&] {
    if (!dir || (index_bytes != 4 && index_bytes != 8)) throw Error(TOFE_ERR_ARG, "bad arguments");
    if (!c->design_set) throw Error(TOFE_ERR_STATE, "no design set");
    const std::string d(dir), sfx(suffix ? suffix : "");
    if (U) {
      const int rc = tofe_petsc_vec_write((d + "/U" + sfx + ".bin").c_str(), U, c->ndof, index_bytes);
      if (rc) throw Error(rc, "cannot write U");
    }
    write_vec(c, d + "/x" + sfx + ".bin", c->x.p, c->ne, index_bytes);
    write_vec(c, d + "/V" + sfx + ".bin", c->rho.p, c->ne, index_bytes);
    write_vec(c, d + "/Es" + sfx + ".bin", c->Es.p, c->ne, index_bytes);
    write_vec(c, d + "/E" + sfx + ".bin", c->E.p, c->ne, index_bytes);
  });
}

int tofe_set_precision(tofe_ctx *c, int precision)
{
  if (!c) return TOFE_ERR_ARG;
  if (precision != TOFE_PREC_FP64 && precision != TOFE_PREC_MIXED) {
    c->err = "unknown precision";
    return TOFE_ERR_ARG;
  }
  c->precision = precision;
  return TOFE_OK;
}

int tofe_get_stats(const tofe_ctx *c, tofe_stats *st)
{
  if (!c || !st) return TOFE_ERR_ARG;
  *st = c->stats;
  st->fine_form = c->fine_mfma ? 4 : c->ke_wht ? (c->ke_wht_cube ? 3 : 2) : c->ke_sym ? 1 : 0;
  st->coarse_form = c->direct ? 1 : 0;
  return TOFE_OK;
}

int tofe_get_level_stencil(tofe_ctx *c, int level, double *out, int64_t *nnode_l)
{
  return guarded(c, "tofe_get_level_stencil", [&] {
    if (level < 0 || level >= c->L) throw Error(TOFE_ERR_ARG, "bad level");
    if (!c->assembled) throw Error(TOFE_ERR_STATE, "not assembled");
    const int64_t nn = c->grids[level].nn();
    if (nnode_l) *nnode_l = nn;
    if (level == 0 && out) {
      c->fineS.alloc((size_t)243 * c->nn);
      fine_stencil_build(c->stream, fine_of<double>(c), c->fineS.p);
    }
    if (out) d2h(c, out, level == 0 ? c->fineS.p : c->lv64[level]->S.p, (size_t)243 * nn);
  });
}

int tofe_time_operator(tofe_ctx *c, int level, int precision, int epilogue, int nrep, double *ms)
{
  return guarded(c, "tofe_time_operator", [&] {
    if (level < 0 || level >= c->L || nrep < 1 || !ms) throw Error(TOFE_ERR_ARG, "bad arguments");
    need_assembled(c);
    const bool f32 = precision == TOFE_PREC_MIXED;
    if (f32 && c->built_precision != TOFE_PREC_MIXED)
      throw Error(TOFE_ERR_STATE, "FP32 kernels need an assemble in mixed precision");
    *ms = time_level_op(c, level, f32, epilogue == 1 ? EPI_CHEB : EPI_APPLY, nrep);
  });
}

int tofe_time_kernels(tofe_ctx *c, int start, tofe_kernel_time *out, int max_kinds, int *nkinds)
{
  return guarded(c, "tofe_time_kernels", [&] {
    if (start) {
      constexpr size_t kLaunches = 4096;
      if (c->kt_ev.empty()) {
        c->kt_ev.assign(2 * kLaunches, nullptr);
        c->kt_kind.assign(kLaunches, 0);
        for (auto &e : c->kt_ev) TOFE_HIP(hipEventCreate(&e));
      }
      c->kt_n = 0;
      c->kt_on = true;
      return;
    }
    if (!nkinds || (max_kinds > 0 && !out)) throw Error(TOFE_ERR_ARG, "bad arguments");
    c->kt_on = false;
    std::vector<int> kinds;
    std::vector<int> cnt;
    std::vector<double> tot;
    for (size_t k = 0; k < c->kt_n; ++k) {
      TOFE_HIP(hipEventSynchronize(c->kt_ev[2 * k + 1]));
      float t = 0;
      TOFE_HIP(hipEventElapsedTime(&t, c->kt_ev[2 * k], c->kt_ev[2 * k + 1]));
      size_t j = 0;
      while (j < kinds.size() && kinds[j] != c->kt_kind[k]) ++j;
      if (j == kinds.size()) {
        kinds.push_back(c->kt_kind[k]);
        cnt.push_back(0);
        tot.push_back(0.0);
      }
      ++cnt[j];
      tot[j] += t;
    }
    *nkinds = (int)kinds.size();
    for (int j = 0; j < (int)kinds.size() && j < max_kinds; ++j) {
      tofe_kernel_time &o = out[j];
      o.level = kinds[j] / 16;
      o.precision = (kinds[j] / 8) % 2 ? TOFE_PREC_MIXED : TOFE_PREC_FP64;
      o.epilogue = kinds[j] % 8;
      o.count = cnt[j];
      o.total_ms = tot[j];
    }
  });
}

int tofe_get_level_lambda(const tofe_ctx *c, int level, double *lambda)
{
  if (!c || !lambda || level < 0 || level >= c->L) return TOFE_ERR_ARG;
  if (!c->assembled) return TOFE_ERR_STATE;
  *lambda = c->lam[level];
  return TOFE_OK;
}

int tofe_autotune(tofe_ctx *c, int nrep, int *fine_zlen, int64_t *split_nodes)
{
  return guarded(c, "tofe_autotune", [&] {
    if (nrep < 1) throw Error(TOFE_ERR_ARG, "nrep must be positive");
    autotune(c, nrep, fine_zlen, split_nodes);
    TOFE_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"

// ---- batches of independent contexts --------------------------------------------------

struct tofe_batch {
  std::vector<tofe_ctx *> ctx;
  hipStream_t origin = nullptr;  // the stream the graph is captured from and launched on
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  // how the chunks of iterations run: one graph of batched launches, each step of the
  // iteration one launch over every unfinished instance (TOFE_BATCH_GRAPH=merged, the
  // default); one graph holding every unfinished instance's launches as a branch (=single);
  // one graph per instance on the instance's own stream (=instance); or no graph, the chunk's
  // kernels enqueued on the instances' streams in turn (=streams)
  enum Form { MERGED, SINGLE, INSTANCE, STREAMS } form = MERGED;
  // of the first merged program of the last solve, per iteration: launches, those serving one
  // instance, and the launches the instances' own iterations make
  int program_launches = 0, program_unbatched = 0, program_recorded = 0;
  tofe_batch_stats stats{};
  std::string err;

  tofe_batch() = default;
  tofe_batch(const tofe_batch &) = delete;
  tofe_batch &operator=(const tofe_batch &) = delete;
  ~tofe_batch()
  {
    if (origin) (void)hipStreamSynchronize(origin);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    for (hipEvent_t e : ev_join)
      if (e) (void)hipEventDestroy(e);
    if (origin) (void)hipStreamDestroy(origin);
  }
};

namespace {

template <typename F>
int batch_guarded(tofe_batch *b, const char *name, F &&f)
{
  if (!b) return TOFE_ERR_ARG;
  Range range(name);
  try {
    f();
    return TOFE_OK;
  } catch (const Error &e) {
    b->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    b->err = "host allocation failed";
    return TOFE_ERR_NOMEM;
  } catch (const std::exception &e) {
    b->err = e.what();
    return TOFE_ERR_HIP;
  }
}

// Graph of `chunk` solve iterations of every instance that still has columns to solve: each
// instance's kernels on its own stream, forked from and joined back to the origin stream.
hipGraphExec_t capture_iterations(tofe_batch *b, const std::vector<double *> &U,
                                  const std::vector<int> &live, int nvec, int chunk)
{
  hipGraph_t g = nullptr;
  TOFE_HIP(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
  try {
    TOFE_HIP(hipEventRecord(b->ev_fork, b->origin));
    for (size_t i = 0; i < b->ctx.size(); ++i) {
      if (!live[i]) continue;
      tofe_ctx *c = b->ctx[i];
      TOFE_HIP(hipStreamWaitEvent(c->stream, b->ev_fork, 0));
      for (int k = 0; k < chunk; ++k) iteration_on_device(c, U[i], nvec);
      TOFE_HIP(hipEventRecord(b->ev_join[i], c->stream));
      TOFE_HIP(hipStreamWaitEvent(b->origin, b->ev_join[i], 0));
    }
  } catch (...) {
    (void)hipStreamEndCapture(b->origin, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  TOFE_HIP(hipStreamEndCapture(b->origin, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    throw Error(TOFE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return exec;
}

// Graph of `chunk` solve iterations of one instance, captured from its own stream.
hipGraphExec_t capture_instance(tofe_ctx *c, double *U, int nvec, int chunk)
{
  hipGraph_t g = nullptr;
  TOFE_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int k = 0; k < chunk; ++k) iteration_on_device(c, U, nvec);
  } catch (...) {
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  TOFE_HIP(hipStreamEndCapture(c->stream, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    throw Error(TOFE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return exec;
}

// One solve iteration of several instances as a program of batched launches.  Each
// instance's launches are recorded (iteration_on_device in recording mode); the sequences are
// then merged step by step: the instance with the most launches left leads, and every
// instance whose next launch has the lead's key joins that step's launch.  Each instance's
// launches keep their order and each block of a batched launch does what the instance's own
// launch does with that block, so the results are those of the instances' own iterations,
// bit for bit.  Launches without a batched form run on their own, in their place.
struct BatchProgram {
  struct Step {
    int key = BK_CLOSURE;
    int n = 0, total = 0, ld_max = 0;
    size_t tab = 0, b0 = 0;  // offsets of the parameter table and block offsets in the arena
    std::function<void(hipStream_t)> run;
  };
  std::vector<Step> steps;
  DBuf<unsigned char> arena;
  int nvec = 1;
  int batched = 0, single = 0;  // launches per iteration with several instances / one
  int recorded = 0;             // the instances' own launches per iteration, summed

  void build(const std::vector<tofe_ctx *> &ctx, const std::vector<double *> &U,
             const std::vector<int> &live, int nv)
  {
    nvec = nv;
    steps.clear();
    batched = single = recorded = 0;
    const size_t n = ctx.size();
    std::vector<Recorder> rec(n);
    for (size_t i = 0; i < n; ++i) {
      if (!live[i]) continue;
      struct Stop {
        tofe_ctx *c;
        ~Stop() { c->rec = nullptr; }
      } stop{ctx[i]};
      ctx[i]->rec = &rec[i];
      iteration_on_device(ctx[i], U[i], nvec);
      recorded += (int)rec[i].ops.size();
    }
    std::vector<size_t> pos(n, 0);
    std::vector<unsigned char> host;
    auto align = [&host] { host.resize((host.size() + 255) / 256 * 256); };
    for (;;) {
      size_t lead = n, most = 0;
      for (size_t i = 0; i < n; ++i) {
        const size_t left = rec[i].ops.size() - pos[i];
        if (left > most) {
          most = left;
          lead = i;
        }
      }
      if (lead == n) break;
      const RecOp &head = rec[lead].ops[pos[lead]];
      Step st;
      st.key = head.key;
      if (head.key == BK_CLOSURE) {
        st.run = head.run;
        ++pos[lead];
        steps.push_back(std::move(st));
        ++single;
        continue;
      }
      std::vector<size_t> grp;
      for (size_t i = 0; i < n; ++i)
        if (pos[i] < rec[i].ops.size() && rec[i].ops[pos[i]].key == head.key) grp.push_back(i);
      align();
      st.tab = host.size();
      std::vector<int> b0;
      int64_t total = 0;
      for (size_t i : grp) {
        const RecOp &o = rec[i].ops[pos[i]];
        host.insert(host.end(), o.desc.begin(), o.desc.end());
        b0.push_back((int)total);
        total += o.nblocks;
        st.ld_max = std::max(st.ld_max, o.ld);
        ++pos[i];
      }
      if (total >= (int64_t(1) << 31)) throw Error(TOFE_ERR_ARG, "batched launch of 2^31 blocks or more");
      b0.push_back((int)total);
      align();
      st.b0 = host.size();
      const unsigned char *pb = reinterpret_cast<const unsigned char *>(b0.data());
      host.insert(host.end(), pb, pb + b0.size() * sizeof(int));
      st.n = (int)grp.size();
      st.total = (int)total;
      steps.push_back(std::move(st));
      ++(grp.size() > 1 ? batched : single);
    }
    arena.alloc(host.size());
    if (!host.empty()) TOFE_HIP(hipMemcpy(arena.p, host.data(), host.size(), hipMemcpyHostToDevice));
  }

  template <typename D>
  const D *tab(const Step &st) const
  {
    return reinterpret_cast<const D *>(arena.p + st.tab);
  }

  void enqueue(hipStream_t s) const
  {
    for (const Step &st : steps) {
      if (st.key == BK_CLOSURE) {
        st.run(s);
        continue;
      }
      const int kind = st.key >> 8, epi = st.key & 63;
      const bool f32 = (st.key >> 7) & 1, cube = (st.key >> 6) & 1;
      const int *b0 = reinterpret_cast<const int *>(arena.p + st.b0);
      switch (kind) {
        case BK_FINE:
          if (f32) fine_wht_batch<float>(s, (Epi)epi, cube, tab<FineLaunch<float>>(st), b0, st.n, st.total, nvec);
          else fine_wht_batch<double>(s, (Epi)epi, cube, tab<FineLaunch<double>>(st), b0, st.n, st.total, nvec);
          break;
        case BK_STENCIL:
          if (f32) stencil_sym_batch<float>(s, (Epi)epi, tab<StencilLaunch<float>>(st), b0, st.n, st.total, nvec);
          else stencil_sym_batch<double>(s, (Epi)epi, tab<StencilLaunch<double>>(st), b0, st.n, st.total, nvec);
          break;
        case BK_RESTRICT:
          if (f32) restrict_batch<float>(s, tab<RestrictLaunch<float>>(st), b0, st.n, st.total, nvec);
          else restrict_batch<double>(s, tab<RestrictLaunch<double>>(st), b0, st.n, st.total, nvec);
          break;
        case BK_PROLONG:
          if (f32) prolong_batch<float>(s, tab<ProlongLaunch<float>>(st), b0, st.n, st.total, nvec);
          else prolong_batch<double>(s, tab<ProlongLaunch<double>>(st), b0, st.n, st.total, nvec);
          break;
        case BK_TRI_LO:
        case BK_TRI_UP: {
          const bool lower = kind == BK_TRI_LO;
          if (f32)
            trigemv_batch<float>(s, lower, tab<TrigemvLaunch<float>>(st), b0, st.n, st.total, nvec, st.ld_max);
          else
            trigemv_batch<double>(s, lower, tab<TrigemvLaunch<double>>(st), b0, st.n, st.total, nvec, st.ld_max);
          break;
        }
        case BK_FCG_XR: fcg_update_r_batch(s, tab<FcgXrLaunch>(st), b0, st.n, st.total, nvec); break;
        case BK_FCG_P: fcg_update_xp_batch(s, tab<FcgPLaunch>(st), b0, st.n, st.total, nvec); break;
        case BK_REDUCE: reduce_final_batch(s, tab<ReduceLaunch>(st), b0, st.n, nvec); break;
        default: throw Error(TOFE_ERR_STATE, "batched launch of an unknown kind");
      }
    }
  }
};

// Graph of `chunk` iterations of a merged program, captured from the batch's stream.
hipGraphExec_t capture_program(tofe_batch *b, const BatchProgram &prog, int chunk)
{
  hipGraph_t g = nullptr;
  TOFE_HIP(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
  try {
    for (int k = 0; k < chunk; ++k) prog.enqueue(b->origin);
  } catch (...) {
    (void)hipStreamEndCapture(b->origin, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  TOFE_HIP(hipStreamEndCapture(b->origin, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess)
    throw Error(TOFE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return exec;
}

}  // namespace

extern "C" {

int tofe_batch_create(tofe_ctx *const *ctxs, int n, tofe_batch **out)
{
  if (!ctxs || n < 1 || !out) return TOFE_ERR_ARG;
  *out = nullptr;
  std::unique_ptr<tofe_batch> b(new tofe_batch());
  try {
    for (int i = 0; i < n; ++i) {
      if (!ctxs[i]) return TOFE_ERR_ARG;
      for (int j = 0; j < i; ++j)
        if (ctxs[j] == ctxs[i]) return TOFE_ERR_ARG;  // one stream per instance
      if (ctxs[i]->device != ctxs[0]->device) return TOFE_ERR_ARG;
      b->ctx.push_back(ctxs[i]);
    }
    TOFE_HIP(hipSetDevice(ctxs[0]->device));
    TOFE_HIP(hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking));
    TOFE_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    b->ev_join.assign(n, nullptr);
    for (int i = 0; i < n; ++i) TOFE_HIP(hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming));
    if (const char *m = std::getenv("TOFE_BATCH_GRAPH"))
      b->form = std::strcmp(m, "instance") == 0  ? tofe_batch::INSTANCE
                : std::strcmp(m, "streams") == 0 ? tofe_batch::STREAMS
                : std::strcmp(m, "single") == 0  ? tofe_batch::SINGLE
                                                 : tofe_batch::MERGED;
  } catch (const Error &e) {
    return e.code;
  } catch (...) {
    return TOFE_ERR_HIP;
  }
  *out = b.release();
  return TOFE_OK;
}

int tofe_batch_destroy(tofe_batch *b)
{
  delete b;
  return TOFE_OK;
}

const char *tofe_batch_last_error(const tofe_batch *b) { return b ? b->err.c_str() : "null batch"; }

int tofe_batch_solve_dev(tofe_batch *b, const double *const *d_F, double *const *d_U, int nrhs,
                         double rtol, int maxit, int chunk, int *its, int *reasons)
{
  return batch_guarded(b, "tofe_batch_solve_dev", [&] {
    const int n = (int)b->ctx.size();
    if (!d_F || !d_U || nrhs < 1 || nrhs > kMaxBlockVec || chunk < 1 || maxit < 0)
      throw Error(TOFE_ERR_ARG, "bad batch solve arguments");
    for (int i = 0; i < n; ++i) {
      tofe_ctx *c = b->ctx[i];
      TOFE_HIP(hipSetDevice(c->device));
      coarse_settle(c);  // before any recording or capture of its iteration
      need_assembled(c);
      if (!d_F[i] || !d_U[i]) throw Error(TOFE_ERR_ARG, "null right-hand side or solution");
      ensure_nvec(c, nrhs);
    }
    const double t0 = now_s();
    b->program_launches = b->program_unbatched = b->program_recorded = 0;
    // start of every instance (each on its own stream), then the graph of `chunk` steps
    std::vector<int> live(n), rs((size_t)n * nrhs);
    std::vector<double *> U(d_U, d_U + n);
    int any = 0;
    for (int i = 0; i < n; ++i) {
      live[i] = start_on_device(b->ctx[i], d_F[i], d_U[i], nrhs, rtol, &rs[(size_t)i * nrhs]);
      any += live[i] > 0;
    }
    for (int i = 0; i < n; ++i) TOFE_HIP(hipStreamSynchronize(b->ctx[i]->stream));
    // Replays: the graph covers the instances with columns still to solve; when an instance
    // has finished, the next replay's graph is captured without it, so finished instances
    // cost nothing while the others go on.
    int done_its = 0;
    double t_graph = 0, t_replay = 0;
    int replays = 0;
    if (any && maxit > 0) {
      hipGraphExec_t exec = nullptr;           // single-graph and merged forms
      std::vector<hipGraphExec_t> execs(n);    // per-instance form
      BatchProgram prog;                       // merged form
      std::vector<int> captured;
      auto release = [&] {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        for (hipGraphExec_t &e : execs)
          if (e) (void)hipGraphExecDestroy(e);
      };
      const double tr = now_s();
      try {
        if (b->form == tofe_batch::INSTANCE) {
          const double tg = now_s();
          for (int i = 0; i < n; ++i)
            if (live[i]) execs[i] = capture_instance(b->ctx[i], U[i], nrhs, chunk);
          t_graph += now_s() - tg;
        }
        while (done_its < maxit) {
          if (b->form == tofe_batch::STREAMS) {
            // the chunk's iterations on every unfinished instance's stream, one iteration of
            // each instance in turn, joined on the batch stream
            TOFE_HIP(hipEventRecord(b->ev_fork, b->origin));
            for (int i = 0; i < n; ++i)
              if (live[i]) TOFE_HIP(hipStreamWaitEvent(b->ctx[i]->stream, b->ev_fork, 0));
            for (int k = 0; k < chunk; ++k)
              for (int i = 0; i < n; ++i)
                if (live[i]) iteration_on_device(b->ctx[i], U[i], nrhs);
            for (int i = 0; i < n; ++i) {
              if (!live[i]) continue;
              TOFE_HIP(hipEventRecord(b->ev_join[i], b->ctx[i]->stream));
              TOFE_HIP(hipStreamWaitEvent(b->origin, b->ev_join[i], 0));
            }
          } else if (b->form == tofe_batch::INSTANCE) {
            // each unfinished instance's graph on its own stream, joined on the batch stream
            TOFE_HIP(hipEventRecord(b->ev_fork, b->origin));
            for (int i = 0; i < n; ++i) {
              if (!live[i]) continue;
              tofe_ctx *c = b->ctx[i];
              TOFE_HIP(hipStreamWaitEvent(c->stream, b->ev_fork, 0));
              TOFE_HIP(hipGraphLaunch(execs[i], c->stream));
              TOFE_HIP(hipEventRecord(b->ev_join[i], c->stream));
              TOFE_HIP(hipStreamWaitEvent(b->origin, b->ev_join[i], 0));
            }
          } else {
            // the graph covers the instances with columns still to solve (merged: the program
            // is rebuilt for them)
            std::vector<int> alive(n);
            for (int i = 0; i < n; ++i) alive[i] = live[i] > 0;
            if (alive != captured) {
              if (exec) (void)hipGraphExecDestroy(exec);
              exec = nullptr;
              const double tg = now_s();
              if (b->form == tofe_batch::MERGED) {
                prog.build(b->ctx, U, alive, nrhs);
                if (replays == 0) {  // the first program, over every instance still to solve
                  b->program_launches = (int)prog.steps.size();
                  b->program_unbatched = prog.single;
                  b->program_recorded = prog.recorded;
                }
                exec = capture_program(b, prog, chunk);
              } else {
                exec = capture_iterations(b, U, alive, nrhs, chunk);
              }
              t_graph += now_s() - tg;
              captured = alive;
            }
            TOFE_HIP(hipGraphLaunch(exec, b->origin));
          }
          ++replays;
          done_its += chunk;
          for (int i = 0; i < n; ++i)
            if (live[i])
              TOFE_HIP(hipMemcpyAsync(b->ctx[i]->hscal, b->ctx[i]->scal.p,
                                      (size_t)SL_COUNT * nrhs * sizeof(double),
                                      hipMemcpyDeviceToHost, b->origin));
          TOFE_HIP(hipStreamSynchronize(b->origin));
          int remaining = 0;
          for (int i = 0; i < n; ++i) {
            if (!live[i]) continue;
            int open = 0;
            for (int v = 0; v < nrhs; ++v)
              open += b->ctx[i]->hscal[(size_t)v * SL_COUNT + SL_FROZEN] == 0.0;
            live[i] = open;
            remaining += open > 0;
          }
          if (!remaining) break;
        }
      } catch (...) {
        release();
        throw;
      }
      release();
      t_replay = now_s() - tr - t_graph;
    }
    // per-column iteration counts and reasons from the device state
    for (int i = 0; i < n; ++i) {
      tofe_ctx *c = b->ctx[i];
      std::vector<double> h((size_t)SL_COUNT * nrhs);
      TOFE_HIP(hipMemcpy(h.data(), c->scal.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
      int kmax = 0, rmin = 0;
      for (int v = 0; v < nrhs; ++v) {
        const size_t o = (size_t)i * nrhs + v;
        int k = 0, r = rs[o];
        if (!r) {
          k = (int)h[(size_t)v * SL_COUNT + SL_ITS];
          r = (int)h[(size_t)v * SL_COUNT + SL_REASON];
          if (!r) r = TOFE_DIVERGED_ITS;
        }
        if (its) its[o] = k;
        if (reasons) reasons[o] = r;
        kmax = std::max(kmax, k);
        if (rmin >= 0) rmin = r;
      }
      c->stats.its = kmax;
      c->stats.reason = rmin;
    }
    b->stats.seconds = now_s() - t0;
    b->stats.capture_seconds = t_graph;
    b->stats.replay_seconds = t_replay;
    b->stats.replays = replays;
    b->stats.iterations = done_its;
    b->stats.launches = b->form == tofe_batch::MERGED ? b->program_launches : 0;
    b->stats.unbatched = b->form == tofe_batch::MERGED ? b->program_unbatched : 0;
    b->stats.instance_launches = b->form == tofe_batch::MERGED ? b->program_recorded : 0;
  });
}

int tofe_batch_get_stats(const tofe_batch *b, tofe_batch_stats *st)
{
  if (!b || !st) return TOFE_ERR_ARG;
  *st = b->stats;
  return TOFE_OK;
}

}  // extern "C"
