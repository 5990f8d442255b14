// Compliance and element sensitivities without a context and without allocating device
// memory: for callers that cannot call hipMalloc (code under HIP graph capture, serving loops
// that forbid allocation).  The result is the same as tofe_compliance_dev's c (with the
// dcdx output left out) and tofe_sensitivity's g, from the same kernels in the same order.
//
// Device memory beyond the caller's inputs and outputs -- the element matrix and the
// reduction partials of c = U.F -- lives in one file-scope array of kScratchDoubles doubles
// (16 KiB) per GPU, held until the process ends.  Host code never passes the array's name to
// a kernel or a copy (on the host it only stands in for the symbol): the device address is
// looked up once per device with hipGetSymbolAddress and that address is used instead.
// Every call on one GPU shares that GPU's array and overwrites it; calls that overlap on one
// GPU (several host threads, or a call made while another is in flight) race.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "host_setup.hpp"
#include "tofe_internal.hpp"

namespace {

constexpr int kScratchKE = 0;                        // 576: element stiffness, row-major
constexpr int kScratchPart = 576;                    // kRedBlocks: partial sums of U.F
constexpr int kScratchC = kScratchPart + tofe::kRedBlocks;  // 1: c
constexpr int kScratchDoubles = 2048;
static_assert(kScratchC + 1 <= kScratchDoubles, "scratch too small");

static __device__ __align__(256) double scratch[kScratchDoubles];

// The scratch does not grow with the mesh (the reduction has a fixed number of partials);
// the largest mesh is set by the element kernel's one-thread-per-element grid.
constexpr int64_t kNoallocMaxElements = 2147483647;

constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;
thread_local std::string g_err;

// Device address of this device's scratch array, looked up on the first call on the device
// (with that device current).
double *scratch_of(int dev)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    TOFE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw tofe::Error(TOFE_ERR_HIP, "hipGetSymbolAddress returned no address");
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

// message of this thread's last failing call, read through tofe_last_error(NULL)
const char *tofe::noalloc_last_error() { return g_err.c_str(); }

extern "C" int tofe_compliance_noalloc(int nx, int ny, int nz, double hx, double hy, double hz,
                                       double E0, double nu, const double *d_U,
                                       const double *d_F, const double *d_dEdz, double *c,
                                       double *d_g)
{
  roctxRangePushA("tofe_compliance_noalloc");
  int rc = TOFE_OK;
  try {
    if (nx < 1 || ny < 1 || nz < 1)
      throw tofe::Error(TOFE_ERR_ARG, "tofe_compliance_noalloc: mesh dimensions must be positive");
    const int64_t ne = (int64_t)nx * ny * nz;
    const int64_t ndof = 3 * (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    if (ne > kNoallocMaxElements)
      throw tofe::Error(TOFE_ERR_ARG,
                        "tofe_compliance_noalloc: " + std::to_string(ne) +
                            " elements exceed the maximum of " +
                            std::to_string(kNoallocMaxElements));
    if (!(hx > 0) || !(hy > 0) || !(hz > 0) || !(E0 > 0) || !(nu > -1.0 && nu < 0.5))
      throw tofe::Error(TOFE_ERR_ARG, "tofe_compliance_noalloc: bad element size or material");
    if (!d_U || (c && !d_F) || (d_g && !d_dEdz) || (!c && !d_g))
      throw tofe::Error(TOFE_ERR_ARG, "tofe_compliance_noalloc: missing input or output");
    int dev = 0;
    TOFE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
      throw tofe::Error(TOFE_ERR_ARG, "tofe_compliance_noalloc: device index out of range");
    double *s = scratch_of(dev);
    const hipStream_t st = nullptr;  // the null stream: ordered after the caller's blocking work
    if (c) {
      // partials are written by every block of the dot kernel and c by the final reduction:
      // nothing in the scratch needs clearing between calls
      tofe::dot_partial<double>(st, d_U, d_F, ndof, s + kScratchPart);
      tofe::reduce_final(st, s + kScratchPart, tofe::kRedBlocks, s + kScratchC,
                         tofe::POST_NONE, s + kScratchC);
    }
    if (d_g) {
      double KE[576];
      tofe::compute_ke(E0, nu, hx, hy, hz, KE);
      TOFE_HIP(hipMemcpy(s + kScratchKE, KE, sizeof(KE), hipMemcpyHostToDevice));
      tofe::sensitivity(st, d_U, d_dEdz, s + kScratchKE, nx, ny, nz, d_g);
    }
    if (c) TOFE_HIP(hipMemcpy(c, s + kScratchC, sizeof(double), hipMemcpyDeviceToHost));
    TOFE_HIP(hipStreamSynchronize(st));
  } catch (const tofe::Error &e) {
    g_err = e.what();
    rc = e.code;
  } catch (const std::exception &e) {
    g_err = e.what();
    rc = TOFE_ERR_HIP;
  }
  roctxRangePop();
  return rc;
}
