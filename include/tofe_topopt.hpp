// tofe_topopt.hpp -- header-only C++ host facade over the C ABI of tofe.h.
//
// Keeps the method names, order and meaning of the reference's FE-analysis surface
// (TopOpt.h:282-292, Functions.h:65-77) so host code written against TopOpt --
// the design loop of Main.cpp:131-174, MMA, the eigensolvers -- can drive the GPU path:
//
//   tofe::TopOpt t;
//   t.CreateMesh({0,48, 0,24, 0,24}, {48,24,24}, 1.5);             // RecMesh.cpp:426
//   t.Set_BC(center, radius, limits, values, tofe::SUPPORT);        // Inputs.cpp:105
//   t.FEInitialize();                                               // FEAnalysis.cpp:215
//   for (...) {
//     t.MatIntFnc(x);                                              // Material_Interpolation.cpp:20
//     t.FEAssemble();                                              // FEAnalysis.cpp:432
//     t.FESolve();                                                 // FEAnalysis.cpp:507
//     tofe::Compliance c; c.Function(&t);                          // Compliance.cpp:15
//   }
//
// Every method returns 0 or a negative TOFE_ERR_* code (the PetscErrorCode convention);
// LastError() gives the message.  PETSc Vec/Mat/KSP members become std::vector<double>.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "tofe.h"

namespace tofe {

enum BCTYPE { SUPPORT = 0, LOAD = 1 };

class TopOpt {
 public:
  // material and solver parameters (TopOpt.h: E0, Nu0, penal; KSP rtol 1e-8 FEAnalysis.cpp:390)
  double E0 = 1.0, Nu0 = 0.3, penal = 3.0, eps = 1e-10;
  // maximum length scale and interpolation (TopOpt.h:167-178): vdMin <= 0 disables it
  double vdPenal = 1.0, vdMin = -1.0;
  int interpolation = TOFE_INTERP_SIMP;
  double interp_param[2] = {0, 0};
  int mg_levels = 0, smoother = TOFE_SMOOTHER_CHEBYSHEV, precision = TOFE_PREC_FP64;
  double rtol = 1e-8;
  int maxit = 10000;

  // mesh (single rank, x-fastest numbering)
  int Nel[3] = {0, 0, 0};
  double dimensions[6] = {0, 0, 0, 0, 0, 0};
  double dx[3] = {0, 0, 0};
  double Rmin = 0, Rmax = 0;
  int64_t nElem = 0, nNode = 0;

  // boundary conditions (suppNode/supports, loadNode/loads)
  std::vector<int64_t> suppNode, loadNode;
  std::vector<uint8_t> supports;  // 3 flags per supported node
  std::vector<double> loads;      // 3 values per loaded node

  // FE state
  std::vector<uint8_t> fixed_dof;
  std::vector<double> F, U, x, rho, y, rhoq, E, dEdz, Es, dEsdz, V;
  int KUF_reason = 0, its = 0;

  TopOpt() = default;
  TopOpt(const TopOpt &) = delete;
  TopOpt &operator=(const TopOpt &) = delete;
  ~TopOpt() { Clear(); }

  int Clear()
  {
    if (ctx_) tofe_destroy(ctx_);
    ctx_ = nullptr;
    return 0;
  }

  int CreateMesh(const double dims[6], const int nel[3], double rmin, double rmax = 0)
  {
    Rmax = rmax;
    for (int i = 0; i < 6; ++i) dimensions[i] = dims[i];
    for (int i = 0; i < 3; ++i) {
      Nel[i] = nel[i];
      dx[i] = (dims[2 * i + 1] - dims[2 * i]) / nel[i];
    }
    Rmin = rmin;
    nElem = (int64_t)nel[0] * nel[1] * nel[2];
    nNode = (int64_t)(nel[0] + 1) * (nel[1] + 1) * (nel[2] + 1);
    return 0;
  }

  double NodeCoord(int64_t n, int d) const
  {
    const int64_t Nx = Nel[0] + 1, Ny = Nel[1] + 1;
    const int64_t idx[3] = {n % Nx, (n / Nx) % Ny, n / (Nx * Ny)};
    return dimensions[2 * d] + dx[d] * (double)idx[d];
  }

  // Nodes with sum(((x_i - c_i)/r_i)^2) <= 1 inside the box limits (Inputs.cpp:114-125).
  int Set_BC(const double center[3], const double radius[3], const double limits[6],
             const double values[3], BCTYPE type)
  {
    for (int64_t n = 0; n < nNode; ++n) {
      double dist = 0;
      bool valid = true;
      for (int i = 0; i < 3; ++i) {
        const double xi = NodeCoord(n, i);
        dist += ((xi - center[i]) / radius[i]) * ((xi - center[i]) / radius[i]);
        valid = valid && xi >= limits[2 * i] && xi <= limits[2 * i + 1];
      }
      if (!valid || dist > 1) continue;
      if (type == SUPPORT) {
        suppNode.push_back(n);
        for (int i = 0; i < 3; ++i) supports.push_back(values[i] != 0);
      } else {
        loadNode.push_back(n);
        for (int i = 0; i < 3; ++i) loads.push_back(values[i]);
      }
    }
    return 0;
  }

  // fixedDof = 3 gNode + j (FEAnalysis.cpp:236-246); F[3 loadNode + j] += loads (349-356)
  int FEInitialize()
  {
    Clear();
    const int64_t ndof = 3 * nNode;
    fixed_dof.assign(ndof, 0);
    for (size_t i = 0; i < suppNode.size(); ++i)
      for (int j = 0; j < 3; ++j)
        if (supports[3 * i + j]) fixed_dof[3 * suppNode[i] + j] = 1;
    F.assign(ndof, 0.0);
    for (size_t i = 0; i < loadNode.size(); ++i)
      for (int j = 0; j < 3; ++j) F[3 * loadNode[i] + j] += loads[3 * i + j];
    U.assign(ndof, 0.0);
    tofe_desc d{};
    d.nx = Nel[0]; d.ny = Nel[1]; d.nz = Nel[2];
    d.hx = dx[0]; d.hy = dx[1]; d.hz = dx[2];
    d.E0 = E0; d.nu = Nu0; d.Emin = eps; d.penal = penal; d.rmin = Rmin;
    d.mg_levels = mg_levels; d.smoother = smoother;
    d.nu_pre = 2; d.nu_post = 2;
    d.precision = precision;
    d.fixed_dof = fixed_dof.data();
    d.rmax = Rmax; d.vd_min = vdMin; d.vd_penal = vdPenal;
    d.interpolation = interpolation;
    d.interp_param[0] = interp_param[0]; d.interp_param[1] = interp_param[1];
    return tofe_create(&d, &ctx_);
  }

  // rho = P x, y, and the SIMP / SIMP_CUT / SIMP_LOGISTIC interpolation with the current
  // penal and vdPenal (Material_Interpolation.cpp:20-140; Main.cpp:131-133)
  int MatIntFnc(const std::vector<double> &design)
  {
    x = design;
    int ierr = tofe_set_penalty(ctx_, penal, vdPenal);
    if (ierr || (ierr = tofe_set_design(ctx_, x.data()))) return ierr;
    const struct { int f; std::vector<double> *v; } out[] = {
        {TOFE_FIELD_RHO, &rho}, {TOFE_FIELD_Y, &y}, {TOFE_FIELD_RHOQ, &rhoq}, {TOFE_FIELD_E, &E},
        {TOFE_FIELD_DEDZ, &dEdz}, {TOFE_FIELD_ES, &Es}, {TOFE_FIELD_DESDZ, &dEsdz}};
    for (const auto &o : out) {
      o.v->resize(nElem);
      if ((ierr = tofe_get_field(ctx_, o.f, o.v->data()))) return ierr;
    }
    V = rho;
    return 0;
  }

  int FEAssemble() { return tofe_assemble(ctx_); }

  // KSPSolve(KUF, F, U) warm-started from U (FEAnalysis.cpp:389, 673)
  int FESolve() { return tofe_solve(ctx_, F.data(), U.data(), 1, rtol, maxit, &its, &KUF_reason); }

  // dfdV <- P^T dfdV, dfdE <- P^T[dfdE .* y - rhoq .* (y < 1) .* R^T(dfdE .* rho)]
  // in place (Material_Interpolation.cpp:150-187)
  int Chain_Filter(std::vector<double> *dfdV, std::vector<double> *dfdE)
  {
    return tofe_chain_filter(ctx_, dfdV ? dfdV->data() : nullptr, dfdE ? dfdE->data() : nullptr);
  }

  std::string LastError() const { return ctx_ ? tofe_last_error(ctx_) : "no context"; }
  tofe_ctx *Context() { return ctx_; }

 private:
  tofe_ctx *ctx_ = nullptr;
};

// c = U.F and dc/dx = Chain_Filter(-dEdz u_e^T KE u_e) (Compliance.cpp:15-57)
class Compliance {
 public:
  bool calc_gradient = true;
  std::vector<double> values = std::vector<double>(1, 0.0);
  std::vector<double> gradients;

  int Function(TopOpt *t)
  {
    if (calc_gradient) gradients.resize(t->nElem);
    return tofe_compliance(t->Context(), t->U.data(), t->F.data(), values.data(),
                           calc_gradient ? gradients.data() : nullptr);
  }
};

// Volume fraction sum(rho)/nElem with gradient P^T(1/nElem) (Volume.cpp:15-41)
class Volume {
 public:
  std::vector<double> values = std::vector<double>(1, 0.0);
  std::vector<double> gradients;

  int Function(TopOpt *t)
  {
    double s = 0;
    for (double v : t->V) s += v;
    values[0] = s / (double)t->nElem;
    gradients.assign(t->nElem, 1.0 / (double)t->nElem);
    return t->Chain_Filter(&gradients, nullptr);
  }
};

}  // namespace tofe
