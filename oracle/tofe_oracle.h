/*
 * tofe_oracle.h -- FP64 CPU restatement of TopOpt3D's finite-element analysis path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle for the HIP
 * implementation in topopt3d_amd/ and the CPU baseline timed by bench.py.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product path never links or calls it.
 *
 * It restates, in plain C with OpenMP, the single-rank behaviour of:
 *   FEAnalysis.cpp:869-1081   element stiffness KE (2x2x2 Gauss, B, D)
 *   RecMesh.cpp:443-512       constitutive matrix, element/node numbering
 *   Filter.cpp:21-150         density filter (pattern, weights, row scaling)
 *   Material_Interpolation.cpp:20-187   SIMP interpolation and filter chain rule
 *   FEAnalysis.cpp:432-490    assembly, MatZeroRowsColumns Dirichlet semantics
 *   Interpolation.cpp:31-355  trilinear transfer operators with row scaling
 *   FEAnalysis.cpp:387-416,507-732  CG + Galerkin geometric multigrid (PCMG)
 *   Compliance.cpp:15-57      compliance and element sensitivities
 *
 * The reference itself cannot be built here (PETSc, SLEPc, ParMETIS and Eigen are
 * absent), so PETSc's internal choices are restated from its documentation:
 * KSPCG with unpreconditioned-norm convergence test ||r|| <= rtol*||b||, PCMG
 * V-cycle with Galerkin coarse operators, Chebyshev/Jacobi level smoothers with
 * eigenvalue-estimate transform (0.1, 1.1), and a dense direct coarse solve.
 */
#ifndef TOFE_ORACLE_H
#define TOFE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_csr {
  int64_t n, m;        /* rows, cols */
  int64_t nnz;
  int64_t *rowptr;     /* n+1 */
  int32_t *col;        /* nnz, ascending within a row */
  double  *val;        /* nnz */
} orc_csr;

/* ---- element and mesh ------------------------------------------------- */
void orc_dmat(double E0, double nu, double D[36]);
void orc_ke(double E0, double nu, double hx, double hy, double hz, double KE[576]);
void orc_element_nodes(int nx, int ny, int nz, int64_t el, int64_t nodes[8]);

/* ---- filter -------------------------------------------------------------- */
int64_t orc_filter_build(int nx, int ny, int nz, double dx, double dy, double dz,
                         double R, int scale, int64_t *rowptr, int32_t *col, double *w);
orc_csr *orc_filter_csr(int nx, int ny, int nz, double dx, double dy, double dz,
                        double R, int scale);

orc_csr *orc_rfilter_csr(int nx, int ny, int nz, double dx, double dy, double dz, double Rmax,
                         double *REdge);

/* ---- SIMP ----------------------------------------------------------------- */
double orc_vecpow1(double x, double p);
void orc_simp(int64_t ne, const double *rho, double penal, double eps,
              double *E, double *dEdz, double *Es, double *dEsdz);
void orc_matint(const orc_csr *P, const orc_csr *R, const double *REdge, const double *x,
                double penal, double vdPenal, double vdMin, int interp, const double *param,
                double eps, double *rho, double *y, double *rhoq, double *E, double *dEdz,
                double *Es, double *dEsdz);
void orc_chain_filter(const orc_csr *P, const orc_csr *R, const double *rho, const double *y,
                      const double *rhoq, double *dfdV, double *dfdE);

/* ---- sparse helpers ------------------------------------------------------- */
void orc_csr_free(orc_csr *A);
void orc_csr_info(const orc_csr *A, int64_t *n, int64_t *m, int64_t *nnz);
void orc_csr_export(const orc_csr *A, int64_t *rowptr, int32_t *col, double *val);
orc_csr *orc_csr_import(int64_t n, int64_t m, const int64_t *rowptr,
                        const int32_t *col, const double *val);
void orc_spmv(const orc_csr *A, const double *x, double *y);
void orc_spmv_t(const orc_csr *A, const double *x, double *y);
orc_csr *orc_transpose(const orc_csr *A);
orc_csr *orc_spgemm(const orc_csr *A, const orc_csr *B);
orc_csr *orc_ptap(const orc_csr *A, const orc_csr *P);

/* ---- stiffness and transfers --------------------------------------------- */
orc_csr *orc_assemble_K(int nx, int ny, int nz, const double KE[576],
                        const double *E, const uint8_t *fixed_dof);
orc_csr *orc_build_P(int Nfx, int Nfy, int Nfz);

/* ---- multigrid-preconditioned CG ------------------------------------------- */
typedef struct orc_mg orc_mg;
orc_mg *orc_mg_create(const orc_csr *K, int Nx, int Ny, int Nz, int levels,
                      int smoother, int nu1, int nu2, double lo, double hi,
                      int eig_its);
int  orc_mg_levels(const orc_mg *mg);
const orc_csr *orc_mg_A(const orc_mg *mg, int l);
const orc_csr *orc_mg_P(const orc_mg *mg, int l);
double orc_mg_lambda(const orc_mg *mg, int l);
void orc_mg_dims(const orc_mg *mg, int l, int dims[3]);
void orc_mg_vcycle(orc_mg *mg, const double *r, double *z);
int  orc_pcg(orc_mg *mg, const double *b, double *x, double rtol, int maxit,
             int *its, double *rnorm);
void orc_mg_free(orc_mg *mg);
double orc_mg_setup_seconds(const orc_mg *mg);

/* ---- functions ------------------------------------------------------------ */
double orc_dot(int64_t n, const double *a, const double *b);
/* Lumped mass diagonal, Frequency::DiagMassFnc (Dynamic.cpp:181-236) and the Dirichlet
 * treatment MatZeroRowsColumns(M, fixed, 1e-8) (Dynamic.cpp:44-45): element blocks
 * V_e * m0 * I(24), m0 = 1/2^3/3 * elemSize * density, added element by element, then
 * MLump (may be NULL), then 1e-8 on fixed DOFs (fixed may be NULL). */
void orc_mass_diag(int nx, int ny, int nz, const double *V, double m0, const double *mlump,
                   const uint8_t *fixed, double *M);

void orc_sensitivity(int nx, int ny, int nz, const double KE[576], const double *U,
                     const double *dEdz, double *g);
void orc_power_start(int64_t n, double *v);
double orc_tridiag_max_eig(int k, const double *a, const double *b);
double orc_lanczos_lambda(const orc_csr *A, const double *dinv, int its);
/* the PCMG level smoother: Chebyshev (smoother 0) or damped Jacobi (1) on D^-1 A */
void orc_smooth(const orc_csr *A, const double *dinv, double lam, double lo, double hi,
                int smoother, int steps, int zero_guess, const double *b, double *x, double *d,
                double *r);

int orc_num_threads(void);

/* ---- linearised buckling (Buckling.cpp:15-414) ---- */
void orc_shape_grad(double hx, double hy, double hz, double dNdx[8][3][8], double detJ[8]);
/* Ks = sum_e (-Es_e) ks_e(u_e), MatZeroRowsColumns(Ks, fixed, 0.0) */
orc_csr *orc_assemble_Ks(int nx, int ny, int nz, double E0, double nu, double hx, double hy,
                         double hz, const double *U, const double *Es, const uint8_t *fixed);
/* out[j*ne + e] = phi_j,e^T ks_e(u_e) phi_j,e (without the -Es_e factor) */
void orc_ks_quadratic(int nx, int ny, int nz, double E0, double nu, double hx, double hy,
                      double hz, const double *U, const double *phi, int nvec, double *out);
/* adjoint right-hand sides dKsdU (nvec x ndof) */
void orc_stability_rhs(int nx, int ny, int nz, double E0, double nu, double hx, double hy,
                       double hz, const double *Es, const double *phi, int nvec,
                       const uint8_t *fixed, double *out);

#ifdef __cplusplus
}
#endif
#endif
