/*
 * tofe.h -- C ABI of the MI355X-native TopOpt3D finite-element analysis path.
 *
 * One context owns a structured trilinear-hexahedral mesh, its density filter,
 * its geometric-multigrid hierarchy and all device memory for one GPU.  The
 * entry points replace the TopOpt member functions and PETSc handles that the
 * reference's design loop, functions and eigensolvers call; each declaration
 * names the reference interface it stands in for.
 *
 * Conventions (shared with the reference):
 *   - nodes and elements are numbered x-fastest, then y, then z
 *     (RecMesh.cpp:486-512, 565-597); DOF = 3*node + component
 *     (FEAnalysis.cpp:241, Compliance.cpp:40);
 *   - element-local node order (0,0,0),(1,0,0),(1,1,0),(0,1,0), then +z
 *     (RecMesh.cpp:486-512).
 * Errors: every function returns 0 on success and a negative TOFE_ERR_* code on
 * failure; tofe_last_error() gives the message (the CHKERRQ/SETERRQ convention
 * of the reference, with no exceptions crossing the ABI).  A solve that does
 * not converge is not an error: it is reported through `reason`, which uses
 * KSPConvergedReason values (FEAnalysis.cpp:677-684).
 * Host-pointer entry points copy in and out; *_dev entry points take device
 * pointers that must stay valid for the call.  One context per host thread;
 * each context runs on its own HIP stream and every call returns with its
 * stream synchronised.
 */
#ifndef TOFE_H
#define TOFE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tofe_ctx tofe_ctx;

enum {
  TOFE_OK = 0,
  TOFE_ERR_ARG = -1,       /* invalid argument */
  TOFE_ERR_HIP = -2,       /* HIP runtime or device failure */
  TOFE_ERR_STATE = -3,     /* call out of order (e.g. solve before assemble) */
  TOFE_ERR_NUMERIC = -4,   /* factorisation failed (coarse operator not SPD) */
  TOFE_ERR_NOMEM = -5
};

/* KSPConvergedReason values reported by tofe_solve (PETSc semantics). */
enum {
  TOFE_CONVERGED_RTOL = 2,
  TOFE_CONVERGED_ATOL = 3,
  TOFE_DIVERGED_ITS = -3,
  TOFE_DIVERGED_DTOL = -4,
  TOFE_DIVERGED_INDEFINITE_PC = -8,
  TOFE_DIVERGED_NANORINF = -9,
  TOFE_DIVERGED_INDEFINITE_MAT = -10
};

enum { TOFE_SMOOTHER_CHEBYSHEV = 0, TOFE_SMOOTHER_JACOBI = 1 };
/* Material interpolation (TopOpt::interpolation, TopOpt.h:27; Inputs.cpp:339-358). */
enum { TOFE_INTERP_SIMP = 0, TOFE_INTERP_SIMP_CUT = 1, TOFE_INTERP_SIMP_LOGISTIC = 2 };
enum { TOFE_PREC_FP64 = 0, TOFE_PREC_MIXED = 1 };
/* Mass matrix forms (tofe_set_mass_form). */
enum { TOFE_MASS_LUMPED = 0, TOFE_MASS_CONSISTENT = 1 };

/* Fields readable with tofe_get_field (the TopOpt Vecs of the same names,
 * TopOpt.h:180-184). */
enum {
  TOFE_FIELD_X = 0,      /* design variables x            (nelem) */
  TOFE_FIELD_RHO = 1,    /* filtered densities rho = P x  (nelem) */
  TOFE_FIELD_E = 2,      /* element stiffness scale E     (nelem) */
  TOFE_FIELD_DEDZ = 3,   /* dE/dz                         (nelem) */
  TOFE_FIELD_ES = 4,     /* Es = z^p                      (nelem) */
  TOFE_FIELD_DESDZ = 5,  /* dEs/dz                        (nelem) */
  TOFE_FIELD_V = 6,      /* volume interpolation V = rho  (nelem) */
  TOFE_FIELD_DIAG = 7,   /* Jacobi diagonal of fine K     (ndof)  */
  TOFE_FIELD_Y = 8,      /* max-length-scale factor y     (nelem) */
  TOFE_FIELD_RHOQ = 9,   /* rhoq = (q/vdMin)(1-rho)^(q-1) (nelem) */
  TOFE_FIELD_REDGE = 10, /* REdge = max row count of R - row count (nelem) */
  TOFE_FIELD_MDIAG = 11  /* lumped mass diagonal           (ndof)  */
};

typedef struct tofe_desc {
  int nx, ny, nz;          /* elements per direction (Nel, RecMesh.cpp:426) */
  double hx, hy, hz;       /* element edge lengths (dx, RecMesh.cpp:528-533) */
  double E0, nu;           /* material (TopOpt::E0, Nu0) */
  double Emin;             /* minimum stiffness eps (Material_Interpolation.cpp:28: 1e-10) */
  double penal;            /* SIMP exponent p */
  double rmin;             /* density filter radius (absolute length) */
  int mg_levels;           /* levels incl. the fine one; 0 = floor(log2(min nel)) (RecMesh.cpp:552) */
  int smoother;            /* TOFE_SMOOTHER_* */
  int nu_pre, nu_post;     /* smoothing steps (V(nu_pre, nu_post)); 0 -> 2 */
  double cheb_lo, cheb_hi; /* Chebyshev interval as fractions of lambda_max(D^-1 A); 0 -> 0.1, 1.1 */
  int eig_its;             /* Jacobi-CG (Lanczos) steps of each level's lambda_max estimate,
                              1..64 (the reference's runtime smoother options,
                              FEAnalysis.cpp:497-499); 0 -> 10 */
  int precision;           /* TOFE_PREC_*: FP64, or FP32 V-cycle inside FP64 flexible CG */
  const uint8_t *fixed_dof;/* 3*nnode flags, 1 = Dirichlet DOF (fixedDof, FEAnalysis.cpp:236-246); may be NULL */
  /* maximum length scale and interpolation (Material_Interpolation.cpp:37-116) */
  double rmax;             /* radius of the unit-weight filter R (RecMesh.cpp:544); used when vd_min > 0 */
  double vd_min;           /* Void_Minimum; <= 0: no maximum length scale, y = 1 */
  double vd_penal;         /* void penalty q (vdPenal); 0 -> 1 */
  int interpolation;       /* TOFE_INTERP_* */
  double interp_param[2];  /* SIMP_CUT threshold; SIMP_LOGISTIC (a, b) */
} tofe_desc;

typedef struct tofe_stats {
  int levels;              /* multigrid levels used */
  int64_t coarse_size;     /* DOFs of the coarsest (directly solved) level */
  int its;                 /* CG iterations of the last solve */
  int reason;              /* KSPConvergedReason of the last solve */
  double rnorm;            /* final unpreconditioned residual 2-norm */
  double bnorm;            /* ||F|| of the last solve */
  double setup_seconds;    /* last tofe_assemble (Galerkin, diagonals, eigen bounds; the coarse
                              factorisation it starts may end inside the next solve) */
  double solve_seconds;    /* last tofe_solve wall time */
  int fine_form;           /* fine operator kernel: 0 node-centric with KE rows, 1 node-centric
                              with KE's cube-orbit values, 2 element-centric Walsh-Hadamard
                              (box element), 3 the same with the cubic element's 9 values,
                              4 element-centric on the matrix cores (TOFE_FINE_MFMA=1) */
  int coarse_form;         /* coarsest operator: 0 the Galerkin chain's last level, 1 formed
                              straight from E (TOFE_DIRECT_COARSE=1 at tofe_create; element
                              counts multiples of 2^(levels-1) <= 16) */
} tofe_stats;

/* Create the context: mesh numbering, KE, density filter, transfer operators and
 * device buffers.  Replaces TopOpt::CreateMesh's filter/P setup (RecMesh.cpp:426-728,
 * Filter.cpp:21-150, Interpolation.cpp:31-355) and TopOpt::FEInitialize
 * (FEAnalysis.cpp:215-424). */
int tofe_create(const tofe_desc *desc, tofe_ctx **ctx);

/* Release all device memory (TopOpt::Clear, TopOpt.cpp:59-89). */
int tofe_destroy(tofe_ctx *ctx);

/* Message of the last failing call on this context (PETSc error text). */
const char *tofe_last_error(const tofe_ctx *ctx);

/* Mesh and hierarchy sizes (TopOpt::nElem, nNode, PR.size()+1). */
int tofe_sizes(const tofe_ctx *ctx, int64_t *nelem, int64_t *nnode, int64_t *ndof,
               int *levels);

/* Element stiffness matrix actually used, 24x24 row-major (TopOpt::ke[0] =
 * LocalK(0), FEAnalysis.cpp:340-341, 869-892). */
int tofe_get_ke(const tofe_ctx *ctx, double *ke576);

/* rho = P x; SIMP: E, dEdz, Es, dEsdz, V (TopOpt::MatIntFnc,
 * Material_Interpolation.cpp:20-140, plain density filter y = 1). */
int tofe_set_design(tofe_ctx *ctx, const double *x);
int tofe_set_design_dev(tofe_ctx *ctx, const double *d_x);

/* Penalty continuation: the SIMP exponent and the void penalty of the next design
 * iteration (Main.cpp:131-133 sets topOpt->penal and vdPenal per penalty step). */
int tofe_set_penalty(tofe_ctx *ctx, double penal, double vd_penal);

/* Filter chain rule in place on host arrays (either may be NULL):
 * dfdV <- P^T dfdV, dfdE <- P^T[dfdE .* y - rhoq .* (y < 1) .* R^T(dfdE .* rho)]
 * (TopOpt::Chain_Filter, Material_Interpolation.cpp:150-187). */
int tofe_chain_filter(tofe_ctx *ctx, double *dfdV, double *dfdE);

/* Override the element stiffness scale E directly (dEdz is left as is). */
int tofe_set_E(tofe_ctx *ctx, const double *E);

/* Copy out one of the TOFE_FIELD_* vectors. */
int tofe_get_field(const tofe_ctx *ctx, int field, double *out);

/* Spring supports: K + diag(spk) on the free DOFs (3*nnode values; NULL removes them), the
 * reference's MatDiagonalSet(K, spKVec, ADD_VALUES) before its Dirichlet rows
 * (FEAnalysis.cpp:360-368, 471), so Dirichlet DOFs keep their unit row.  The operator, the
 * Jacobi diagonal and the Galerkin hierarchy (P^T diag(spk) P on level 1) include it from the
 * next tofe_assemble. */
int tofe_set_springs(tofe_ctx *ctx, const double *spk);

/* Rebuild the solver for the current E: fine Jacobi diagonal, Galerkin coarse
 * operators P^T A P, Chebyshev eigenvalue bounds and the coarsest factorisation.
 * Replaces TopOpt::FEAssemble (FEAnalysis.cpp:432-490) plus PCSetUp of PCMG
 * (FEAnalysis.cpp:541, 555-573).  Returns when the eigenvalue bounds are known; the
 * coarsest factorisation may still be running (PETSc factors inside KSPSolve's PCSetUp):
 * tofe_solve_dev overlaps its start with it and reports a failed factorisation
 * (TOFE_ERR_NUMERIC) when it ends, every other call before it starts. */
int tofe_assemble(tofe_ctx *ctx);

/* Solve K U = F for nrhs right-hand sides stored one after another (ndof each),
 * warm-started from U, with CG + V-cycle to ||r|| <= rtol*||F||.  its/reason
 * receive the last right-hand side's iteration count and KSPConvergedReason
 * (or may be NULL).  Replaces TopOpt::FESolve's KSPSolve(KUF, F, U)
 * (FEAnalysis.cpp:507-732) and the adjoint KSPSolve of Buckling.cpp:199. */
int tofe_solve(tofe_ctx *ctx, const double *F, double *U, int nrhs, double rtol,
               int maxit, int *its, int *reason);
int tofe_solve_dev(tofe_ctx *ctx, const double *d_F, double *d_U, int nrhs, double rtol,
                   int maxit, int *its, int *reason);

/* Y = K X with the Dirichlet rows/columns of MatZeroRowsColumns(K, fixed, 1.0);
 * MatMult(K) as used by PRINVIT/LOPGMRES (PRINVIT.cpp:353-354). */
int tofe_apply_K(tofe_ctx *ctx, const double *X, double *Y, int nrhs);
int tofe_apply_K_dev(tofe_ctx *ctx, const double *d_X, double *d_Y, int nrhs);

/* Lumped mass matrix of the current design (Frequency::DiagMassFnc, Dynamic.cpp:181-236):
 * M_dd = sum over the node's elements of V_e * (elemSize * density / 24), plus mlump_d
 * (the MASS boundary condition's MLump, FEAnalysis.cpp:374-384; may be NULL), with the
 * Dirichlet rows and columns of MatZeroRowsColumns(M, fixed, 1e-8) (Dynamic.cpp:44-45). */
int tofe_assemble_mass(tofe_ctx *ctx, double density, const double *mlump);

/* Mass form used by the next tofe_assemble_mass: TOFE_MASS_LUMPED (default) is the
 * reference's DiagMassFnc diagonal above; TOFE_MASS_CONSISTENT has no counterpart in the
 * reference: M = sum_e V_e m0 Q (x) I_3 + diag(mlump) with the same m0 and Q_ab =
 * 24 prod_d (a_d == b_d ? 1/3 : 1/6), the exact integral of N_a N_b over the box element
 * divided by m0 / density (each row of Q sums to 3: the consistent matrix holds three times
 * the reference's lumped element mass, elemSize * density per displacement component).  Dirichlet
 * rows and columns as MatZeroRowsColumns(M, fixed, 1e-8).  Applied matrix-free. */
int tofe_set_mass_form(tofe_ctx *ctx, int form);

/* Y = M X for nrhs vectors: MatMult(M) of the block eigensolvers (LOPGMRES/PRINVIT
 * operators set at Dynamic.cpp:67). */
int tofe_apply_M(tofe_ctx *ctx, const double *X, double *Y, int nrhs);
int tofe_apply_M_dev(tofe_ctx *ctx, const double *d_X, double *d_Y, int nrhs);

/* Element quadratic forms of nvec vectors phi (ndof each): kq[j*nelem+e] = phi_e^T KE phi_e
 * and, if mq is not NULL, mq[j*nelem+e] = sum of phi^2 over the element's 24 DOFs -- the
 * per-element terms of the eigenvalue sensitivities (Dynamic.cpp:113-154: dKdy = dEdz KE,
 * dMdy = dVdrho m0 I) -- or, with the consistent mass form, phi_e^T (Q (x) I_3) phi_e
 * (dMdy = dVdrho m0 Q (x) I_3). */
int tofe_element_quadratic(tofe_ctx *ctx, const double *phi, int nvec, double *kq, double *mq);

/* Linearised buckling (Stability::Function, Buckling.cpp:15-414).
 * tofe_set_stress stores the stress state of the displacement U (StressFnc, 283-345):
 * per element H_e = sum_q W detJ grad N^T s(D B_q u_e) grad N, so that the stress
 * stiffness of the current design is Ks = sum_e (-Es_e) H_e (x) I_3 (Es from the last
 * tofe_set_design). */
int tofe_set_stress(tofe_ctx *ctx, const double *U);
int tofe_set_stress_dev(tofe_ctx *ctx, const double *d_U);

/* Y = Ks X with the Dirichlet rows and columns of MatZeroRowsColumns(Ks, fixed, 0.0)
 * (Buckling.cpp:28-29): MatMult(Ks) of the eigensolvers (operators set at Buckling.cpp:67). */
int tofe_apply_Ks(tofe_ctx *ctx, const double *X, double *Y, int nrhs);
int tofe_apply_Ks_dev(tofe_ctx *ctx, const double *d_X, double *d_Y, int nrhs);

/* Eigenvalue-sensitivity terms of nvec buckling modes phi (ndof each), either output may be
 * NULL: pksp[j*nelem+e] = phi_e^T ks_e phi_e (times -dEsdz_e this is the pdKsp term,
 * Buckling.cpp:239-242, 324); dksdu[j*ndof+d] = sum over the elements holding d of
 * Es_e phi_e^T (d ks_e / d u_d) phi_e with fixed rows zero -- the right-hand sides of the
 * adjoint solves (Buckling.cpp:130-191). */
int tofe_stability_terms(tofe_ctx *ctx, const double *phi, int nvec, double *pksp,
                         double *dksdu);

/* out[j*nelem+e] = a_j,e^T KE b_e for nvec vectors a against one vector b (the
 * v^T dK/dy U term of the buckling sensitivity without dEdz, Buckling.cpp:245-246). */
int tofe_element_bilinear(tofe_ctx *ctx, const double *A, int nvec, const double *B,
                          double *out);

/* Z = M^-1 R, one multigrid V-cycle; PCApply of the KUF preconditioner
 * (LOPGMRES.cpp:125). */
int tofe_pc_apply(tofe_ctx *ctx, const double *R, double *Z, int nrhs);
int tofe_pc_apply_dev(tofe_ctx *ctx, const double *d_R, double *d_Z, int nrhs);

/* c = U.F and dc/dx = Chain_Filter(-dEdz .* u_e^T KE u_e): Compliance::Function
 * (Compliance.cpp:15-57) including TopOpt::Chain_Filter
 * (Material_Interpolation.cpp:150-187).  dcdx may be NULL. */
int tofe_compliance(tofe_ctx *ctx, const double *U, const double *F, double *c,
                    double *dcdx);
int tofe_compliance_dev(tofe_ctx *ctx, const double *d_U, const double *d_F, double *c,
                        double *d_dcdx);

/* Context-free compliance c = U.F and element sensitivities g_e = -dEdz_e u_e^T KE u_e (KE of
 * an nx x ny x nz mesh of hx x hy x hz elements with E0, nu) on device pointers, allocating no
 * device memory.  No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, serving loops that forbid hipMalloc).  It gives the c of
 * tofe_compliance_dev and the g of tofe_sensitivity bit for bit (the filter chain rule is left
 * to the caller).  c or d_g may be NULL (not both).  Scratch: a fixed 16 KiB device array per
 * GPU, whatever the mesh size; meshes above 2^31 - 1 elements are rejected with
 * TOFE_ERR_ARG.  Work runs on the null stream and the call returns synchronised.  Calls
 * overlapping on one GPU share that scratch and race.  The message of a failure is
 * tofe_last_error(NULL). */
int tofe_compliance_noalloc(int nx, int ny, int nz, double hx, double hy, double hz, double E0,
                            double nu, const double *d_U, const double *d_F,
                            const double *d_dEdz, double *c, double *d_g);

/* Element sensitivities g_e = -dEdz_e u_e^T KE u_e before the filter chain rule
 * (Compliance.cpp:37-46). */
int tofe_sensitivity(tofe_ctx *ctx, const double *U, double *g);

/* rho = P x and gx = P^T g with the density filter (MatMult / MatMultTranspose of
 * TopOpt::P, Material_Interpolation.cpp:35, 159-162). */
int tofe_filter(tofe_ctx *ctx, const double *x, double *rho);
int tofe_filter_T(tofe_ctx *ctx, const double *g, double *gx);

/* The density filter as CSR computed on the device: rows in element order,
 * columns ascending, row-scaled weights (TopOpt::P after Assemble_Filter,
 * Filter.cpp:96-150).  rowptr has nelem+1 entries; with col == NULL only rowptr
 * is filled (rowptr[nelem] = nnz). */
int tofe_filter_pattern(tofe_ctx *ctx, int64_t *rowptr, int32_t *col, double *w);

/* Switch the solve precision (TOFE_PREC_*); takes effect at the next assemble. */
int tofe_set_precision(tofe_ctx *ctx, int precision);

/* Statistics of the last assemble/solve: the quantities of FESolve's verbose
 * timing line (FEAnalysis.cpp:687-711). */
int tofe_get_stats(const tofe_ctx *ctx, tofe_stats *stats);

/* Galerkin operator of level l (1 <= l < levels) as a dense 27-point block stencil:
 * out[(o*9 + a*3 + b) * nnode_l + n] is the (a,b) entry of the 3x3 block coupling
 * node n with its neighbour at offset o = (dx+1) + 3(dy+1) + 9(dz+1).  Level 0 is
 * the matrix-free fine operator (returns the assembled stencil of K).
 * Used to check the coarse operators against P^T A P. */
int tofe_get_level_stencil(tofe_ctx *ctx, int level, double *out, int64_t *nnode_l);

/* lambda_max(D^-1 A) estimate of level l used for the Chebyshev bounds. */
int tofe_get_level_lambda(const tofe_ctx *ctx, int level, double *lambda);

/* Average duration (ms) of nrep back-to-back launches of one level operator kernel,
 * timed with HIP events on the context's stream: level 0 is the matrix-free fine
 * operator, level >= 1 the Galerkin stencil; precision TOFE_PREC_FP64 or
 * TOFE_PREC_MIXED (FP32 kernel); epilogue 0 = y = K u, 1 = Chebyshev smoothing step.
 * No reference counterpart: it instruments the kernels that FESolve's timing line
 * (FEAnalysis.cpp:687-711) only sees in aggregate.  Requires tofe_assemble (FP32
 * kernels need an assemble in mixed precision). */
int tofe_time_operator(tofe_ctx *ctx, int level, int precision, int epilogue, int nrep,
                       double *ms);

/* Durations of the level operator kernels as they run inside the calls that follow.
 * start != 0 begins timing: an HIP event pair is recorded around every operator launch on the
 * context's stream (up to 4096 launches; the fine operator is level 0, a Galerkin stencil
 * level >= 1).  start == 0 stops and aggregates them by kind: *nkinds receives the number of
 * kinds, and up to max_kinds of them are written to out.  No reference counterpart: it
 * instruments the kernels that FESolve's timing line (FEAnalysis.cpp:687-711) only sees in
 * aggregate. */
typedef struct tofe_kernel_time {
  int level;        /* 0 fine operator, l >= 1 Galerkin stencil of level l */
  int precision;    /* TOFE_PREC_FP64, or TOFE_PREC_MIXED for an FP32 kernel */
  int epilogue;     /* 0 y = Au; 1 with u.Au; 2 residual b - Au; 3 Chebyshev step;
                       4 with u.Au and u.b; 5 last Chebyshev step with FP64 output */
  int count;        /* launches timed */
  double total_ms;  /* their summed durations */
} tofe_kernel_time;
int tofe_time_kernels(tofe_ctx *ctx, int start, tofe_kernel_time *out, int max_kinds,
                      int *nkinds);

/* Launch shapes from measurement (no reference counterpart: PETSc has no such knob): for
 * the precision of the last assemble, times the fine operator's Chebyshev step over
 * candidate z-chunk lengths and every coarse level's stencil kernel one thread per node
 * and one per DOF (nrep launches each), keeps the fastest chunk length and the node-count
 * threshold below which levels run per DOF that minimises their summed time.  Outputs
 * (may be NULL): the chunk length and the threshold. */
int tofe_autotune(tofe_ctx *ctx, int nrep, int *fine_zlen, int64_t *split_nodes);

/* ---- batches of independent instances ---------------------------------------------
 * Solves of several contexts (uneven meshes, nrhs right-hand sides each) as one graph:
 * after each instance's start (r = F - K U, norms, first preconditioner application),
 * `chunk` iterations of every instance are captured from the instances' streams into one
 * HIP graph, with the convergence test on the device (a converged right-hand side freezes
 * itself), and the graph is replayed until every right-hand side has finished or maxit
 * iterations have run; once an instance has finished, the graph is captured again without
 * it.  Results equal tofe_solve_dev on each instance bit for bit.  No
 * reference counterpart: it replaces a host loop of FESolve's KSPSolve
 * (FEAnalysis.cpp:673) over independent problems.  A context belongs to at most one batch
 * call at a time; the divergence test ||r|| >= 1e5 ||b|| of tofe_solve is not applied. */
typedef struct tofe_batch tofe_batch;
typedef struct tofe_batch_stats {
  double seconds;          /* last tofe_batch_solve_dev, wall time */
  double capture_seconds;  /* of which: capturing and instantiating the graph */
  double replay_seconds;   /* of which: graph replays and their convergence reads */
  int replays;             /* graph launches */
  int iterations;          /* iterations run by the longest-running instance (chunk x replays) */
  int launches;            /* merged form: launches per iteration in the first graph of the solve */
  int unbatched;           /* ... of which serve a single instance */
  int instance_launches;   /* ... and the launches the instances' own iterations make, summed */
} tofe_batch_stats;
int tofe_batch_create(tofe_ctx *const *ctxs, int n, tofe_batch **batch);
int tofe_batch_destroy(tofe_batch *batch);
const char *tofe_batch_last_error(const tofe_batch *batch);
/* d_F[i], d_U[i]: nrhs vectors of instance i's ndof back to back (U is the initial guess);
 * its / reasons (may be NULL): n * nrhs values, instance-major. */
int tofe_batch_solve_dev(tofe_batch *batch, const double *const *d_F, double *const *d_U,
                         int nrhs, double rtol, int maxit, int chunk, int *its, int *reasons);
int tofe_batch_get_stats(const tofe_batch *batch, tofe_batch_stats *stats);

/* ---- on-disk formats ------------------------------------------------------------
 * PETSc binary files as PetscViewerBinaryOpen + VecView / MatView write them (big-endian;
 * Vec: classid 1211214, n, n doubles; sequential AIJ: classid 1211216, rows, cols, nnz,
 * row lengths, column indices, values).  index_bytes is sizeof(PetscInt): 4, or 8 for a
 * PETSc configured with 64-bit indices. */
int tofe_petsc_vec_write(const char *path, const double *v, int64_t n, int index_bytes);
/* v == NULL: *n <- length; otherwise *n is the capacity of v and becomes the length. */
int tofe_petsc_vec_read(const char *path, double *v, int64_t *n, int index_bytes);
int tofe_petsc_aij_write(const char *path, int64_t rows, int64_t cols, const int64_t *rowptr,
                         const int32_t *col, const double *val, int index_bytes);
/* rowptr == NULL: sizes only; otherwise *rows / *nnz are the capacities of rowptr (rows+1)
 * and col/val, and col/val may be NULL to read the row pointers alone. */
int tofe_petsc_aij_read(const char *path, int64_t *rows, int64_t *cols, int64_t *nnz,
                        int64_t *rowptr, int32_t *col, double *val, int index_bytes);

/* TopOpt::MeshOut for one rank (TopOpt.cpp:97-360): Element_Distribution.bin,
 * Node_Distribution.bin, elements.bin, nodes.bin, active.bin (raw native-endian),
 * Filter.bin, Max_Filter.bin and Void_Edge_Volume.bin when rmax > 0, and P<l>.bin for the
 * multigrid interpolations (PETSc binary) under dir.  origin = the mesh's lower corner
 * (may be NULL for 0).  The boundary-condition files belong to the host facade. */
int tofe_mesh_out(tofe_ctx *ctx, const char *dir, const double origin[3], int index_bytes);

/* TopOpt::PrintVals (TopOpt.cpp:442-480): U<suffix>.bin (if U is not NULL), x, V, Es and E
 * of the current design as PETSc binary Vecs under dir. */
int tofe_print_vals(tofe_ctx *ctx, const char *dir, const char *suffix, const double *U,
                    int index_bytes);

#ifdef __cplusplus
}
#endif
#endif /* TOFE_H */
