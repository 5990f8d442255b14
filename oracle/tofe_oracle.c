/*
 * tofe_oracle.c -- FP64 CPU restatement of TopOpt3D's finite-element analysis path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle for the HIP path and the CPU
 * baseline timed by bench.py.  Never linked into or called by the product path.
 * Each function cites the reference file:line it restates.  Built with
 * -ffp-contract=off so no fused multiply-adds change its rounding.
 */
#include "tofe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_seconds(void)
{
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return 0.0;
#endif
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ======================================================================
 * Element stiffness
 * ==================================================================== */

/* Isotropic 3-D constitutive matrix, Voigt order [xx,yy,zz,xy,yz,xz].
 * RecMesh.cpp:453-462. */
void orc_dmat(double E0, double nu, double D[36])
{
  double c = E0 / ((1 + nu) * (1 - 2 * nu));
  double G = E0 / (2 * (1 + nu));
  double t1 = (1 - nu) * c, t2 = G, t3 = nu * c;
  memset(D, 0, 36 * sizeof(double));
  D[0 * 6 + 0] = t1; D[0 * 6 + 1] = t3; D[0 * 6 + 2] = t3;
  D[1 * 6 + 0] = t3; D[1 * 6 + 1] = t1; D[1 * 6 + 2] = t3;
  D[2 * 6 + 0] = t3; D[2 * 6 + 1] = t3; D[2 * 6 + 2] = t1;
  D[3 * 6 + 3] = t2; D[4 * 6 + 4] = t2; D[5 * 6 + 5] = t2;
}

/* Element-local node order (RecMesh.cpp:486-512): (0,0,0),(1,0,0),(1,1,0),(0,1,0),
 * then the same four on the +z layer. */
static const int LN[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                             {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

/* 24x24 Q1 hexahedron stiffness by 2x2x2 Gauss quadrature with unit weights:
 * KE = sum_q B_q^T D B_q det(J).  Row-major, DOF index 3*localnode + component.
 * LocalK FEAnalysis.cpp:869-892, GaussPoints 900-926, dN 936-990, AssignB 1001-1050. */
void orc_ke(double E0, double nu, double hx, double hy, double hz, double KE[576])
{
  double D[36];
  orc_dmat(E0, nu, D);
  const double g = 1 / sqrt(3.0);
  const double h[3] = {hx, hy, hz};
  double coords[8][3];
  for (int n = 0; n < 8; n++)
    for (int d = 0; d < 3; d++) coords[n][d] = h[d] * LN[n][d];
  memset(KE, 0, 576 * sizeof(double));
  for (int q = 0; q < 8; q++) {
    /* Gauss point q has the same sign pattern as local node q (FEAnalysis.cpp:913-922) */
    double xi = (2 * LN[q][0] - 1) * g, eta = (2 * LN[q][1] - 1) * g,
           zeta = (2 * LN[q][2] - 1) * g;
    double dNdxi[3][8];
    for (int n = 0; n < 8; n++) {
      double sx = 2 * LN[n][0] - 1, sy = 2 * LN[n][1] - 1, sz = 2 * LN[n][2] - 1;
      dNdxi[0][n] = sx / 8 * (1 + sy * eta) * (1 + sz * zeta);
      dNdxi[1][n] = sy / 8 * (1 + sx * xi) * (1 + sz * zeta);
      dNdxi[2][n] = sz / 8 * (1 + sx * xi) * (1 + sy * eta);
    }
    double J[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double s = 0;
        for (int n = 0; n < 8; n++) s += dNdxi[a][n] * coords[n][b];
        J[a][b] = s;
      }
    double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                 J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double inv[3][3];
    inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    double dNdx[3][8];
    for (int a = 0; a < 3; a++)
      for (int n = 0; n < 8; n++) {
        double s = 0;
        for (int b = 0; b < 3; b++) s += inv[a][b] * dNdxi[b][n];
        dNdx[a][n] = s;
      }
    /* B (6x24): rows [xx,yy,zz,xy,yz,xz] (AssignB, FEAnalysis.cpp:1024-1046) */
    double B[6][24];
    memset(B, 0, sizeof(B));
    for (int n = 0; n < 8; n++) {
      B[0][3 * n + 0] = dNdx[0][n];
      B[1][3 * n + 1] = dNdx[1][n];
      B[2][3 * n + 2] = dNdx[2][n];
      B[3][3 * n + 0] = dNdx[1][n];
      B[3][3 * n + 1] = dNdx[0][n];
      B[4][3 * n + 1] = dNdx[2][n];
      B[4][3 * n + 2] = dNdx[1][n];
      B[5][3 * n + 0] = dNdx[2][n];
      B[5][3 * n + 2] = dNdx[0][n];
    }
    double DB[6][24];
    for (int a = 0; a < 6; a++)
      for (int c = 0; c < 24; c++) {
        double s = 0;
        for (int b = 0; b < 6; b++) s += D[a * 6 + b] * B[b][c];
        DB[a][c] = s;
      }
    for (int r = 0; r < 24; r++)
      for (int c = 0; c < 24; c++) {
        double s = 0;
        for (int a = 0; a < 6; a++) s += B[a][r] * DB[a][c];
        KE[r * 24 + c] += s * det;
      }
  }
}

/* Node ids of element el, element-local order; x-fastest numbering of both
 * elements and nodes (RecMesh.cpp:486-512, 565-597). */
void orc_element_nodes(int nx, int ny, int nz, int64_t el, int64_t nodes[8])
{
  (void)nz;
  int64_t ex = el % nx, ey = (el / nx) % ny, ez = el / ((int64_t)nx * ny);
  int64_t Nx = nx + 1, Ny = ny + 1;
  for (int l = 0; l < 8; l++)
    nodes[l] = (ex + LN[l][0]) + Nx * ((ey + LN[l][1]) + Ny * (ez + LN[l][2]));
}

/* ======================================================================
 * Density filter: RecFilter (Filter.cpp:21-94) + Assemble_Filter (96-150)
 * ==================================================================== */

typedef struct {
  int n;            /* template entries with dist < R */
  int di[1024], dj[1024], dk[1024];
  double w[1024];   /* 1 - dist/R */
} filt_template;

static int make_template(double dx, double dy, double dz, double R, filt_template *t)
{
  short N[3];
  N[0] = (short)(R / dx);   /* Filter.cpp:32-35: truncating conversion to short */
  N[1] = (short)(R / dy);
  N[2] = (short)(R / dz);
  t->n = 0;
  for (int k = -N[2]; k < N[2] + 1; k++)
    for (int j = -N[1]; j < N[1] + 1; j++)
      for (int i = -N[0]; i < N[0] + 1; i++) {
        /* sqrt(pow(i*dx,2) + pow(j*dx,2) + pow(k*dx,2)) (47); compilers fold pow(v,2)
         * to the correctly rounded v*v, written out here */
        double vi = i * dx, vj = j * dy, vk = k * dz;
        double dist = sqrt(vi * vi + vj * vj + vk * vk);
        if (dist < R) {                                                       /* 48 */
          if (t->n >= 1024) return -1;
          t->di[t->n] = i; t->dj[t->n] = j; t->dk[t->n] = k;
          t->w[t->n] = 1 - dist / R;                                          /* 80 */
          t->n++;
        }
      }
  return 0;
}

/* Filter rows in element order, columns ascending (the template is walked k,j,i).
 * With scale != 0 each row is multiplied by the reciprocal of its sequential
 * row sum (MatGetRowSum, VecPointwiseDivide(ones/rowsum), MatDiagonalScale,
 * Filter.cpp:136-147).  Pass NULL arrays to count; returns nnz (or -1). */
int64_t orc_filter_build(int nx, int ny, int nz, double dx, double dy, double dz,
                         double R, int scale, int64_t *rowptr, int32_t *col, double *w)
{
  filt_template t;
  if (make_template(dx, dy, dz, R, &t)) return -1;
  int64_t ne = (int64_t)nx * ny * nz;
  int64_t *cnt = (int64_t *)malloc((ne + 1) * sizeof(int64_t));
#pragma omp parallel for schedule(static)
  for (int64_t el = 0; el < ne; el++) {
    int ex = (int)(el % nx), ey = (int)((el / nx) % ny), ez = (int)(el / ((int64_t)nx * ny));
    int64_t c = 0;
    for (int q = 0; q < t.n; q++) {
      int i = ex + t.di[q], j = ey + t.dj[q], k = ez + t.dk[q];
      if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) c++;
    }
    cnt[el + 1] = c;
  }
  cnt[0] = 0;
  for (int64_t el = 0; el < ne; el++) cnt[el + 1] += cnt[el];
  int64_t nnz = cnt[ne];
  if (rowptr) memcpy(rowptr, cnt, (ne + 1) * sizeof(int64_t));
  if (col && w) {
#pragma omp parallel for schedule(static)
    for (int64_t el = 0; el < ne; el++) {
      int ex = (int)(el % nx), ey = (int)((el / nx) % ny), ez = (int)(el / ((int64_t)nx * ny));
      int64_t p = cnt[el];
      for (int q = 0; q < t.n; q++) {
        int i = ex + t.di[q], j = ey + t.dj[q], k = ez + t.dk[q];
        if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
          col[p] = (int32_t)(el + t.di[q] + (int64_t)t.dj[q] * nx + (int64_t)t.dk[q] * nx * ny);
          w[p] = t.w[q];
          p++;
        }
      }
      if (scale) {
        double s = 0.0;
        for (int64_t a = cnt[el]; a < cnt[el + 1]; a++) s += w[a];
        double inv = 1.0 / s;
        for (int64_t a = cnt[el]; a < cnt[el + 1]; a++) w[a] = w[a] * inv;
      }
    }
  }
  free(cnt);
  return nnz;
}

orc_csr *orc_filter_csr(int nx, int ny, int nz, double dx, double dy, double dz,
                        double R, int scale)
{
  int64_t ne = (int64_t)nx * ny * nz;
  int64_t nnz = orc_filter_build(nx, ny, nz, dx, dy, dz, R, scale, NULL, NULL, NULL);
  if (nnz < 0) return NULL;
  orc_csr *A = (orc_csr *)calloc(1, sizeof(orc_csr));
  A->n = A->m = ne;
  A->nnz = nnz;
  A->rowptr = (int64_t *)malloc((ne + 1) * sizeof(int64_t));
  A->col = (int32_t *)malloc((nnz ? nnz : 1) * sizeof(int32_t));
  A->val = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
  orc_filter_build(nx, ny, nz, dx, dy, dz, R, scale, A->rowptr, A->col, A->val);
  return A;
}

/* ======================================================================
 * SIMP interpolation (Material_Interpolation.cpp:20-75, plain-filter branch y=1)
 * ==================================================================== */

/* VecPow with PETSc's special exponents; pow() otherwise. */
double orc_vecpow1(double x, double p)
{
  if (p == 1.0) return x;
  if (p == 0.0) return 1.0;
  if (p == -1.0) return 1.0 / x;
  if (p == 0.5) return sqrt(x);
  if (p == -0.5) return 1.0 / sqrt(x);
  if (p == 2.0) return x * x;
  if (p == -2.0) return 1.0 / (x * x);
  return pow(x, p);
}

void orc_simp(int64_t ne, const double *rho, double penal, double eps,
              double *E, double *dEdz, double *Es, double *dEsdz)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ne; i++) {
    double z = rho[i] * 1.0;                 /* z = rho .* y, y = 1   (61) */
    double d = orc_vecpow1(z, penal - 1);    /* dEsdz = z^(p-1)       (67-68) */
    double es = d * z;                       /* Es = dEsdz .* z       (69) */
    d = d * penal;                           /* dEsdz *= p            (70) */
    if (Es) Es[i] = es;
    if (dEsdz) dEsdz[i] = d;
    if (dEdz) dEdz[i] = d * (1 - eps);       /* dEdz = (1-eps) dEsdz  (72-73) */
    if (E) E[i] = eps + (1 - eps) * es;      /* E = eps + (1-eps) Es  (74-75) */
  }
}

/* Maximum-length-scale filter R: the same neighbourhood walk with radius Rmax and unit
 * weights (RecFilter(..., nonzeros = 1), RecMesh.cpp:544; assembled unscaled, 706), and
 * REdge = max_i rowsum_i - rowsum_i (RecMesh.cpp:710-718). */
orc_csr *orc_rfilter_csr(int nx, int ny, int nz, double dx, double dy, double dz, double Rmax,
                         double *REdge)
{
  orc_csr *A = orc_filter_csr(nx, ny, nz, dx, dy, dz, Rmax, 0);
  if (!A) return NULL;
  for (int64_t p = 0; p < A->nnz; p++) A->val[p] = 1.0;
  if (REdge) {
    double mx = 0;
    for (int64_t i = 0; i < A->n; i++) {
      double s = 0.0;
      for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; p++) s += A->val[p];
      REdge[i] = s;
      if (s > mx) mx = s;
    }
    for (int64_t i = 0; i < A->n; i++) REdge[i] = -1 * REdge[i] + mx; /* VecAYPX(REdge,-1,max) */
  }
  return A;
}

/* TopOpt::MatIntFnc (Material_Interpolation.cpp:20-140), all branches:
 *   rho = P x; with vdMin > 0: rho = min(rho,1), y = min(1, (REdge + R (1-rho)^q) / vdMin),
 *   otherwise y = 1; rhoq = (q/vdMin) (1-rho)^(q-1); z = rho y; Es = z^p, dEsdz = p z^(p-1),
 *   E = eps + (1-eps) Es, dEdz = (1-eps) dEsdz; then SIMP_CUT (interp 1) zeroes Es, dEsdz
 *   where z <= c, SIMP_LOGISTIC (interp 2) divides by 1 + exp(a (b - z)).
 * R/REdge may be NULL when vdMin <= 0. */
void orc_matint(const orc_csr *P, const orc_csr *R, const double *REdge, const double *x,
                double penal, double vdPenal, double vdMin, int interp, const double *param,
                double eps, double *rho, double *y, double *rhoq, double *E, double *dEdz,
                double *Es, double *dEsdz)
{
  int64_t ne = P->n;
  orc_spmv(P, x, rho);
  if (vdMin > 0) {
    double *zt = (double *)malloc(ne * sizeof(double));
    for (int64_t i = 0; i < ne; i++) {
      if (rho[i] > 1.0) rho[i] = 1.0;                       /* VecPointwiseMin(rho,rho,1) */
      zt[i] = orc_vecpow1(1.0 + -1 * rho[i], vdPenal);       /* z = (1-rho)^q */
    }
    double inv = 1 / vdMin;
    for (int64_t i = 0; i < ne; i++) {
      double s = REdge[i];                                   /* MatMultAdd(R, z, REdge, y) */
      for (int64_t p = R->rowptr[i]; p < R->rowptr[i + 1]; p++) s += R->val[p] * zt[R->col[p]];
      s = s * inv;                                           /* VecScale(y, 1/vdMin) */
      y[i] = s < 1.0 ? s : 1.0;                              /* VecPointwiseMin(y, y, 1) */
    }
    free(zt);
  } else {
    for (int64_t i = 0; i < ne; i++) y[i] = 1.0;
  }
  double sc = vdPenal / vdMin;
  for (int64_t i = 0; i < ne; i++) {
    rhoq[i] = orc_vecpow1(1.0 + -1 * rho[i], vdPenal - 1) * sc;
    double z = rho[i] * y[i];
    double d = orc_vecpow1(z, penal - 1);
    double es = d * z;
    d = d * penal;
    dEdz[i] = d * (1 - eps);
    E[i] = eps + (1 - eps) * es;
    if (interp == 1) {
      double cut = (z > param[0]) ? 1.0 : 0.0;
      es = es * cut;
      d = d * cut;
    } else if (interp == 2) {
      double denom = 1 + exp(param[0] * (param[1] - z));
      es = es / denom;
      d = (d + (es * param[0]) * (1 - 1 / denom)) / denom;
    }
    Es[i] = es;
    dEsdz[i] = d;
  }
}

/* TopOpt::Chain_Filter (Material_Interpolation.cpp:150-187), in place:
 *   dfdV <- P^T dfdV;
 *   dfdE <- P^T [dfdE .* y - rhoq .* (y < 1) .* R^T (dfdE .* rho)]
 * exactly as written there (the (y < 1) mask multiplies R^T(...) entrywise). */
void orc_chain_filter(const orc_csr *P, const orc_csr *R, const double *rho, const double *y,
                      const double *rhoq, double *dfdV, double *dfdE)
{
  int64_t ne = P->n;
  double *t1 = (double *)malloc(ne * sizeof(double));
  double *t2 = (double *)malloc(ne * sizeof(double));
  if (dfdV) {
    orc_spmv_t(P, dfdV, t1);
    memcpy(dfdV, t1, ne * sizeof(double));
  }
  if (dfdE) {
    for (int64_t i = 0; i < ne; i++) t2[i] = dfdE[i] * rho[i];
    if (R) orc_spmv_t(R, t2, t1);
    else memset(t1, 0, ne * sizeof(double));
    for (int64_t i = 0; i < ne; i++) {
      double mask = (y[i] < 1.0) ? 1.0 : 0.0;
      double a = t1[i] * mask;
      a = rhoq[i] * a;
      double b = dfdE[i] * y[i];
      t2[i] = b - a;                                   /* VecAYPX(temp1, -1, dfdE) */
    }
    orc_spmv_t(P, t2, dfdE);
  }
  free(t1);
  free(t2);
}

/* ======================================================================
 * Sparse matrices
 * ==================================================================== */

static orc_csr *csr_alloc(int64_t n, int64_t m, int64_t nnz)
{
  orc_csr *A = (orc_csr *)calloc(1, sizeof(orc_csr));
  A->n = n; A->m = m; A->nnz = nnz;
  A->rowptr = (int64_t *)calloc(n + 1, sizeof(int64_t));
  A->col = (int32_t *)malloc((nnz ? nnz : 1) * sizeof(int32_t));
  A->val = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
  return A;
}

void orc_csr_free(orc_csr *A)
{
  if (!A) return;
  free(A->rowptr); free(A->col); free(A->val); free(A);
}

void orc_csr_info(const orc_csr *A, int64_t *n, int64_t *m, int64_t *nnz)
{
  *n = A->n; *m = A->m; *nnz = A->nnz;
}

void orc_csr_export(const orc_csr *A, int64_t *rowptr, int32_t *col, double *val)
{
  if (rowptr) memcpy(rowptr, A->rowptr, (A->n + 1) * sizeof(int64_t));
  if (col) memcpy(col, A->col, A->nnz * sizeof(int32_t));
  if (val) memcpy(val, A->val, A->nnz * sizeof(double));
}

orc_csr *orc_csr_import(int64_t n, int64_t m, const int64_t *rowptr,
                        const int32_t *col, const double *val)
{
  orc_csr *A = csr_alloc(n, m, rowptr[n]);
  memcpy(A->rowptr, rowptr, (n + 1) * sizeof(int64_t));
  memcpy(A->col, col, A->nnz * sizeof(int32_t));
  memcpy(A->val, val, A->nnz * sizeof(double));
  return A;
}

/* y = A x, sequential sum per row from 0.0 in ascending column order
 * (PETSc MatMult_SeqAIJ). */
void orc_spmv(const orc_csr *A, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < A->n; i++) {
    double s = 0.0;
    for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; p++) s += A->val[p] * x[A->col[p]];
    y[i] = s;
  }
}

/* y = A^T x, rows visited in ascending order (MatMultTranspose_SeqAIJ). */
void orc_spmv_t(const orc_csr *A, const double *x, double *y)
{
  memset(y, 0, A->m * sizeof(double));
  for (int64_t i = 0; i < A->n; i++)
    for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; p++) y[A->col[p]] += A->val[p] * x[i];
}

orc_csr *orc_transpose(const orc_csr *A)
{
  orc_csr *T = csr_alloc(A->m, A->n, A->nnz);
  for (int64_t p = 0; p < A->nnz; p++) T->rowptr[A->col[p] + 1]++;
  for (int64_t i = 0; i < A->m; i++) T->rowptr[i + 1] += T->rowptr[i];
  int64_t *fill = (int64_t *)malloc(A->m * sizeof(int64_t));
  memcpy(fill, T->rowptr, A->m * sizeof(int64_t));
  for (int64_t i = 0; i < A->n; i++)
    for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; p++) {
      int64_t q = fill[A->col[p]]++;
      T->col[q] = (int32_t)i;
      T->val[q] = A->val[p];
    }
  free(fill);
  return T;
}

static void sort_row(int32_t *c, double *v, int64_t n)
{
  for (int64_t i = 1; i < n; i++) {
    int32_t ci = c[i]; double vi = v[i];
    int64_t j = i - 1;
    while (j >= 0 && c[j] > ci) { c[j + 1] = c[j]; v[j + 1] = v[j]; j--; }
    c[j + 1] = ci; v[j + 1] = vi;
  }
}

/* C = A B by Gustavson's row-by-row algorithm (two passes, sorted rows). */
orc_csr *orc_spgemm(const orc_csr *A, const orc_csr *B)
{
  int64_t n = A->n, m = B->m;
  int64_t *cnt = (int64_t *)calloc(n + 1, sizeof(int64_t));
#pragma omp parallel
  {
    int64_t *mark = (int64_t *)malloc(m * sizeof(int64_t));
    for (int64_t j = 0; j < m; j++) mark[j] = -1;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; i++) {
      int64_t c = 0;
      for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; p++) {
        int64_t k = A->col[p];
        for (int64_t q = B->rowptr[k]; q < B->rowptr[k + 1]; q++) {
          int64_t j = B->col[q];
          if (mark[j] != i) { mark[j] = i; c++; }
        }
      }
      cnt[i + 1] = c;
    }
    free(mark);
  }
  for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  orc_csr *C = csr_alloc(n, m, cnt[n]);
  memcpy(C->rowptr, cnt, (n + 1) * sizeof(int64_t));
  free(cnt);
#pragma omp parallel
  {
    int64_t *pos = (int64_t *)malloc(m * sizeof(int64_t));
    for (int64_t j = 0; j < m; j++) pos[j] = -1;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; i++) {
      int64_t base = C->rowptr[i], c = 0;
      for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; p++) {
        int64_t k = A->col[p];
        double a = A->val[p];
        for (int64_t q = B->rowptr[k]; q < B->rowptr[k + 1]; q++) {
          int64_t j = B->col[q];
          if (pos[j] < base) {
            pos[j] = base + c;
            C->col[base + c] = (int32_t)j;
            C->val[base + c] = a * B->val[q];
            c++;
          } else {
            C->val[pos[j]] += a * B->val[q];
          }
        }
      }
      sort_row(C->col + base, C->val + base, c);
    }
    free(pos);
  }
  return C;
}

/* Galerkin coarse operator P^T A P (PCMGSetGalerkin PC_MG_GALERKIN_BOTH,
 * FEAnalysis.cpp:412). */
orc_csr *orc_ptap(const orc_csr *A, const orc_csr *P)
{
  orc_csr *AP = orc_spgemm(A, P);
  orc_csr *PT = orc_transpose(P);
  orc_csr *C = orc_spgemm(PT, AP);
  orc_csr_free(AP);
  orc_csr_free(PT);
  return C;
}

/* ======================================================================
 * Stiffness assembly with Dirichlet rows/columns (FEAnalysis.cpp:432-490)
 * ==================================================================== */

static int local_index(int lx, int ly, int lz)
{
  for (int l = 0; l < 8; l++)
    if (LN[l][0] == lx && LN[l][1] == ly && LN[l][2] == lz) return l;
  return -1;
}

/* K = sum_e E_e KE scattered by 3x3 blocks, each entry summed in ascending
 * element order from 0 (MatZeroEntries + ADD_VALUES, 446-464).  Then
 * MatZeroRowsColumns(K, fixed, 1.0): fixed rows and columns zero, unit
 * diagonal (473); any diagonal <= 0 becomes 1 (476-484).  fixed_dof may be NULL. */
orc_csr *orc_assemble_K(int nx, int ny, int nz, const double KE[576],
                        const double *E, const uint8_t *fixed_dof)
{
  int64_t Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
  int64_t nn = Nx * Ny * Nz, ndof = 3 * nn;
  int64_t *cnt = (int64_t *)calloc(ndof + 1, sizeof(int64_t));
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < nn; n++) {
    int64_t i = n % Nx, j = (n / Nx) % Ny, k = n / (Nx * Ny);
    int64_t c = 0;
    for (int dk = -1; dk <= 1; dk++)
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++)
          if (i + di >= 0 && i + di < Nx && j + dj >= 0 && j + dj < Ny && k + dk >= 0 &&
              k + dk < Nz)
            c++;
    for (int a = 0; a < 3; a++) cnt[3 * n + a + 1] = 3 * c;
  }
  for (int64_t r = 0; r < ndof; r++) cnt[r + 1] += cnt[r];
  orc_csr *K = csr_alloc(ndof, ndof, cnt[ndof]);
  memcpy(K->rowptr, cnt, (ndof + 1) * sizeof(int64_t));
  free(cnt);
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < nn; n++) {
    int64_t i = n % Nx, j = (n / Nx) % Ny, k = n / (Nx * Ny);
    /* neighbour slots in ascending node order */
    int64_t nb[27]; int slot[27]; int cntnb = 0;
    for (int dk = -1; dk <= 1; dk++)
      for (int dj = -1; dj <= 1; dj++)
        for (int di = -1; di <= 1; di++) {
          int o = (di + 1) + 3 * (dj + 1) + 9 * (dk + 1);
          if (i + di >= 0 && i + di < Nx && j + dj >= 0 && j + dj < Ny && k + dk >= 0 &&
              k + dk < Nz) {
            nb[cntnb] = n + di + Nx * (dj + Ny * dk);
            slot[o] = cntnb++;
          } else {
            slot[o] = -1;
          }
        }
    for (int a = 0; a < 3; a++) {
      int64_t base = K->rowptr[3 * n + a];
      for (int s = 0; s < cntnb; s++)
        for (int b = 0; b < 3; b++) {
          K->col[base + 3 * s + b] = (int32_t)(3 * nb[s] + b);
          K->val[base + 3 * s + b] = 0.0;
        }
    }
    /* elements touching node n in ascending element order: (ez, ey, ex) ascending */
    for (int ez = (int)k - 1; ez <= (int)k; ez++)
      for (int ey = (int)j - 1; ey <= (int)j; ey++)
        for (int ex = (int)i - 1; ex <= (int)i; ex++) {
          if (ex < 0 || ex >= nx || ey < 0 || ey >= ny || ez < 0 || ez >= nz) continue;
          double Ee = E[ex + (int64_t)nx * (ey + (int64_t)ny * ez)];
          int ln = local_index((int)i - ex, (int)j - ey, (int)k - ez);
          for (int lm = 0; lm < 8; lm++) {
            int mx = ex + LN[lm][0], my = ey + LN[lm][1], mz = ez + LN[lm][2];
            int o = (mx - (int)i + 1) + 3 * (my - (int)j + 1) + 9 * (mz - (int)k + 1);
            int s = slot[o];
            for (int a = 0; a < 3; a++)
              for (int b = 0; b < 3; b++) {
                double ke = Ee * KE[(3 * ln + a) * 24 + 3 * lm + b];
                K->val[K->rowptr[3 * n + a] + 3 * s + b] += ke;
              }
          }
        }
  }
  if (fixed_dof) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < ndof; r++) {
      for (int64_t p = K->rowptr[r]; p < K->rowptr[r + 1]; p++) {
        int64_t c = K->col[p];
        if (fixed_dof[r] || fixed_dof[c]) K->val[p] = (r == c) ? 1.0 : 0.0;
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < ndof; r++)
    for (int64_t p = K->rowptr[r]; p < K->rowptr[r + 1]; p++)
      if (K->col[p] == r && !(K->val[p] > 0)) K->val[p] = 1.0;
  return K;
}

/* ======================================================================
 * Transfer operators (Interpolation.cpp:123-169, 322-347)
 * ==================================================================== */

/* Interpolation from the coarse node grid Nc = (Nf+1)/2 (every second fine node)
 * to the fine node grid, 3x3 identity blocks per node pair.  Unscaled weights are
 * vals(fi+1)*vals(fj+1)*vals(fk+1), vals = (0.5, 1, 0.5) (135, 157); each row is
 * then multiplied by the reciprocal of its sequential row sum (331-347). */
orc_csr *orc_build_P(int Nfx, int Nfy, int Nfz)
{
  int64_t Ncx = (Nfx + 1) / 2, Ncy = (Nfy + 1) / 2, Ncz = (Nfz + 1) / 2;
  int64_t nf = (int64_t)Nfx * Nfy * Nfz, nc = Ncx * Ncy * Ncz;
  static const double vals[3] = {0.5, 1.0, 0.5};
  /* per fine node: parents in each dim */
  int64_t *cnt = (int64_t *)calloc(3 * nf + 1, sizeof(int64_t));
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < nf; f++) {
    int64_t ix = f % Nfx, iy = (f / Nfx) % Nfy, iz = f / ((int64_t)Nfx * Nfy);
    int px = 0, py = 0, pz = 0;
    for (int64_t J = ix / 2 - 1; J <= ix / 2 + 1; J++) if (J >= 0 && J < Ncx && llabs(ix - 2 * J) <= 1) px++;
    for (int64_t J = iy / 2 - 1; J <= iy / 2 + 1; J++) if (J >= 0 && J < Ncy && llabs(iy - 2 * J) <= 1) py++;
    for (int64_t J = iz / 2 - 1; J <= iz / 2 + 1; J++) if (J >= 0 && J < Ncz && llabs(iz - 2 * J) <= 1) pz++;
    for (int a = 0; a < 3; a++) cnt[3 * f + a + 1] = px * py * pz;
  }
  for (int64_t r = 0; r < 3 * nf; r++) cnt[r + 1] += cnt[r];
  orc_csr *P = csr_alloc(3 * nf, 3 * nc, cnt[3 * nf]);
  memcpy(P->rowptr, cnt, (3 * nf + 1) * sizeof(int64_t));
  free(cnt);
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < nf; f++) {
    int64_t ix = f % Nfx, iy = (f / Nfx) % Nfy, iz = f / ((int64_t)Nfx * Nfy);
    int64_t cols[8]; double w[8]; int c = 0;
    for (int64_t Jz = iz / 2 - 1; Jz <= iz / 2 + 1; Jz++) {
      int64_t tz = iz - 2 * Jz;
      if (Jz < 0 || Jz >= Ncz || tz < -1 || tz > 1) continue;
      for (int64_t Jy = iy / 2 - 1; Jy <= iy / 2 + 1; Jy++) {
        int64_t ty = iy - 2 * Jy;
        if (Jy < 0 || Jy >= Ncy || ty < -1 || ty > 1) continue;
        for (int64_t Jx = ix / 2 - 1; Jx <= ix / 2 + 1; Jx++) {
          int64_t tx = ix - 2 * Jx;
          if (Jx < 0 || Jx >= Ncx || tx < -1 || tx > 1) continue;
          cols[c] = Jx + Ncx * (Jy + Ncy * Jz);
          /* coarse node at fine index 2J sees fine offset fi = i - 2J */
          w[c] = vals[tx + 1] * vals[ty + 1] * vals[tz + 1];
          c++;
        }
      }
    }
    double s = 0.0;
    for (int q = 0; q < c; q++) s += w[q];
    double inv = 1.0 / s;
    for (int a = 0; a < 3; a++) {
      int64_t base = P->rowptr[3 * f + a];
      for (int q = 0; q < c; q++) {
        P->col[base + q] = (int32_t)(3 * cols[q] + a);
        P->val[base + q] = w[q] * inv;
      }
    }
  }
  return P;
}

/* ======================================================================
 * Multigrid-preconditioned CG
 * ==================================================================== */

struct orc_mg {
  int L;
  orc_csr *A[16], *P[16], *PT[16];
  int dims[16][3];
  double *dinv[16], *diag[16];
  double lam[16];
  double *b[16], *x[16], *r[16], *d[16], *t[16];
  double *chol;     /* dense lower Cholesky factor of the coarsest operator */
  int64_t nc;
  int smoother, nu1, nu2, eig_its;
  double lo, hi;
  double setup_s;
};

int orc_mg_levels(const orc_mg *mg) { return mg->L; }
const orc_csr *orc_mg_A(const orc_mg *mg, int l) { return mg->A[l]; }
const orc_csr *orc_mg_P(const orc_mg *mg, int l) { return mg->P[l]; }
double orc_mg_lambda(const orc_mg *mg, int l) { return mg->lam[l]; }
double orc_mg_setup_seconds(const orc_mg *mg) { return mg->setup_s; }
void orc_mg_dims(const orc_mg *mg, int l, int dims[3])
{
  dims[0] = mg->dims[l][0]; dims[1] = mg->dims[l][1]; dims[2] = mg->dims[l][2];
}

double orc_dot(int64_t n, const double *a, const double *b)
{
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* Deterministic power-iteration start vector shared with the GPU path:
 * v_i = 0.5 + frac(i * 2654435761 / 2^32). */
void orc_power_start(int64_t n, double *v)
{
  for (int64_t i = 0; i < n; i++) {
    uint32_t h = (uint32_t)((uint64_t)i * 2654435761ull);
    v[i] = 0.5 + (double)h / 4294967296.0;
  }
}

/* Largest eigenvalue of the symmetric tridiagonal matrix (diag a, off-diagonal b) by
 * Sturm-sequence bisection on the Gershgorin interval. */
double orc_tridiag_max_eig(int k, const double *a, const double *b)
{
  double lo = a[0], hi = a[0];
  for (int j = 0; j < k; j++) {
    double r = (j > 0 ? fabs(b[j - 1]) : 0) + (j + 1 < k ? fabs(b[j]) : 0);
    if (a[j] - r < lo) lo = a[j] - r;
    if (a[j] + r > hi) hi = a[j] + r;
  }
  for (int it = 0; it < 200; it++) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    /* number of eigenvalues < mid */
    int cnt = 0;
    double q = a[0] - mid;
    if (q < 0) cnt++;
    for (int j = 1; j < k; j++) {
      if (q == 0) q = 1e-300;
      q = a[j] - mid - b[j - 1] * b[j - 1] / q;
      if (q < 0) cnt++;
    }
    if (cnt == k) hi = mid; else lo = mid;
  }
  return hi;
}

/* lambda_max(D^-1 A) from `its` steps of Jacobi-preconditioned CG started from the
 * deterministic vector of orc_power_start: the CG coefficients give the Lanczos
 * tridiagonal T (T_jj = 1/alpha_j + beta_{j-1}/alpha_{j-1}, T_j,j+1 = sqrt(beta_j)/alpha_j)
 * whose largest eigenvalue estimates lambda_max -- the estimate PETSc's Chebyshev
 * smoother forms with KSPChebyshevEstEigSet (transform 0, 0.1, 0, 1.1). */
double orc_lanczos_lambda(const orc_csr *A, const double *dinv, int its)
{
  int64_t n = A->n;
  double *r = (double *)malloc(n * sizeof(double)), *z = (double *)malloc(n * sizeof(double));
  double *p = (double *)malloc(n * sizeof(double)), *q = (double *)malloc(n * sizeof(double));
  double alpha[256], beta[256], ta[256], tb[256];
  if (its > 256) its = 256;
  orc_power_start(n, r);
  for (int64_t i = 0; i < n; i++) { z[i] = dinv[i] * r[i]; p[i] = z[i]; }
  double rz = orc_dot(n, r, z);
  int k = 0;
  for (; k < its; k++) {
    orc_spmv(A, p, q);
    double pq = orc_dot(n, p, q);
    if (!(pq > 0) || !(rz > 0)) break;
    alpha[k] = rz / pq;
    for (int64_t i = 0; i < n; i++) { r[i] -= alpha[k] * q[i]; z[i] = dinv[i] * r[i]; }
    double rz_new = orc_dot(n, r, z);
    beta[k] = rz_new / rz;
    rz = rz_new;
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta[k] * p[i];
  }
  free(r); free(z); free(p); free(q);
  if (k == 0) return 1.0;
  for (int j = 0; j < k; j++) {
    ta[j] = 1.0 / alpha[j] + (j > 0 ? beta[j - 1] / alpha[j - 1] : 0.0);
    tb[j] = sqrt(beta[j]) / alpha[j];
  }
  return orc_tridiag_max_eig(k, ta, tb);
}

static int dense_cholesky(double *A, int64_t n)
{
  for (int64_t k = 0; k < n; k++) {
    double akk = A[k * n + k];
    if (!(akk > 0)) return (int)k + 1;
    akk = sqrt(akk);
    A[k * n + k] = akk;
    for (int64_t i = k + 1; i < n; i++) A[i * n + k] /= akk;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = k + 1; i < n; i++) {
      double lik = A[i * n + k];
      if (lik == 0.0) continue;
      for (int64_t j = k + 1; j <= i; j++) A[i * n + j] -= lik * A[j * n + k];
    }
  }
  return 0;
}

static void cholesky_solve(const double *Lf, int64_t n, const double *b, double *x)
{
  for (int64_t i = 0; i < n; i++) {
    double s = b[i];
    for (int64_t j = 0; j < i; j++) s -= Lf[i * n + j] * x[j];
    x[i] = s / Lf[i * n + i];
  }
  for (int64_t i = n - 1; i >= 0; i--) {
    double s = x[i];
    for (int64_t j = i + 1; j < n; j++) s -= Lf[j * n + i] * x[j];
    x[i] = s / Lf[i * n + i];
  }
}

orc_mg *orc_mg_create(const orc_csr *K, int Nx, int Ny, int Nz, int levels,
                      int smoother, int nu1, int nu2, double lo, double hi, int eig_its)
{
  if (levels < 1 || levels > 16) return NULL;
  double t0 = now_seconds();
  orc_mg *mg = (orc_mg *)calloc(1, sizeof(orc_mg));
  mg->L = levels;
  mg->smoother = smoother; mg->nu1 = nu1; mg->nu2 = nu2;
  mg->lo = lo; mg->hi = hi; mg->eig_its = eig_its;
  mg->A[0] = orc_csr_import(K->n, K->m, K->rowptr, K->col, K->val);
  mg->dims[0][0] = Nx; mg->dims[0][1] = Ny; mg->dims[0][2] = Nz;
  for (int l = 0; l + 1 < levels; l++) {
    mg->P[l] = orc_build_P(mg->dims[l][0], mg->dims[l][1], mg->dims[l][2]);
    mg->PT[l] = orc_transpose(mg->P[l]);
    mg->A[l + 1] = orc_ptap(mg->A[l], mg->P[l]);
    for (int d = 0; d < 3; d++) mg->dims[l + 1][d] = (mg->dims[l][d] + 1) / 2;
  }
  for (int l = 0; l < levels; l++) {
    int64_t n = mg->A[l]->n;
    mg->b[l] = (double *)calloc(n, sizeof(double));
    mg->x[l] = (double *)calloc(n, sizeof(double));
    mg->r[l] = (double *)calloc(n, sizeof(double));
    mg->d[l] = (double *)calloc(n, sizeof(double));
    mg->t[l] = (double *)calloc(n, sizeof(double));
    mg->diag[l] = (double *)calloc(n, sizeof(double));
    mg->dinv[l] = (double *)calloc(n, sizeof(double));
    const orc_csr *A = mg->A[l];
    for (int64_t i = 0; i < n; i++)
      for (int64_t p = A->rowptr[i]; p < A->rowptr[i + 1]; p++)
        if (A->col[p] == i) mg->diag[l][i] = A->val[p];
    for (int64_t i = 0; i < n; i++) mg->dinv[l][i] = 1.0 / mg->diag[l][i];
    if (l + 1 < levels) mg->lam[l] = orc_lanczos_lambda(A, mg->dinv[l], eig_its);
  }
  /* coarse direct solve: dense Cholesky (PCBJACOBI + PCLU, FEAnalysis.cpp:555-573) */
  const orc_csr *Ac = mg->A[levels - 1];
  int64_t nc = Ac->n;
  mg->nc = nc;
  mg->chol = (double *)calloc(nc * nc, sizeof(double));
  for (int64_t i = 0; i < nc; i++)
    for (int64_t p = Ac->rowptr[i]; p < Ac->rowptr[i + 1]; p++) mg->chol[i * nc + Ac->col[p]] = Ac->val[p];
  if (dense_cholesky(mg->chol, nc)) {
    orc_mg_free(mg);
    return NULL;
  }
  mg->setup_s = now_seconds() - t0;
  return mg;
}

void orc_mg_free(orc_mg *mg)
{
  if (!mg) return;
  for (int l = 0; l < mg->L; l++) {
    orc_csr_free(mg->A[l]);
    if (l + 1 < mg->L) { orc_csr_free(mg->P[l]); orc_csr_free(mg->PT[l]); }
    free(mg->b[l]); free(mg->x[l]); free(mg->r[l]); free(mg->d[l]); free(mg->t[l]);
    free(mg->diag[l]); free(mg->dinv[l]);
  }
  free(mg->chol);
  free(mg);
}

/* `steps` steps of the Chebyshev iteration on D^-1 A over [lo*lam, hi*lam] (smoother 0) or
 * of damped Jacobi with omega = 2/(emin+emax) (smoother 1), updating x in place; step 0 from
 * x = 0 (zero_guess) needs no operator apply.  d and r are n-vectors of scratch.  This is the
 * level smoother of the PCMG V-cycle (FEAnalysis.cpp:397-416 selects PCMG; PETSc's
 * KSPChebyshev with the eigenvalue transform 0.1/1.1 is restated from its documentation). */
void orc_smooth(const orc_csr *A, const double *dinv, double lam, double lo, double hi,
                int smoother, int steps, int zero_guess, const double *b, double *x, double *d,
                double *r)
{
  int64_t n = A->n;
  double emax = hi * lam, emin = lo * lam;
  double theta = 0.5 * (emax + emin), delta = 0.5 * (emax - emin);
  double sigma = theta / delta, rho_old = 1.0 / sigma;
  for (int k = 0; k < steps; k++) {
    double c1, c2;
    if (smoother == 1) { c1 = 0.0; c2 = 1.0 / theta; }
    else if (k == 0) { c1 = 0.0; c2 = 1.0 / theta; }
    else {
      double rho = 1.0 / (2 * sigma - rho_old);
      c1 = rho * rho_old; c2 = 2 * rho / delta;
      rho_old = rho;
    }
    if (k == 0 && zero_guess) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) { d[i] = c2 * (dinv[i] * b[i]); x[i] = d[i]; }
    } else {
      orc_spmv(A, x, r);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) {
        double ri = dinv[i] * (b[i] - r[i]);
        d[i] = c1 * d[i] + c2 * ri;
        x[i] = x[i] + d[i];
      }
    }
  }
}

static void smooth(orc_mg *mg, int l, int steps, int zero_guess)
{
  orc_smooth(mg->A[l], mg->dinv[l], mg->lam[l], mg->lo, mg->hi, mg->smoother, steps, zero_guess,
             mg->b[l], mg->x[l], mg->d[l], mg->r[l]);
}

static void vcycle(orc_mg *mg, int l)
{
  if (l == mg->L - 1) {
    cholesky_solve(mg->chol, mg->nc, mg->b[l], mg->x[l]);
    return;
  }
  const orc_csr *A = mg->A[l];
  int64_t n = A->n;
  smooth(mg, l, mg->nu1, 1);
  orc_spmv(A, mg->x[l], mg->r[l]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) mg->t[l][i] = mg->b[l][i] - mg->r[l][i];
  orc_spmv(mg->PT[l], mg->t[l], mg->b[l + 1]);
  vcycle(mg, l + 1);
  orc_spmv(mg->P[l], mg->x[l + 1], mg->t[l]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) mg->x[l][i] += mg->t[l][i];
  smooth(mg, l, mg->nu2, 0);
}

void orc_mg_vcycle(orc_mg *mg, const double *r, double *z)
{
  int64_t n = mg->A[0]->n;
  memcpy(mg->b[0], r, n * sizeof(double));
  vcycle(mg, 0);
  memcpy(z, mg->x[0], n * sizeof(double));
}

/* KSPCG with KSP_NORM_UNPRECONDITIONED and KSPConvergedDefault semantics:
 * converged when ||r|| <= max(rtol*||b||, 1e-50) (reason 2/3), diverged when
 * ||r|| >= 1e5*||b|| (-4), NaN (-9), indefinite PC (-8) or matrix (-10), or
 * maxit reached (-3).  x holds the initial guess (FEAnalysis.cpp:389). */
int orc_pcg(orc_mg *mg, const double *b, double *x, double rtol, int maxit, int *its,
            double *rnorm)
{
  const orc_csr *A = mg->A[0];
  int64_t n = A->n;
  double *r = (double *)malloc(n * sizeof(double));
  double *z = (double *)malloc(n * sizeof(double));
  double *p = (double *)malloc(n * sizeof(double));
  double *q = (double *)malloc(n * sizeof(double));
  orc_spmv(A, x, q);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) r[i] = b[i] - q[i];
  double bnorm = sqrt(orc_dot(n, b, b));
  double rn = sqrt(orc_dot(n, r, r));
  double rnorm0 = bnorm > 0 ? bnorm : rn;
  double ttol = rtol * rnorm0 > 1e-50 ? rtol * rnorm0 : 1e-50;
  int reason = 0, k = 0;
  if (rn <= ttol) reason = (rtol * rnorm0 > 1e-50) ? 2 : 3;
  double rz = 0;
  if (!reason) {
    orc_mg_vcycle(mg, r, z);
    rz = orc_dot(n, r, z);
    if (rz < 0) reason = -8;
    memcpy(p, z, n * sizeof(double));
  }
  while (!reason) {
    if (k >= maxit) { reason = -3; break; }
    orc_spmv(A, p, q);
    double pq = orc_dot(n, p, q);
    if (!(pq > 0)) { reason = isnan(pq) ? -9 : -10; break; }
    double alpha = rz / pq;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    k++;
    rn = sqrt(orc_dot(n, r, r));
    if (isnan(rn) || isinf(rn)) { reason = -9; break; }
    if (rn <= ttol) { reason = (rtol * rnorm0 > 1e-50) ? 2 : 3; break; }
    if (rn >= 1e5 * rnorm0) { reason = -4; break; }
    orc_mg_vcycle(mg, r, z);
    double rz_new = orc_dot(n, r, z);
    if (rz_new < 0) { reason = -8; break; }
    double beta = rz_new / rz;
    rz = rz_new;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  if (its) *its = k;
  if (rnorm) *rnorm = rn;
  free(r); free(z); free(p); free(q);
  return reason;
}

/* ======================================================================
 * Compliance sensitivities (Compliance.cpp:37-46)
 * ==================================================================== */

/* g_e = -dEdz_e * u_e^T (KE u_e) */
void orc_mass_diag(int nx, int ny, int nz, const double *V, double m0, const double *mlump,
                   const uint8_t *fixed, double *M)
{
  int64_t Nx = nx + 1, Ny = ny + 1;
  int64_t ndof = 3 * Nx * Ny * (nz + 1);
  for (int64_t i = 0; i < ndof; i++) M[i] = 0.0;
  for (int ez = 0; ez < nz; ez++)
    for (int ey = 0; ey < ny; ey++)
      for (int ex = 0; ex < nx; ex++) {
        int64_t el = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
        double v = V[el] * m0;                  /* nodeMat = p_V[el] * mMat block */
        for (int l = 0; l < 8; l++) {
          int64_t n = (ex + LN[l][0]) + Nx * ((ey + LN[l][1]) + Ny * (ez + LN[l][2]));
          for (int c = 0; c < 3; c++) M[3 * n + c] += v;
        }
      }
  if (mlump)
    for (int64_t i = 0; i < ndof; i++) M[i] = M[i] + mlump[i];
  if (fixed)
    for (int64_t i = 0; i < ndof; i++)
      if (fixed[i]) M[i] = 1e-8;
}

void orc_sensitivity(int nx, int ny, int nz, const double KE[576], const double *U,
                     const double *dEdz, double *g)
{
  int64_t ne = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int64_t el = 0; el < ne; el++) {
    int64_t nodes[8];
    orc_element_nodes(nx, ny, nz, el, nodes);
    double u[24], t[24];
    for (int j = 0; j < 8; j++)
      for (int i = 0; i < 3; i++) u[3 * j + i] = U[3 * nodes[j] + i];
    for (int r = 0; r < 24; r++) {
      double s = 0;
      for (int c = 0; c < 24; c++) s += KE[r * 24 + c] * u[c];
      t[r] = s;
    }
    double s = 0;
    for (int r = 0; r < 24; r++) s += u[r] * t[r];
    g[el] = -dEdz[el] * s;
  }
}

/* ======================================================================
 * Linearised buckling: stress stiffness and its sensitivities
 * (Stability::Function / StressFnc / sigtos, Buckling.cpp:15-414)
 * ==================================================================== */

/* dN/dx (3 x 8) at the 8 Gauss points and det(J) of the axis-aligned box element
 * (LocalK FEAnalysis.cpp:869-892, GaussPoints 900-926, dN 936-990). */
void orc_shape_grad(double hx, double hy, double hz, double dNdx[8][3][8], double detJ[8])
{
  const double g = 1 / sqrt(3.0);
  const double h[3] = {hx, hy, hz};
  for (int q = 0; q < 8; q++) {
    double xi = (2 * LN[q][0] - 1) * g, eta = (2 * LN[q][1] - 1) * g,
           zeta = (2 * LN[q][2] - 1) * g;
    double dNdxi[3][8];
    for (int n = 0; n < 8; n++) {
      double sx = 2 * LN[n][0] - 1, sy = 2 * LN[n][1] - 1, sz = 2 * LN[n][2] - 1;
      dNdxi[0][n] = sx / 8 * (1 + sy * eta) * (1 + sz * zeta);
      dNdxi[1][n] = sy / 8 * (1 + sx * xi) * (1 + sz * zeta);
      dNdxi[2][n] = sz / 8 * (1 + sx * xi) * (1 + sy * eta);
    }
    /* J = dNdxi * coords is diagonal (h/2) for the box; invert it the general way */
    double J[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double s = 0;
        for (int n = 0; n < 8; n++) s += dNdxi[a][n] * (h[b] * LN[n][b]);
        J[a][b] = s;
      }
    double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                 J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double inv[3][3];
    inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int a = 0; a < 3; a++)
      for (int n = 0; n < 8; n++) {
        double s = 0;
        for (int b = 0; b < 3; b++) s += inv[a][b] * dNdxi[b][n];
        dNdx[q][a][n] = s;
      }
    detJ[q] = det;
  }
}

/* B (6x24, AssignB FEAnalysis.cpp:1024-1046) and G (9x24, AssignG 1062-1081:
 * G[3i+j][3n+i] = dN_n/dx_j) of one Gauss point. */
static void bg_matrices(const double dNdx[3][8], double B[6][24], double G[9][24])
{
  memset(B, 0, 6 * 24 * sizeof(double));
  memset(G, 0, 9 * 24 * sizeof(double));
  for (int n = 0; n < 8; n++) {
    B[0][3 * n + 0] = dNdx[0][n];
    B[1][3 * n + 1] = dNdx[1][n];
    B[2][3 * n + 2] = dNdx[2][n];
    B[3][3 * n + 0] = dNdx[1][n];
    B[3][3 * n + 1] = dNdx[0][n];
    B[4][3 * n + 1] = dNdx[2][n];
    B[4][3 * n + 2] = dNdx[1][n];
    B[5][3 * n + 0] = dNdx[2][n];
    B[5][3 * n + 2] = dNdx[0][n];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) G[3 * i + j][3 * n + i] = dNdx[j][n];
  }
}

/* ks(u) = sum_q W G_q^T S(D B_q u) G_q detJ, W = 1 (StressFnc Buckling.cpp:313-317); S is
 * the 9x9 block diagonal of three copies of the 3x3 stress tensor built from the Voigt
 * vector [xx,yy,zz,xy,yz,xz] (sigtos, 371-414). */
static void stress_stiffness(const double D[36], const double dNdx[8][3][8], const double detJ[8],
                             const double u[24], double ks[576])
{
  memset(ks, 0, 576 * sizeof(double));
  for (int q = 0; q < 8; q++) {
    double B[6][24], G[9][24];
    bg_matrices(dNdx[q], B, G);
    double eps[6], sig[6];
    for (int a = 0; a < 6; a++) {
      double s = 0;
      for (int c = 0; c < 24; c++) s += B[a][c] * u[c];
      eps[a] = s;
    }
    for (int a = 0; a < 6; a++) {
      double s = 0;
      for (int b = 0; b < 6; b++) s += D[a * 6 + b] * eps[b];
      sig[a] = s;
    }
    double S3[3][3] = {{sig[0], sig[3], sig[5]}, {sig[3], sig[1], sig[4]}, {sig[5], sig[4], sig[2]}};
    double S[9][9];
    memset(S, 0, sizeof(S));
    for (int blk = 0; blk < 3; blk++)
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) S[3 * blk + a][3 * blk + b] = S3[a][b];
    double SG[9][24];
    for (int a = 0; a < 9; a++)
      for (int c = 0; c < 24; c++) {
        double s = 0;
        for (int b = 0; b < 9; b++) s += S[a][b] * G[b][c];
        SG[a][c] = s;
      }
    for (int r = 0; r < 24; r++)
      for (int c = 0; c < 24; c++) {
        double s = 0;
        for (int a = 0; a < 9; a++) s += G[a][r] * SG[a][c];
        ks[r * 24 + c] += 1.0 * s * detJ[q];
      }
  }
}

static void element_dofs(int nx, int ny, int64_t e, int64_t dof[24])
{
  int64_t Nx = nx + 1, Ny = ny + 1;
  int64_t ex = e % nx, ey = (e / nx) % ny, ez = e / ((int64_t)nx * ny);
  for (int l = 0; l < 8; l++) {
    int64_t n = (ex + LN[l][0]) + Nx * ((ey + LN[l][1]) + Ny * (ez + LN[l][2]));
    for (int c = 0; c < 3; c++) dof[3 * l + c] = 3 * n + c;
  }
}

orc_csr *orc_assemble_Ks(int nx, int ny, int nz, double E0, double nu, double hx, double hy,
                         double hz, const double *U, const double *Es, const uint8_t *fixed)
{
  double D[36], dNdx[8][3][8], detJ[8];
  orc_dmat(E0, nu, D);
  orc_shape_grad(hx, hy, hz, dNdx, detJ);
  int64_t ne = (int64_t)nx * ny * nz;
  int64_t ndof = 3 * (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  /* dense accumulation per row in a CSR with the 81-entry node pattern: build through
   * triplets into a dense-per-row map keyed by column */
  orc_csr *K = orc_assemble_K(nx, ny, nz, (double[576]){0}, Es, NULL); /* pattern only */
  for (int64_t p = 0; p < K->nnz; p++) K->val[p] = 0.0;
  for (int64_t e = 0; e < ne; e++) {
    int64_t dof[24];
    element_dofs(nx, ny, e, dof);
    double u[24], ks[576];
    for (int c = 0; c < 24; c++) u[c] = U[dof[c]];
    stress_stiffness(D, dNdx, detJ, u, ks);
    for (int r = 0; r < 24; r++)
      for (int c = 0; c < 24; c++) {
        double v = -Es[e] * ks[r * 24 + c];    /* ks *= -p_Es[el] (Buckling.cpp:331) */
        int64_t row = dof[r];
        for (int64_t p = K->rowptr[row]; p < K->rowptr[row + 1]; p++)
          if (K->col[p] == dof[c]) { K->val[p] += v; break; }
      }
  }
  if (fixed)  /* MatZeroRowsColumns(Ks, fixed, 0.0) (Buckling.cpp:28-29) */
    for (int64_t r = 0; r < ndof; r++)
      for (int64_t p = K->rowptr[r]; p < K->rowptr[r + 1]; p++)
        if (fixed[r] || fixed[K->col[p]]) K->val[p] = 0.0;
  return K;
}

void orc_ks_quadratic(int nx, int ny, int nz, double E0, double nu, double hx, double hy,
                      double hz, const double *U, const double *phi, int nvec, double *out)
{
  double D[36], dNdx[8][3][8], detJ[8];
  orc_dmat(E0, nu, D);
  orc_shape_grad(hx, hy, hz, dNdx, detJ);
  int64_t ne = (int64_t)nx * ny * nz;
  int64_t ndof = 3 * (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  for (int64_t e = 0; e < ne; e++) {
    int64_t dof[24];
    element_dofs(nx, ny, e, dof);
    double u[24], ks[576];
    for (int c = 0; c < 24; c++) u[c] = U[dof[c]];
    stress_stiffness(D, dNdx, detJ, u, ks);
    for (int j = 0; j < nvec; j++) {
      const double *ph = phi + (int64_t)j * ndof;
      double s = 0;
      for (int r = 0; r < 24; r++) {
        double t = 0;
        for (int c = 0; c < 24; c++) t += ks[r * 24 + c] * ph[dof[c]];
        s += ph[dof[r]] * t;
      }
      out[(int64_t)j * ne + e] = s;
    }
  }
}

/* dKsdu[d] = sum_q W G^T S(D B e_d) G detJ for the 24 unit element displacements
 * (Buckling.cpp:132-147); the adjoint right-hand side of eigenvector j is
 * dKsdU[d_global] += Es_e phi_e^T dKsdu[d_local] phi_e over the elements holding the DOF
 * (155-172), with fixed rows zeroed (187-191). */
void orc_stability_rhs(int nx, int ny, int nz, double E0, double nu, double hx, double hy,
                       double hz, const double *Es, const double *phi, int nvec,
                       const uint8_t *fixed, double *out)
{
  double D[36], dNdx[8][3][8], detJ[8];
  orc_dmat(E0, nu, D);
  orc_shape_grad(hx, hy, hz, dNdx, detJ);
  static double dks[24][576];
  for (int d = 0; d < 24; d++) {
    double du[24] = {0};
    du[d] = 1;
    stress_stiffness(D, dNdx, detJ, du, dks[d]);
  }
  int64_t ne = (int64_t)nx * ny * nz;
  int64_t ndof = 3 * (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  memset(out, 0, (size_t)nvec * ndof * sizeof(double));
  for (int64_t e = 0; e < ne; e++) {
    int64_t dof[24];
    element_dofs(nx, ny, e, dof);
    for (int j = 0; j < nvec; j++) {
      const double *ph = phi + (int64_t)j * ndof;
      double pl[24];
      for (int c = 0; c < 24; c++) pl[c] = ph[dof[c]];
      for (int d = 0; d < 24; d++) {
        double s = 0;
        for (int r = 0; r < 24; r++) {
          double t = 0;
          for (int c = 0; c < 24; c++) t += Es[e] * dks[d][r * 24 + c] * pl[c];
          s += pl[r] * t;
        }
        out[(int64_t)j * ndof + dof[d]] += s;
      }
    }
  }
  if (fixed)
    for (int j = 0; j < nvec; j++)
      for (int64_t r = 0; r < ndof; r++)
        if (fixed[r]) out[(int64_t)j * ndof + r] = 0.0;
}
