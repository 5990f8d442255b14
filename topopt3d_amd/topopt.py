"""Host-side mirror of the reference's FE-analysis call surface over the C ABI.

The method names, their order in the design loop and their meaning follow the
reference's ``TopOpt`` class and ``Compliance`` function (TopOpt.h:282-292,
Functions.h:65-77): CreateMesh -> Set_BC -> FEInitialize -> (MatIntFnc -> FEAssemble ->
FESolve -> Compliance.Function -> Chain_Filter) per design iteration.  PETSc Vec/Mat/
KSP handles are replaced by numpy arrays and one tofe context; every numerical step
runs in the HIP library.
"""
from __future__ import annotations

from enum import IntEnum

import numpy as np

from . import _lib
from ._lib import Context


class BCType(IntEnum):
    """Boundary-condition kinds of the input file's [BC] section (Inputs.cpp:148-217)."""
    SUPPORT = 0
    LOAD = 1
    PRESSURE = 2
    MASS = 3
    SPRING = 4
    EIGEN = 5


# Node lists of the six faces of a hexahedral element in its local node order
# (Get_Faces, Inputs.cpp:62-91).
FACES = np.array([[0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 5, 4], [3, 2, 6, 7], [0, 4, 7, 3],
                  [1, 5, 6, 2]])


class TopOpt:
    """Single-GPU stand-in for the reference TopOpt object's FE-analysis members."""

    def __init__(self, E0=1.0, Nu0=0.3, penal=3.0, verbose=0):
        self.E0 = E0
        self.Nu0 = Nu0
        self.penal = penal
        self.vdPenal = 1.0          # void penalty q (TopOpt.h:167)
        self.vdMin = -1.0           # minimum void count within Rmax; <= 0 disables (TopOpt.h:175)
        self.interpolation = _lib.INTERP_SIMP
        self.interp_param = [0.0, 0.0]
        self.verbose = verbose
        self.eps = 1e-10  # Material_Interpolation.cpp:28
        self.ctx = None
        self.suppNode = np.zeros(0, np.int64)
        self.supports = np.zeros((0, 3), bool)
        self.loadNode = np.zeros(0, np.int64)
        self.loads = np.zeros((0, 3))
        self.massNode = np.zeros(0, np.int64)
        self.masses = np.zeros((0, 3))
        self.springNode = np.zeros(0, np.int64)
        self.springs = np.zeros((0, 3))
        self.eigenSuppNode = np.zeros(0, np.int64)
        self.eigenSupports = np.zeros((0, 3), bool)
        self.mg_levels = 0
        self.smoother = _lib.SMOOTHER_CHEBYSHEV
        self.precision = _lib.PREC_FP64
        self.rtol = 1e-8  # KSPSetTolerances(KUF, 1e-8, ...) FEAnalysis.cpp:390
        self.maxit = 10000
        self.KUF_reason = 0
        self.its = 0

    # -- mesh (RecMesh.cpp:426-597, single rank, no reordering) ----------------
    def CreateMesh(self, dimensions, Nel, Rmin, Rmax=0.0):
        """dimensions = (x0, x1, y0, y1, z0, z1); Nel = elements per direction; Rmin and
        Rmax are the density-filter and maximum-length-scale radii in units of length."""
        self.dimensions = np.asarray(dimensions, np.float64)
        self.Nel = tuple(int(n) for n in Nel)
        nx, ny, nz = self.Nel
        self.dx = tuple((self.dimensions[2 * i + 1] - self.dimensions[2 * i]) / self.Nel[i]
                        for i in range(3))
        self.nElem = nx * ny * nz
        self.nNode = (nx + 1) * (ny + 1) * (nz + 1)
        k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1),
                              indexing="ij")
        self.node = np.stack([self.dimensions[0] + self.dx[0] * i.reshape(-1),
                              self.dimensions[2] + self.dx[1] * j.reshape(-1),
                              self.dimensions[4] + self.dx[2] * k.reshape(-1)], axis=1)
        ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        base = (ex + (nx + 1) * (ey + (ny + 1) * ez)).reshape(-1)
        Nx, NxNy = nx + 1, (nx + 1) * (ny + 1)
        self.element = np.stack([base, base + 1, base + 1 + Nx, base + Nx,
                                 base + NxNy, base + 1 + NxNy, base + 1 + Nx + NxNy,
                                 base + Nx + NxNy], axis=1)
        self.Rmin = float(Rmin)
        self.Rmax = float(Rmax)

    # -- boundary conditions (Inputs.cpp:105-223) ----------------------------------
    def Set_BC(self, center, radius, limits, values, TYPE):
        """Select nodes with sum(((x_i - c_i)/r_i)^2) <= 1 inside the box limits, or, for
        PRESSURE, element faces whose centre (mean of the face's four node coordinates)
        satisfies the same test; every node of every selected face then receives the full
        value vector as a load, so a node shared by several selected faces is loaded once per
        face (Inputs.cpp:114-147, 174-199).  Returns the selected nodes (PRESSURE: with
        repetitions, in face order)."""
        center = np.asarray(center, np.float64)
        radius = np.asarray(radius, np.float64)
        limits = np.asarray(limits, np.float64).reshape(3, 2)
        values = np.asarray(values, np.float64)
        TYPE = BCType(TYPE)
        if TYPE != BCType.PRESSURE:
            pts = self.node
        else:
            # face centres, element-major then face (index el * 6 + face)
            pts = np.zeros((self.nElem * 6, 3))
            for nd in range(4):
                pts += self.node[self.element[:, FACES[:, nd]]].reshape(-1, 3)
            pts /= 4
        d = np.zeros(pts.shape[0])
        valid = np.ones(pts.shape[0], bool)
        for i in range(3):
            d += ((pts[:, i] - center[i]) / radius[i]) ** 2
            valid &= (pts[:, i] >= limits[i, 0]) & (pts[:, i] <= limits[i, 1])
        valid &= d <= 1
        sel = np.flatnonzero(valid)
        if TYPE == BCType.PRESSURE:
            el, face = sel // 6, sel % 6
            newNode = self.element[el[:, None], FACES[face]].reshape(-1)
            self.loads = np.vstack([self.loads, np.tile(values, (newNode.size, 1))])
            self.loadNode = np.concatenate([self.loadNode, newNode])
            return newNode
        newNode = sel
        if TYPE == BCType.SUPPORT:
            self.supports = np.vstack([self.supports, np.tile(values.astype(bool), (newNode.size, 1))])
            self.suppNode = np.concatenate([self.suppNode, newNode])
        elif TYPE == BCType.EIGEN:
            self.eigenSupports = np.vstack([self.eigenSupports,
                                            np.tile(values.astype(bool), (newNode.size, 1))])
            self.eigenSuppNode = np.concatenate([self.eigenSuppNode, newNode])
        elif TYPE == BCType.LOAD:
            self.loads = np.vstack([self.loads, np.tile(values, (newNode.size, 1))])
            self.loadNode = np.concatenate([self.loadNode, newNode])
        elif TYPE == BCType.MASS:
            self.masses = np.vstack([self.masses, np.tile(values, (newNode.size, 1))])
            self.massNode = np.concatenate([self.massNode, newNode])
        elif TYPE == BCType.SPRING:
            self.springs = np.vstack([self.springs, np.tile(values, (newNode.size, 1))])
            self.springNode = np.concatenate([self.springNode, newNode])
        return newNode

    # -- FE set-up (FEAnalysis.cpp:215-424) -----------------------------------
    def FEInitialize(self):
        ndof = 3 * self.nNode
        self.fixed_dof = np.zeros(ndof, np.uint8)
        for n, s in zip(self.suppNode, self.supports):
            for j in range(3):
                if s[j]:
                    self.fixed_dof[3 * n + j] = 1
        self.fixedDof = np.flatnonzero(self.fixed_dof)
        self.F = np.zeros(ndof)
        for n, l in zip(self.loadNode, self.loads):
            for j in range(3):
                self.F[3 * n + j] += l[j]   # F[3 loadNode + j] += loads (349-356)
        self.U = np.zeros(ndof)
        # lumped point masses MLump (FEAnalysis.cpp:374-384), spring stiffnesses spKVec
        # (360-368), extra eigenproblem supports eigenFixedDof (249-255)
        self.MLump = np.zeros(ndof)
        for n, m in zip(self.massNode, self.masses):
            self.MLump[3 * n:3 * n + 3] += m
        self.spKVec = np.zeros(ndof)
        for n, k in zip(self.springNode, self.springs):
            self.spKVec[3 * n:3 * n + 3] += k
        self.eigen_fixed_dof = np.zeros(ndof, np.uint8)
        for n, s in zip(self.eigenSuppNode, self.eigenSupports):
            for j in range(3):
                if s[j]:
                    self.eigen_fixed_dof[3 * n + j] = 1
        self.eigenFixedDof = np.flatnonzero(self.eigen_fixed_dof)
        if self.ctx is not None:
            self.ctx.close()
        self.ctx = Context(self.Nel, h=self.dx, E0=self.E0, nu=self.Nu0, Emin=self.eps,
                           penal=self.penal, rmin=self.Rmin, mg_levels=self.mg_levels,
                           smoother=self.smoother, precision=self.precision,
                           fixed_dof=self.fixed_dof, rmax=self.Rmax, vd_min=self.vdMin,
                           vd_penal=self.vdPenal, interpolation=int(self.interpolation),
                           interp_param=tuple(self.interp_param))
        if self.spKVec.any():  # MatDiagonalSet(K, spKVec) in FEAssemble (FEAnalysis.cpp:471)
            self.ctx.set_springs(self.spKVec)
        self.ke = [self.ctx.ke()]

    # -- design iteration ---------------------------------------------------------
    def MatIntFnc(self, design):
        """rho = P x, the maximum-length-scale factor y and the SIMP / SIMP_CUT /
        SIMP_LOGISTIC interpolation (Material_Interpolation.cpp:20-140).  The current
        penal and vdPenal apply (penalty continuation, Main.cpp:131-133)."""
        self.ctx.set_penalty(self.penal, self.vdPenal)
        self.ctx.set_design(design)
        self.x = np.asarray(design, np.float64)
        self.rho = self.ctx.field(_lib.FIELD_RHO)
        self.y = self.ctx.field(_lib.FIELD_Y)
        self.rhoq = self.ctx.field(_lib.FIELD_RHOQ)
        self.E = self.ctx.field(_lib.FIELD_E)
        self.dEdz = self.ctx.field(_lib.FIELD_DEDZ)
        self.Es = self.ctx.field(_lib.FIELD_ES)
        self.dEsdz = self.ctx.field(_lib.FIELD_DESDZ)
        self.V = self.rho
        self.dVdrho = np.ones_like(self.rho)

    def FEAssemble(self):
        self.ctx.assemble()

    def FESolve(self):
        """KSPSolve(KUF, F, U), warm-started from the previous U (FEAnalysis.cpp:389, 673)."""
        self.U, self.its, self.KUF_reason = self.ctx.solve(self.F, self.U, rtol=self.rtol,
                                                           maxit=self.maxit)
        return self.KUF_reason

    def Chain_Filter(self, dfdV=None, dfdE=None):
        """dfdV <- P^T dfdV; dfdE <- P^T [dfdE .* y - rhoq .* (y < 1) .* R^T(dfdE .* rho)]
        (Material_Interpolation.cpp:150-187); returns the two results."""
        return self.ctx.chain_filter(dfdV, dfdE)


class Compliance:
    """Compliance objective c = U.F and its sensitivities (Compliance.cpp:15-57)."""

    def __init__(self, calc_gradient=True):
        self.calc_gradient = calc_gradient
        self.values = np.zeros(1)
        self.gradients = None

    def Function(self, topOpt: TopOpt):
        if self.calc_gradient:
            c, g = topOpt.ctx.compliance(topOpt.U, topOpt.F, gradient=True)
            self.gradients = g
        else:
            c = topOpt.ctx.compliance(topOpt.U, topOpt.F, gradient=False)
        self.values[0] = c
        return 0


class Volume:
    """Volume fraction sum(rho)/nElem with gradient P^T (1/nElem) (Volume.cpp:15-41)."""

    def __init__(self):
        self.values = np.zeros(1)
        self.gradients = None

    def Function(self, topOpt: TopOpt):
        self.values[0] = topOpt.V.sum() / topOpt.nElem
        self.gradients = topOpt.ctx.filter_T(np.full(topOpt.nElem, 1.0 / topOpt.nElem))
        return 0


class Stability:
    """Linearised buckling (Stability::Function, Buckling.cpp:15-262): the nvals largest
    eigenvalues mu of Ks phi = mu K phi for the current design and displacement,
    aggregated as (sum mu^p)^(1/p), with the sensitivity of the reference: per element
    sum_j mu_j^(p-1) (phi^T dKs/dy phi - mu_j phi^T dK/dy phi + v_j^T dK/dy U), K v_j = dKsdU_j,
    scaled by value^(1-p) and passed through Chain_Filter.

    The eigenproblem is solved on the host with scipy's LOBPCG over the GPU operators
    (tofe_apply_Ks, tofe_apply_K, tofe_pc_apply as preconditioner); the reference uses its
    own LOPGMRES over the same three operators."""

    def __init__(self, nvals=4, p=8.0, calc_gradient=True, tol=1e-9, maxit=500, seed=0):
        self.nvals = nvals
        self.p = p
        self.calc_gradient = calc_gradient
        self.tol = tol
        self.maxit = maxit
        self.seed = seed
        self.values = np.zeros(1)
        self.gradients = None
        self.lam = None
        self.phi = None

    def Function(self, topOpt: TopOpt):
        from scipy.sparse.linalg import LinearOperator, lobpcg

        ctx = topOpt.ctx
        n = ctx.ndof
        ctx.set_stress(topOpt.U)
        Ks = LinearOperator((n, n), matvec=lambda x: ctx.apply_Ks(np.ravel(x)),
                            matmat=lambda X: ctx.apply_Ks(np.ascontiguousarray(X.T)).T)
        K = LinearOperator((n, n), matvec=lambda x: ctx.apply_K(np.ravel(x)),
                           matmat=lambda X: ctx.apply_K(np.ascontiguousarray(X.T)).T)
        M = LinearOperator((n, n), matvec=lambda x: ctx.pc_apply(np.ravel(x)),
                           matmat=lambda X: ctx.pc_apply(np.ascontiguousarray(X.T)).T)
        X0 = np.random.default_rng(self.seed).standard_normal((n, self.nvals))
        X0[topOpt.fixed_dof.astype(bool)] = 0
        mu, V = lobpcg(Ks, X0, B=K, M=M, largest=True, tol=self.tol, maxiter=self.maxit)
        order = np.argsort(mu)[::-1]
        mu, V = mu[order], V[:, order]
        self.lam, self.phi = mu, V.T.copy()
        p = self.p
        self.values[0] = np.sum(mu ** p) ** (1 / p)
        if not self.calc_gradient:
            return 0
        phi = self.phi
        pksp, rhs = ctx.stability_terms(phi)
        pkp, _ = ctx.element_quadratic(phi, mass=False)
        v = np.stack([ctx.solve(rhs[j], rtol=1e-10, maxit=topOpt.maxit)[0]
                      for j in range(len(mu))])
        vku = ctx.element_bilinear(v, topOpt.U)
        dEdz, dEsdz = ctx.field(_lib.FIELD_DEDZ), ctx.field(_lib.FIELD_DESDZ)
        g = np.zeros(topOpt.nElem)
        for j, m in enumerate(mu):
            g += m ** (p - 1) * (-dEsdz * pksp[j] - m * dEdz * pkp[j] + dEdz * vku[j])
        g *= self.values[0] ** (1 - p)
        _, self.gradients = topOpt.Chain_Filter(None, g)
        return 0
