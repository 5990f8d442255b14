"""Block eigensolver over the C ABI, resident in HBM: the eigenproblems of the reference's
Stability (linearised buckling, Buckling.cpp:15-270) and Frequency (free vibration,
Dynamic.cpp:15-166) functions.

The reference solves them with its own LOPGMRES (locally optimal preconditioned block
method, LOPGMRES.cpp) over MatMult(A), MatMult(B) and PCApply of the displacement solver's
multigrid preconditioner, one vector at a time.  Here the same three operators are applied
to a whole block per launch (tofe_apply_Ks_dev / tofe_apply_K_dev / tofe_apply_M_dev /
tofe_pc_apply_dev with nrhs = block size), inside a locally optimal block preconditioned
iteration (LOBPCG: Rayleigh-Ritz on [X, T R, P] with SVQB orthonormalisation in the B inner
product).  Vectors are rows of (k, n) float64 tensors on the GPU; the Gram matrices are
FP64 GEMMs on the device and the small projected eigenproblem is solved in FP64 on the host.

Adjoint sensitivities of the buckling eigenvalues follow Stability::Function
(Buckling.cpp:130-262): for a K-normalised mode phi_j with eigenvalue mu_j of
Ks phi = mu K phi,
    dmu_j/drho_e = -dEsdz_e phi_e^T ks_e phi_e - mu_j dEdz_e phi_e^T KE phi_e
                   + dEdz_e v_je^T KE u_e,      K v_j = dKs/du[phi_j, phi_j],
and dmu_j/dx = Chain_Filter(dmu_j/drho).  The adjoint solves are one block solve.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib


def _torch():
    import torch

    return torch


@dataclass
class EigResult:
    values: np.ndarray      # (k,) eigenvalues, wanted end first
    vectors: object         # (k, n) torch float64 tensor, B-orthonormal rows
    iterations: int
    residuals: np.ndarray   # (k,) ||A x - theta B x|| / (|theta| ||B x||)
    converged: bool


def _svqb(G, drop=1e-12):
    """Z with Z^T G Z = I for the Gram matrix G of a basis (SVQB: diagonal scaling, then the
    eigenvectors of the scaled Gram matrix with the directions below drop * largest removed)."""
    d = np.sqrt(np.clip(np.diag(G), 1e-300, None))
    Gs = G / d[:, None] / d[None, :]
    lam, V = np.linalg.eigh(0.5 * (Gs + Gs.T))
    keep = lam > drop * lam.max()
    return (V[:, keep] / np.sqrt(lam[keep])[None, :]) / d[:, None]


def lobpcg(apply_A, apply_B, apply_T, X0, largest, tol=1e-8, maxit=500, mask=None,
           log=None, nwant=None):
    """Eigenpairs of A x = theta B x (B symmetric positive definite): the block of
    k = X0.shape[0] vectors iterates, and the nwant (default k) largest or smallest theta are
    tested and returned -- the other k - nwant are guard vectors that speed up convergence
    near clustered eigenvalues.  apply_* map a (m, n) tensor of row vectors to their images
    (m <= 16 per call is one block pass); mask (n,) zeroes the Dirichlet DOFs of every new
    search direction."""
    torch = _torch()
    k = X0.shape[0]
    nw = k if nwant is None else min(nwant, k)

    def clean(Y):
        return Y * mask if mask is not None else Y

    X = clean(X0.clone())
    BX = apply_B(X)
    Z = _svqb((X @ BX.T).cpu().numpy())
    Zt = torch.from_numpy(np.ascontiguousarray(Z.T)).to(X)
    X, BX = Zt @ X, Zt @ BX
    AX = apply_A(X)
    P = AP = BP = None
    theta = None
    res = np.full(k, np.inf)
    it = 0
    for it in range(1, maxit + 1):
        # Rayleigh-Ritz on span[X, W, P]
        if theta is None:
            GA = (X @ AX.T).cpu().numpy()
            GA = 0.5 * (GA + GA.T)
            w, Y = np.linalg.eigh(GA)
            sel = np.argsort(w)[::-1][:k] if largest else np.argsort(w)[:k]
            theta = w[sel]
            C = torch.from_numpy(np.ascontiguousarray(Y[:, sel].T)).to(X)
            X, AX, BX = C @ X, C @ AX, C @ BX
        th = torch.from_numpy(theta.copy()).to(X)
        R = AX - th[:, None] * BX
        rn = torch.linalg.vector_norm(R, dim=1).cpu().numpy()
        bn = torch.linalg.vector_norm(BX, dim=1).cpu().numpy()
        res = rn / np.maximum(np.abs(theta) * bn, 1e-300)
        if log:
            log(it, theta[:nw], res[:nw])
        if np.all(res[:nw] < tol):
            return EigResult(theta[:nw], X[:nw], it, res[:nw], True)
        W = clean(apply_T(R))
        AW, BW = apply_A(W), apply_B(W)
        blocks = [(X, AX, BX), (W, AW, BW)] + ([(P, AP, BP)] if P is not None else [])
        S = torch.cat([b[0] for b in blocks])
        AS = torch.cat([b[1] for b in blocks])
        BS = torch.cat([b[2] for b in blocks])
        GB = (S @ BS.T).cpu().numpy()
        GB = 0.5 * (GB + GB.T)
        Zb = _svqb(GB)
        GA = (S @ AS.T).cpu().numpy()
        GA = 0.5 * (GA + GA.T)
        H = Zb.T @ GA @ Zb
        w, Y = np.linalg.eigh(0.5 * (H + H.T))
        sel = np.argsort(w)[::-1][:k] if largest else np.argsort(w)[:k]
        theta = w[sel]
        Cf = Zb @ Y[:, sel]                      # (m, k) coefficients over the basis rows
        C = torch.from_numpy(np.ascontiguousarray(Cf.T)).to(X)
        Xn, AXn, BXn = C @ S, C @ AS, C @ BS
        # new direction: the part of the update outside span(X)
        Cp = Cf.copy()
        Cp[:k] = 0
        Cpt = torch.from_numpy(np.ascontiguousarray(Cp.T)).to(X)
        P, AP, BP = Cpt @ S, Cpt @ AS, Cpt @ BS
        X, AX, BX = Xn, AXn, BXn
    return EigResult(theta[:nw], X[:nw], it, res[:nw], False)


class _Ops:
    """Block operators of one context on (m, n) CUDA tensors, in passes of at most 16."""

    def __init__(self, ctx):
        self.ctx = ctx

    def _run(self, fn, X):
        torch = _torch()
        X = X.contiguous()
        Y = torch.empty_like(X)
        fn(X.data_ptr(), Y.data_ptr(), X.shape[0])
        return Y

    def K(self, X):
        return self._run(self.ctx.apply_K_dev, X)

    def Ks(self, X):
        return self._run(self.ctx.apply_Ks_dev, X)

    def M(self, X):
        return self._run(self.ctx.apply_M_dev, X)

    def T(self, X):
        return self._run(self.ctx.pc_apply_dev, X)


def _start_block(n, k, fixed, seed, device):
    torch = _torch()
    g = torch.Generator(device="cpu").manual_seed(seed)
    X = torch.randn((k, n), generator=g, dtype=torch.float64)
    mask = torch.from_numpy((1 - fixed.astype(np.float64))) if fixed is not None else None
    X = X.to(device)
    return X, (mask.to(device) if mask is not None else None)


def block_size(nev):
    """Vectors iterated for nev wanted eigenpairs: up to four guard vectors, at most one block
    pass (16)."""
    return max(nev, min(16, nev + 4))


def buckling_modes(ctx, U, fixed_dof, nev=6, tol=1e-8, maxit=500, seed=0, log=None):
    """The nev largest mu of Ks(U) phi = mu K phi (Stability::Function's eigenproblem,
    Set_Operators(Ks, K) and target LR at Buckling.cpp:67-70); phi K-orthonormal.
    U is a (n,) CUDA tensor; the context must be assembled for the current design."""
    ctx.set_stress_dev(U.data_ptr())
    ops = _Ops(ctx)
    X0, mask = _start_block(ctx.ndof, block_size(nev), fixed_dof, seed, U.device)
    return lobpcg(ops.Ks, ops.K, ops.T, X0, largest=True, tol=tol, maxit=maxit, mask=mask,
                  log=log, nwant=nev)


def vibration_modes(ctx, fixed_dof, nev=12, tol=1e-8, maxit=500, seed=0, device="cuda",
                    log=None):
    """The nev lowest free-vibration modes: as Frequency::Function (Set_Operators(M, K) and
    target LR at Dynamic.cpp:67-70), the nev largest lambda of M phi = lambda K phi, with
    omega^2 = 1 / lambda.  Returns the EigResult of (M, K) (phi K-orthonormal; values are
    lambda, largest first).  Needs tofe_assemble and tofe_assemble_mass."""
    ops = _Ops(ctx)
    X0, mask = _start_block(ctx.ndof, block_size(nev), fixed_dof, seed, device)
    return lobpcg(ops.M, ops.K, ops.T, X0, largest=True, tol=tol, maxit=maxit, mask=mask,
                  log=log, nwant=nev)


def buckling_sensitivities(ctx, U, mu, phi, rtol=1e-10, maxit=2000):
    """dmu_j/dx for the modes (mu, phi) of buckling_modes (K-normalised rows of phi, a CUDA
    tensor; U the displacement, a CUDA tensor): (nev, nelem) array, Buckling.cpp:130-262
    per eigenvalue, filter chain rule included."""
    torch = _torch()
    phi_h = phi.cpu().numpy()
    U_h = U.cpu().numpy()
    nev = phi_h.shape[0]
    pksp, rhs = ctx.stability_terms(phi_h)
    pkp, _ = ctx.element_quadratic(phi_h, mass=False)
    rhs_d = torch.from_numpy(np.ascontiguousarray(rhs)).to(U.device)
    V = torch.zeros_like(rhs_d)
    its, reason = ctx.solve_dev(rhs_d.data_ptr(), V.data_ptr(), nrhs=nev, rtol=rtol, maxit=maxit)
    if reason < 0:
        raise RuntimeError(f"adjoint solves did not converge (reason {reason})")
    vku = ctx.element_bilinear(V.cpu().numpy(), U_h)
    dEdz = ctx.field(_lib.FIELD_DEDZ)
    dEsdz = ctx.field(_lib.FIELD_DESDZ)
    g = np.empty((nev, ctx.nelem))
    for j in range(nev):
        gj = -dEsdz * pksp[j] - mu[j] * dEdz * pkp[j] + dEdz * vku[j]
        g[j] = ctx.chain_filter(None, gj)[1]
    return g
