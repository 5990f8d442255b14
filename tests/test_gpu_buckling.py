"""Linearised buckling end to end on the GPU (Stability::Function, Buckling.cpp:15-270): the
column workload (z = 0 face clamped, unit pressure on the top face with the reference's
per-face-node load, Inputs.cpp:126-147, 174-199; filter R = 2.0), the 6 largest mu of
Ks phi = mu K phi from the block eigensolver over the GPU operators, the adjoint
sensitivities of the compliance and of every eigenvalue, wrapped as a torch.autograd.Function.

Checks, in FP64 and in mixed precision (FP32 V-cycle inside FP64 CG):
- eigenvalues and per-eigenvalue gradients against the dense oracle (generalized eigh of the
  oracle's assembled Ks and K, the reference's formula with dense adjoint solves);
- both gradients against central finite differences at step 1e-6 on 32 seeded elements,
  the forward evaluated on the GPU from warm starts with fixed iteration counts (a smooth map
  of x; a convergence test could stop x + h and x - h at different iterations);
- mixed against FP64.
The mesh is small enough that the step-1e-6 difference quotient is not lost in rounding:
mu changes by about 2e-6 * mu / nelem per step, against rounding near 1e-15 * mu.
"""
import numpy as np
import pytest

from topopt3d_amd import configs

pytestmark = pytest.mark.gpu

NEL = (8, 4, 8)
NEV = 6


def dense_reference(oracle, pb):
    """(c, dc/dx, mu, dmu/dx) from dense oracle pieces (Buckling.cpp:80-262 per eigenvalue)."""
    import scipy.linalg as sla

    nel = pb.nel
    P = oracle.filter_csr(*nel, pb.rmin)
    rho = P.matvec(pb.x)
    E, dEdz, Es, dEsdz = oracle.simp(rho, pb.penal, pb.Emin)
    KE = oracle.ke()
    K = oracle.assemble_K(*nel, KE, E, pb.fixed_dof).to_scipy().toarray()
    U = np.linalg.solve(K, pb.F)
    c = pb.F @ U
    ones = np.ones(pb.nelem)
    dc = P.rmatvec(oracle.sensitivity(*nel, KE, U, dEdz))
    Ks = oracle.assemble_Ks(*nel, U, Es, pb.fixed_dof).to_scipy().toarray()
    mu, V = sla.eigh(Ks, K)
    mu, V = mu[::-1][:NEV], V[:, ::-1][:, :NEV]
    phi = V.T.copy()
    q = oracle.ks_quadratic(*nel, U, phi)
    rhs = oracle.stability_rhs(*nel, Es, phi, pb.fixed_dof)
    dmu = np.zeros((NEV, pb.nelem))
    for j in range(NEV):
        v = np.linalg.solve(K, rhs[j])
        pKp = -oracle.sensitivity(*nel, KE, phi[j], ones)
        vKu = (oracle.sensitivity(*nel, KE, v - U, ones) - oracle.sensitivity(*nel, KE, v + U, ones)) / 4
        dmu[j] = P.rmatvec(-dEsdz * q[j] - mu[j] * dEdz * pKp + dEdz * vKu)
    return c, dc, mu, dmu


class Evaluator:
    """Compliance and the NEV buckling eigenvalues of a design on one context, with the
    adjoint gradients; warm starts from the last full evaluation."""

    def __init__(self, tofe, pb, precision):
        import torch

        self.torch = torch
        self.tofe = tofe
        self.pb = pb
        self.ctx = tofe.Context(pb.nel, rmin=pb.rmin, penal=pb.penal, Emin=pb.Emin,
                                mg_levels=pb.mg_levels, fixed_dof=pb.fixed_dof,
                                precision=precision)
        self.F = torch.from_numpy(pb.F).cuda()
        self.U0 = None
        self.X0 = None

    def full(self, x):
        from topopt3d_amd import eigen

        torch = self.torch
        ctx = self.ctx
        ctx.set_design(x)
        ctx.assemble()
        U = torch.zeros_like(self.F)
        its, reason = ctx.solve_dev(self.F.data_ptr(), U.data_ptr(), rtol=1e-13, maxit=500)
        assert reason == self.tofe.CONVERGED_RTOL, reason
        c, g = ctx.compliance(U.cpu().numpy(), self.pb.F)
        res = eigen.buckling_modes(ctx, U, self.pb.fixed_dof, nev=NEV, tol=1e-11, maxit=400)
        assert res.converged, res.residuals
        dmu = eigen.buckling_sensitivities(ctx, U, res.values, res.vectors, rtol=1e-12)
        self.U0, self.X0 = U, res.vectors
        return c, g, res.values, dmu

    def value(self, x, cg_its=40, eig_its=4):
        """(c, mu) at x from the last full evaluation's U and modes, fixed iteration counts."""
        from topopt3d_amd import eigen

        torch = self.torch
        ctx = self.ctx
        ctx.set_design(x)
        ctx.assemble()
        U = self.U0.clone()
        ctx.solve_dev(self.F.data_ptr(), U.data_ptr(), rtol=0.0, maxit=cg_its)
        c = float(self.F @ U)
        ctx.set_stress_dev(U.data_ptr())
        ops = eigen._Ops(ctx)
        mask = torch.from_numpy(1.0 - self.pb.fixed_dof.astype(np.float64)).cuda()
        res = eigen.lobpcg(ops.Ks, ops.K, ops.T, self.X0.clone(), largest=True, tol=0.0,
                           maxit=eig_its, mask=mask)
        return c, res.values


@pytest.fixture(scope="module")
def column():
    return configs.column(NEL, seed=3, lo=0.1, hi=1.0, rmin=2.0, Emin=1e-9, mg_levels=2)


@pytest.mark.parametrize("precision", [0, 1])
def test_buckling_modes_and_sensitivities_match_dense_oracle(oracle, tofe, column, precision):
    print(f"case buckling dense nel={NEL} precision={precision}", flush=True)
    pb = column
    c0, dc0, mu0, dmu0 = dense_reference(oracle, pb)
    assert (mu0 > 0).all()
    ev = Evaluator(tofe, pb, precision)
    c, dc, mu, dmu = ev.full(pb.x)
    assert c == pytest.approx(c0, rel=1e-10)
    assert np.abs(dc - dc0).max() <= 1e-9 * np.abs(dc0).max()
    np.testing.assert_allclose(mu, mu0, rtol=1e-9)
    for j in range(NEV):
        assert np.abs(dmu[j] - dmu0[j]).max() <= 1e-6 * np.abs(dmu0[j]).max(), j
    ev.ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_gradients_match_central_differences(tofe, column, precision):
    """torch.autograd.Function over the GPU path: backward = the adjoint gradients; checked
    against central differences (step 1e-6) on 32 seeded elements for the compliance and
    each of the 6 eigenvalues."""
    import torch

    pb = column
    print(f"case buckling FD nel={NEL} precision={precision}", flush=True)
    ev = Evaluator(tofe, pb, precision)

    class BucklingFn(torch.autograd.Function):
        @staticmethod
        def forward(fctx, x):
            c, g, mu, dmu = ev.full(x.detach().cpu().numpy())
            fctx.save_for_backward(torch.from_numpy(np.vstack([g[None, :], dmu])))
            return torch.from_numpy(np.concatenate([[c], mu]))

        @staticmethod
        def backward(fctx, gout):
            (J,) = fctx.saved_tensors
            return gout @ J

    x = torch.from_numpy(pb.x.copy()).requires_grad_(True)
    y = BucklingFn.apply(x)
    rows = [torch.autograd.grad(y[i], x, retain_graph=True)[0].numpy() for i in range(NEV + 1)]
    J = np.stack(rows)
    elems = np.random.default_rng(3).choice(pb.nelem, 32, replace=False)
    h = 1e-6
    worst = 0.0
    for e in elems:
        xp, xm = pb.x.copy(), pb.x.copy()
        xp[e] += h
        xm[e] -= h
        cp, mup = ev.value(xp)
        cm, mum = ev.value(xm)
        fd = np.concatenate([[(cp - cm) / (2 * h)], (mup - mum) / (2 * h)])
        scale = np.abs(J).max(axis=1)
        err = np.abs(fd - J[:, e]) / np.maximum(np.abs(J[:, e]), 1e-3 * scale)
        worst = max(worst, err.max())
    print(f"buckling FD worst relative error {worst:.2e}", flush=True)
    assert worst <= 1e-5
    ev.ctx.close()


def test_mixed_matches_fp64(tofe, column):
    pb = column
    e64, emx = Evaluator(tofe, pb, 0), Evaluator(tofe, pb, 1)
    c64, g64, mu64, dmu64 = e64.full(pb.x)
    cmx, gmx, mumx, dmumx = emx.full(pb.x)
    assert cmx == pytest.approx(c64, rel=1e-11)
    np.testing.assert_allclose(mumx, mu64, rtol=1e-10)
    assert np.abs(dmumx - dmu64).max() <= 1e-8 * np.abs(dmu64).max()
    e64.ctx.close()
    emx.ctx.close()
