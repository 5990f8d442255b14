"""Precision and differentiation forms on the GPU.

Mixed precision: FP32 V-cycle inside the FP64 flexible CG must reach the FP64 relative
residual 1e-8 (checked with the oracle's assembled K) and agree with the FP64 solve.
Differentiation: the compliance and its adjoint sensitivity (element sensitivity
kernel composed with the filter transpose) wrapped as a torch.autograd.Function, checked
against central finite differences with tight solves.
"""
import numpy as np
import pytest

from topopt3d_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nel,L", [((16, 8, 8), 3), ((48, 24, 24), 3), ((30, 14, 18), 3)])
def test_mixed_precision_reaches_fp64_tolerance(oracle, tofe, nel, L):
    pb = configs.cantilever(nel, seed=2, lo=0.01, hi=1.0, mg_levels=L)
    print(f"case mixed nel={nel}", flush=True)
    ctx = tofe.Context(nel, mg_levels=L, fixed_dof=pb.fixed_dof, precision=tofe.PREC_MIXED)
    ctx.set_design(pb.x)
    ctx.assemble()
    Um, its_m, reason_m = ctx.solve(pb.F, rtol=1e-8)
    assert reason_m == 2
    E = ctx.field(tofe.FIELD_E)
    K = oracle.assemble_K(*nel, oracle.ke(), E, pb.fixed_dof)
    res = np.linalg.norm(pb.F - K.matvec(Um)) / np.linalg.norm(pb.F)
    assert res <= 1e-8 * 1.01
    ctx.set_precision(tofe.PREC_FP64)
    ctx.assemble()
    Ud, its_d, reason_d = ctx.solve(pb.F, rtol=1e-8)
    assert reason_d == 2
    assert its_m <= its_d + 3
    # both are 1e-8-residual solutions of the same system
    assert np.abs(Um - Ud).max() <= 1e-6 * np.abs(Ud).max()
    c_m = ctx.compliance(Um, pb.F, gradient=False)
    c_d = ctx.compliance(Ud, pb.F, gradient=False)
    assert c_m == pytest.approx(c_d, rel=1e-7)
    ctx.close()


def test_adjoint_sensitivity_matches_finite_differences(tofe):
    torch = pytest.importorskip("torch")
    nel = (12, 6, 6)
    pb = configs.cantilever(nel, seed=9, lo=0.2, hi=1.0, mg_levels=2)
    ctx = tofe.Context(nel, mg_levels=2, fixed_dof=pb.fixed_dof)

    class ComplianceFn(torch.autograd.Function):
        """c(x) = F.U(x); backward returns dc/dx = P^T(-dEdz u_e^T KE u_e) from the
        HIP adjoint kernels (self-adjoint problem: the adjoint solution is -U)."""

        @staticmethod
        def forward(fctx, x):
            ctx.set_design(x.detach().cpu().numpy())
            ctx.assemble()
            U, its, reason = ctx.solve(pb.F, rtol=1e-13, maxit=2000)
            assert reason == 2
            c, g = ctx.compliance(U, pb.F)
            # residual-corrected compliance F.U + U.r: its error is quadratic in the solve
            # residual r, so finite differences with step 1e-6 see only rounding
            r = pb.F - ctx.apply_K(U)
            fctx.save_for_backward(torch.from_numpy(g))
            return torch.tensor(c + U @ r, dtype=torch.float64)

        @staticmethod
        def backward(fctx, gout):
            (g,) = fctx.saved_tensors
            return gout * g

    x = torch.tensor(pb.x, dtype=torch.float64, requires_grad=True)
    c = ComplianceFn.apply(x)
    c.backward()
    grad = x.grad.numpy()
    assert np.all(grad < 0)
    # Step 1e-4: with c ~ 50 and gradient entries ~ 1e-2, a 1e-6 step leaves only
    # eps*c/h ~ 1e-8 of resolution (1e-5 relative) -- the oracle shows the same -- while
    # the O(h^2) truncation of central differences at 1e-4 is below 1e-7 relative.
    h = 1e-4
    rng = np.random.default_rng(4)
    for e in rng.choice(pb.nelem, 8, replace=False):
        xp, xm = pb.x.copy(), pb.x.copy()
        xp[e] += h
        xm[e] -= h
        cp = ComplianceFn.apply(torch.tensor(xp)).item()
        cm = ComplianceFn.apply(torch.tensor(xm)).item()
        fd = (cp - cm) / (2 * h)
        assert fd == pytest.approx(grad[e], rel=1e-6), f"element {e}"
    ctx.close()
