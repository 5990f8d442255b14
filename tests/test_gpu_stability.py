"""Linearised-buckling pieces on the GPU against the oracle (Stability::StressFnc /
Function, Buckling.cpp:15-414): the stress-stiffness apply against the oracle's assembled
Ks (<= 1e-13 relative), the element quadratic forms phi^T ks phi (<= 1e-13), the adjoint
right-hand sides dKsdU (<= 1e-12; the GPU uses the closed form (D B e_d) . tau, the oracle
the 24 unit-displacement matrices), and a_e^T KE b_e (<= 1e-13, by polarisation of the
oracle's element kernel)."""
import numpy as np
import pytest

from topopt3d_amd import configs

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("nel,h", [((8, 4, 4), (1.0, 1.0, 1.0)), ((9, 5, 4), (0.5, 0.75, 1.25))])
def test_stress_stiffness_pieces(oracle, tofe, nel, h):
    print(f"case stability nel={nel} h={h}", flush=True)
    pb = configs.cantilever(nel, seed=6, lo=0.05, hi=1.0)
    ctx = tofe.Context(nel, h=h, mg_levels=2, fixed_dof=pb.fixed_dof)
    ctx.set_design(pb.x)
    Es = ctx.field(tofe.FIELD_ES)
    rng = np.random.default_rng(9)
    U = rng.standard_normal(pb.ndof)
    U[pb.fixed_dof.astype(bool)] = 0
    ctx.set_stress(U)
    Ks = oracle.assemble_Ks(*nel, U, Es, pb.fixed_dof, h=h).to_scipy()
    X = rng.standard_normal((3, pb.ndof))
    Y = ctx.apply_Ks(X)
    for k in range(3):
        assert rel(Y[k], Ks @ X[k]) <= 1e-13
    phi = rng.standard_normal((2, pb.ndof))
    phi[:, pb.fixed_dof.astype(bool)] = 0
    q, r = ctx.stability_terms(phi)
    assert rel(q, oracle.ks_quadratic(*nel, U, phi, h=h)) <= 1e-13
    assert rel(r, oracle.stability_rhs(*nel, Es, phi, pb.fixed_dof, h=h)) <= 1e-12
    A = rng.standard_normal((2, pb.ndof))
    b = rng.standard_normal(pb.ndof)
    out = ctx.element_bilinear(A, b)
    KE = ctx.ke()
    ones = np.ones(pb.nelem)
    for j in range(2):
        pol = (oracle.sensitivity(*nel, KE, A[j] - b, ones) -
               oracle.sensitivity(*nel, KE, A[j] + b, ones)) / 4
        assert rel(out[j], pol) <= 1e-13
    ctx.close()


def test_stress_state_errors(tofe):
    nel = (4, 2, 2)
    ctx = tofe.Context(nel, mg_levels=1)
    ctx.set_design(np.full(16, 0.5))
    with pytest.raises(tofe.TofeError):
        ctx.apply_Ks(np.zeros(ctx.ndof))
    ctx.close()
