"""Lumped mass operator and the batched element quadratic forms on the GPU against the
oracle (Frequency::DiagMassFnc, Dynamic.cpp:181-236; eigen-sensitivity element terms,
Dynamic.cpp:113-154).  Bars: the mass diagonal and M X are bit-exact; phi_e^T KE phi_e is
bit-exact against the oracle's element kernel (the compliance sensitivity with dE/dz = 1);
sum phi_e^2 to 1e-15 relative against numpy."""
import numpy as np
import pytest

from topopt3d_amd import configs
from topopt3d_amd.topopt import TopOpt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nel,h", [((8, 4, 4), (1.0, 1.0, 1.0)), ((13, 7, 5), (0.5, 0.25, 2.0))])
def test_mass_bit_exact(oracle, tofe, nel, h):
    pb = configs.cantilever(nel, seed=4, lo=0.05, hi=1.0)
    print(f"case mass nel={nel} h={h}", flush=True)
    ctx = tofe.Context(nel, h=h, mg_levels=2, fixed_dof=pb.fixed_dof)
    ctx.set_design(pb.x)
    rho = ctx.field(tofe.FIELD_RHO)
    ml = np.zeros(pb.ndof)
    ml[-6:] = [0.5, 0.5, 0.5, 1.5, 1.5, 1.5]
    for mlump in (None, ml):
        ctx.assemble_mass(7.85, mlump)
        Mo = oracle.mass_diag(*nel, rho, 7.85, h, mlump=mlump, fixed_dof=pb.fixed_dof)
        M = ctx.field(tofe.FIELD_MDIAG)
        assert np.array_equal(M, Mo)
        X = np.random.default_rng(2).standard_normal((3, pb.ndof))
        assert np.array_equal(ctx.apply_M(X), Mo * X)
    ctx.close()


def test_element_quadratic(oracle, tofe):
    nel = (12, 6, 6)
    pb = configs.cantilever(nel, seed=4, lo=0.05, hi=1.0)
    ctx = tofe.Context(nel, mg_levels=2, fixed_dof=pb.fixed_dof)
    ctx.set_design(pb.x)
    phi = np.random.default_rng(5).standard_normal((4, pb.ndof))
    kq, mq = ctx.element_quadratic(phi)
    KE = oracle.ke()
    ones = np.ones(pb.nelem)
    mesh = TopOpt()
    mesh.CreateMesh([0, nel[0], 0, nel[1], 0, nel[2]], nel, 1.5)
    dofs = (3 * mesh.element[:, :, None] + np.arange(3)).reshape(pb.nelem, 24)
    for j in range(4):
        assert np.array_equal(kq[j], -oracle.sensitivity(*nel, KE, phi[j], ones))
        assert np.allclose(mq[j], (phi[j][dofs] ** 2).sum(1), rtol=1e-15, atol=0)
    # rigid translation: zero strain energy in every element
    t = np.zeros(pb.ndof)
    t[0::3] = 1.0
    kq0, _ = ctx.element_quadratic(t, mass=False)
    assert np.abs(kq0).max() <= 1e-14
    ctx.close()


def consistent_mass_csr(nel, h, V, density, fixed=None, mlump=None):
    """Consistent mass assembled from its closed form: per element V_e * density * the exact
    integral of N_a N_b over the box (a tensor product of the 1-D matrices h/6 [[2, 1],
    [1, 2]]) times I_3; fixed rows and columns zero with 1e-8 on the diagonal."""
    import scipy.sparse as sp

    m1 = [hd / 6 * np.array([[2.0, 1.0], [1.0, 2.0]]) for hd in h]
    LN = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    Me = np.array([[m1[0][a[0], b[0]] * m1[1][a[1], b[1]] * m1[2][a[2], b[2]] for b in LN]
                   for a in LN])
    Me = np.kron(Me, np.eye(3)) * density
    mesh = TopOpt()
    mesh.CreateMesh([0, nel[0] * h[0], 0, nel[1] * h[1], 0, nel[2] * h[2]], nel, 1.5)
    dofs = (3 * mesh.element[:, :, None] + np.arange(3)).reshape(-1, 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    vals = (V[:, None, None] * Me[None]).reshape(len(V), -1).ravel()
    n = 3 * mesh.nNode
    M = sp.csr_matrix((vals, (rows, cols)), shape=(n, n)).tolil()
    if mlump is not None:
        M.setdiag(M.diagonal() + mlump)
    if fixed is not None:
        f = np.flatnonzero(fixed)
        M[f, :] = 0
        M[:, f] = 0
        for d in f:
            M[d, d] = 1e-8
    return M.tocsr(), Me, dofs


@pytest.mark.parametrize("nel,h", [((8, 4, 4), (1.0, 1.0, 1.0)), ((7, 5, 3), (0.5, 0.25, 2.0))])
def test_consistent_mass_matches_closed_form(tofe, nel, h):
    """tofe_set_mass_form(CONSISTENT): M X against the closed-form assembly (1e-15
    relative), the element forms phi_e^T (M_e / m0) phi_e, the block apply bit for bit
    against single vectors, and the total mass sum_e V_e elemSize density per component."""
    pb = configs.cantilever(nel, seed=4, lo=0.05, hi=1.0)
    print(f"case consistent mass nel={nel} h={h}", flush=True)
    ctx = tofe.Context(nel, h=h, mg_levels=2, fixed_dof=pb.fixed_dof)
    ctx.set_design(pb.x)
    rho = ctx.field(tofe.FIELD_RHO)
    ctx.set_mass_form(tofe.MASS_CONSISTENT)
    density = 7.85
    ml = np.zeros(pb.ndof)
    ml[-6:] = [0.5, 0.5, 0.5, 1.5, 1.5, 1.5]
    X = np.random.default_rng(2).standard_normal((5, pb.ndof))
    for mlump in (None, ml):
        ctx.assemble_mass(density, mlump)
        M, Me, dofs = consistent_mass_csr(nel, h, rho, density, pb.fixed_dof, mlump)
        Y = ctx.apply_M(X)
        for j in range(5):
            ref = M @ X[j]
            assert np.abs(Y[j] - ref).max() <= 1e-14 * np.abs(ref).max()
            np.testing.assert_array_equal(Y[j], ctx.apply_M(X[j]))
    # element forms: m0 = elemSize * density / 24
    m0 = h[0] * h[1] * h[2] * density / 24
    _, mq = ctx.element_quadratic(X[:2])
    for j in range(2):
        ref = np.einsum("ea,ab,eb->e", X[j][dofs], Me / m0, X[j][dofs])
        assert np.allclose(mq[j], ref, rtol=1e-13, atol=0)
    ctx.close()
    # total mass per component without supports: e_c^T M e_c = sum_e V_e elemSize density
    free = tofe.Context(nel, h=h, mg_levels=1)
    free.set_design(pb.x)
    free.set_mass_form(tofe.MASS_CONSISTENT)
    free.assemble_mass(density)
    rho_f = free.field(tofe.FIELD_RHO)
    for c in range(3):
        e = np.zeros(pb.ndof)
        e[c::3] = 1.0
        total = e @ free.apply_M(e)
        assert total == pytest.approx(rho_f.sum() * h[0] * h[1] * h[2] * density, rel=1e-13)
    free.close()
