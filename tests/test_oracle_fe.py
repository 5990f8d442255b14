"""Pins of the oracle's assembly, Dirichlet semantics, transfer operators, Galerkin
coarse operators, MG-preconditioned CG and compliance sensitivities.

Independent checks: dense numpy assembly by np.add.at on tiny meshes, rigid-body
null space, dense numpy solves, SciPy sparse-direct solve of the 48x24x24 cantilever,
exact prolongation of trilinear fields, dense P^T A P, and central finite differences
of the compliance with respect to the design variables.
"""
import numpy as np
import pytest

from topopt3d_amd import configs


def dense_K(oracle, nel, KE, E, fixed=None):
    nx, ny, nz = nel
    nn = (nx + 1) * (ny + 1) * (nz + 1)
    K = np.zeros((3 * nn, 3 * nn))
    LN = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    for ez in range(nz):
        for ey in range(ny):
            for ex in range(nx):
                e = ex + nx * (ey + ny * ez)
                nodes = [(ex + a) + (nx + 1) * ((ey + b) + (ny + 1) * (ez + c)) for a, b, c in LN]
                dofs = np.array([3 * n + d for n in nodes for d in range(3)])
                np.add.at(K, (dofs[:, None], dofs[None, :]), E[e] * KE)
    if fixed is not None:
        f = np.flatnonzero(fixed)
        K[f, :] = 0
        K[:, f] = 0
        K[f, f] = 1
    return K


def rbm(nel):
    nx, ny, nz = nel
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    x, y, z = i.ravel().astype(float), j.ravel().astype(float), k.ravel().astype(float)
    nn = x.size
    R = np.zeros((3 * nn, 6))
    R[0::3, 0] = 1
    R[1::3, 1] = 1
    R[2::3, 2] = 1
    R[0::3, 3], R[1::3, 3] = -y, x
    R[1::3, 4], R[2::3, 4] = -z, y
    R[0::3, 5], R[2::3, 5] = z, -x
    return R


@pytest.mark.parametrize("nel", [(2, 2, 2), (3, 2, 4), (4, 3, 2)])
def test_assembly_matches_dense(oracle, nel):
    rng = np.random.default_rng(0)
    ne = nel[0] * nel[1] * nel[2]
    E = rng.uniform(0.1, 1.0, ne)
    KE = oracle.ke()
    Kd = dense_K(oracle, nel, KE, E)
    K = oracle.assemble_K(*nel, KE, E).to_scipy().toarray()
    assert np.abs(K - Kd).max() <= 1e-15 * np.abs(Kd).max()
    assert np.abs(K - K.T).max() <= 1e-16
    assert np.abs(K @ rbm(nel)).max() < 1e-13
    pb = configs.cantilever(nel)
    Kf = oracle.assemble_K(*nel, KE, E, pb.fixed_dof).to_scipy().toarray()
    Kfd = dense_K(oracle, nel, KE, E, pb.fixed_dof)
    assert np.abs(Kf - Kfd).max() <= 1e-15 * np.abs(Kd).max()
    f = np.flatnonzero(pb.fixed_dof)
    assert np.all(Kf[f][:, f] == np.eye(f.size))


def test_uniform_density_scales_K(oracle):
    nel = (4, 3, 3)
    KE = oracle.ke()
    ne = 36
    K1 = oracle.assemble_K(*nel, KE, np.ones(ne)).to_scipy()
    K3 = oracle.assemble_K(*nel, KE, np.full(ne, 0.027)).to_scipy()
    assert abs(K3 - 0.027 * K1).max() <= 1e-15


@pytest.mark.parametrize("Nf", [(9, 5, 5), (8, 6, 5), (4, 3, 7)])
def test_transfer_operator(oracle, Nf):
    P = oracle.build_P(*Nf).to_scipy()
    Nc = tuple((n + 1) // 2 for n in Nf)
    nf, nc = np.prod(Nf), np.prod(Nc)
    assert P.shape == (3 * nf, 3 * nc)
    assert np.allclose(np.asarray(P.sum(axis=1)).ravel(), 1.0, rtol=0, atol=0)
    assert set(np.unique(P.data)) <= {0.125, 0.25, 0.5, 1.0}
    # trilinear field sampled on the coarse grid is reproduced on the fine grid where the
    # fine node lies between two coarse nodes (odd node counts); with an even count the
    # last fine node copies its single parent (row scaling)
    def field(ix, iy, iz):
        return 1.5 + 0.3 * ix - 0.7 * iy + 0.2 * iz + 0.05 * ix * iy * iz

    Kc, Jc, Ic = np.meshgrid(np.arange(Nc[2]), np.arange(Nc[1]), np.arange(Nc[0]), indexing="ij")
    uc = np.zeros(3 * nc)
    uc[0::3] = field(2.0 * Ic.ravel(), 2.0 * Jc.ravel(), 2.0 * Kc.ravel())
    uf = P @ uc
    Kf, Jf, If = np.meshgrid(np.arange(Nf[2]), np.arange(Nf[1]), np.arange(Nf[0]), indexing="ij")
    X = [If.ravel().astype(float), Jf.ravel().astype(float), Kf.ravel().astype(float)]
    for d in range(3):
        if Nf[d] % 2 == 0:
            X[d] = np.minimum(X[d], Nf[d] - 2)
    assert np.allclose(uf[0::3], field(*X), rtol=1e-14, atol=1e-14)
    assert np.all(uf[1::3] == 0)


def test_galerkin_and_mg_solve_small(oracle):
    nel = (8, 4, 4)
    pb = configs.cantilever(nel, seed=11, lo=0.05)
    KE = oracle.ke()
    rho = oracle.filter_csr(*nel, 1.5).matvec(pb.x)
    E = oracle.simp(rho)[0]
    K = oracle.assemble_K(*nel, KE, E, pb.fixed_dof)
    mg = oracle.MG(K, (9, 5, 5), 3)
    A0 = K.to_scipy().toarray()
    P0 = mg.P(0).to_scipy().toarray()
    A1 = mg.A(1).to_scipy().toarray()
    assert np.abs(A1 - P0.T @ A0 @ P0).max() <= 1e-14 * np.abs(A1).max()
    P1 = mg.P(1).to_scipy().toarray()
    A2 = mg.A(2).to_scipy().toarray()
    assert np.abs(A2 - P1.T @ A1 @ P1).max() <= 1e-14 * np.abs(A2).max()
    # lambda estimate below the true lambda_max(D^-1 A) and close to it
    d = np.diag(A0)
    lmax = np.linalg.eigvalsh(A0 / np.sqrt(np.outer(d, d))).max()
    assert mg.lam(0) <= lmax * (1 + 1e-12)
    assert mg.lam(0) >= 0.8 * lmax
    x, its, reason, rn = mg.solve(pb.F, rtol=1e-12)
    assert reason == 2
    xd = np.linalg.solve(A0, pb.F)
    assert np.abs(x - xd).max() <= 1e-10 * np.abs(xd).max()
    assert rn <= 1e-12 * np.linalg.norm(pb.F)
    # compliance identities c = U.F = U.K.U
    assert x @ pb.F == pytest.approx(x @ A0 @ x, rel=1e-12)


def test_coarsest_operator_as_element_sum(oracle):
    """The identity the library's direct coarsest operator (k_direct.hip) rests on: when every
    element count is a multiple of r = 2^(L-1), the Galerkin chain's coarsest operator equals
    sum_e E_e W_e^T KE_masked,e W_e + sum over fixed DOFs of w w^T + sum over free spring DOFs of
    s w w^T, with W_e the trilinear weights of the element's nodes on the coarsest cell's nodes."""
    nel, L = (8, 4, 4), 3
    r = 2 ** (L - 1)
    nx, ny, nz = nel
    pb = configs.cantilever(nel, seed=11, lo=0.05, mg_levels=L)
    KE = oracle.ke()
    E = oracle.simp(oracle.filter_csr(*nel, 1.5).matvec(pb.x))[0]
    fixed = pb.fixed_dof.astype(bool)
    Nf = (nx + 1, ny + 1, nz + 1)
    spk = np.zeros(3 * Nf[0] * Nf[1] * Nf[2])
    spk[3 * np.arange(Nf[0] * Nf[1] * Nf[2])[::7] + 2] = 0.3  # springs, some on fixed DOFs
    free_spk = spk * ~fixed
    import scipy.sparse as sp
    K = (oracle.assemble_K(nx, ny, nz, KE, E, pb.fixed_dof).to_scipy() + sp.diags(free_spk)).tocsr()
    mg = oracle.MG(oracle.CSR.from_scipy(K), Nf, L)
    chain = mg.A(L - 1).to_scipy().toarray()
    Nc = (nx // r + 1, ny // r + 1, nz // r + 1)
    nc = 3 * Nc[0] * Nc[1] * Nc[2]

    def weights(fx, fy, fz):  # (coarse node, weight) of a fine node
        out = []
        for Iz in range(Nc[2]):
            for Iy in range(Nc[1]):
                for Ix in range(Nc[0]):
                    wt = (max(0.0, 1 - abs(fx - r * Ix) / r) * max(0.0, 1 - abs(fy - r * Iy) / r)
                          * max(0.0, 1 - abs(fz - r * Iz) / r))
                    if wt:
                        out.append((Ix + Nc[0] * (Iy + Nc[1] * Iz), wt))
        return out

    LN = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    A = np.zeros((nc, nc))
    for ez in range(nz):
        for ey in range(ny):
            for ex in range(nx):
                e = ex + nx * (ey + ny * ez)
                W = np.zeros((24, nc))
                dofs = []
                for k, (a, b, c) in enumerate(LN):
                    n = (ex + a) + Nf[0] * ((ey + b) + Nf[1] * (ez + c))
                    dofs += [3 * n, 3 * n + 1, 3 * n + 2]
                    for J, wt in weights(ex + a, ey + b, ez + c):
                        for d in range(3):
                            W[3 * k + d, 3 * J + d] = wt
                Km = KE.copy()
                m = fixed[dofs]
                Km[m, :] = 0
                Km[:, m] = 0
                A += E[e] * W.T @ Km @ W
    for n in range(Nf[0] * Nf[1] * Nf[2]):
        f = (n % Nf[0], (n // Nf[0]) % Nf[1], n // (Nf[0] * Nf[1]))
        for d in range(3):
            s = 1.0 if fixed[3 * n + d] else free_spk[3 * n + d]
            if s:
                w = np.zeros(nc)
                for J, wt in weights(*f):
                    w[3 * J + d] = wt
                A += s * np.outer(w, w)
    assert np.abs(A - chain).max() <= 1e-14 * np.abs(chain).max()


def test_pcg_reason_codes(oracle):
    nel = (8, 4, 4)
    pb = configs.cantilever(nel)
    K = oracle.assemble_K(*nel, oracle.ke(), np.full(128, 0.027), pb.fixed_dof)
    mg = oracle.MG(K, (9, 5, 5), 2)
    x, its, reason, _ = mg.solve(pb.F, rtol=1e-12, maxit=2)
    assert reason == -3 and its == 2
    x, its, reason, _ = mg.solve(pb.F, rtol=1e-8)
    assert reason == 2
    x2, its2, reason2, _ = mg.solve(pb.F, x0=x, rtol=1e-8)
    assert its2 == 0 and reason2 == 2


def test_cantilever48_against_sparse_direct(oracle):
    """The 48x24x24 cantilever (91,875 DOF) by the oracle's 3-level MG-CG to 1e-12
    against the committed SciPy sparse-direct fixture (tests/golden/make_cantilever48_direct.py,
    independent of the oracle): displacements, compliance and element sensitivities."""
    import os

    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "cantilever48_direct.npz"))
    pb = configs.cantilever_48x24x24()
    nel = pb.nel
    KE = oracle.ke()
    rho = oracle.filter_csr(*nel, 1.5).matvec(pb.x)
    assert np.allclose(rho, 0.3, rtol=1e-15)
    E, dEdz, _, _ = oracle.simp(rho)
    K = oracle.assemble_K(*nel, KE, E, pb.fixed_dof)
    mg = oracle.MG(K, tuple(n + 1 for n in nel), 3)
    x, its, reason, rn = mg.solve(pb.F, rtol=1e-12)
    assert reason == 2 and its < 60
    scale = float(fx["U_norm"]) / np.sqrt(x.size)
    assert np.abs(x[fx["dof_samples"]] - fx["U_samples"]).max() <= 1e-9 * np.abs(fx["U_samples"]).max()
    assert np.linalg.norm(x) == pytest.approx(float(fx["U_norm"]), rel=1e-10)
    assert x @ pb.F == pytest.approx(float(fx["compliance"]), rel=1e-10)
    assert scale > 0
    g = oracle.sensitivity(*nel, KE, x, dEdz)
    assert np.allclose(g[fx["el_samples"]], fx["g_samples"], rtol=1e-9, atol=0)
    assert g.sum() == pytest.approx(float(fx["g_sum"]), rel=1e-10)


def test_compliance_sensitivity_finite_differences(oracle):
    """dc/dx from the oracle (sensitivity + filter transpose) against central differences
    (h = 1e-4, truncation ~1e-8 relative) of the residual-corrected compliance on a 6x3x3
    cantilever."""
    nel = (6, 3, 3)
    pb = configs.cantilever(nel, seed=5, lo=0.2, hi=1.0)
    KE = oracle.ke()
    Pf = oracle.filter_csr(*nel, 1.5)
    Nn = tuple(n + 1 for n in nel)

    def comp(x):
        rho = Pf.matvec(x)
        E, dEdz, _, _ = oracle.simp(rho)
        K = oracle.assemble_K(*nel, KE, E, pb.fixed_dof)
        mg = oracle.MG(K, Nn, 2)
        u, its, reason, _ = mg.solve(pb.F, rtol=1e-13, maxit=2000)
        # u.F + u.(F - K u) differs from F.K^-1 F only to second order in the solve error
        return u @ pb.F + u @ (pb.F - K.matvec(u)), u, dEdz

    c0, u, dEdz = comp(pb.x)
    g = oracle.sensitivity(*nel, KE, u, dEdz)
    assert np.all(g <= 0)
    dc = Pf.rmatvec(g)
    h = 1e-4
    for e in [0, 7, 23, 41, 53]:
        xp, xm = pb.x.copy(), pb.x.copy()
        xp[e] += h
        xm[e] -= h
        fd = (comp(xp)[0] - comp(xm)[0]) / (2 * h)
        assert fd == pytest.approx(dc[e], rel=1e-6)
