"""Spring supports (FEAnalysis.cpp:360-368, 471): K + diag(spKVec) on the free DOFs, the
Dirichlet rows staying unit rows.  The GPU operator, its Galerkin hierarchy and the solve
against the oracle's assembled K with the spring diagonal added (scipy), and P^T K P from the
oracle's transfer operators."""
import numpy as np
import pytest
import scipy.sparse as sp

from topopt3d_amd import configs

pytestmark = pytest.mark.gpu


def _springs(pb, nel):
    """Springs on the x = max face (z stiffness 0.05, x 0.02) and on the clamped face (to be
    ignored there)."""
    nx, ny, nz = nel
    spk = np.zeros(pb.ndof)
    for j in range(ny + 1):
        for k in range(nz + 1):
            n = configs.node_index(nel, nx, j, k)
            spk[3 * n + 2] = 0.05
            spk[3 * n] = 0.02
            spk[3 * configs.node_index(nel, 0, j, k) + 1] = 7.0
    free = spk * (1 - pb.fixed_dof)
    return spk, free


@pytest.mark.parametrize("precision", [0, 1])
def test_springs_operator_hierarchy_and_solve(oracle, tofe, precision):
    nel, L = (12, 6, 6), 3
    print(f"case springs nel={nel} precision={precision}", flush=True)
    pb = configs.cantilever(nel, seed=7, lo=0.05, hi=1.0, mg_levels=L)
    spk, free = _springs(pb, nel)
    ctx = tofe.Context(nel, mg_levels=L, fixed_dof=pb.fixed_dof, precision=precision)
    ctx.set_design(pb.x)
    ctx.set_springs(spk)
    ctx.assemble()
    E = ctx.field(tofe.FIELD_E)
    K0 = oracle.assemble_K(*nel, oracle.ke(), E, pb.fixed_dof).to_scipy()
    Ks = (K0 + sp.diags(free)).tocsr()
    X = np.random.default_rng(3).standard_normal((3, pb.ndof))
    Y = ctx.apply_K(X)
    for j in range(3):
        ref = Ks @ X[j]
        assert np.abs(Y[j] - ref).max() <= 1e-14 * np.abs(ref).max()
    np.testing.assert_allclose(ctx.field(tofe.FIELD_DIAG), Ks.diagonal(), rtol=1e-15)
    if precision == 0:
        mg = oracle.MG(oracle.CSR.from_scipy(Ks), tuple(n + 1 for n in nel), L)
        for l in range(1, L):
            S = ctx.level_stencil(l)
            A = mg.A(l).to_scipy().tocoo()
            d = mg.dims(l)
            n, m = A.row // 3, A.col // 3
            ni, nj, nk = n % d[0], (n // d[0]) % d[1], n // (d[0] * d[1])
            mi, mj, mk = m % d[0], (m // d[0]) % d[1], m // (d[0] * d[1])
            o = (mi - ni + 1) + 3 * (mj - nj + 1) + 9 * (mk - nk + 1)
            vals = S[o, A.row % 3, A.col % 3, n]
            assert np.abs(vals - A.data).max() <= 1e-14 * np.abs(A.data).max()
    U, its, reason = ctx.solve(pb.F, rtol=1e-10)
    assert reason == tofe.CONVERGED_RTOL
    res = np.linalg.norm(pb.F - Ks @ U) / np.linalg.norm(pb.F)
    assert res <= 1.01e-10
    # the springs stiffen the structure: compliance drops against the same design without them
    ctx.set_springs(None)
    ctx.assemble()
    U0, _, _ = ctx.solve(pb.F, rtol=1e-10)
    assert pb.F @ U < pb.F @ U0
    ctx.close()
