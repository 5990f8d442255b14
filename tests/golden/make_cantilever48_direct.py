"""Writes tests/golden/cantilever48_direct.npz: an independent sparse-direct solution of the
48x24x24 cantilever (91,875 DOF, x = 0.3, R = 1.5, p = 3, eps = 1e-10).

Nothing here uses the oracle or the HIP library: the element matrix comes from exact
integration of the Q1 shape-function products (1-D stiffness/mass factors), the
stiffness matrix is assembled with scipy.sparse COO, the Dirichlet rows/columns follow
MatZeroRowsColumns(K, fixed, 1.0), and the system is solved with SciPy's SuperLU.
The fixture stores scalars and seeded samples only (a few KB).

Run:  python tests/golden/make_cantilever48_direct.py      (about 4 minutes on 8 cores)
"""
import itertools
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
LN = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def ke_exact(E0=1.0, nu=0.3):
    lam = E0 * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E0 / (2 * (1 + nu))
    st = lambda a, b: 1.0 if a == b else -1.0
    ma = lambda a, b: 1 / 3 if a == b else 1 / 6
    sg = lambda a: 1.0 if a else -1.0
    K = np.zeros((24, 24))
    for A, B in itertools.product(range(8), repeat=2):
        a, b = LN[A], LN[B]
        I = [np.prod([st(a[d], b[d]) if d == k else ma(a[d], b[d]) for d in range(3)]) for k in range(3)]
        for i in range(3):
            K[3 * A + i, 3 * B + i] = (lam + 2 * mu) * I[i] + mu * sum(I[k] for k in range(3) if k != i)
            for j in range(3):
                if j != i:
                    k = 3 - i - j
                    m = ma(a[k], b[k])
                    K[3 * A + i, 3 * B + j] = lam * sg(a[i]) * sg(b[j]) * m / 4 + mu * sg(a[j]) * sg(b[i]) * m / 4
    return K


def main():
    nx, ny, nz = 48, 24, 24
    Nx, Ny, Nz = nx + 1, ny + 1, nz + 1
    nn = Nx * Ny * Nz
    ndof = 3 * nn
    eps, p = 1e-10, 3.0
    rho = 0.3                       # uniform design: every filter row sums to one
    Ee = eps + (1 - eps) * rho ** p
    dEdz = (1 - eps) * p * rho ** (p - 1)
    KE = ke_exact()
    ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    base = (ex + Nx * (ey + Ny * ez)).ravel()
    offs = [a + Nx * (b + Ny * c) for a, b, c in LN]
    nodes = base[:, None] + np.array(offs)[None, :]
    dofs = (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    vals = np.tile((Ee * KE).ravel(), dofs.shape[0])
    K = sp.coo_matrix((vals, (rows, cols)), shape=(ndof, ndof)).tocsr()
    fixed = np.zeros(ndof, bool)
    xnodes = np.arange(0, nn, Nx)          # i = 0 face
    for c in range(3):
        fixed[3 * xnodes + c] = True
    keep = sp.diags((~fixed).astype(float))
    K = (keep @ K @ keep + sp.diags(fixed.astype(float))).tocsc()
    F = np.zeros(ndof)
    load = nx + Nx * (ny // 2 + Ny * 0)
    F[3 * load + 2] = -1.0
    U = spla.spsolve(K, F, permc_spec="MMD_AT_PLUS_A")
    res = np.linalg.norm(K @ U - F)
    c = U @ F
    uel = U[dofs]
    g = -dEdz * np.einsum("ei,ij,ej->e", uel, KE, uel)
    rng = np.random.default_rng(1)
    dof_s = np.sort(rng.choice(ndof, 64, replace=False))
    el_s = np.sort(rng.choice(nx * ny * nz, 32, replace=False))
    out = os.path.join(HERE, "cantilever48_direct.npz")
    np.savez(out, compliance=c, residual=res, U_load=U[3 * load:3 * load + 3], dof_samples=dof_s,
             U_samples=U[dof_s], U_norm=np.linalg.norm(U), el_samples=el_s, g_samples=g[el_s],
             g_sum=g.sum())
    print("wrote", out, "compliance", c, "residual", res)
    return 0


if __name__ == "__main__":
    sys.exit(main())
