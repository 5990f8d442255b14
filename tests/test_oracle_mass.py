"""Pins of the oracle's lumped mass (orc_mass_diag; Frequency::DiagMassFnc,
Dynamic.cpp:181-236 and the Dirichlet treatment at 44-45): 24 m0 = elemSize * density, so
the free diagonal sums to the total mass sum_e V_e elemSize density; an interior node of a
uniform design carries 8 V m0 per direction; lumped masses add on the diagonal; fixed DOFs
hold 1e-8."""
import numpy as np
import pytest

from topopt3d_amd import configs


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.25, 2.0)])
def test_mass_conservation_and_closed_forms(oracle, h):
    nel = (7, 5, 4)
    ne = int(np.prod(nel))
    V = configs.uniform(9, ne, 0.01, 1.0)
    rho_m = 7.85
    M = oracle.mass_diag(*nel, V, rho_m, h)
    elem = h[0] * h[1] * h[2]
    assert M.sum() == pytest.approx(V.sum() * elem * rho_m, rel=1e-13)
    Mu = oracle.mass_diag(*nel, np.full(ne, 0.4), rho_m, h)
    Nx, Ny = nel[0] + 1, nel[1] + 1
    inner = 3 + Nx * (2 + Ny * 2)
    assert Mu[3 * inner:3 * inner + 3] == pytest.approx([8 * 0.4 * elem * rho_m / 24] * 3, rel=1e-15)
    assert Mu[0] == pytest.approx(0.4 * elem * rho_m / 24, rel=1e-15)   # corner: one element
    ml = np.zeros(M.size)
    ml[30:33] = [1.0, 2.0, 3.0]
    Ml = oracle.mass_diag(*nel, V, rho_m, h, mlump=ml)
    assert np.array_equal(Ml - M, ml)
    pb = configs.cantilever(nel)
    Mf = oracle.mass_diag(*nel, V, rho_m, h, fixed_dof=pb.fixed_dof)
    fx = pb.fixed_dof.astype(bool)
    assert (Mf[fx] == 1e-8).all() and np.array_equal(Mf[~fx], M[~fx])
