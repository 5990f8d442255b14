"""Pins of the oracle's element stiffness matrix (FEAnalysis.cpp:869-1050).

The oracle integrates KE with 2x2x2 Gauss quadrature.  Here it is checked against
(a) exact rational values for E0 = 1, nu = 0.3 on the unit cube, (b) an independent
closed form obtained by integrating the Q1 shape-function products exactly with 1-D
stiffness/mass factors, and (c) the rigid-body null space and scaling laws.
"""
import itertools

import numpy as np
import pytest

LN = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def ke_closed_form(E0, nu, h):
    """KE from exact 1-D integrals: int N_a' N_b' = +-1/h, int N_a N_b = h/3 or h/6,
    int N_a' N_b = sgn(a)/2; lam, mu Lame constants.  No quadrature is involved."""
    lam = E0 * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E0 / (2 * (1 + nu))

    def stiff(a, b, hd):
        return (1.0 if a == b else -1.0) / hd

    def mass(a, b, hd):
        return hd / 3 if a == b else hd / 6

    def sgn(a):
        return 1.0 if a else -1.0

    K = np.zeros((24, 24))
    for A, B in itertools.product(range(8), repeat=2):
        a, b = LN[A], LN[B]
        # I[k] = int dN_a/dx_k dN_b/dx_k
        I = []
        for k in range(3):
            v = 1.0
            for d in range(3):
                v *= stiff(a[d], b[d], h[d]) if d == k else mass(a[d], b[d], h[d])
            I.append(v)
        for i in range(3):
            K[3 * A + i, 3 * B + i] = (lam + 2 * mu) * I[i] + mu * sum(I[k] for k in range(3) if k != i)
            for j in range(3):
                if j == i:
                    continue
                k = 3 - i - j
                # int dN_a/dx_i dN_b/dx_j = (sgn a_i / 2)(sgn b_j / 2) mass_k * (h_i h_j h_k)/(h_i h_j h_k)
                m = mass(a[k], b[k], h[k])
                t1 = sgn(a[i]) / 2 * sgn(b[j]) / 2 * m
                t2 = sgn(a[j]) / 2 * sgn(b[i]) / 2 * m
                K[3 * A + i, 3 * B + j] = lam * t1 + mu * t2
    return K


def rigid_modes():
    X = np.array(LN, dtype=float)
    R = np.zeros((24, 6))
    for n in range(8):
        x, y, z = X[n]
        R[3 * n:3 * n + 3, 0] = [1, 0, 0]
        R[3 * n:3 * n + 3, 1] = [0, 1, 0]
        R[3 * n:3 * n + 3, 2] = [0, 0, 1]
        R[3 * n:3 * n + 3, 3] = [-y, x, 0]
        R[3 * n:3 * n + 3, 4] = [0, -z, y]
        R[3 * n:3 * n + 3, 5] = [z, 0, -x]
    return R


def test_ke_rational_anchors(oracle):
    KE = oracle.ke(1.0, 0.3)
    assert KE[0, 0] == pytest.approx(55 / 234, rel=1e-15)
    assert KE[0, 1] == pytest.approx(25 / 312, rel=1e-14)
    assert KE[0, 3] == pytest.approx(-25 / 234, rel=1e-14)
    assert np.trace(KE) == pytest.approx(220 / 39, rel=1e-15)


def test_ke_matches_exact_integration(oracle):
    for E0, nu, h in [(1.0, 0.3, (1, 1, 1)), (2.5, 0.2, (1, 1, 1)), (1.0, 0.3, (0.5, 1.0, 2.0)),
                      (7.0, 0.45, (1.3, 0.7, 0.9))]:
        KE = oracle.ke(E0, nu, h)
        Kc = ke_closed_form(E0, nu, h)
        assert np.abs(KE - Kc).max() <= 1e-13 * np.abs(Kc).max()


def test_ke_spectrum_and_nullspace(oracle):
    KE = oracle.ke(1.0, 0.3)
    assert np.abs(KE - KE.T).max() < 1e-15
    w = np.linalg.eigvalsh(KE)
    assert np.all(np.abs(w[:6]) < 1e-14)
    assert w[6] > 0.05
    assert w[-1] == pytest.approx(1.25, rel=1e-13)
    R = rigid_modes()
    assert np.abs(KE @ R).max() < 1e-14


def test_ke_scaling(oracle):
    K1 = oracle.ke(1.0, 0.3, (1, 1, 1))
    assert np.allclose(oracle.ke(3.0, 0.3), 3 * K1, rtol=1e-14, atol=1e-16)
    # stiffness of a cube scales with its edge length in 3-D
    assert np.allclose(oracle.ke(1.0, 0.3, (2, 2, 2)), 2 * K1, rtol=1e-14, atol=1e-16)


def test_dmat(oracle):
    D = oracle.dmat(1.0, 0.3)
    lam = 0.3 / (1.3 * 0.4)
    mu = 1 / 2.6
    assert D[0, 0] == pytest.approx(lam + 2 * mu, rel=1e-15)
    assert D[0, 1] == pytest.approx(lam, rel=1e-15)
    assert D[3, 3] == pytest.approx(mu, rel=1e-15)
    assert D[0, 3] == 0
