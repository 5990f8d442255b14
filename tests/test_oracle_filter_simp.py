"""Pins of the oracle's density filter (Filter.cpp:21-150) and SIMP interpolation
(Material_Interpolation.cpp:20-75).

Filter: neighbourhood counts and unnormalised row sums fixed by geometry, a
brute-force set-based construction on small boxes (pattern and weights), row sums of
one, pattern symmetry, and the clipped corner neighbourhood.
SIMP: closed forms E = eps + (1-eps) rho^p, dE/drho = (1-eps) p rho^(p-1).
"""
import itertools
import math

import numpy as np
import pytest


def brute_filter(nx, ny, nz, R, h=(1.0, 1.0, 1.0)):
    """Independent construction: for every pair of element centres at distance < R
    (offsets limited to |i| <= R/h), weight 1 - d/R, then divide by the row sum."""
    ne = nx * ny * nz
    rows = []
    for ez, ey, ex in itertools.product(range(nz), range(ny), range(nx)):
        ents = {}
        for k, j, i in itertools.product(range(nz), range(ny), range(nx)):
            di, dj, dk = i - ex, j - ey, k - ez
            if abs(di) > int(R / h[0]) or abs(dj) > int(R / h[1]) or abs(dk) > int(R / h[2]):
                continue
            d = math.sqrt((di * h[0]) ** 2 + (dj * h[1]) ** 2 + (dk * h[2]) ** 2)
            if d < R:
                ents[i + nx * (j + ny * k)] = 1 - d / R
        rows.append(ents)
    return rows


def test_interior_neighbourhoods(oracle):
    F = oracle.filter_csr(10, 10, 10, 1.5, scale=False)
    rp, col, val = F.arrays()
    el = 5 + 10 * (5 + 10 * 5)
    assert rp[el + 1] - rp[el] == 19
    assert sum(val[rp[el]:rp[el + 1]]) == 3.6862915010152384
    F = oracle.filter_csr(10, 10, 10, 2.0, scale=False)
    rp, col, val = F.arrays()
    assert rp[el + 1] - rp[el] == 27
    assert sum(val[rp[el]:rp[el + 1]]) == 8.5865153954859181
    # closed form of the R = 1.5 row sum: 1 + 6(1-1/1.5) + 12(1-sqrt2/1.5)
    assert 1 + 6 * (1 - 1 / 1.5) + 12 * (1 - math.sqrt(2) / 1.5) == pytest.approx(3.6862915010152384, rel=1e-15)


def test_corner_clipping(oracle):
    F = oracle.filter_csr(6, 5, 4, 1.5, scale=False)
    rp, col, val = F.arrays()
    assert rp[1] - rp[0] == 7           # self, 3 faces, 3 edges; the body diagonal is sqrt3 > 1.5
    assert list(col[rp[0]:rp[1]]) == sorted(col[rp[0]:rp[1]])


@pytest.mark.parametrize("nel,R,h", [((5, 4, 3), 1.5, (1, 1, 1)), ((4, 4, 4), 2.0, (1, 1, 1)),
                                      ((6, 3, 4), 2.3, (1, 1, 1)), ((5, 4, 4), 1.2, (0.5, 0.6, 0.7))])
def test_filter_against_brute_force(oracle, nel, R, h):
    nx, ny, nz = nel
    F = oracle.filter_csr(nx, ny, nz, R, h=h)
    rp, col, val = F.arrays()
    rows = brute_filter(nx, ny, nz, R, h)
    for e, ents in enumerate(rows):
        cols = sorted(ents)
        assert list(col[rp[e]:rp[e + 1]]) == cols
        w = np.array([ents[c] for c in cols])
        assert np.allclose(val[rp[e]:rp[e + 1]], w / w.sum(), rtol=1e-14, atol=0)


def test_filter_rows_and_symmetry(oracle):
    F = oracle.filter_csr(9, 7, 5, 2.0)
    A = F.to_scipy()
    assert np.allclose(np.asarray(A.sum(axis=1)).ravel(), 1.0, rtol=0, atol=4e-16)
    pattern = (A != 0).astype(int)
    assert (pattern - pattern.T).nnz == 0
    assert A.data.min() > 0


def test_vecpow_special_exponents(oracle):
    x = 0.37
    assert oracle.vecpow(x, 2.0) == x * x
    assert oracle.vecpow(x, 1.0) == x
    assert oracle.vecpow(x, 0.5) == math.sqrt(x)
    assert oracle.vecpow(x, -1.0) == 1 / x
    assert oracle.vecpow(x, 2.5) == pytest.approx(x ** 2.5, rel=1e-15)


@pytest.mark.parametrize("p", [3.0, 1.0, 3.5])
def test_simp_closed_form(oracle, p):
    rho = np.linspace(0.01, 1.0, 101)
    eps = 1e-10
    E, dEdz, Es, dEsdz = oracle.simp(rho, p, eps)
    assert np.allclose(Es, rho ** p, rtol=1e-14, atol=0)
    assert np.allclose(E, eps + (1 - eps) * rho ** p, rtol=1e-14, atol=0)
    assert np.allclose(dEsdz, p * rho ** (p - 1), rtol=1e-14, atol=0)
    assert np.allclose(dEdz, (1 - eps) * p * rho ** (p - 1), rtol=1e-14, atol=0)
    # derivative consistency by central differences
    hstep = 1e-6
    Ep, _, _, _ = oracle.simp(rho + hstep, p, eps)
    Em, _, _, _ = oracle.simp(rho - hstep, p, eps)
    assert np.allclose((Ep - Em) / (2 * hstep), dEdz, rtol=1e-7, atol=1e-9)
