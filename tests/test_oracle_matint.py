"""Pins of the oracle's maximum-length-scale interpolation and filter chain rule
(orc_rfilter_csr, orc_matint, orc_chain_filter; Material_Interpolation.cpp:20-187,
RecMesh.cpp:544,706-718) against what the mathematics fixes:

- R: lattice-point counts of the ball (brute force), symmetry, REdge = max count - count;
- MatIntFnc with vdMin <= 0 reduces to the plain SIMP routine (pinned in test_oracle_filter_simp);
- the chain rule against central finite differences of E(x): exactly where the (y < 1) mask
  is constant over an element's R neighbourhood (all y < 1, all y = 1, and the interior of
  each region in a mixed design); where the mask changes inside the neighbourhood the
  reference applies it after R^T (Material_Interpolation.cpp:174-176), which the oracle
  keeps, and the test checks that form's closed value there (y = 1 elements pass dfdE
  through unchanged);
- SIMP_CUT keeps the SIMP values above the threshold and zeroes them below;
  SIMP_LOGISTIC divides Es by 1 + exp(a (b - z)).
"""
import numpy as np
import pytest

from topopt3d_amd import configs


def ball_count(R, clip=None):
    """Integer offsets with |v| < R (unit spacing), optionally only those with v >= clip."""
    N = int(R)
    n = 0
    for k in range(-N, N + 1):
        for j in range(-N, N + 1):
            for i in range(-N, N + 1):
                if np.sqrt(i * i + j * j + k * k) < R:
                    if clip is None or (i >= clip[0] and j >= clip[1] and k >= clip[2]):
                        n += 1
    return n


@pytest.mark.parametrize("Rmax", [2.0, 2.5, 3.0])
def test_rfilter_counts_and_redge(oracle, Rmax):
    nel = (12, 11, 10)
    R, REdge = oracle.rfilter_csr(*nel, Rmax)
    S = R.to_scipy()
    assert (S.data == 1.0).all()
    assert abs(S - S.T).max() == 0
    counts = np.diff(S.indptr)
    nx, ny, nz = nel
    inner = 5 + nx * (5 + ny * 5)
    assert counts[inner] == ball_count(Rmax) == counts.max()
    assert counts[0] == ball_count(Rmax, clip=(0, 0, 0))
    assert np.array_equal(REdge, counts.max() - counts.astype(float))
    if Rmax == 3.0:
        assert ball_count(3.0) == 93


def test_matint_without_max_length_is_simp(oracle):
    nel = (9, 7, 5)
    x = configs.uniform(4, int(np.prod(nel)), 0.05, 1.0)
    P = oracle.filter_csr(*nel, 1.5)
    for p in (3.0, 3.5):
        o = oracle.matint(P, None, None, x, p, 1.0, -1.0)
        assert (o["y"] == 1.0).all()
        E, dEdz, Es, dEsdz = oracle.simp(P.matvec(x), p, 1e-10)
        for a, b in ((o["E"], E), (o["dEdz"], dEdz), (o["Es"], Es), (o["dEsdz"], dEsdz)):
            assert np.array_equal(a, b)


def _E(oracle, P, R, RE, x, vd, q):
    return oracle.matint(P, R, RE, x, 3.0, q, vd)


def two_region_design(nel):
    """Solid (0.9) for ex < nx/2, void (0.1) beyond, with +-0.05 noise: y < 1 in one region."""
    ne = int(np.prod(nel))
    ex = np.arange(ne) % nel[0]
    return np.where(ex < nel[0] // 2, 0.9, 0.1) + configs.uniform(6, ne, -0.05, 0.05)


@pytest.mark.parametrize("case,vd,q", [("random", 400.0, 1.0), ("random", 400.0, 2.0),
                                       ("random", 12.0, 1.0), ("regions", 20.0, 1.0),
                                       ("regions", 20.0, 2.0)])
def test_chain_rule_against_finite_differences(oracle, case, vd, q):
    if case == "random":
        nel, Rmax = (9, 8, 7), 3.0
        x = configs.uniform(6, int(np.prod(nel)), 0.05, 0.95)
    else:
        nel, Rmax = (24, 6, 6), 2.0
        x = two_region_design(nel)
    ne = int(np.prod(nel))
    P = oracle.filter_csr(*nel, 1.5)
    R, RE = oracle.rfilter_csr(*nel, Rmax)
    o = _E(oracle, P, R, RE, x, vd, q)
    w = np.random.default_rng(2).standard_normal(ne)
    # f(x) = sum_e w_e E_e(x): df/dz = w dE/dz
    _, g = oracle.chain_filter(P, R, o["rho"], o["y"], o["rhoq"], None, w * o["dEdz"])
    mask = o["y"] < 1
    S = R.to_scipy()
    Pm = P.to_scipy()
    # elements j whose filter footprint only reaches R neighbourhoods of one mask value
    reach = (abs(S) @ abs(Pm)).T.tocsr()     # reach[j, i] != 0: x_j influences y_i
    uniform = np.array([mask[reach.indices[reach.indptr[j]:reach.indptr[j + 1]]].all() or
                        (~mask[reach.indices[reach.indptr[j]:reach.indptr[j + 1]]]).all()
                        for j in range(ne)])
    if vd == 400.0:
        assert mask.all() and uniform.all()
    if vd == 12.0:
        assert (~mask).all()
    if case == "regions":
        assert 0.3 < mask.mean() < 0.7
        assert (uniform & mask).sum() >= 100 and (uniform & ~mask).sum() >= 100
    h = 1e-6
    checked = 0
    for j in np.flatnonzero(uniform)[:: max(1, uniform.sum() // 40)]:
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        op, om = _E(oracle, P, R, RE, xp, vd, q), _E(oracle, P, R, RE, xm, vd, q)
        if not (np.array_equal(op["y"] < 1, mask) and np.array_equal(om["y"] < 1, mask)):
            continue
        fd = (w @ op["E"] - w @ om["E"]) / (2 * h)
        assert fd == pytest.approx(g[j], rel=2e-6, abs=1e-9 * np.abs(g).max())
        checked += 1
    assert checked >= 10


def test_chain_mask_form_on_y_equal_one(oracle):
    """Before P^T the chain rule is dfdE .* y - rhoq .* (y < 1) .* R^T(dfdE .* rho): with the
    identity as P (radius 0.5) it passes dfdE through wherever y = 1, and with R^T(...)
    removed (vdMin <= 0 form) it is dfdE .* y."""
    nel = (9, 8, 7)
    ne = int(np.prod(nel))
    x = configs.uniform(6, ne, 0.05, 0.95)
    I = oracle.filter_csr(*nel, 0.5)
    assert np.array_equal(I.to_scipy().toarray(), np.eye(ne))
    R, RE = oracle.rfilter_csr(*nel, 3.0)
    o = oracle.matint(I, R, RE, x, 3.0, 1.0, 55.0)
    m = o["y"] < 1
    assert 0.05 < m.mean() < 0.95
    g = np.random.default_rng(8).standard_normal(ne)
    _, out = oracle.chain_filter(I, R, o["rho"], o["y"], o["rhoq"], None, g)
    assert np.array_equal(out[~m], g[~m])
    _, out0 = oracle.chain_filter(I, None, o["rho"], o["y"], o["rhoq"], None, g)
    assert np.array_equal(out0, g * o["y"])
    # on y < 1 the correction is negative-definite in the sense rhoq >= 0, R >= 0
    gp = np.abs(g)
    _, outp = oracle.chain_filter(I, R, o["rho"], o["y"], o["rhoq"], None, gp)
    assert (outp[m] <= gp[m] * o["y"][m]).all()


def test_dfdV_is_filter_transpose(oracle):
    nel = (7, 6, 5)
    ne = int(np.prod(nel))
    P = oracle.filter_csr(*nel, 1.5)
    R, RE = oracle.rfilter_csr(*nel, 2.0)
    x = configs.uniform(1, ne, 0.1, 0.9)
    o = oracle.matint(P, R, RE, x, 3.0, 1.0, 20.0)
    g = np.random.default_rng(1).standard_normal(ne)
    V, _ = oracle.chain_filter(P, R, o["rho"], o["y"], o["rhoq"], g, None)
    assert np.allclose(V, P.to_scipy().T @ g, rtol=1e-14, atol=1e-15)


def test_cut_and_logistic(oracle):
    nel = (9, 8, 7)
    ne = int(np.prod(nel))
    x = configs.uniform(3, ne, 0.05, 0.95)
    P = oracle.filter_csr(*nel, 1.5)
    R, RE = oracle.rfilter_csr(*nel, 2.5)
    base = oracle.matint(P, R, RE, x, 3.0, 1.0, 45.0)
    z = base["rho"] * base["y"]
    c = 0.3
    cut = oracle.matint(P, R, RE, x, 3.0, 1.0, 45.0, interp=1, param=(c, 0.0))
    above = z > c
    assert above.any() and (~above).any()
    for k in ("E", "dEdz", "rho", "y", "rhoq"):
        assert np.array_equal(cut[k], base[k])
    assert np.array_equal(cut["Es"][above], base["Es"][above]) and (cut["Es"][~above] == 0).all()
    assert np.array_equal(cut["dEsdz"][above], base["dEsdz"][above])
    assert (cut["dEsdz"][~above] == 0).all()
    a, b = 12.0, 0.4
    lg = oracle.matint(P, R, RE, x, 3.0, 1.0, 45.0, interp=2, param=(a, b))
    D = 1 + np.exp(a * (b - z))
    assert np.allclose(lg["Es"] * D, base["Es"], rtol=1e-14, atol=0)
    assert np.array_equal(lg["E"], base["E"])
    # far above b the logistic factor is 1 and the SIMP derivative is recovered
    big = D < 1 + 1e-17
    if big.any():
        assert np.allclose(lg["dEsdz"][big], base["dEsdz"][big], rtol=1e-15)
