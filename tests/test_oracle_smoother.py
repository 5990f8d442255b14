"""Pins of the oracle's level smoother and eigenvalue estimate by closed forms.

The oracle's Chebyshev smoother (orc_smooth) runs a three-term recurrence; its result is
fixed by the mathematics: k steps from any x0 leave the error p_k(D^-1 A) e0 with the
scaled Chebyshev polynomial p_k(t) = T_k((theta - t)/delta) / T_k(theta/delta),
theta/delta the centre/half-width of [lo*lam, hi*lam].  Here p_k is evaluated from the
cos / cosh closed form of T_k on the eigenvalues of D^-1/2 A D^-1/2, never by a recurrence,
so a swapped coefficient, a wrong damping or a dropped term fails.  The damped Jacobi
smoother leaves (1 - t/theta)^k.

The Lanczos estimate (orc_lanczos_lambda) forms the tridiagonal of Jacobi-preconditioned
CG; its top eigenvalue must equal the top Ritz value of D^-1/2 A D^-1/2 on the Krylov
space spanned from D^-1/2 r0, computed here by Arnoldi with full reorthogonalisation
(and equal lambda_max itself once the Krylov space is the whole space).
"""
import numpy as np
import pytest

from topopt3d_amd import configs


def cheb_T(k, s):
    """Chebyshev polynomial T_k(s) from its closed forms (no recurrence)."""
    s = np.asarray(s, np.float64)
    out = np.empty_like(s)
    inside = np.abs(s) <= 1
    out[inside] = np.cos(k * np.arccos(s[inside]))
    big = s > 1
    out[big] = np.cosh(k * np.arccosh(s[big]))
    neg = s < -1
    out[neg] = (-1.0) ** k * np.cosh(k * np.arccosh(-s[neg]))
    return out


def error_poly(t, k, lam, lo, hi, smoother):
    emax, emin = hi * lam, lo * lam
    theta, delta = 0.5 * (emax + emin), 0.5 * (emax - emin)
    if smoother == 1:
        return (1.0 - t / theta) ** k
    return cheb_T(k, (theta - t) / delta) / cheb_T(k, np.array([theta / delta]))[0]


def small_K(oracle, nel=(3, 2, 2), seed=7):
    pb = configs.cantilever(nel, seed=seed, lo=0.1, hi=1.0)
    E = oracle.simp(oracle.filter_csr(*nel, 1.5).matvec(pb.x))[0]
    return oracle.assemble_K(*nel, oracle.ke(), E, pb.fixed_dof)


@pytest.mark.parametrize("smoother", [0, 1])
@pytest.mark.parametrize("steps", [1, 2, 3, 5])
def test_smoother_on_diagonal_matches_closed_form(oracle, smoother, steps):
    """A = diag(lambda_i), D = I: x_k = (1 - p_k(lambda)) / lambda * b from x = 0, and
    x* - p_k(lambda) (x* - x0) from a nonzero start; eigenvalues inside and below the
    Chebyshev interval."""
    import scipy.sparse as sp

    rng = np.random.default_rng(3)
    lam_max = 2.0
    t = np.concatenate([rng.uniform(0.1 * lam_max, 1.1 * lam_max, 30),
                        rng.uniform(1e-3, 0.1 * lam_max, 10)])
    A = oracle.CSR.from_scipy(sp.diags(t).tocsr())
    dinv = np.ones(t.size)
    b = rng.standard_normal(t.size)
    xs = b / t
    pk = error_poly(t, steps, lam_max, 0.1, 1.1, smoother)
    x = oracle.smooth(A, dinv, lam_max, b, steps=steps, smoother=smoother)
    np.testing.assert_allclose(x, xs - pk * xs, rtol=0, atol=1e-13 * np.abs(xs).max())
    x0 = rng.standard_normal(t.size)
    x = oracle.smooth(A, dinv, lam_max, b, x0=x0, steps=steps, smoother=smoother)
    np.testing.assert_allclose(x, xs - pk * (xs - x0), rtol=0, atol=1e-13 * np.abs(xs).max())


@pytest.mark.parametrize("smoother", [0, 1])
@pytest.mark.parametrize("steps", [2, 4])
def test_smoother_on_stiffness_matches_closed_form(oracle, smoother, steps):
    """A = the Dirichlet-masked stiffness of a 3x2x2 cantilever with its Jacobi diagonal:
    D^1/2 e_k = p_k(S) D^1/2 e0 with S = D^-1/2 A D^-1/2, through S's eigendecomposition."""
    K = small_K(oracle)
    Ad = K.to_scipy().toarray()
    d = np.diag(Ad)
    S = Ad / np.sqrt(np.outer(d, d))
    w, Q = np.linalg.eigh(S)
    lam = 0.97 * w.max()  # an estimate slightly below lambda_max, as Lanczos gives
    rng = np.random.default_rng(11)
    b = rng.standard_normal(d.size)
    xs = np.linalg.solve(Ad, b)
    for x0 in (None, rng.standard_normal(d.size)):
        e0 = xs - (0 if x0 is None else x0)
        ek = (Q @ (error_poly(w, steps, lam, 0.1, 1.1, smoother) * (Q.T @ (np.sqrt(d) * e0)))) / np.sqrt(d)
        x = oracle.smooth(K, 1.0 / d, lam, b, x0=x0, steps=steps, smoother=smoother)
        np.testing.assert_allclose(x, xs - ek, rtol=0, atol=1e-12 * np.abs(xs).max())


def test_mg_vcycle_uses_the_pinned_smoother(oracle):
    """A two-level V(1,0) cycle equals one zero-guess smoother step followed by the exact
    coarse correction with P and the coarse operator's dense inverse."""
    nel = (4, 2, 2)
    pb = configs.cantilever(nel, seed=2, lo=0.2)
    E = oracle.simp(oracle.filter_csr(*nel, 1.5).matvec(pb.x))[0]
    K = oracle.assemble_K(*nel, oracle.ke(), E, pb.fixed_dof)
    mg = oracle.MG(K, (5, 3, 3), 2, nu1=1, nu2=0)
    A0 = K.to_scipy().toarray()
    P = mg.P(0).to_scipy().toarray()
    A1 = mg.A(1).to_scipy().toarray()
    r = np.random.default_rng(5).standard_normal(A0.shape[0])
    d = np.diag(A0)
    lam = mg.lam(0)
    x1 = error_poly(np.zeros(1), 1, lam, 0.1, 1.1, 0)  # p_1(0) = 1
    assert x1[0] == pytest.approx(1.0)
    theta = 0.5 * (0.1 + 1.1) * lam
    xs = r / d / theta                      # first Chebyshev step from zero
    xs = xs + P @ np.linalg.solve(A1, P.T @ (r - A0 @ xs))
    z = mg.vcycle(r)
    np.testing.assert_allclose(z, xs, rtol=0, atol=1e-12 * np.abs(xs).max())


def test_tridiagonal_bisection(oracle):
    rng = np.random.default_rng(1)
    for k in (1, 2, 5, 17):
        a = rng.uniform(-1, 3, k)
        b = rng.uniform(-1, 1, max(k - 1, 0))
        T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        assert oracle.tridiag_max_eig(a, np.append(b, 0.0)) == pytest.approx(
            np.linalg.eigvalsh(T).max(), rel=1e-13, abs=1e-13)


def krylov_top_ritz(S, w, k):
    """Top eigenvalue of V^T S V, V an orthonormal basis of span{w, S w, ..., S^(k-1) w}
    (Arnoldi with two passes of classical Gram-Schmidt per vector)."""
    V = np.zeros((w.size, k))
    V[:, 0] = w / np.linalg.norm(w)
    for j in range(1, k):
        v = S @ V[:, j - 1]
        for _ in range(2):
            v -= V[:, :j] @ (V[:, :j].T @ v)
        V[:, j] = v / np.linalg.norm(v)
    return np.linalg.eigvalsh(V.T @ S @ V).max()


@pytest.mark.parametrize("its", [2, 4, 7])
def test_lanczos_equals_krylov_ritz_value(oracle, its):
    K = small_K(oracle)
    Ad = K.to_scipy().toarray()
    d = np.diag(Ad)
    S = Ad / np.sqrt(np.outer(d, d))
    r0 = oracle.power_start(d.size)
    ritz = krylov_top_ritz(S, r0 / np.sqrt(d), its)
    assert oracle.lanczos_lambda(K, 1.0 / d, its) == pytest.approx(ritz, rel=1e-12)


def test_lanczos_full_space_gives_lambda_max(oracle):
    """With as many steps as the space has dimensions the estimate is lambda_max(D^-1 A)."""
    import scipy.sparse as sp

    rng = np.random.default_rng(4)
    n = 24
    B = rng.standard_normal((n, n))
    Ad = B @ B.T + n * np.eye(n)
    A = oracle.CSR.from_scipy(sp.csr_matrix(Ad))
    d = np.diag(Ad)
    lmax = np.linalg.eigvalsh(Ad / np.sqrt(np.outer(d, d))).max()
    for its in (n, n + 6):
        assert oracle.lanczos_lambda(A, 1.0 / d, its) == pytest.approx(lmax, rel=1e-12)
