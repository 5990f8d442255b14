"""Parity of the HIP path with the CPU oracle, through the C ABI, on seeded inputs.

Bars: bit-exact for the filter pattern/weights, rho = P x, P^T g, the SIMP fields
(p = 3) and element sensitivities for a given U; <= 1e-14 relative for the matrix-free
operator and the Galerkin coarse operators; solves driven to 1e-12 agree to 1e-10
relative in U, compliance and sensitivities (the FP64 tolerance of the workload).
"""
import numpy as np
import pytest

from topopt3d_amd import configs

pytestmark = pytest.mark.gpu

CASES = [((8, 4, 4), 2), ((7, 5, 3), 2), ((9, 6, 5), 3), ((16, 8, 8), 3), ((48, 24, 24), 3),
         ((32, 16, 16), 5)]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup_pair(oracle, tofe, nel, L, seed=7, R=1.5, p=3.0, fixed=True):
    pb = configs.cantilever(nel, seed=seed, lo=0.05, hi=1.0, rmin=R, penal=p, mg_levels=L)
    print(f"case nel={nel} L={L} seed={seed} R={R} p={p}", flush=True)
    ctx = tofe.Context(nel, rmin=R, penal=p, mg_levels=L,
                       fixed_dof=pb.fixed_dof if fixed else None)
    ctx.set_design(pb.x)
    Pf = oracle.filter_csr(*nel, R)
    rho = Pf.matvec(pb.x)
    E, dEdz, Es, dEsdz = oracle.simp(rho, p, 1e-10)
    K = oracle.assemble_K(*nel, oracle.ke(), E, pb.fixed_dof if fixed else None)
    return pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K


@pytest.mark.parametrize("nel,L", CASES)
def test_filter_and_simp_bit_exact(oracle, tofe, nel, L):
    for R, p in [(1.5, 3.0), (2.0, 3.0), (2.4, 3.5)]:
        pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, L, R=R, p=p)
        rp, col, w = ctx.filter_pattern()
        rpo, colo, wo = Pf.arrays()
        assert np.array_equal(rp, rpo) and np.array_equal(col, colo)
        assert np.array_equal(w, wo)
        assert np.array_equal(ctx.field(tofe.FIELD_RHO), rho)
        assert np.array_equal(ctx.filter(pb.x), rho)
        g = np.random.default_rng(3).standard_normal(pb.nelem)
        assert np.array_equal(ctx.filter_T(g), Pf.rmatvec(g))
        if p == 3.0:
            assert np.array_equal(ctx.field(tofe.FIELD_E), E)
            assert np.array_equal(ctx.field(tofe.FIELD_DEDZ), dEdz)
            assert np.array_equal(ctx.field(tofe.FIELD_ES), Es)
            assert np.array_equal(ctx.field(tofe.FIELD_DESDZ), dEsdz)
        else:  # general exponent: device pow() against glibc pow()
            assert rel(ctx.field(tofe.FIELD_E), E) <= 1e-15
            assert rel(ctx.field(tofe.FIELD_DEDZ), dEdz) <= 1e-15
        ctx.close()


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.75, 1.25)])
def test_element_matrix_bit_exact(oracle, tofe, h):
    """The library's KE (host quadrature, det J per Gauss point) is the oracle's bit for bit."""
    ctx = tofe.Context((2, 2, 2), h=h, E0=2.5, nu=0.27, mg_levels=1)
    assert np.array_equal(ctx.ke(), oracle.ke(2.5, 0.27, h))
    ctx.close()


@pytest.mark.parametrize("nel,L", CASES)
def test_operator_matches_assembled_K(oracle, tofe, nel, L):
    for fixed in (True, False):
        pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, L, fixed=fixed)
        rng = np.random.default_rng(11)
        X = rng.standard_normal((2, pb.ndof))
        Y = ctx.apply_K(X.ravel()).reshape(2, -1)
        for k in range(2):
            assert rel(Y[k], K.matvec(X[k])) <= 1e-14
        # symmetry u.Kv = v.Ku
        assert X[0] @ Y[1] == pytest.approx(X[1] @ Y[0], rel=1e-12)
        ctx.close()


@pytest.mark.parametrize("mode", ["wht", "whtbox", "sym", "rows", "mfma"])
@pytest.mark.parametrize("h,E0,nu", [((1.0, 1.0, 1.0), 1.0, 0.3),
                                     ((0.5, 0.5, 0.5), 2.5, 0.27),
                                     ((0.5, 0.75, 1.25), 1.0, 0.3),
                                     ((1.0, 1.0, 2.0), 1.7, 0.35)])
def test_operator_element_shapes(oracle, tofe, h, E0, nu, mode, monkeypatch):
    """The three forms of the fine operator give the oracle's K u with the element matrix of
    (E0, nu, h): the element-centric Walsh-Hadamard form (default; 9 block values for cubic
    elements, 33 for box elements or with TOFE_FINE_WHT_CUBE=0), the node-centric form with
    KE from its 12 cube-symmetry orbit values (cubic elements; TOFE_FINE_WHT=0) and the
    node-centric form reading KE's rows (TOFE_FINE_WHT=0 TOFE_FINE_SYM=0, and box elements), and
    the element-centric form on the matrix cores (TOFE_FINE_MFMA=1: KE x (E_e u_e) as 16-element
    v_mfma_{f32,f64}_16x16x4 GEMMs)."""
    monkeypatch.setenv("TOFE_FINE_MFMA", "1" if mode == "mfma" else "0")
    monkeypatch.setenv("TOFE_FINE_WHT", "1" if mode.startswith("wht") else "0")
    monkeypatch.setenv("TOFE_FINE_WHT_CUBE", "0" if mode == "whtbox" else "1")
    monkeypatch.setenv("TOFE_FINE_SYM", "1" if mode == "sym" else "0")
    nel = (13, 9, 7)
    pb = configs.cantilever(nel, seed=4, lo=0.05, hi=1.0, mg_levels=2)
    print(f"h={h} E0={E0} nu={nu} mode={mode}", flush=True)
    ctx = tofe.Context(nel, h=h, E0=E0, nu=nu, mg_levels=2, fixed_dof=pb.fixed_dof)
    cube = h[0] == h[1] == h[2]
    want = {"wht": 3 if cube else 2, "whtbox": 2, "sym": 1 if cube else 0, "rows": 0, "mfma": 4}[mode]
    assert ctx.stats()["fine_form"] == want
    ctx.set_design(pb.x)
    E = ctx.field(tofe.FIELD_E)
    K = oracle.assemble_K(*nel, oracle.ke(E0, nu, h), E, pb.fixed_dof)
    X = np.random.default_rng(8).standard_normal((2, pb.ndof))
    Y = ctx.apply_K(X.ravel()).reshape(2, -1)
    for k in range(2):
        assert rel(Y[k], K.matvec(X[k])) <= 1e-14
    ctx.set_precision(tofe.PREC_MIXED)
    ctx.assemble()
    U, its, reason = ctx.solve(pb.F, rtol=1e-10, maxit=200)
    assert reason == tofe.CONVERGED_RTOL
    assert np.linalg.norm(pb.F - K.matvec(U)) <= 1e-10 * np.linalg.norm(pb.F) * 1.01
    ctx.close()


@pytest.mark.parametrize("tile", ["0:3", "1:8", "2:5", "3:16", "1:1", "wht1:2", "wht5:16", "wht24:3"])
@pytest.mark.parametrize("nel", [(30, 17, 9), (7, 5, 3)])
def test_operator_tile_shapes(oracle, tofe, nel, tile, monkeypatch):
    """Every compiled column-tile shape and z-chunk length of the node-centric fine operator
    (picked with TOFE_FINE_TILE), and z-chunk lengths of the element-centric one
    (TOFE_FINE_WHT_Z), give the oracle's K u, in FP64 and inside the mixed solve."""
    if tile.startswith("wht"):
        monkeypatch.setenv("TOFE_FINE_WHT_Z", tile[3:])
    else:
        monkeypatch.setenv("TOFE_FINE_WHT", "0")
        monkeypatch.setenv("TOFE_FINE_TILE", tile)
    print(f"tile {tile}", flush=True)
    pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, 2)
    X = np.random.default_rng(5).standard_normal(pb.ndof)
    assert rel(ctx.apply_K(X), K.matvec(X)) <= 1e-14
    ctx.set_precision(tofe.PREC_MIXED)
    ctx.assemble()
    U, its, reason = ctx.solve(pb.F, rtol=1e-10, maxit=200)
    assert reason == tofe.CONVERGED_RTOL
    assert np.linalg.norm(pb.F - K.matvec(U)) <= 1e-10 * np.linalg.norm(pb.F) * 1.01
    ctx.close()


@pytest.mark.parametrize("nel,L", CASES)
def test_galerkin_levels_match_ptap(oracle, tofe, nel, L):
    pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, L)
    ctx.assemble()
    mg = oracle.MG(K, tuple(n + 1 for n in nel), L)
    # level 0 stencil = assembled K, entry by entry (same summation order, no FMA)
    for l in range(0, L):
        S = ctx.level_stencil(l)
        A = (K if l == 0 else mg.A(l)).to_scipy().tocoo()
        d = mg.dims(l)
        n = A.row // 3
        m = A.col // 3
        dn = m - n
        di = dn % d[0]
        # offset of m relative to n as (dx, dy, dz)
        ni, nj, nk = n % d[0], (n // d[0]) % d[1], n // (d[0] * d[1])
        mi, mj, mk = m % d[0], (m // d[0]) % d[1], m // (d[0] * d[1])
        o = (mi - ni + 1) + 3 * (mj - nj + 1) + 9 * (mk - nk + 1)
        vals = S[o, A.row % 3, A.col % 3, n]
        if l == 0:
            assert np.array_equal(vals, A.data)
        else:
            assert np.abs(vals - A.data).max() <= 1e-14 * np.abs(A.data).max()
        assert np.abs(S).sum() == pytest.approx(np.abs(A.data).sum(), rel=1e-13)
        if l < L - 1:
            assert ctx.level_lambda(l) == pytest.approx(mg.lam(l), rel=1e-10)
    ctx.close()


@pytest.mark.parametrize("nel,L", [((16, 8, 8), 3), ((48, 24, 24), 4)])
def test_mixed_lambdas_from_fp32_operators(oracle, tofe, nel, L):
    """The mixed-precision build estimates lambda_max(D^-1 A) with the same Jacobi-CG steps
    on its FP32 operators: within FP32 rounding of the FP64 (oracle-matched) estimates."""
    pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, L)
    mg = oracle.MG(K, tuple(n + 1 for n in nel), L)
    ctx.set_precision(tofe.PREC_MIXED)
    ctx.assemble()
    for l in range(L - 1):
        assert ctx.level_lambda(l) == pytest.approx(mg.lam(l), rel=1e-4)
    ctx.close()


@pytest.mark.parametrize("nel,L", CASES)
def test_solve_compliance_sensitivity(oracle, tofe, nel, L):
    pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, L)
    ctx.assemble()
    U, its, reason = ctx.solve(pb.F, rtol=1e-12, maxit=500)
    mg = oracle.MG(K, tuple(n + 1 for n in nel), L)
    Uo, itso, reasono, _ = mg.solve(pb.F, rtol=1e-12, maxit=500)
    assert reason == 2 and reasono == 2
    assert abs(its - itso) <= 1
    assert rel(U, Uo) <= 1e-10
    r = pb.F - K.matvec(U)
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(pb.F) * 1.01
    c, g = ctx.compliance(U, pb.F)
    co = Uo @ pb.F
    assert c == pytest.approx(co, rel=1e-10)
    go = Pf.rmatvec(oracle.sensitivity(*nel, oracle.ke(), Uo, dEdz))
    assert rel(g, go) <= 1e-10
    # sensitivities from identical U are bit-exact
    assert np.array_equal(ctx.sensitivity(Uo), oracle.sensitivity(*nel, oracle.ke(), Uo, dEdz))
    st = ctx.stats()
    assert st["its"] == its and st["reason"] == 2 and st["levels"] == L
    ctx.close()


def test_reason_codes_and_warm_start(oracle, tofe):
    nel = (16, 8, 8)
    pb, ctx, *_ = setup_pair(oracle, tofe, nel, 3)
    ctx.assemble()
    U, its, reason = ctx.solve(pb.F, rtol=1e-12, maxit=2)
    assert reason == tofe.DIVERGED_ITS and its == 2
    U, its, reason = ctx.solve(pb.F, rtol=1e-8)
    assert reason == tofe.CONVERGED_RTOL
    U2, its2, reason2 = ctx.solve(pb.F, U, rtol=1e-8)
    assert its2 == 0 and reason2 == tofe.CONVERGED_RTOL and np.array_equal(U, U2)
    # rtol = 0 falls back to the absolute tolerance 1e-50 and runs into maxit
    U3, its3, reason3 = ctx.solve(pb.F, rtol=0.0, maxit=5)
    assert reason3 == tofe.DIVERGED_ITS
    # zero right-hand side with zero guess converges at iteration 0
    U4, its4, reason4 = ctx.solve(np.zeros(pb.ndof), rtol=1e-8)
    assert its4 == 0 and reason4 in (2, 3) and not U4.any()
    ctx.close()


def test_state_errors(tofe):
    nel = (8, 4, 4)
    pb = configs.cantilever(nel)
    ctx = tofe.Context(nel, mg_levels=2, fixed_dof=pb.fixed_dof)
    with pytest.raises(tofe.TofeError) as e:
        ctx.solve(pb.F)
    assert e.value.code == tofe.TOFE_ERR_STATE
    ctx.set_design(pb.x)
    with pytest.raises(tofe.TofeError):
        ctx.solve(pb.F)            # not assembled for this design
    ctx.assemble()
    ctx.solve(pb.F)
    ctx.set_precision(tofe.PREC_MIXED)
    with pytest.raises(tofe.TofeError):
        ctx.solve(pb.F)            # precision changed: needs a new assemble
    with pytest.raises(tofe.TofeError):
        tofe.Context((0, 4, 4))
    ctx.close()


def test_multiple_rhs(oracle, tofe):
    nel = (16, 8, 8)
    pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, 3)
    ctx.assemble()
    rng = np.random.default_rng(5)
    Fs = rng.standard_normal((3, pb.ndof))
    Fs[:, pb.fixed_dof.astype(bool)] = 0
    U, its, reason = ctx.solve(Fs.ravel(), rtol=1e-12)
    U = U.reshape(3, -1)
    for k in range(3):
        Uk, _, rk = ctx.solve(Fs[k], rtol=1e-12)
        assert rk == 2
        assert np.array_equal(U[k], Uk)
    ctx.close()


def test_pc_apply_is_symmetric(oracle, tofe):
    nel = (16, 8, 8)
    pb, ctx, *_ = setup_pair(oracle, tofe, nel, 3)
    ctx.assemble()
    rng = np.random.default_rng(2)
    a, b = rng.standard_normal((2, pb.ndof))
    Ma, Mb = ctx.pc_apply(a), ctx.pc_apply(b)
    assert b @ Ma == pytest.approx(a @ Mb, rel=1e-10)
    assert a @ Ma > 0
    ctx.close()


def test_cantilever48_against_sparse_direct_fixture(tofe):
    """48x24x24 cantilever (x = 0.3) on the GPU against the committed SciPy sparse-direct
    fixture (tests/golden/make_cantilever48_direct.py, independent of the oracle and the GPU)."""
    import os

    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", "cantilever48_direct.npz"))
    pb = configs.cantilever_48x24x24()
    ctx = tofe.Context(pb.nel, mg_levels=3, fixed_dof=pb.fixed_dof)
    ctx.set_design(pb.x)
    ctx.assemble()
    U, its, reason = ctx.solve(pb.F, rtol=1e-12)
    assert reason == 2
    assert np.abs(U[fx["dof_samples"]] - fx["U_samples"]).max() <= 1e-9 * np.abs(fx["U_samples"]).max()
    assert np.linalg.norm(U) == pytest.approx(float(fx["U_norm"]), rel=1e-10)
    c, g = ctx.compliance(U, pb.F)
    assert c == pytest.approx(float(fx["compliance"]), rel=1e-10)
    gs = ctx.sensitivity(U)
    assert np.allclose(gs[fx["el_samples"]], fx["g_samples"], rtol=1e-9, atol=0)
    ctx.close()


@pytest.mark.parametrize("nel,L", [((8, 4, 4), 3), ((9, 6, 5), 3), ((16, 8, 8), 3),
                                   ((32, 16, 16), 5), ((48, 24, 24), 5)])
def test_pc_apply_matches_oracle_vcycle(oracle, tofe, nel, L):
    """One V-cycle on the GPU equals the oracle's V-cycle (same Galerkin levels, Chebyshev
    bounds from the same Lanczos estimate, exact coarse solve), up to the benchmark's five
    levels (coarsest grids of 3x2x2 and 4x2x2 nodes)."""
    print(f"case V-cycle {nel} levels {L}", flush=True)
    pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, L)
    ctx.assemble()
    mg = oracle.MG(K, tuple(n + 1 for n in nel), L)
    for lvl in range(L - 1):
        assert ctx.level_lambda(lvl) == pytest.approx(mg.lam(lvl), rel=1e-10)
    r = np.random.default_rng(4).standard_normal(pb.ndof)
    r[pb.fixed_dof.astype(bool)] = 0
    z = ctx.pc_apply(r)
    zo = mg.vcycle(r)
    assert rel(z, zo) <= 1e-11
    ctx.close()


def test_stencil_thread_modes(oracle, tofe, monkeypatch):
    """The coarse-level stencil kernels run one thread per node, or one per DOF on grids below
    TOFE_STENCIL_SPLIT_NODES nodes (the 192x96x96 cantilever's level 1 runs per node, these small
    grids per DOF by default).  Both modes sum each DOF in the same order; only the block
    partial sums of the dot products (Lanczos estimate) are grouped differently.  FP64 and
    mixed V-cycles in either mode against the oracle's V-cycle."""
    nel, L = (16, 8, 8), 3
    out = {}
    for mode, threshold in [("node", "0"), ("dof", "1000000000")]:
        monkeypatch.setenv("TOFE_STENCIL_SPLIT_NODES", threshold)
        pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, L)
        r = np.random.default_rng(4).standard_normal(pb.ndof)
        r[pb.fixed_dof.astype(bool)] = 0
        ctx.set_precision(tofe.PREC_FP64)
        ctx.assemble()
        z64 = ctx.pc_apply(r)
        ctx.set_precision(tofe.PREC_MIXED)
        ctx.assemble()
        z32 = ctx.pc_apply(r)
        out[mode] = (z64, z32)
        ctx.close()
    zo = oracle.MG(K, tuple(n + 1 for n in nel), L).vcycle(r)
    for mode in out:
        assert rel(out[mode][0], zo) <= 1e-11, mode
        assert rel(out[mode][1], zo) <= 1e-4, mode
    assert rel(out["node"][0], out["dof"][0]) <= 1e-12
    assert rel(out["node"][1], out["dof"][1]) <= 1e-5


def test_fp64_level_copies_follow_the_design(oracle, tofe):
    """The FP64 levels read an upper-half copy of their Galerkin stencil that every FP64
    set-up forms anew: after a new design and a mixed-precision set-up in between, one FP64
    V-cycle equals the oracle's for the new design and a fresh context's."""
    nel, L = (16, 8, 8), 3
    pbA, ctx = setup_pair(oracle, tofe, nel, L, seed=7)[:2]
    pbB, fresh, Pf, rho, E, dEdz, Es, dEsdz, KB = setup_pair(oracle, tofe, nel, L, seed=8)
    r = np.random.default_rng(5).standard_normal(pbA.ndof)
    r[pbA.fixed_dof.astype(bool)] = 0
    ctx.set_precision(tofe.PREC_FP64)
    ctx.assemble()
    zA = ctx.pc_apply(r)
    ctx.set_design(pbB.x)
    ctx.set_precision(tofe.PREC_MIXED)
    ctx.assemble()
    ctx.set_precision(tofe.PREC_FP64)
    ctx.assemble()
    z = ctx.pc_apply(r)
    fresh.set_precision(tofe.PREC_FP64)
    fresh.assemble()
    zf = fresh.pc_apply(r)
    zo = oracle.MG(KB, tuple(n + 1 for n in nel), L).vcycle(r)
    assert rel(z, zo) <= 1e-11
    assert rel(z, zf) <= 1e-14
    assert rel(zA, zo) > 1e-6  # the first design's V-cycle is a different operator
    ctx.close()
    fresh.close()


@pytest.mark.parametrize("nel", [(8, 4, 4), (10, 6, 5)])
def test_single_level_is_exact_coarse_solve(oracle, tofe, nel):
    """With one level the preconditioner is the blocked Cholesky inverse of K (several
    64-column diagonal blocks, the last one partial): M^-1 r = K^-1 r."""
    pb, ctx, Pf, rho, E, dEdz, Es, dEsdz, K = setup_pair(oracle, tofe, nel, 1)
    ctx.assemble()
    assert ctx.levels == 1 and pb.ndof % 64 != 0 and pb.ndof > 3 * 64
    r = np.random.default_rng(6).standard_normal(pb.ndof)
    z = ctx.pc_apply(r)
    x = np.linalg.solve(K.to_scipy().toarray(), r)
    assert rel(z, x) <= 1e-10
    ctx.close()
