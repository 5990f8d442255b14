"""Maximum-length-scale interpolation and the full filter chain rule on the GPU, against the
oracle (TopOpt::MatIntFnc, Material_Interpolation.cpp:20-116; TopOpt::Chain_Filter,
150-187).

Bars: with the special exponents (p = 3, q in {1, 2}) every field (rho, y, rhoq, E, dE/dz,
Es, dEs/dz, REdge) and both chain-rule outputs are bit-identical to the oracle; a general
exponent goes through the device pow() and SIMP_LOGISTIC through the device exp(), so those
are held to 4e-16 relative; compliance gradients with y < 1 after a 1e-12 solve to 1e-9.
"""
import numpy as np
import pytest

from topopt3d_amd import configs

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "y", "rhoq", "E", "dEdz", "Es", "dEsdz")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup(oracle, tofe, nel, rmax, vd_min, q=1.0, p=3.0, interp=0, param=(0.0, 0.0), seed=5,
          L=2, fixed=None):
    print(f"case nel={nel} rmax={rmax} vd_min={vd_min} q={q} p={p} interp={interp}", flush=True)
    x = configs.uniform(seed, int(np.prod(nel)), 0.05, 1.0)
    ctx = tofe.Context(nel, rmin=1.5, penal=p, mg_levels=L, fixed_dof=fixed, rmax=rmax,
                       vd_min=vd_min, vd_penal=q, interpolation=interp, interp_param=param)
    ctx.set_design(x)
    P = oracle.filter_csr(*nel, 1.5)
    R, REdge = oracle.rfilter_csr(*nel, rmax) if vd_min > 0 else (None, None)
    o = oracle.matint(P, R, REdge, x, p, q, vd_min, interp, param)
    return x, ctx, P, R, REdge, o


def gpu_fields(tofe, ctx):
    ids = dict(rho=tofe.FIELD_RHO, y=tofe.FIELD_Y, rhoq=tofe.FIELD_RHOQ, E=tofe.FIELD_E,
               dEdz=tofe.FIELD_DEDZ, Es=tofe.FIELD_ES, dEsdz=tofe.FIELD_DESDZ)
    return {k: ctx.field(v) for k, v in ids.items()}


@pytest.mark.parametrize("nel", [(10, 8, 6), (13, 7, 9), (24, 12, 12)])
@pytest.mark.parametrize("vd_min", [-1.0, 12.0, 55.0, 400.0])
@pytest.mark.parametrize("q", [1.0, 2.0])
def test_matint_and_chain_bit_exact(oracle, tofe, nel, vd_min, q):
    x, ctx, P, R, REdge, o = setup(oracle, tofe, nel, 3.0, vd_min, q=q)
    g = gpu_fields(tofe, ctx)
    for k in FIELDS:
        assert np.array_equal(g[k], o[k]), k
    if vd_min > 0:
        assert np.array_equal(ctx.field(tofe.FIELD_REDGE), REdge)
        frac = (o["y"] < 1).mean()
        print(f"  fraction y < 1: {frac:.3f}", flush=True)
    rng = np.random.default_rng(11)
    dV, dE = rng.standard_normal(x.size), rng.standard_normal(x.size)
    gV, gE = ctx.chain_filter(dV, dE)
    oV, oE = oracle.chain_filter(P, R, o["rho"], o["y"], o["rhoq"], dV, dE)
    assert np.array_equal(gV, oV)
    assert np.array_equal(gE, oE)
    ctx.close()


@pytest.mark.parametrize("interp,param", [(1, (0.3, 0.0)), (2, (12.0, 0.4))])
def test_cut_and_logistic(oracle, tofe, interp, param):
    nel = (12, 10, 8)
    x, ctx, P, R, REdge, o = setup(oracle, tofe, nel, 2.5, 45.0, interp=interp, param=param)
    g = gpu_fields(tofe, ctx)
    for k in ("rho", "y", "rhoq", "E", "dEdz"):
        assert np.array_equal(g[k], o[k]), k
    if interp == 1:
        assert np.array_equal(g["Es"], o["Es"]) and np.array_equal(g["dEsdz"], o["dEsdz"])
        assert (g["Es"] == 0).any() and (g["Es"] > 0).any()
    else:  # device exp() against glibc exp()
        assert rel(g["Es"], o["Es"]) <= 4e-16
        assert rel(g["dEsdz"], o["dEsdz"]) <= 4e-16
    ctx.close()


def test_general_exponents(oracle, tofe):
    nel = (12, 10, 8)
    x, ctx, P, R, REdge, o = setup(oracle, tofe, nel, 3.0, 55.0, q=2.5, p=3.5)
    g = gpu_fields(tofe, ctx)
    assert np.array_equal(g["rho"], o["rho"])
    for k in ("y", "rhoq", "E", "dEdz"):
        assert rel(g[k], o[k]) <= 4e-16, k
    ctx.close()


def test_penalty_continuation(oracle, tofe):
    nel = (10, 8, 6)
    x, ctx, P, R, REdge, o = setup(oracle, tofe, nel, 3.0, 55.0, q=1.0)
    ctx.set_penalty(3.0, 2.0)
    ctx.set_design(x)
    o2 = oracle.matint(P, R, REdge, x, 3.0, 2.0, 55.0)
    g = gpu_fields(tofe, ctx)
    for k in FIELDS:
        assert np.array_equal(g[k], o2[k]), k
    assert not np.array_equal(o2["rhoq"], o["rhoq"])
    ctx.close()


def test_compliance_gradient_with_max_length(oracle, tofe):
    nel = (16, 8, 8)
    pb = configs.cantilever(nel, seed=3, lo=0.05, hi=1.0, mg_levels=3)
    ctx = tofe.Context(nel, rmin=1.5, mg_levels=3, fixed_dof=pb.fixed_dof, rmax=3.0, vd_min=55.0)
    ctx.set_design(pb.x)
    P = oracle.filter_csr(*nel, 1.5)
    R, REdge = oracle.rfilter_csr(*nel, 3.0)
    o = oracle.matint(P, R, REdge, pb.x, 3.0, 1.0, 55.0)
    assert 0.05 < (o["y"] < 1).mean() < 0.95
    ctx.assemble()
    U, its, reason = ctx.solve(pb.F, rtol=1e-12)
    assert reason == 2
    c, grad = ctx.compliance(U, pb.F)
    K = oracle.assemble_K(*nel, oracle.ke(), o["E"], pb.fixed_dof)
    mg = oracle.MG(K, tuple(n + 1 for n in nel), ctx.levels)
    Uo, _, r2, _ = mg.solve(pb.F, rtol=1e-12)
    assert c == pytest.approx(Uo @ pb.F, rel=1e-10)
    dfdE = oracle.sensitivity(*nel, oracle.ke(), Uo, o["dEdz"])
    _, go = oracle.chain_filter(P, R, o["rho"], o["y"], o["rhoq"], None, dfdE)
    assert rel(grad, go) <= 1e-9
    ctx.close()
