"""The host mirrors of the reference's call surface on the GPU: the Python TopOpt/
Compliance/Volume classes and the C++ facade (include/tofe_topopt.hpp, driven by
topopt3d_amd/examples/cantilever) give the same compliance as the C ABI used directly,
and both agree with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from topopt3d_amd import _lib, configs
from topopt3d_amd.topopt import BCType, Compliance, TopOpt, Volume

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def facade_cantilever(nel, precision=_lib.PREC_FP64):
    t = TopOpt()
    t.precision = precision
    nx, ny, nz = nel
    t.CreateMesh([0, nx, 0, ny, 0, nz], nel, 1.5)
    big = 1e10
    lim = [-big, big] * 3
    t.Set_BC([0, ny / 2, nz / 2], [1e-10, big, big], lim, [1, 1, 1], BCType.SUPPORT)
    t.Set_BC([nx, ny // 2, 0], [1e-10, 1e-10, 1e-10], lim, [0, 0, -1], BCType.LOAD)
    t.FEInitialize()
    return t


def test_python_facade_matches_api_and_oracle(oracle, tofe):
    nel = (24, 12, 12)
    t = facade_cantilever(nel)
    pb = configs.cantilever(nel, x=0.3)
    assert np.array_equal(t.fixed_dof, pb.fixed_dof)
    assert np.array_equal(t.F, pb.F)
    x = np.full(t.nElem, 0.3)
    t.MatIntFnc(x)
    t.FEAssemble()
    assert t.FESolve() == 2
    comp, vol = Compliance(), Volume()
    comp.Function(t)
    vol.Function(t)
    assert vol.values[0] == pytest.approx(0.3, rel=1e-14)
    # oracle
    KE = oracle.ke()
    Pf = oracle.filter_csr(*nel, 1.5)
    # dV/dx = P^T (1/nElem): the filter is row- not column-normalised (Volume.cpp:35-41)
    assert np.array_equal(vol.gradients, Pf.rmatvec(np.full(t.nElem, 1.0 / t.nElem)))
    E, dEdz, _, _ = oracle.simp(Pf.matvec(x))
    K = oracle.assemble_K(*nel, KE, E, pb.fixed_dof)
    mg = oracle.MG(K, tuple(n + 1 for n in nel), t.ctx.levels)
    Uo, _, reason, _ = mg.solve(pb.F, rtol=1e-12)
    co = Uo @ pb.F
    assert comp.values[0] == pytest.approx(co, rel=1e-7)   # both at rtol <= 1e-8
    go = Pf.rmatvec(oracle.sensitivity(*nel, KE, Uo, dEdz))
    assert np.abs(comp.gradients - go).max() <= 1e-6 * np.abs(go).max()
    # Chain_Filter is the filter transpose
    g = np.random.default_rng(0).standard_normal(t.nElem)
    dV, dE = t.Chain_Filter(g, g)
    assert np.array_equal(dV, Pf.rmatvec(g)) and np.array_equal(dE, dV)
    t.ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_cpp_facade_design_loop(tofe, precision):
    exe = os.path.join(ROOT, "topopt3d_amd", "examples", "cantilever")
    assert os.path.exists(exe), "build the examples with topopt3d_amd/build.py"
    nel = (24, 12, 12)
    print(f"case cpp facade {nel} precision {precision}", flush=True)
    out = subprocess.run([exe, *map(str, nel), "4", str(precision)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [l.split() for l in out.stdout.splitlines()]
    rows = [l for l in lines if l[0] != "checksum"]
    assert len(rows) == 4
    c0 = float(rows[0][1])
    assert all(int(r[4]) == 2 for r in rows)
    # compliance decreases under the OC updates at fixed volume
    cs = [float(r[1]) for r in rows]
    assert cs[-1] < cs[0]
    assert all(abs(float(r[2]) - 0.3) < 2e-3 for r in rows)
    # iteration 0 equals the Python facade on the same problem
    t = facade_cantilever(nel)
    t.MatIntFnc(np.full(t.nElem, 0.3))
    t.FEAssemble()
    t.FESolve()
    comp = Compliance()
    comp.Function(t)
    assert c0 == pytest.approx(comp.values[0], rel=1e-7)
    t.ctx.close()


def test_python_facade_max_length_scale(oracle, tofe):
    """MatIntFnc/Chain_Filter with Rmax, vdMin and the void-penalty continuation."""
    nel = (12, 10, 8)
    t = TopOpt()
    t.CreateMesh([0, 12, 0, 10, 0, 8], nel, 1.5, 3.0)
    t.vdMin = 55.0
    t.FEInitialize()
    x = configs.uniform(5, t.nElem, 0.05, 1.0)
    P = oracle.filter_csr(*nel, 1.5)
    R, REdge = oracle.rfilter_csr(*nel, 3.0)
    g = np.random.default_rng(4).standard_normal(t.nElem)
    for q in (1.0, 2.0):
        t.vdPenal = q
        t.MatIntFnc(x)
        o = oracle.matint(P, R, REdge, x, 3.0, q, 55.0)
        for k in ("rho", "y", "rhoq", "E", "dEdz", "Es", "dEsdz"):
            assert np.array_equal(getattr(t, k), o[k]), (q, k)
        dV, dE = t.Chain_Filter(g, g)
        oV, oE = oracle.chain_filter(P, R, o["rho"], o["y"], o["rhoq"], g, g)
        assert np.array_equal(dV, oV) and np.array_equal(dE, oE)
    t.ctx.close()


def test_stability_function_matches_oracle(oracle, tofe):
    """Stability.Function (LOBPCG over the GPU K, Ks and V-cycle) gives the oracle's
    aggregated buckling eigenvalues and sensitivities (dense generalized eigh)."""
    from test_oracle_buckling import buckling_gradient

    from topopt3d_amd.topopt import Stability

    nel, nev, p = (6, 2, 2), 2, 8.0
    pb = configs.cantilever(nel, seed=4, lo=0.3, hi=1.0)
    t = facade_cantilever(nel)
    t.MatIntFnc(pb.x)
    t.FEAssemble()
    t.rtol = 1e-12
    assert t.FESolve() == 2
    st = Stability(nvals=nev, p=p, tol=1e-10)
    st.Function(t)
    val, grad, s = buckling_gradient(oracle, nel, pb, pb.x, nev, p)
    assert np.allclose(st.lam, s["mu"], rtol=1e-7)
    assert st.values[0] == pytest.approx(val, rel=1e-7)
    assert np.abs(st.gradients - grad).max() <= 1e-5 * np.abs(grad).max()
    t.ctx.close()
