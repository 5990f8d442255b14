"""The coarsest operator formed straight from E (TOFE_DIRECT_COARSE=1: one GEMM over the coarsest
cells of the element scales and a table of the element matrix seen from a cell's nodes, plus
Dirichlet and spring terms; for meshes whose element counts are multiples of 2^(L-1)) against the
Galerkin chain's (the default, PCMGSetGalerkin, FEAnalysis.cpp:412): the V-cycles and solves agree, with
the clamped face and with springs, in FP64 and with the FP32 V-cycle.  The identity itself is
pinned against the oracle's chain in tests/test_oracle_fe.py."""
import numpy as np
import pytest

from topopt3d_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nel,L,springs", [((16, 8, 8), 3, False), ((12, 6, 6), 2, True),
                                           ((32, 16, 16), 5, True)])
def test_direct_coarsest_matches_chain(tofe, nel, L, springs, precision, monkeypatch):
    print(f"case direct coarse nel={nel} L={L} springs={springs} precision={precision}", flush=True)
    pb = configs.cantilever(nel, seed=5, lo=0.05, hi=1.0, mg_levels=L)
    spk = None
    if springs:
        spk = np.zeros(pb.ndof)
        spk[2::15] = 0.04
    ctxs = []
    for mode in ("1", "0"):
        monkeypatch.setenv("TOFE_DIRECT_COARSE", mode)
        c = tofe.Context(nel, mg_levels=L, fixed_dof=pb.fixed_dof, precision=precision)
        c.set_design(pb.x)
        if spk is not None:
            c.set_springs(spk)
        c.assemble()
        ctxs.append(c)
    assert ctxs[0].stats()["coarse_form"] == 1 and ctxs[1].stats()["coarse_form"] == 0
    X = np.random.default_rng(2).standard_normal((3, pb.ndof))
    X[:, pb.fixed_dof.astype(bool)] = 0
    Zd, Zc = ctxs[0].pc_apply(X), ctxs[1].pc_apply(X)
    assert np.abs(Zd - Zc).max() <= 1e-9 * np.abs(Zc).max()
    Ud, itd, rd = ctxs[0].solve(pb.F, rtol=1e-10)
    Uc, itc, rc = ctxs[1].solve(pb.F, rtol=1e-10)
    assert rd == rc == tofe.CONVERGED_RTOL and abs(itd - itc) <= 1
    assert np.abs(Ud - Uc).max() <= 1e-8 * np.abs(Uc).max()
    for c in ctxs:
        c.close()
