"""Batches of independent instances of uneven size (the reference solves each problem with
its own KSPSolve, FEAnalysis.cpp:673): tofe_batch_solve_dev captures `chunk` iterations of
every instance into one HIP graph, with the convergence test on the device, and replays it.
In the merged form each step of the iteration is one launch over every instance (table-driven
batched kernels); in the single form the graph holds every instance's own launches.  Every
replay must leave each instance exactly where the eager block solve stands after the same
number of iterations (bit for bit), and the converged batch must equal the eager solves and
meet the oracle's residual; the autotuner must keep the operator's results unchanged."""
import numpy as np
import pytest

from topopt3d_amd import configs

pytestmark = pytest.mark.gpu


def _instances(tofe, precision, n=5, sizes=(8, 10, 12, 14, 16), nrhs=3):
    import torch

    pbs = configs.batch_instances(n, sizes, seed=4, nrhs=nrhs)
    ctxs, F = [], []
    for pb in pbs:
        c = tofe.Context(pb.nel, Emin=pb.Emin, fixed_dof=pb.fixed_dof, precision=precision)
        c.set_design(pb.x)
        c.assemble()
        ctxs.append(c)
        F.append(torch.from_numpy(pb.F).cuda())
    return pbs, ctxs, F


@pytest.mark.parametrize("form", ["merged", "single"])
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_graph_matches_eager_every_replay(oracle, tofe, precision, form, monkeypatch):
    import torch

    print(f"case batch precision={precision} form={form}", flush=True)
    monkeypatch.setenv("TOFE_BATCH_GRAPH", form)
    pbs, ctxs, F = _instances(tofe, precision)
    nrhs, chunk = 3, 2
    batch = tofe.Batch(ctxs)
    for reps in (1, 2, 3, None):
        maxit = 1000 if reps is None else reps * chunk
        Ub = [torch.zeros_like(f) for f in F]
        its, rs = batch.solve_dev([f.data_ptr() for f in F], [u.data_ptr() for u in Ub], nrhs,
                                  rtol=1e-10, maxit=maxit, chunk=chunk)
        for i, c in enumerate(ctxs):
            Ue = torch.zeros_like(F[i])
            c.solve_dev(F[i].data_ptr(), Ue.data_ptr(), nrhs=nrhs, rtol=1e-10, maxit=maxit)
            assert torch.equal(Ub[i], Ue), (reps, i)
        if reps is None:
            assert (rs == tofe.CONVERGED_RTOL).all()
            st = batch.stats()
            assert st["replays"] >= 1 and st["iterations"] >= its.max()
            if form == "merged":
                # the five instances' launches share far fewer launches (in FP64 the CG updates
                # and the dot products of the preconditioned residual have no batched form)
                bound = st["instance_launches"] / (3 if precision else 1)
                assert 0 < st["launches"] < bound, st
            for i, pb in enumerate(pbs):
                E = ctxs[i].field(tofe.FIELD_E)
                K = oracle.assemble_K(*pb.nel, oracle.ke(), E, pb.fixed_dof).to_scipy()
                U = Ub[i].cpu().numpy()
                for k in range(nrhs):
                    res = np.linalg.norm(pb.F[k] - K @ U[k]) / np.linalg.norm(pb.F[k])
                    assert res <= 1.01e-10
        else:
            assert (rs == tofe.DIVERGED_ITS).all() and (its == maxit).all()
    batch.close()
    for c in ctxs:
        c.close()


def test_autotune_keeps_results(tofe):
    pb = configs.cantilever((24, 12, 12), seed=3, lo=0.05, hi=1.0, mg_levels=3)
    for precision in (0, 1):
        c = tofe.Context(pb.nel, mg_levels=3, fixed_dof=pb.fixed_dof, precision=precision)
        c.set_design(pb.x)
        c.assemble()
        X = np.random.default_rng(1).standard_normal((2, pb.ndof))
        X[:, pb.fixed_dof.astype(bool)] = 0
        KX, ZX = c.apply_K(X), c.pc_apply(X)
        U0, _, r0 = c.solve(pb.F, rtol=1e-10)
        zlen, split = c.autotune(nrep=3)
        assert 4 <= zlen <= 24 and split >= 0
        np.testing.assert_array_equal(c.apply_K(X), KX)  # shapes change the order of nothing
        ZX2 = c.pc_apply(X)
        assert np.abs(ZX2 - ZX).max() <= 1e-12 * np.abs(ZX).max()
        U1, _, r1 = c.solve(pb.F, rtol=1e-10)
        assert r1 == r0 and np.abs(U1 - U0).max() <= 1e-8 * np.abs(U0).max()
        c.close()
