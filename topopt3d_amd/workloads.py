"""Runners of the eigen workloads on one GPU: each builds its synthetic problem, runs it in
FP64 and with the FP32 V-cycle (mixed), and returns a dict of timings, results and checks.

- buckling (the 128x64x64 column, 1,635,075 DOF): compliance, the 6 largest mu of
  Ks phi = mu K phi (Stability::Function, Buckling.cpp:15-270), the adjoint gradients of the
  compliance and of each eigenvalue, central differences on 32 seeded elements;
- vibration (the 160x80x80 cantilever, 3,168,963 DOF): the 12 lowest modes with the
  consistent mass (Frequency::Function, Dynamic.cpp:15-166, posed as M phi = lambda K phi),
  and the cost of one block pass of 12 vectors against 12 single-vector passes.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib, configs, eigen


def _log(*a):
    import sys

    print(*a, file=sys.stderr, flush=True)


def _context(pb, precision):
    return _lib.Context(pb.nel, Emin=pb.Emin, penal=pb.penal, rmin=pb.rmin,
                        mg_levels=pb.mg_levels, fixed_dof=pb.fixed_dof, precision=precision)


def _sync():
    import torch

    torch.cuda.synchronize()


def buckling(nel=None, nev=6, fd_elements=32, fd_steps=(1e-6, 1e-3), eig_tol=1e-9):
    """The 128x64x64 column (or a smaller column with nel): both precisions, then the FD check in
    FP64."""
    import torch

    pb = configs.column_128x64x64() if nel is None else configs.column(nel, seed=3, rmin=2.0, mg_levels=0)
    dev = torch.device("cuda:0")
    F = torch.from_numpy(pb.F).to(dev)
    out = {"workload": f"{pb.name}: {pb.nelem} elements, {pb.ndof} DOF, R={pb.rmin}, "
                       f"{nev} buckling modes", "precisions": {}}
    keep = {}
    for prec, name in [(_lib.PREC_FP64, "fp64"), (_lib.PREC_MIXED, "mixed")]:
        ctx = _context(pb, prec)
        _sync()
        t0 = time.perf_counter()
        ctx.set_design_dev(torch.from_numpy(pb.x).to(dev).data_ptr())
        ctx.assemble()
        U = torch.zeros_like(F)
        its, reason = ctx.solve_dev(F.data_ptr(), U.data_ptr(), rtol=1e-12, maxit=1000)
        g = torch.empty(pb.nelem, dtype=torch.float64, device=dev)
        c = ctx.compliance_dev(U.data_ptr(), F.data_ptr(), g.data_ptr())
        _sync()
        t1 = time.perf_counter()
        res = eigen.buckling_modes(ctx, U, pb.fixed_dof, nev=nev, tol=eig_tol, maxit=1000)
        _sync()
        t2 = time.perf_counter()
        dmu = eigen.buckling_sensitivities(ctx, U, res.values, res.vectors, rtol=1e-11)
        _sync()
        t3 = time.perf_counter()
        _log(f"buckling {name}: solve {its} its {t1 - t0:.3f} s, eigen {res.iterations} its "
             f"{t2 - t1:.3f} s (converged {res.converged}), gradients {t3 - t2:.3f} s, "
             f"mu {np.array2string(res.values, precision=6)}")
        out["precisions"][name] = {
            "solve_s": t1 - t0, "solve_its": its, "solve_reason": reason, "compliance": c,
            "eigen_s": t2 - t1, "eigen_its": res.iterations, "eigen_converged": bool(res.converged),
            "eigen_residual_max": float(np.max(res.residuals)), "mu": res.values.tolist(),
            "gradients_s": t3 - t2,
        }
        keep[name] = (ctx, U, res, g.cpu().numpy(), dmu)
    c64, cmx = out["precisions"]["fp64"]["compliance"], out["precisions"]["mixed"]["compliance"]
    _, U64, r64, g64, d64 = keep["fp64"]
    _, Umx, rmx, gmx, dmx = keep["mixed"]
    out["mixed_vs_fp64"] = {
        "compliance_rel_err": abs(cmx - c64) / abs(c64),
        "u_rel_err": float(torch.linalg.vector_norm(Umx - U64) / torch.linalg.vector_norm(U64)),
        "dc_rel_err_max": float(np.abs(gmx - g64).max() / np.abs(g64).max()),
        "mu_rel_err_max": float(np.max(np.abs(rmx.values - r64.values) / np.abs(r64.values))),
        "dmu_rel_err_max": float(np.abs(dmx - d64).max() / np.abs(d64).max()),
    }
    # central differences at fd_step on seeded elements, FP64, forward from warm starts with
    # fixed iteration counts (a smooth function of x)
    ctx64 = keep["fp64"][0]
    elems = np.random.default_rng(3).choice(pb.nelem, fd_elements, replace=False)
    J = np.vstack([g64[None, :], d64])
    mask = torch.from_numpy(1.0 - pb.fixed_dof.astype(np.float64)).to(dev)
    ops = eigen._Ops(ctx64)
    xd = torch.empty(pb.nelem, dtype=torch.float64, device=dev)

    def value(x):
        xd.copy_(torch.from_numpy(x))
        ctx64.set_design_dev(xd.data_ptr())
        ctx64.assemble()
        U = U64.clone()
        ctx64.solve_dev(F.data_ptr(), U.data_ptr(), rtol=0.0, maxit=40)
        cval = float(F @ U)
        ctx64.set_stress_dev(U.data_ptr())
        r = eigen.lobpcg(ops.Ks, ops.K, ops.T, r64.vectors.clone(), largest=True, tol=0.0,
                         maxit=4, mask=mask)
        return cval, r.values

    # relative error per element and function, against max(|J_e|, 1e-3 max_e |J|) so that
    # elements whose derivative is nearly zero do not divide by it
    scale = np.maximum(np.abs(J[:, elems]), 1e-3 * np.abs(J).max(axis=1)[:, None])
    out["fd"] = []
    for h in fd_steps:
        t0 = time.perf_counter()
        fd = np.empty((J.shape[0], len(elems)))
        for j, e in enumerate(elems):
            xp, xm = pb.x.copy(), pb.x.copy()
            xp[e] += h
            xm[e] -= h
            cp, mup = value(xp)
            cm, mum = value(xm)
            fd[:, j] = np.concatenate([[(cp - cm) / (2 * h)], (mup - mum) / (2 * h)])
            if (j + 1) % 16 == 0:
                _log(f"buckling FD step {h}: {j + 1} of {len(elems)} elements")
        err = np.abs(fd - J[:, elems]) / scale
        rec = {
            "step": h, "elements": len(elems), "seconds": time.perf_counter() - t0,
            "compliance_rel_err_max": float(err[0].max()),
            "compliance_rel_err_median": float(np.median(err[0])),
            "mu_rel_err_max": float(err[1:].max()),
            "mu_rel_err_median": float(np.median(err[1:])),
        }
        out["fd"].append(rec)
        _log(f"buckling FD step {h}: compliance max {rec['compliance_rel_err_max']:.2e} median "
             f"{rec['compliance_rel_err_median']:.2e}; mu max {rec['mu_rel_err_max']:.2e} "
             f"median {rec['mu_rel_err_median']:.2e}")
    for name in keep:
        keep[name][0].close()
    return out


def vibration(nel=None, nev=12, eig_tol=1e-8, mass_form=_lib.MASS_CONSISTENT, reps=10):
    """The 160x80x80 cantilever (or a smaller cantilever with nel): 12 lowest modes in both precisions, and the
    block-pass timings of K, M and the V-cycle on 12 vectors."""
    import torch

    pb = configs.vibration_160x80x80() if nel is None else configs.cantilever(nel, seed=5, lo=0.05,
                                                                    Emin=1e-9, mg_levels=0)
    dev = torch.device("cuda:0")
    out = {"workload": f"{pb.name}: {pb.nelem} elements, {pb.ndof} DOF, {nev} modes, "
                       f"{'consistent' if mass_form == _lib.MASS_CONSISTENT else 'lumped'} mass",
           "precisions": {}}
    keep = {}
    for prec, name in [(_lib.PREC_FP64, "fp64"), (_lib.PREC_MIXED, "mixed")]:
        ctx = _context(pb, prec)
        ctx.set_design_dev(torch.from_numpy(pb.x).to(dev).data_ptr())
        _sync()
        t0 = time.perf_counter()
        ctx.assemble()
        ctx.set_mass_form(mass_form)
        ctx.assemble_mass(1.0)
        _sync()
        t1 = time.perf_counter()
        res = eigen.vibration_modes(ctx, pb.fixed_dof, nev=nev, tol=eig_tol, maxit=1000)
        _sync()
        t2 = time.perf_counter()
        omega2 = 1.0 / res.values
        _log(f"vibration {name}: set-up {t1 - t0:.3f} s, eigen {res.iterations} its "
             f"{t2 - t1:.3f} s (converged {res.converged}), omega^2 "
             f"{np.array2string(omega2[:4], precision=6)} ...")
        out["precisions"][name] = {
            "setup_s": t1 - t0, "eigen_s": t2 - t1, "eigen_its": res.iterations,
            "eigen_converged": bool(res.converged),
            "eigen_residual_max": float(np.max(res.residuals)), "omega2": omega2.tolist(),
        }
        keep[name] = (ctx, res)
    o64 = np.array(out["precisions"]["fp64"]["omega2"])
    omx = np.array(out["precisions"]["mixed"]["omega2"])
    out["mixed_vs_fp64"] = {"omega2_rel_err_max": float(np.max(np.abs(omx - o64) / o64))}
    # one block pass of nev vectors against nev single-vector passes (FP64 context)
    ctx = keep["fp64"][0]
    X = keep["fp64"][1].vectors.contiguous()
    Y = torch.empty_like(X)
    timings = {}
    for opname, fn in [("K", ctx.apply_K_dev), ("M", ctx.apply_M_dev), ("V-cycle", ctx.pc_apply_dev)]:
        fn(X.data_ptr(), Y.data_ptr(), nev)
        _sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn(X.data_ptr(), Y.data_ptr(), nev)
        _sync()
        t_block = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            for j in range(nev):
                fn(X[j].data_ptr(), Y[j].data_ptr(), 1)
        _sync()
        t_single = (time.perf_counter() - t0) / reps
        timings[opname] = {"block_ms": t_block * 1e3, "single_ms": t_single * 1e3 / nev,
                           "block_over_single": t_block / (t_single / nev)}
        _log(f"vibration block pass {opname}: {nev} vectors {t_block * 1e3:.3f} ms, one vector "
             f"{t_single * 1e3 / nev:.3f} ms")
    out["block_pass"] = timings
    for name in keep:
        keep[name][0].close()
    return out


def batch(n=32, nrhs=4, rtol=1e-8, chunk=2, autotune=True,
          forms=("merged", "single", "instance", "streams")):
    """32 cantilevers of uneven size (nx in {16..96}), 4 load cases each, FP32
    V-cycle inside FP64 flexible CG to 1e-8.  Times the eager solves one instance after the
    other against graph-captured batch solves of all instances in each batch form, and checks
    that every form gives the eager solves' bits."""
    import torch

    pbs = configs.cantilevers_batch32(nrhs=nrhs)[:n]
    dev = torch.device("cuda:0")
    out = {"workload": f"{len(pbs)} cantilevers nx in {{16..96}}, ny = nz = nx/2, "
                       f"{sum(p.ndof for p in pbs)} DOF in total, {nrhs} load cases each, "
                       "mixed precision to 1e-8"}
    ctxs, F = [], []
    _sync()
    t0 = time.perf_counter()
    for pb in pbs:
        c = _context(pb, _lib.PREC_MIXED)
        c.set_design(pb.x)
        c.assemble()
        ctxs.append(c)
        F.append(torch.from_numpy(pb.F).to(dev))
    _sync()
    out["setup_s"] = time.perf_counter() - t0
    if autotune:
        t0 = time.perf_counter()
        shapes = [c.autotune(nrep=5) for c in ctxs]
        for c in ctxs:
            c.assemble()  # the V-cycle's eigenvalue estimates with the chosen shapes
        _sync()
        out["autotune_s"] = time.perf_counter() - t0
        out["autotuned_shapes"] = [{"nx": pb.nel[0], "fine_zlen": z, "split_nodes": t}
                                   for pb, (z, t) in zip(pbs, shapes)]
    _log(f"batch: {len(pbs)} instances set up in {out['setup_s']:.2f} s")
    # eager: one instance after the other
    Ue = [torch.zeros_like(f) for f in F]
    for c, f, u in zip(ctxs, F, Ue):  # warm-up
        c.solve_dev(f.data_ptr(), u.data_ptr(), nrhs=nrhs, rtol=rtol, maxit=500)
    its_e = []
    for u in Ue:
        u.zero_()
    _sync()
    t0 = time.perf_counter()
    for c, f, u in zip(ctxs, F, Ue):
        its_e.append(c.solve_dev(f.data_ptr(), u.data_ptr(), nrhs=nrhs, rtol=rtol, maxit=500))
    _sync()
    t_eager = time.perf_counter() - t0
    # graph batch (TOFE_BATCH_GRAPH): one launch per step over every instance (merged), one
    # graph with every instance's own launches as a branch (single), one graph per instance on
    # the instances' streams (instance), no graph (streams)
    import os

    timed = {}
    for form in forms:
        os.environ["TOFE_BATCH_GRAPH"] = form
        b = _lib.Batch(ctxs)
        Ub = [torch.zeros_like(f) for f in F]
        b.solve_dev([f.data_ptr() for f in F], [u.data_ptr() for u in Ub], nrhs, rtol=rtol,
                    maxit=500, chunk=chunk)  # warm-up
        for u in Ub:
            u.zero_()
        _sync()
        t0 = time.perf_counter()
        its, rs = b.solve_dev([f.data_ptr() for f in F], [u.data_ptr() for u in Ub], nrhs,
                              rtol=rtol, maxit=500, chunk=chunk)
        _sync()
        t_batch = time.perf_counter() - t0
        st = b.stats()
        same = all(bool(torch.equal(a, e)) for a, e in zip(Ub, Ue))
        timed[form] = {"batch_s": t_batch, "stats": st, "identical_to_eager": same,
                       "all_converged": bool((rs == _lib.CONVERGED_RTOL).all()),
                       "its_max": int(its.max()), "its_min": int(its.min())}
        _log(f"batch ({form} graph): {t_batch * 1e3:.1f} ms (capture "
             f"{st['capture_seconds'] * 1e3:.1f} ms, replays {st['replays']}), identical {same}, "
             f"its {its.min()}..{its.max()}")
        b.close()
    os.environ.pop("TOFE_BATCH_GRAPH", None)
    best = min(timed, key=lambda k: timed[k]["batch_s"])
    out.update({
        "eager_s": t_eager, "batch_s": timed[best]["batch_s"], "batch_form": best,
        "forms": timed, "batch_over_eager": timed[best]["batch_s"] / t_eager,
        "identical_to_eager": all(f["identical_to_eager"] for f in timed.values()),
        "eager_its_max": int(max(k for k, _ in its_e)),
    })
    _log(f"batch: eager {t_eager * 1e3:.1f} ms")
    for c in ctxs:
        c.close()
    return out
