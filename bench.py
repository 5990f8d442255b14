"""Benchmark of the FE-analysis hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): the 192x96x96
cantilever, 1,769,472 elements / 5,447,811 DOF, x = 0 face clamped, unit downward
point load at the bottom-centre node of the x = max face, densities i.i.d. U[0.01,1]
(SplitMix64 seed 2) passed through the R = 1.5 density filter, SIMP p = 3,
Emin = 1e-9, 5-level V(2,2) Chebyshev-Jacobi Galerkin multigrid, CG to 1e-8 FP64
relative residual.

One step = one FESolve: Galerkin set-up for the current design (tofe_assemble) plus the
solve from a zero initial guess (tofe_solve_dev), with F and U resident in HBM.  The
headline value is the mixed-precision path (FP32 V-cycle inside FP64 flexible CG);
the FP64 path, the operator rate and the per-kernel roofline are reported beside it,
with the oracle (a C restatement of the reference's PETSc path) timed on host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s/solve to 1e-8 res at 5.45M DOF, mixed vs FP64 on GPU; operator MDOF/s"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    copy-kernel calibrated), committed in profiles/traffic.txt as 'name: bytes'."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(name + ":"):
            try:
                return float(line.split(":", 1)[1].split()[0])
            except ValueError:
                return None
    return None


def cpu_baseline(pb, its_ref, n_iter=3):
    """Oracle (C + OpenMP restatement of the reference PETSc path) on the same workload:
    full set-up (filter, SIMP, CSR assembly, Galerkin PtAP hierarchy, dense Cholesky)
    plus n_iter timed PCG iterations, extrapolated to its_ref iterations."""
    from oracle import oracle as O

    O.build()
    t0 = time.perf_counter()
    rho = O.filter_csr(*pb.nel, pb.rmin).matvec(pb.x)
    E = O.simp(rho, pb.penal, pb.Emin)[0]
    K = O.assemble_K(*pb.nel, O.ke(), E, pb.fixed_dof)
    mg = O.MG(K, tuple(n + 1 for n in pb.nel), pb.mg_levels)
    t1 = time.perf_counter()
    _, its, reason, _ = mg.solve(pb.F, rtol=1e-30, maxit=n_iter)
    t2 = time.perf_counter()
    setup = t1 - t0
    per_it = (t2 - t1) / max(its, 1)
    return {
        "value": setup + per_it * its_ref,
        "unit": "s/solve",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": (f"same 5.45M-DOF workload on host cores: full CPU set-up measured "
                   f"({setup:.1f} s: filter, SIMP, CSR assembly, Galerkin PtAP, dense "
                   f"Cholesky) + {its} PCG iterations measured ({per_it:.2f} s/it), "
                   f"extrapolated to the {its_ref} iterations of the FP64 GPU solve"),
        "setup_s": setup,
        "s_per_iteration": per_it,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--small", action="store_true", help="48x24x24 instead of 192x96x96")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only launch the dominant kernel 5 times and a 512 MiB device copy, "
                         "for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); every box in this "
              "setting has a single MI355X and the path is not sharded here", file=sys.stderr)
        sys.exit(2)

    import torch

    from topopt3d_amd import _lib, configs

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    pb = configs.cantilever((48, 24, 24), seed=2, lo=0.01, hi=1.0, Emin=1e-9, mg_levels=3) \
        if args.small else configs.config2()
    log(f"bench: {pb.name} ndof={pb.ndof} nelem={pb.nelem}")
    ctx = _lib.Context(pb.nel, Emin=pb.Emin, penal=pb.penal, rmin=pb.rmin,
                       mg_levels=pb.mg_levels, fixed_dof=pb.fixed_dof)
    ctx.set_design(pb.x)
    if args.pmc_probe:
        # calibration copy: 512 MiB read + 512 MiB written by one copy kernel
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        ctx.set_precision(_lib.PREC_MIXED)
        ctx.assemble()
        ms = ctx.time_operator(0, _lib.PREC_MIXED, 1, 5)
        log(f"bench: pmc probe: copy 536870912 B read + written; fine operator fp32 cheb "
            f"x5 {ms * 1e3:.1f} us each, algorithmic bytes "
            f"{4.0 * (2 * pb.ndof + pb.nelem) + 16.0 * pb.ndof:.0f}")
        ctx.close()
        return
    F = torch.from_numpy(pb.F).to(dev)
    U = torch.zeros_like(F)
    r = torch.zeros_like(F)
    torch.cuda.synchronize()

    def step(prec):
        U.zero_()
        torch.cuda.synchronize()
        ctx.assemble()
        its, reason = ctx.solve_dev(F.data_ptr(), U.data_ptr(), rtol=1e-8, maxit=500)
        st = ctx.stats()
        return its, reason, st["setup_seconds"], st["solve_seconds"]

    results = {}
    for prec, name in [(_lib.PREC_FP64, "fp64"), (_lib.PREC_MIXED, "mixed")]:
        ctx.set_precision(prec)
        for _ in range(args.warmup):
            step(prec)
        torch.cuda.synchronize()
        if name == "mixed":  # HIP events around the dominant kernel's launches in these steps
            ctx.time_solve_kernel(True)
        t0 = time.perf_counter()
        rec = [step(prec) for _ in range(args.steps)]
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        if name == "mixed":
            kt_ms, kt_n = ctx.time_solve_kernel(False)
        # independent FP64 residual check of the last solve
        ctx.apply_K_dev(U.data_ptr(), r.data_ptr())
        res = (torch.linalg.vector_norm(F - r) / torch.linalg.vector_norm(F)).item()
        its = rec[-1][0]
        results[name] = {
            "s_per_solve": wall,
            "setup_s": float(np.mean([x[2] for x in rec])),
            "solve_s": float(np.mean([x[3] for x in rec])),
            "its": its,
            "reason": rec[-1][1],
            "relres_fp64": res,
        }
        log(f"bench: {name} {wall:.4f} s/solve its={its} relres={res:.2e} "
            f"setup={results[name]['setup_s']:.4f} solve={results[name]['solve_s']:.4f}")
        assert rec[-1][1] == 2 and res <= 1.01e-8, "solve did not reach 1e-8"
        if name == "fp64":
            U64 = U.clone()
    err_u = (torch.linalg.vector_norm(U - U64) / torch.linalg.vector_norm(U64)).item()
    c64 = ctx.compliance_dev(U64.data_ptr(), F.data_ptr())
    cmx = ctx.compliance_dev(U.data_ptr(), F.data_ptr())

    # operator rate and the dominant kernel's roofline (HIP events on the library stream)
    ndof, ne = pb.ndof, pb.nelem
    # isolated back-to-back launches (operands largely cache-resident) for the operator rates;
    # the dominant kernel's roofline uses its launches inside the timed solves (t32c)
    t64 = ctx.time_operator(0, _lib.PREC_FP64, 0, 30)
    t32 = ctx.time_operator(0, _lib.PREC_MIXED, 0, 30)
    t1_32 = ctx.time_operator(1, _lib.PREC_MIXED, 1, 30)
    t32c = kt_ms
    log(f"bench: fine op fp64 {t64 * 1e3:.1f} us, fp32 {t32 * 1e3:.1f} us (isolated); "
        f"fp32 + Chebyshev in the timed solves {t32c * 1e3:.1f} us ({kt_n} launches); "
        f"level-1 stencil fp32 cheb {t1_32 * 1e3:.1f} us (isolated)")
    # Dominant solve kernel: the FP32 fine operator fused with a Chebyshev step.  Its
    # algorithmic HBM bytes per launch are u, b, D^-1, d read and d, x written (4 B per DOF
    # each) plus E (4 B per element).  The element-centric Walsh-Hadamard form needs about
    # 260 FLOP per element (two 8-point transforms per component and 45 block terms), under
    # 4 FLOP per HBM byte, far below the FP32 ridge point: the launch is bounded by HBM.  The
    # node-centric forms apply KE densely (1152 FLOP per element) and are compute-bound.
    form = ctx.stats()["fine_form"]
    abytes = 4.0 * (2 * ndof + ne) + 4.0 * 4 * ndof
    if form >= 2:
        achieved = abytes / (t32c * 1e-3) / 1e9
        roof = {
            "kernel": "k_fine_wht<float, EPI_CHEB> (element-centric fine operator + Chebyshev step)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": read_traffic("k_fine_wht_f32_cheb"),
        }
    else:
        achieved = 1152.0 * ne / (t32c * 1e-3) / 1e12
        roof = {
            "kernel": "k_fine_op<float, EPI_CHEB> (node-centric fine operator + Chebyshev step)",
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": read_traffic("k_fine_op_f32_cheb"),
        }
    roof.update({
        "algorithmic_flops_per_launch_dense_ke": 1152.0 * ne,
        "algorithmic_bytes_per_launch": abytes,
        "ms_per_launch": t32c,
        "launches_timed": kt_n,
        "timing": "HIP events around each launch of the kernel inside the timed mixed solves",
    })

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle) ...")
        cpu = cpu_baseline(pb, results["fp64"]["its"])
        log(f"bench: CPU {cpu['value']:.1f} s/solve on {cpu['cores']} threads")

    mixed = results["mixed"]["s_per_solve"]
    out = {
        "metric": METRIC,
        "value": mixed,
        "unit": "s/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle inside f64 CG",
        "data": "synthetic",
        "config": {
            "workload": (f"{pb.name}: {pb.nelem} Q1 elements, {pb.ndof} DOF, x~U[0.01,1] seed 2 "
                         f"filtered R={pb.rmin}, p={pb.penal}, Emin={pb.Emin}, "
                         f"{pb.mg_levels}-level V(2,2) Chebyshev Galerkin GMG-CG to 1e-8"),
            "global_batch": 1,
            "parallelism": "single GPU",
        },
        "fp64_s_per_solve": results["fp64"]["s_per_solve"],
        "mixed_speedup_vs_fp64": results["fp64"]["s_per_solve"] / mixed,
        "fp64": results["fp64"],
        "mixed": results["mixed"],
        "mixed_vs_fp64": {"u_rel_err": err_u, "compliance_fp64": c64, "compliance_mixed": cmx,
                          "compliance_rel_err": abs(cmx - c64) / abs(c64)},
        "operator_mdof_s": {"fp64": ndof / (t64 * 1e-3) / 1e6, "fp32": ndof / (t32 * 1e-3) / 1e6},
        "operator_us": {"fp64_isolated": t64 * 1e3, "fp32_isolated": t32 * 1e3,
                        "fp32_cheb_in_solve": t32c * 1e3,
                        "level1_stencil_fp32_cheb_isolated": t1_32 * 1e3},
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (cpu["value"] / mixed) if cpu else None,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
