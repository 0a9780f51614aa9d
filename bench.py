"""Benchmark: assembled DOF/s to FP64 rel-res 1e-8 on one MI355X.

Workload (the config the headline metric is quoted on): P1 diffusion on the
unit cube, 64^3 hexes x 6 Kuhn tetrahedra (1,572,864 tets, 274,625 vertices,
n = 250,047 unknowns), k = exp(GRF) with length scale 0.1 and sigma 1.5
(seeded random-Fourier-feature field, 512 features), source -1, homogeneous
Dirichlet boundary, preconditioned CG stopped on the FP64 true residual
||b - A x||_2 / ||b||_2 <= 1e-8.

One step = zero + assemble (all element kernels, colored scatter into the
device CSR) + solve, inputs already resident in HBM.  Two paths are run:
  plain : FP64 assembly + FP64 Jacobi-PCG
  fast  : FP64 assembly (+FP32 copy) + FP32 Jacobi-PCG inner solves with FP64
          iterative refinement
`value` is the fast path's DOF/s; the plain path is reported beside it.

The cpu_baseline leg times the CPU oracle (a single-threaded C restatement of
the reference element loop + CG/Jacobi) on the same problem.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(N, nfeat, seed):
    from feng_amd.api import FN_CONST, FN_ELEM, Fn, Mesh, TET_P1
    from feng_amd.problems import diffusion_system

    t0 = time.perf_counter()
    m = Mesh.cube(N)
    k = m.coeff_grf(0.1, 1.5, nfeat, seed)
    S, num = diffusion_system(m, TET_P1, 1, Fn(FN_ELEM, data=k), Fn(FN_CONST, -1.0))
    log(f"setup: E={m.E} V={m.V} n={num.nInc} nnz={S.nnz} ({time.perf_counter() - t0:.1f}s)")
    return m, k, S, num


def run_step(S, path, rtol, profile=False):
    from feng_amd.api import PREC_FP32_IR, PREC_FP64

    prec = PREC_FP32_IR if path == "fast" else PREC_FP64
    S.set_profiling(profile)
    t0 = time.perf_counter()
    S.setToZero()
    S.assemble(3, prec)
    asm_ms, _ = S.kernel_stats(1)
    r = S.solve(precision=prec, rtol=rtol, inner_rtol=1e-2 if path == "fast" else 1e-3, check_every=32)
    wall = time.perf_counter() - t0
    spmv_ms, launches = S.kernel_stats(0)
    return dict(wall_s=wall, asm_ms=asm_ms, solve_ms=r.ms_solve, iters=r.iterations, outer=r.outer_iterations,
                relres=r.relres, converged=r.converged, spmv_ms=spmv_ms, spmv_launches=launches)


def cpu_baseline(m, k, num, rtol, budget_s=25.0):
    """Oracle (single-threaded C): assembly of the whole problem, then CG
    iterations until the budget; DOF/s extrapolated with the oracle's own
    iteration count when it converges, else with the measured per-iteration
    time times the GPU plain path's iteration count."""
    import ctypes as C

    from tests.oracle_bind import load_oracle

    O = load_oracle()
    n = num.nInc
    t0 = time.perf_counter()
    ia, ja = O.csr(n, [(num.adr[0], 4, num.adr[0], 4)])
    t_csr = time.perf_counter() - t0
    t0 = time.perf_counter()
    val, rhs = O.assemble_diffusion(4, m.xyz, m.conn, num.adr[0], n, ia, ja, 1, (4, 0.0, 0.0, k),
                                    (0, -1.0, 0.0, None), np.zeros(num.nDofs))
    t_asm = time.perf_counter() - t0
    # time a bounded number of CG iterations
    probe = 50
    t0 = time.perf_counter()
    O.cg_jacobi(ia, ja, val, rhs, 1e-30, probe)
    t_it = (time.perf_counter() - t0) / probe
    max_it = int(max(probe, (budget_s - t_asm) / max(t_it, 1e-9)))
    t0 = time.perf_counter()
    x, it, rr = O.cg_jacobi(ia, ja, val, rhs, rtol, max_it)
    t_cg = time.perf_counter() - t0
    converged = rr <= rtol
    return dict(t_csr=t_csr, t_asm=t_asm, t_cg=t_cg, iters=it, converged=converged, t_iter=t_cg / max(it, 1),
                relres=rr)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--nfeat", type=int, default=512)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    rtol = 1e-8
    m, k, S, num = make_problem(args.N, args.nfeat, args.seed)
    n, nnz = num.nInc, S.nnz

    # plain FP64 path (reported beside the headline)
    for _ in range(max(1, args.warmup)):
        run_step(S, "plain", rtol)
    plain = [run_step(S, "plain", rtol, profile=True) for _ in range(args.steps)]
    log("plain", plain[-1])
    # fast path: warmup, then the timed steps
    for _ in range(args.warmup):
        run_step(S, "fast", rtol)
    t0 = time.perf_counter()
    fast = [run_step(S, "fast", rtol, profile=True) for _ in range(args.steps)]
    total = time.perf_counter() - t0
    log("fast", fast[-1])
    for r in fast + plain:
        if not r["converged"] or r["relres"] > rtol:
            raise SystemExit(f"solve did not reach FP64 rel-res {rtol}: {r}")

    ms_step = 1e3 * total / args.steps
    dev_ms = np.mean([r["asm_ms"] + r["solve_ms"] for r in fast])
    value = n / (ms_step / 1e3)
    plain_ms = np.mean([r["wall_s"] for r in plain]) * 1e3
    plain_dofs = n / (plain_ms / 1e3)

    # roofline of the dominant kernel (the fused pipelined-CG iteration,
    # FP32): algorithmic bytes per launch = values 4 nnz + column ids 4 nnz +
    # row pointers 4 (n+1) + own-row reads of D^-1, r, p, w, s_old, x (6 x 4 n)
    # + writes of x, p, s, r_new, w_new (5 x 4 n)
    spmv_ms = float(np.mean([r["spmv_ms"] for r in fast]))
    alg_bytes = 8 * nnz + 4 * (n + 1) + 44 * n
    achieved = alg_bytes / (spmv_ms * 1e-3) / 1e9
    p_spmv_ms = float(np.mean([r["spmv_ms"] for r in plain]))
    p_alg = 12 * nnz + 4 * (n + 1) + 88 * n

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(m, k, num, rtol, args.cpu_budget)
        iters = c["iters"] if c["converged"] else plain[-1]["iters"]
        t_cpu = c["t_asm"] + c["t_iter"] * iters
        cpu = {"value": n / t_cpu, "unit": "DOF/s", "cores": 1, "kind": "port",
               "sample": (f"oracle assembly of all {m.E} tets ({c['t_asm']:.2f}s) + CG/Jacobi "
                          f"{c['iters']} iterations ({c['t_cg']:.1f}s, {'converged' if c['converged'] else 'bounded, extrapolated to'} "
                          f"{iters} iterations)")}
        log("cpu", c)

    out = {
        "metric": "DOF/s assemble+solve to FP64 rel-res 1e-8; learned/FP32 vs plain FP64 PCG",
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-inner/fp64-refine",
        "data": "synthetic (seeded GRF coefficient on a generated Kuhn-tet cube)",
        "config": {"workload": f"P1 tet diffusion, unit cube {args.N}^3 x6 Kuhn tets, k=exp(GRF l=0.1 s=1.5), "
                               f"PCG to FP64 rel-res 1e-8", "n": n, "nnz": nnz, "elements": m.E,
                   "path": "FP32 Jacobi-PCG + FP64 iterative refinement",
                   "device_ms_per_step": dev_ms,
                   "asm_ms": float(np.mean([r["asm_ms"] for r in fast])),
                   "solve_ms": float(np.mean([r["solve_ms"] for r in fast])),
                   "iterations": fast[-1]["iters"], "outer": fast[-1]["outer"], "relres": fast[-1]["relres"],
                   "plain_fp64_dofs": plain_dofs, "plain_fp64_ms": plain_ms,
                   "plain_iterations": plain[-1]["iters"], "speedup_vs_plain_fp64": plain_ms / ms_step,
                   "plain_spmv_roofline_frac": p_alg / (p_spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if p_spmv_ms else None},
        "roofline": {"bound": "hbm", "kernel": "pipeIter<float>", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "avg_launch_ms": spmv_ms, "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    S.close()


if __name__ == "__main__":
    main()
