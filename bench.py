"""Benchmark: assembled DOF/s to FP64 rel-res 1e-8 on one MI355X.

Workload (the config the headline metric is quoted on): P1 diffusion on the
unit cube, 64^3 hexes x 6 Kuhn tetrahedra (1,572,864 tets, 274,625 vertices,
n = 250,047 unknowns), k = exp(GRF) with length scale 0.1 and sigma 1.5
(seeded random-Fourier-feature field, 512 features), source -1, homogeneous
Dirichlet boundary, preconditioned CG stopped on the FP64 true residual
||b - A x||_2 / ||b||_2 <= 1e-8.

One step = zero + assemble (the one-launch row-gather assembly of every
element term into the device CSR and RHS, with the FP32 copy of the values
on the FP32 paths) + solve, inputs already resident in HBM.  Paths:
  plain   : FP64 Jacobi-PCG (pipelined, one kernel/iteration)
  fast    : FP32 Jacobi-PCG inner solves, FP64 iterative refinement
  learned : as fast, with the learned IC(0)-pattern split preconditioner
            (weights from feng_amd/weights/, skipped when absent)
  learned64 : FP64 split PCG with the learned preconditioner
  ic0     : FP64 split IC(0)-pattern PCG (the learned path without the net)
  amg     : FP64 PCG preconditioned by the aggregation-multigrid V-cycle over
            the nested ordering (FP32 level values)
  amg32   : the same in FP32, FP64 refinement by residual replacement
`value` is the DOF/s of the fastest non-plain path (named in config.path);
every path is reported in config.paths, with its speed-up over plain.

--workload batched runs the 256-problem batched workload instead;
--workload transient runs BDF2 steps of the heat equation on the same cube;
--workload ns runs Newton solves of the Kovasznay flow (Taylor-Hood, 512^2);
--copy-probe runs a plain device copy (counter calibration for rocprofv3);
--mfma-probe runs the matrix-core kernels (MFMA counter passes).

The cpu_baseline leg times the CPU oracle (a C restatement of the reference
element loop + Jacobi-CG) on the same problem, on all cores of the process's
affinity (value) and single-threaded (value_1core).
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
# HBM bytes per launch of the dominant kernels from rocprofv3 FETCH_SIZE and
# WRITE_SIZE (scaled by the same counters on a plain copy of known size).
TRAFFIC = {  # HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE (cube 64^3), scaled by the same
    # counters on a 512 MiB copy (x2.00, x1.00): profiles/cube64_hbm_traffic.md (assembly, nested multigrid PCG,
    # plain), profiles/cube64_learned64_hbm_traffic.md (learned FP64 path)
    "spIter<double> (split PCG, A product + updates)": 79.6e6,
    "tileIter<double>": 75.9e6,
    "pcgDown0<double, float> (level-0 CG update + smoothing + restriction)": 76.2e6,
    "pcgDown0<float, float> (level-0 CG update + smoothing + restriction)": 53.0e6,
    "asmGather<double, 4, 3, 16>": 370.5e6,
    # the batched workload's level-0 prolongation product over the whole level
    # (draw_problems(256, 1) as one system): profiles/batched256_hbm_traffic.md
    "amgUp<double, float, 4, 256, true>": 2012.9e6,
}


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


PATHS = {
    # name: (precision, preconditioner, inner rtol)
    "plain": ("fp64", "jacobi", 1e-3),
    "fast": ("fp32", "jacobi", 1e-2),
    "learned": ("fp32", "learned", 1e-2),
    "learned64": ("fp64", "learned", 1e-3),
    "ic0": ("fp64", "ic0", 1e-3),
    "amg": ("fp64", "amg", 1e-3),
    "amg32": ("fp32", "amg", 1e-3),
}

# Timed kernel of each path (the A product of one PCG iteration) and its
# algorithmic bytes per launch: s nnz values + 4 nnz column ids + 4 (n+1) row
# pointers + the vector passes of one iteration (minimal unfused PCG-Jacobi:
# 18 passes of s n), s = 8 (FP64) or 4 (FP32).
ITER_KERNEL = {
    "plain": ("tileIter<double>", 8),
    "fast": ("pipeIter<float>", 4),
    "learned": ("cgSpmv<float> (split PCG, A product)", 4),
    "learned64": ("spIter<double> (split PCG, A product + updates)", 8),
    "ic0": ("spIter<double> (split PCG, A product + updates)", 8),
    "amg": ("pcgDown0<double, float> (level-0 CG update + smoothing + restriction)", 8),
    "amg32": ("pcgDown0<float, float> (level-0 CG update + smoothing + restriction)", 4),
}


def iter_bytes(s, n, nnz):
    return (s + 4) * nnz + 4 * (n + 1) + 18 * s * n


def mg_down0_bytes(s, v, n, nnz, n1):
    """Algorithmic bytes of one pcgDown0 launch (level 0 of the nested
    multigrid PCG): SELL values and columns once ((v + 4) nnz, the two row
    sums share the pass), row descriptors 8 n, the vectors x (read and
    written), p, q, r_old, r_new, u (written), u_old and w (gathered, each
    once): 9 s n, l1 v n, and the level-1 right-hand side plus its aggregate
    bounds (s + 4) n1.  s, v: vector and value bytes."""
    return (v + 4) * nnz + 8 * n + 9 * s * n + v * n + (s + 4) * n1


def mg_iter_bytes(s, v, vprod, rows, nnzs, block_rows=1024):
    """Algorithmic bytes of one nested multigrid PCG iteration: product
    ((vprod + 4) nnz + 8 n + 6 s n + v n: z gathered, p_old and q_old read,
    p, q, w written, l1), down0 (mg_down0_bytes), up0 ((v + 4) nnz + 8 n +
    3 s n + v n), the level-0 prolongation written by level 1 (2 s n), and per
    level l >= 1 below the single-workgroup kernel a restriction and a
    prolongation ((v + 4) nnz_l + 4 n_l + 2 s n_l + v n_l + (s + 4) n_l+1 and
    (v + 4) nnz_l + 8 n_l + 3 s n_l + v n_l + s n_l+1); the levels inside the
    single-workgroup kernel count their values and columns once ((v + 2)
    nnz_l)."""
    n, nnz = rows[0], nnzs[0]
    b = (vprod + 4) * nnz + 8 * n + 6 * s * n + v * n
    b += mg_down0_bytes(s, v, n, nnz, rows[1] if len(rows) > 1 else 0)
    b += (v + 4) * nnz + 8 * n + 3 * s * n + v * n + 2 * s * n
    for l in range(1, len(rows)):
        nl, zl = rows[l], nnzs[l]
        nn = rows[l + 1] if l + 1 < len(rows) else 0
        if nl <= block_rows:
            b += (v + 2) * zl
            continue
        b += (v + 4) * zl + 4 * nl + 2 * s * nl + v * nl + (s + 4) * nn
        b += (v + 4) * zl + 8 * nl + 3 * s * nl + v * nl + s * nn
    return b


def asm_bytes(E, V, n, nnz, nloc=4, dim=3):
    """Algorithmic bytes of one assembly (the metric's definition): geometric
    connectivity 4 nloc E, coordinates 8 dim V, per-element coefficient 8 E,
    CSR values written once 8 nnz, RHS 8 n."""
    return 4 * nloc * E + 8 * dim * V + 8 * E + 8 * nnz + 8 * n


def run_step(S, path, rtol, profile=False):
    from feng_amd.api import PC_AMG, PC_IC0, PC_JACOBI, PC_LEARNED, PREC_FP32_IR, PREC_FP64

    precn, pcn, inner = PATHS[path]
    prec = PREC_FP32_IR if precn == "fp32" else PREC_FP64
    pc = {"jacobi": PC_JACOBI, "learned": PC_LEARNED, "ic0": PC_IC0, "amg": PC_AMG}[pcn]
    S.set_profiling(profile)
    t0 = time.perf_counter()
    S.setToZero()
    S.assemble(3, prec)
    asm_ms, _ = S.kernel_stats(1)
    r = S.solve(pc=pc, precision=prec, rtol=rtol, inner_rtol=inner, check_every=32)
    wall = time.perf_counter() - t0
    spmv_ms, launches = S.kernel_stats(0)
    iter_ms, _ = S.kernel_stats(2)
    return dict(wall_s=wall, asm_ms=asm_ms, solve_ms=r.ms_solve, iters=r.iterations, outer=r.outer_iterations,
                relres=r.relres, converged=r.converged, spmv_ms=spmv_ms, spmv_launches=launches, iter_ms=iter_ms)


def cpu_baseline(m, k, num, rtol, budget_s=25.0):
    """Oracle (C) on the host cores: assembly of the whole problem, then
    Jacobi-PCG until convergence or the budget; DOF/s extrapolated with the
    oracle's own iteration count when it converges, else with the measured
    per-iteration time times the GPU plain path's iteration count.  Timed
    twice: single-threaded, and with the colored OpenMP assembly and threaded
    CG on all cores of this process's affinity."""
    from tests.oracle_bind import load_oracle

    O = load_oracle()
    n = num.nInc
    threads = max(1, len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    ia, ja = O.csr(n, [(num.adr[0], 4, num.adr[0], 4)])
    t_csr = time.perf_counter() - t0
    t0 = time.perf_counter()
    colorOf, ncolors = O.color(m.conn, 4, m.V)
    t_color = time.perf_counter() - t0
    args = (4, m.xyz, m.conn, num.adr[0], n, ia, ja, 1, (4, 0.0, 0.0, k), (0, -1.0, 0.0, None), np.zeros(num.nDofs))
    out = {"t_csr": t_csr, "t_color": t_color, "threads": threads}
    for tag, T in (("1", 1), ("all", threads)):
        t0 = time.perf_counter()
        if T == 1:
            val, rhs = O.assemble_diffusion(*args)
        else:
            val, rhs = O.assemble_diffusion_omp(*args, colorOf, ncolors, T)
        t_asm = time.perf_counter() - t0
        probe = 20
        t0 = time.perf_counter()
        O.cg_jacobi_omp(ia, ja, val, rhs, 1e-30, probe, T)
        t_it = (time.perf_counter() - t0) / probe
        max_it = int(max(probe, (budget_s / 2 - t_asm) / max(t_it, 1e-9)))
        t0 = time.perf_counter()
        if T == 1:
            x, it, rr = O.cg_jacobi(ia, ja, val, rhs, rtol, max_it)
        else:
            x, it, rr = O.cg_jacobi_omp(ia, ja, val, rhs, rtol, max_it, T)
        t_cg = time.perf_counter() - t0
        out[tag] = dict(t_asm=t_asm, t_cg=t_cg, iters=it, converged=rr <= rtol, t_iter=t_cg / max(it, 1), relres=rr)
    return out


def copy_probe(mib=512, reps=5):
    """A plain device copy of known size (calibrates the FETCH_SIZE and
    WRITE_SIZE counters: each launch reads and writes `mib` MiB)."""
    import torch

    x = torch.empty(mib * 2 ** 20 // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "reps": reps}), flush=True)


def mfma_probe():
    """The matrix-core kernels for counter passes (rocprofv3 --pmc): the
    colored MFMA element sweep of the headline cube in FP64 and FP32
    (asm_kernel = 1, fp32_values = 1), the learned net's inference inside one
    learned-PCG solve, and three training steps of the net on a 32^3 cube;
    then the MFMA sweeps of P2 (512^2), P3 and P4 (256^2) triangle diffusion
    and the Taylor-Hood Newton Jacobian (256^2 Kovasznay)."""
    from feng_amd.api import (FN_CONST, FN_ELEM, PC_LEARNED, PREC_FP32_IR, PREC_FP64, TRI_P2, TRI_P3, TRI_P4, Fn,
                              Mesh, essential_values)
    from feng_amd.learned import load_weights, train
    from feng_amd.problems import diffusion_system, kovasznay, navier_stokes_system

    m, k, S, num = make_problem(64, 512, 1)
    S.set_tuning("asm_kernel", 1)
    S.set_tuning("fp32_values", 1)
    for prec in (PREC_FP64, PREC_FP64, PREC_FP32_IR, PREC_FP32_IR):
        S.setToZero()
        S.assemble(3, prec)
    S.set_tuning("asm_kernel", 2)
    load_weights(S)
    S.setToZero()
    S.assemble(3, PREC_FP64)
    r = S.solve(pc=PC_LEARNED, rtol=1e-8)
    S.close()
    train(N=32, instances=3, log_every=1)
    for kind, N in ((TRI_P2, 512), (TRI_P3, 256), (TRI_P4, 256)):
        mt = Mesh.square(N)
        deg = 2 if kind == TRI_P2 else 2 * (kind - 3)  # exact for the stiffness of a straight element
        St, _ = diffusion_system(mt, kind, deg, Fn(FN_ELEM, data=mt.coeff_blocks(8, 0.1, 10.0, 3)), Fn(FN_CONST, -1.0))
        St.set_tuning("asm_kernel", 1)
        for _ in range(3):
            St.setToZero()
            St.assemble(3, PREC_FP64)
        St.close()
    mn = Mesh.square(256, 0.2, 0)
    Sn, numn = navier_stokes_system(mn, 40.0, 5)
    _, ufn, pfn = kovasznay(40.0)
    Sn.set_solution(essential_values(numn, [ufn, pfn]))
    for _ in range(3):  # the Jacobian alone: the matrix-core kernel (ns_mfma 1) assembles matrices only
        Sn.setToZero()
        Sn.assemble(1, PREC_FP64)
    Sn.close()
    print(json.dumps({"mfma_probe": True, "learned_iterations": r.iterations}), flush=True)


def bench_batched(args):
    """256 independent P1/P2 diffusion problems (N_p in [32, 256]) solved to
    FP64 rel-res 1e-8 per problem, two ways: FP32 assembly + batched FP32
    Jacobi-PCG with FP64 refinement (HIP graph, autotuned launch shape), and
    batched AMG-PCG (per-problem scalars, the block-diagonal union V-cycle).
    --batch-layout union (default): every problem of both degrees in ONE
    system; degree: one system per degree.  value = the faster method (sum
    over the systems)."""
    from feng_amd.api import PC_AMG, PREC_FP32_IR, PREC_FP64
    from feng_amd.batched import build_batch, build_union_batch, draw_problems, solve_batch

    probs = draw_problems(256, args.seed)
    out = {"workload": "batched", "problems": len(probs), "layout": args.batch_layout}
    # the multigrid method first: the Jacobi autotuner leaves its block in the context
    methods = {"amg": dict(precision=PREC_FP64, pc=PC_AMG), "jacobi": dict(precision=PREC_FP32_IR)}
    if args.methods:
        methods = {k: v for k, v in methods.items() if k in args.methods.split(",")}
    tot_n = 0
    kern = {}
    tot_ms = {k: 0.0 for k in methods}
    detail = {k: {} for k in methods}
    systems = [(None, "P1+P2")] if args.batch_layout == "union" else [(1, "P1"), (2, "P2")]
    for kind, name in systems:
        b = build_union_batch(probs) if kind is None else build_batch(probs, kind)
        if args.no_batch_graph:
            b.S.set_tuning("batch_graph", 0)  # the multigrid chunks as direct launches
        for meth, kw in methods.items():
            t0 = time.perf_counter()
            warm = solve_batch(b, rtol=1e-8, autotune=True, **kw)  # warmup: autotuners / hierarchy
            setup_s = time.perf_counter() - t0
            runs = [solve_batch(b, rtol=1e-8, autotune=False, **kw) for _ in range(max(1, args.steps))]
            r = runs[-1]
            r["tuned"] = warm["tuned"]
            if not r["converged"].all():
                raise SystemExit(f"batched {name} {meth}: {int((~r['converged']).sum())} problems did not converge")
            ms = float(np.mean([x["asm_ms"] + x["solve_ms"] for x in runs]))
            detail[meth][name] = {"problems": len(b.nInc), "n": r["n"], "nnz": r["nnz"], "asm_ms": r["asm_ms"],
                                  "solve_ms": r["solve_ms"], "max_iterations": int(r["iterations"].max()),
                                  "max_relres": float(r["relres"].max()), "tuned": r["tuned"],
                                  "warmup_s": round(setup_s, 2)}
            tot_ms[meth] += ms
            log(name, meth, detail[meth][name])
        if "amg" in methods:
            # one more solve with events around the level-0 prolongation of the
            # multigrid V-cycle (the dominant kernel of the batched AMG solves),
            # every launch over the whole level (no tiles of stopped problems
            # skipped), so that the bytes of each launch are those below
            b.S.set_tuning("batch_skip", 0)
            rp = solve_batch(b, rtol=1e-8, autotune=False, profile=True, **methods["amg"])
            b.S.set_tuning("batch_skip", 1)
            rows_l, nnz_l = b.S.amg_levels()
            n0, z0 = rows_l[0], nnz_l[0]
            n1 = rows_l[1] if len(rows_l) > 1 else 0
            lanes = b.S.get_tuning("amg_lanes")
            pre = 0 < b.S.get_tuning("amg_pre_rows") <= n0 and lanes in (2, 4)
            if pre:
                # the product over y = u + oc e[agg] stored by the row pass:
                # (4 + 4) nnz + 4 (n + 1) + y, b, x (8 each) and l1 (4) per row
                ub = 8 * z0 + 4 * (n0 + 1) + 28 * n0
            else:
                # (4 + 4) nnz (FP32 values, columns) + 4 (n + 1) + 4 s n + 4 n + s n_1, s = 8
                ub = 8 * z0 + 4 * (n0 + 1) + 32 * n0 + 4 * n0 + 8 * n1
            kern[name] = {"avg_launch_ms": rp["kernel_ms"], "launches": rp["kernel_launches"], "alg_bytes": ub,
                          "kernel": f"amgUp<double, float, {lanes}, 256, true>" if pre
                          else f"amgUp<double, float, {lanes}, 256>"}
        tot_n += r["n"]
        b.S.close()
    best = min(tot_ms, key=tot_ms.get)
    roof = None
    big = systems[-1][1]
    if big in kern and kern[big]["avg_launch_ms"]:
        kp = kern[big]
        ach = kp["alg_bytes"] / (kp["avg_launch_ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": f"{kp['kernel']} (level 0 of the {big} system's V-cycle)",
                "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": TRAFFIC.get(kp["kernel"]) if (args.seed, len(probs), args.batch_layout) == (1, 256, "union")
                else None,
                "avg_launch_ms": kp["avg_launch_ms"], "alg_bytes_per_launch": kp["alg_bytes"], "per_batch": kern}
    cpu = None if args.no_cpu else cpu_batched(probs)
    out.update({"metric": "DOF/s assemble+solve to FP64 rel-res 1e-8 (256 batched problems)",
                "value": tot_n / (tot_ms[best] / 1e3), "unit": "DOF/s", "ms_per_step": tot_ms[best], "n_total": tot_n,
                "method": best, "dofs_per_method": {k: tot_n / (v / 1e3) for k, v in tot_ms.items()},
                "batches": detail, "dtype": "f64" if best == "amg" else "fp32-inner/fp64-refine", "n_gpus": 1,
                "roofline": roof, "cpu_baseline": cpu})
    print(json.dumps(out), flush=True)


def cpu_batched(probs, sample=24):
    """The oracle on the host, on the first `sample` problems of the draw: the
    serial C assembly of each problem and its Jacobi-CG to FP64 rel-res 1e-8
    (the batched path's algorithm), DOF/s = the sample's unknowns over its
    time."""
    from feng_amd.api import Mesh, Numbering
    from tests.oracle_bind import load_oracle

    O = load_oracle()
    n_tot, t_tot = 0, 0.0
    for N_, kind, seed in probs[:sample]:
        m = Mesh.square(N_)
        num = Numbering(m, [kind], [1])
        k = m.coeff_blocks(8, 0.1, 10.0, seed)
        nl = num.adr[0].size // m.E
        ia, ja = O.csr(num.nInc, [(num.adr[0], nl, num.adr[0], nl)])
        t0 = time.perf_counter()
        val, rhs = O.assemble_diffusion(kind, m.xyz, m.conn, num.adr[0], num.nInc, ia, ja, 2, (4, 0.0, 0.0, k),
                                        (0, -1.0, 0.0, None), np.zeros(num.nDofs))
        O.cg_jacobi(ia, ja, val, rhs, 1e-8, 100000)
        t_tot += time.perf_counter() - t0
        n_tot += num.nInc
    return {"value": n_tot / t_tot, "unit": "DOF/s", "cores": 1, "kind": "port",
            "sample": (f"oracle, one thread, on the first {sample} problems of the draw ({n_tot} unknowns): assembly "
                       f"and Jacobi-CG to rel-res 1e-8 of each, {t_tot:.1f}s")}


def transient_alg_bytes(E, V, nDofs, n, nnz, nsol, matrix=True):
    """Algorithmic HBM bytes of one sweep of TRANSIENT_MASS + DIFFUSION +
    SOURCE over P1 tets (the assembly's definition plus the solutions the
    transient term reads): connectivity 16 E, coefficient 8 E, coordinates
    24 V, the current and past solutions 8 nDofs nsol, the RHS 8 n and, with
    the matrix, the CSR values 8 nnz written once."""
    return 16 * E + 8 * E + 24 * V + 8 * nDofs * nsol + 8 * n + (8 * nnz if matrix else 0)


def cpu_transient_step(m, k, num, dt, threads):
    """One BDF1 step of the heat equation by the oracle on the host cores:
    colored OpenMP diffusion + source assembly, the transient-mass terms
    (serial), threaded Jacobi-PCG to rel-res 1e-8."""
    from tests.oracle_bind import load_oracle

    O = load_oracle()
    n = num.nInc
    ia, ja = O.csr(n, [(num.adr[0], 4, num.adr[0], 4)])
    colorOf, ncolors = O.color(m.conn, 4, m.V)
    u = np.zeros(num.nDofs)
    t0 = time.perf_counter()
    val, rhs = O.assemble_diffusion_omp(4, m.xyz, m.conn, num.adr[0], n, ia, ja, 1, (4, 0.0, 0.0, k),
                                        (0, -1.0, 0.0, None), u, colorOf, ncolors, threads)
    O.assemble_transient_mass(4, m.xyz, m.conn, num.adr[0], n, ia, ja, 1, (0, 1.0, 0.0, None), 1.0 / dt,
                              np.zeros(num.nDofs), val, rhs)
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    x, it, rr = O.cg_jacobi_omp(ia, ja, val, rhs, 1e-8, 20000, threads)
    t_cg = time.perf_counter() - t0
    return dict(t_asm=t_asm, t_cg=t_cg, iters=it, relres=rr)


# over-correction factors (x100) the transient AMG path times after its warm-up
TRANSIENT_OC = (120, 140, 160, 180)


def bench_transient(args):
    """Heat equation m u_t - div(k grad u) + S = 0 on the headline cube
    (P1 tets, k = the same exp(GRF) field, m = 1, S = -1, u = 0 on the
    boundary and at t = 0), BDF2 (BDF1 starter) with dt = 1e-3.  One step =
    history rotation + BDF coefficients + Newton: fused residual + matrix
    sweep, PCG to FP64 rel-res 1e-8, correction, residual sweep and check.
    Paths: FP64 Jacobi-PCG and FP64 AMG-PCG; value = DOF-steps/s of the
    faster.  After the warm-up steps, the AMG path times two steps with each
    over-correction factor of TRANSIENT_OC and keeps the fastest (the mass
    term makes the operator better conditioned than the steady one, whose
    best factor, 1.8, is the library default)."""
    from feng_amd.api import FN_CONST, FN_ELEM, PC_AMG, PC_JACOBI, Fn, Mesh, TET_P1
    from feng_amd.problems import NLSolverOptions, heat_system
    from feng_amd.timeint import TimeIntegrator

    dt = 1e-3
    nsteps = max(args.steps, 20)
    m = Mesh.cube(args.N)
    k = m.coeff_grf(0.1, 1.5, args.nfeat, args.seed)
    res = {}
    for name, pc in (("jacobi", PC_JACOBI), ("amg", PC_AMG)):
        S, num, forms = heat_system(m, TET_P1, 1, Fn(FN_CONST, 1.0), Fn(FN_ELEM, data=k), Fn(FN_CONST, -1.0))
        pre = args.warmup + 1 + (2 * len(TRANSIENT_OC) if pc == PC_AMG else 0)  # steps before the timed ones
        ti = TimeIntegrator(S, 2, 0.0, dt * (pre + nsteps), pre + nsteps,
                            np.zeros(num.nDofs), nl=NLSolverOptions(1e-10, 1e-10, 1e4, 10),
                            solve_kw=dict(pc=pc, rtol=1e-8, check_every=32))
        ti.make_steps(args.warmup + 1)  # the BDF2 starting step and warm-up steps
        tuned = None
        if pc == PC_AMG:
            tuned = {}
            for oc in TRANSIENT_OC:
                S.set_tuning("amg_oc_x100", oc)
                t0 = time.perf_counter()
                ti.make_steps(2)  # make_steps(n) runs n more steps
                tuned[oc] = round(1e3 * (time.perf_counter() - t0) / 2, 3)
            S.set_tuning("amg_oc_x100", min(tuned, key=tuned.get))
            tuned = {"amg_oc_x100": min(tuned, key=tuned.get), "ms_per_step": tuned}
        S.set_profiling(False)
        t0 = time.perf_counter()
        ti.make_steps(nsteps)
        wall = (time.perf_counter() - t0) / nsteps
        reps = ti.reports[-nsteps:]
        res[name] = {"ms_per_step": 1e3 * wall, "dof_steps": num.nInc / wall,
                     "asm_ms": float(np.mean([r.asm_ms for r in reps])),
                     "solve_ms": float(np.mean([r.solve_ms for r in reps])),
                     "newton_iterations": float(np.mean([r.newton_iterations for r in reps])),
                     "linear_iterations": float(np.mean([r.linear_iterations for r in reps])),
                     "sweeps_per_step": 2, "tuned": tuned}
        n, nnz = num.nInc, S.nnz
        log("transient", name, res[name])
        S.close()
    best = max(res, key=lambda p: res[p]["dof_steps"])
    B = res[best]
    # the step's two sweeps: residual + matrix (Newton's first iteration), residual (its check)
    alg = transient_alg_bytes(m.E, m.V, num.nDofs, n, nnz, 3) + transient_alg_bytes(m.E, m.V, num.nDofs, n, nnz, 3,
                                                                                    matrix=False)
    achieved = alg / (B["asm_ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        threads = max(1, len(os.sched_getaffinity(0)))
        c = cpu_transient_step(m, k, num, dt, threads)
        cpu = {"value": n / (c["t_asm"] + c["t_cg"]), "unit": "DOF-steps/s", "cores": threads, "kind": "port",
               "sample": (f"one BDF1 step of the same problem by the oracle: OpenMP assembly + serial transient "
                          f"mass {c['t_asm']:.2f}s, threaded Jacobi-PCG {c['iters']} iterations {c['t_cg']:.2f}s")}
        log("cpu", c)
    out = {"metric": "DOF-steps/s of BDF2 heat-equation steps (assemble + solve to FP64 rel-res 1e-8 per step)",
           "value": B["dof_steps"], "unit": "DOF-steps/s", "n_gpus": 1, "steps": nsteps, "warmup": args.warmup,
           "ms_per_step": B["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (seeded GRF coefficient on a generated Kuhn-tet cube)",
           "config": {"workload": f"transient P1 tet heat equation, unit cube {args.N}^3 x6 Kuhn tets, "
                                  f"k=exp(GRF l=0.1 s=1.5), BDF2 dt={dt}", "n": n, "nnz": nnz, "elements": m.E,
                      "path": best, "paths": res},
           "roofline": {"bound": "hbm", "kernel": "asmGather<double, 4, 3, 16> (a step's two sweeps)",
                        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": None, "sweeps_ms_per_step": B["asm_ms"], "alg_bytes_per_step": alg},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def bench_ns(args):
    """Steady Kovasznay flow (Re = 40) on the jittered N^2 Taylor-Hood P2-P1
    mesh (the Navier-Stokes configuration: N = 512, n = 2,354,179).  One step
    = a full Newton solve from the boundary-lifted zero interior state:
    per Newton iteration the residual and matrix sweeps of the four forms and
    GMRES(50) with the block Schur preconditioner to rel-res 1e-8, until
    ||NL(u)||_inf <= 1e-9.  value = unknowns / (wall time of a step).
    Roofline: the dominant kernel by profiled time, the A product of the
    Arnoldi steps (spmv64: FP64 values, column ids, row pointers, the
    gathered input and the output)."""
    from feng_amd.api import KSP_GMRES, PC_FIELDSPLIT, Mesh, essential_values
    from feng_amd.problems import NLSolverOptions, kovasznay, navier_stokes_system, solve_newton

    N = args.N
    m = Mesh.square(N, 0.2, 0)
    S, num = navier_stokes_system(m, 40.0, 5)
    n, nnz = S.n, S.nnz
    _, ufn, pfn = kovasznay(40.0)
    u0 = essential_values(num, [ufn, pfn])
    log(f"ns setup: N={N} n={n} nnz={nnz}")
    restart = 50

    def step(profile=False):
        S.set_solution(u0)
        S.set_profiling(profile)
        lin = []
        t0 = time.perf_counter()
        rep = solve_newton(S, NLSolverOptions(1e-9, 1e-12, 1e4, 12), progress=lambda it, r: lin.append(r),
                           ksp=KSP_GMRES, pc=PC_FIELDSPLIT, rtol=1e-8, restart=restart, max_iter=10000)
        wall = time.perf_counter() - t0
        if not rep.converged or not all(r.converged and r.relres <= 1e-8 for r in lin):
            raise SystemExit(f"ns: Newton or a GMRES solve did not converge: {rep}, {[r.relres for r in lin]}")
        return dict(wall=wall, newton=rep.iterations, gmres=[r.iterations for r in lin],
                    solve_ms=[r.ms_solve for r in lin], asm_ms=rep.asm_ms)

    for _ in range(max(1, args.warmup)):
        step()
    runs = [step() for _ in range(max(1, args.steps))]
    wall = float(np.mean([r["wall"] for r in runs]))
    # one extra solve with events around the dominant kernel (its last linear solve's figures)
    prof = step(profile=True)
    kern_ms, launches = S.kernel_stats(0)
    it_ms, _ = S.kernel_stats(2)
    alg_bytes = 12 * nnz + 4 * (n + 1) + 16 * n  # per launch: the FP64 SpMV
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9 if kern_ms else 0.
    cpu = None
    if not args.no_cpu:
        cpu = cpu_ns(m, num, S, runs[-1], args.cpu_budget)
    last = runs[-1]
    out = {"metric": "DOF/s of a Navier-Stokes Newton solve (assemble + GMRES(50) to rel-res 1e-8 per step)",
           "value": n / wall, "unit": "DOF/s", "n_gpus": 1, "steps": len(runs), "warmup": args.warmup,
           "ms_per_step": 1e3 * wall, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (jittered unit-square mesh, Kovasznay boundary values)",
           "config": {"workload": f"Kovasznay Re=40, Taylor-Hood P2-P1, jittered {N}^2 square, Newton + GMRES(50) "
                                  f"with the block Schur preconditioner (field split)", "n": n, "nnz": nnz,
                      "newton_iterations": last["newton"], "gmres_iterations": last["gmres"],
                      "solve_ms": last["solve_ms"], "asm_ms": last["asm_ms"]},
           "roofline": {"bound": "hbm", "kernel": "spmv64<8> (the Arnoldi steps' A product)", "achieved": achieved,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": None, "avg_launch_ms": kern_ms, "alg_bytes_per_launch": alg_bytes,
                        "launches_timed": launches, "arnoldi_step_ms": it_ms or None},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    S.close()


def cpu_ns(m, num, S, gpu, budget_s):
    """The oracle on the host: the Navier-Stokes assembly of one Newton
    iteration (serial C restatement of the reference's form stack), its
    ILU(0) and GMRES(50)-ILU(0) (the reference's PETSc combination) for a
    bounded number of iterations; the step's time extrapolated with the GPU
    path's Newton and GMRES iteration counts (ILU(0) needs many more
    iterations than the GPU preconditioner, so the figure favours the CPU)."""
    import ctypes as C

    from tests.oracle_bind import P, load_oracle

    O = load_oracle()
    n = S.n
    ia, ja = S.get_csr()
    uU, pP = num.adr
    u = S.get_solution()
    val = np.zeros(ja.size)
    rhs = np.zeros(n)
    t0 = time.perf_counter()
    rc = O.L.or_assemble_ns(m.V, P(m.xyz, C.c_double), m.E, P(m.conn, C.c_int), P(uU, C.c_int), P(pP, C.c_int), n,
                            P(ia, C.c_int), P(ja, C.c_int), 5, 1.0 / 40.0, P(u, C.c_double), None,
                            P(val, C.c_double), P(rhs, C.c_double))
    t_asm = time.perf_counter() - t0
    if rc != 0:
        return None
    t0 = time.perf_counter()
    lu, rc = O.ilu0(ia, ja, val)
    t_ilu = time.perf_counter() - t0
    probe = 5
    t0 = time.perf_counter()
    _, it, rr = O.gmres_ilu0(ia, ja, val, lu, rhs, 50, 1e-30, probe)
    t_it = (time.perf_counter() - t0) / max(1, it)
    total_gmres = int(sum(gpu["gmres"]))
    newton = gpu["newton"]
    t_step = newton * (t_asm + t_ilu) + t_it * total_gmres
    return {"value": n / t_step, "unit": "DOF/s", "cores": 1, "kind": "port",
            "sample": (f"oracle, one thread: Navier-Stokes assembly {t_asm:.2f}s and ILU(0) {t_ilu:.2f}s of one Newton "
                       f"iteration, GMRES(50)-ILU(0) {it} iterations at {1e3 * t_it:.0f} ms each; step = "
                       f"{newton} x (assembly + ILU(0)) + {total_gmres} GMRES iterations (the GPU path's counts)")}


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
    ap.add_argument("--workload", default="cube", choices=["cube", "batched", "transient", "ns"])
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--mfma-probe", action="store_true")
    ap.add_argument("--paths", default=None, help="comma-separated subset of " + ",".join(PATHS))
    ap.add_argument("--methods", default=None, help="batched workload: comma-separated subset of jacobi,amg")
    ap.add_argument("--no-batch-graph", action="store_true",
                    help="batched workload: launch the multigrid method's chunks directly, not from a captured graph")
    ap.add_argument("--batch-layout", default="union", choices=["union", "degree"],
                    help="batched workload: one system for both degrees, or one per degree")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the solver has no multi-GPU path "
              "(row-slab sharding with halo exchange is a design, not code)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        return copy_probe()
    if args.mfma_probe:
        return mfma_probe()
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "transient":
        return bench_transient(args)
    if args.workload == "ns":
        if args.N == 64:
            args.N = 512  # the Navier-Stokes configuration's mesh
        return bench_ns(args)
    rtol = 1e-8
    m, k, S, num = make_problem(args.N, args.nfeat, args.seed)
    n, nnz = num.nInc, S.nnz
    paths = ["plain", "fast", "ic0", "amg", "amg32"]
    try:
        from feng_amd.learned import load_weights

        load_weights(S)
        paths += ["learned", "learned64"]
    except FileNotFoundError as e:
        log("learned path skipped:", e)

    if args.paths:
        want = args.paths.split(",")
        paths = [p for p in paths if p in want]
        if "plain" not in paths:
            paths.insert(0, "plain")
    res = {}
    for path in paths:
        for _ in range(max(1, args.warmup)):
            run_step(S, path, rtol)
        # timed steps without per-launch events (they add queue packets);
        # one extra step with events gives the timed kernel's launch times
        t0 = time.perf_counter()
        steps = [run_step(S, path, rtol, profile=False) for _ in range(args.steps)]
        wall = (time.perf_counter() - t0) / args.steps
        prof = run_step(S, path, rtol, profile=True)
        for r in steps + [prof]:
            if not r["converged"] or r["relres"] > rtol:
                raise SystemExit(f"{path}: solve did not reach FP64 rel-res {rtol}: {r}")
        res[path] = dict(ms=1e3 * wall, dofs=n / wall, steps=steps, last=steps[-1], prof=prof)
        if PATHS[path][1] == "amg":
            rows_l, nnz_l = S.amg_levels()
            res[path]["extra"] = {"amg_rows": rows_l, "amg_nnz": nnz_l}
        log(path, f"{1e3 * wall:.3f} ms/step", steps[-1], res[path].get("extra", ""))

    # FP32 values of the FP32 paths against the FP64 values (those are the
    # oracle's assembly to 1e-12, as the GPU tests pin)
    from feng_amd.api import PREC_FP32_IR

    S.setToZero()
    S.assemble(3, PREC_FP32_IR)
    v64, _ = S.get_system()
    v32 = S.get_system32()
    fp32_err = float(np.abs(v32.astype(np.float64) - v64).max() / np.abs(v64).max())
    fast_paths = [p for p in ("fast", "learned", "learned64", "amg", "amg32") if p in res] or ["plain"]
    best = min(fast_paths, key=lambda p: res[p]["ms"])
    B = res[best]
    plain = res["plain"]

    # roofline of the dominant kernel of the reported path: one PCG iteration
    # launch.  Algorithmic bytes per iteration (minimal unfused PCG-Jacobi):
    # values s nnz + column ids 4 nnz + row pointers 4 (n+1) + 18 vector
    # passes of s n.  FP32 inner solve: s = 4.
    kern_ms = B["prof"]["spmv_ms"]
    kname, sz = ITER_KERNEL[best]
    per_iter = None
    if PATHS[best][1] == "amg":
        # dominant kernel by profiled time: the level-0 restriction pcgDown0
        rows_l, nnz_l = B["extra"]["amg_rows"], B["extra"]["amg_nnz"]
        v = 4  # level values FP32 (pc_values32 = 1)
        alg_bytes = mg_down0_bytes(sz, v, n, nnz, rows_l[1])
        it_bytes = mg_iter_bytes(sz, v, 8 if sz == 8 else 4, rows_l, nnz_l)
        it_ms = B["prof"]["iter_ms"]
        per_iter = {"alg_bytes": it_bytes, "avg_ms": it_ms,
                    "frac": it_bytes / (it_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if it_ms else None}
    else:
        alg_bytes = iter_bytes(sz, n, nnz)
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9 if kern_ms else 0.
    p_kern = plain["prof"]["spmv_ms"]
    p_alg = iter_bytes(8, n, nnz)
    b_asm = asm_bytes(m.E, m.V, n, nnz)
    asm_ms = float(np.mean([x["asm_ms"] for x in B["steps"]]))
    per_step = None
    if per_iter:
        step_bytes = b_asm + per_iter["alg_bytes"] * B["last"]["iters"]
        per_step = {"alg_bytes": step_bytes, "ms": B["ms"],
                    "frac": step_bytes / (B["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}

    cpu = None
    if not args.no_cpu:
        c = cpu_baseline(m, k, num, rtol, args.cpu_budget)

        def dofs(r):
            iters = r["iters"] if r["converged"] else plain["last"]["iters"]
            return n / (r["t_asm"] + r["t_iter"] * iters), iters
        v_all, it_all = dofs(c["all"])
        v_1, _ = dofs(c["1"])
        ca = c["all"]
        cpu = {"value": v_all, "unit": "DOF/s", "cores": c["threads"], "kind": "port",
               "value_1core": v_1,
               "sample": (f"oracle on {c['threads']} threads (all cores of the affinity): colored OpenMP "
                          f"assembly of all {m.E} tets ({ca['t_asm']:.3f}s; coloring {c['t_color']:.2f}s excluded "
                          f"like the CSR build) + threaded Jacobi-CG (FP64 CG preconditioned by the diagonal, the "
                          f"plain path's algorithm) {ca['iters']} iterations ({ca['t_cg']:.2f}s, "
                          f"{'converged' if ca['converged'] else 'bounded, extrapolated to'} {it_all} iterations); "
                          f"value_1core: the serial oracle on the same problem")}
        log("cpu", c)

    def summary(p):
        r = res[p]
        return {"ms": r["ms"], "dofs": r["dofs"], "iterations": r["last"]["iters"], "outer": r["last"]["outer"],
                "relres": r["last"]["relres"], "asm_ms": float(np.mean([x["asm_ms"] for x in r["steps"]])),
                "solve_ms": float(np.mean([x["solve_ms"] for x in r["steps"]])),
                "iter_kernel_ms": r["prof"]["spmv_ms"], "iteration_ms": r["prof"]["iter_ms"] or None,
                **r.get("extra", {})}

    out = {
        "metric": "DOF/s assemble+solve to FP64 rel-res 1e-8; learned/FP32 vs plain FP64 PCG",
        "value": B["dofs"],
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": B["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best in ("plain", "learned64", "ic0", "amg") else "fp32-inner/fp64-refine",
        "data": "synthetic (seeded GRF coefficient on a generated Kuhn-tet cube)",
        "config": {"workload": f"P1 tet diffusion, unit cube {args.N}^3 x6 Kuhn tets, k=exp(GRF l=0.1 s=1.5), "
                               f"PCG to FP64 rel-res 1e-8", "n": n, "nnz": nnz, "elements": m.E,
                   "path": {"fast": "FP32 Jacobi-PCG + FP64 iterative refinement",
                            "learned": "FP32 learned-IC(0) split PCG + FP64 iterative refinement",
                            "learned64": "FP64 learned-IC(0) split PCG",
                            "amg": "FP64 PCG with the aggregation-multigrid V-cycle (FP32 level values)",
                            "amg32": "FP32 PCG with the aggregation-multigrid V-cycle, FP64 refinement by "
                                     "residual replacement",
                            "plain": "FP64 Jacobi-PCG"}[best],
                   "paths": {p: {**summary(p), "speedup_vs_plain_fp64": plain["ms"] / res[p]["ms"]} for p in res},
                   "speedup_vs_plain_fp64": plain["ms"] / B["ms"],
                   "fp32_value_error_vs_fp64": fp32_err,
                   "plain_iter_roofline_frac": p_alg / (p_kern * 1e-3) / 1e9 / HBM_PEAK_GBS if p_kern else None},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC.get(kname),
                     "avg_launch_ms": kern_ms, "alg_bytes_per_launch": alg_bytes,
                     "per_iteration": per_iter, "per_step": per_step,
                     "assembly": {"kernel": "asmGather<double, 4, 3, 16>", "ms": asm_ms, "alg_bytes": b_asm,
                                  "frac": b_asm / (asm_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                  "traffic": TRAFFIC.get("asmGather<double, 4, 3, 16>")}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    S.close()


if __name__ == "__main__":
    main()
