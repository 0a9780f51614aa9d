"""Many independent diffusion problems assembled and solved together.

Problem p is a P1 or P2 Lagrange discretisation of -div(k_p grad u) + S = 0
(S = -1, u = 0 on the boundary) on its own N_p x N_p split-square mesh with
k_p piecewise constant on 8 x 8 blocks (feNG's numbering and forms per
problem: feNumber.cpp:677-843, feSysElm.cpp:16-27,520-589).  All problems of
one element degree are uploaded as ONE disjoint mesh whose DOFs are numbered
problem by problem: problem p's unknowns are the rows [rows[p], rows[p+1]) of
a block-diagonal system (its reference numbering shifted), the essential DOFs
of all problems follow.  One assembly then covers every problem, and
fe_hip_solve_batched runs an independent Jacobi-PCG per problem in shared
launches.  build_union_batch puts the problems of BOTH degrees into one
system (a P1 and a P2 field on the union mesh, each mapping the other
degree's elements to one extra essential DOF), so that one assembly and one
batched solve cover the whole draw; the batched multigrid PCG's V-cycle then
skips the tiles of problems that have stopped.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from . import api as A


@dataclass
class UnionMesh:
    """Duck-typed mesh for LinearSystemHIP.upload_mesh."""
    dim: int
    V: int
    E: int
    nvpe: int
    xyz: np.ndarray
    conn: np.ndarray


@dataclass
class Batch:
    kind: int                 # element kind of every problem; 0: both degrees (build_union_batch)
    sizes: list
    seeds: list
    S: A.LinearSystemHIP
    rows: np.ndarray          # nprob + 1 unknown-row offsets
    nInc: list                # unknowns per problem
    dof_maps: list = field(default_factory=list)  # per problem: union DOF id of each local DOF
    meshes: list = field(default_factory=list)
    coeffs: list = field(default_factory=list)
    setup_s: float = 0.0


# over-correction factors (x100) the multigrid method's autotuner times; the
# last is the library default
AMG_OC_CHOICES = (140, 160, 180)


def draw_problems(nprob=256, seed=0, nmin=32, nmax=256):
    """N_p uniform in [nmin, nmax], degree P1/P2 50/50, coefficient seeds."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, nprob)
    kinds = np.where(rng.random(nprob) < 0.5, A.TRI_P1, A.TRI_P2)
    seeds = rng.integers(1, 2 ** 31, nprob)
    return [(int(N_), int(k), int(s)) for N_, k, s in zip(sizes, kinds, seeds)]


def build_batch(problems, kind, device=0, keep_meshes=False) -> Batch:
    """One device system holding every problem of element kind `kind`."""
    t0 = time.perf_counter()
    probs = [(N_, s) for N_, k, s in problems if k == kind]
    meshes, nums, ks = [], [], []
    for N_, s in probs:
        m = A.Mesh.square(N_)
        meshes.append(m)
        nums.append(A.Numbering(m, [kind], [1]))
        ks.append(m.coeff_blocks(8, 0.1, 10.0, s))
    nInc = [nm.nInc for nm in nums]
    nEss = [nm.nDofs - nm.nInc for nm in nums]
    NU = int(sum(nInc))
    uoff = np.concatenate([[0], np.cumsum(nInc)]).astype(np.int64)
    eoff = np.concatenate([[0], np.cumsum(nEss)]).astype(np.int64)
    voff = np.concatenate([[0], np.cumsum([m.V for m in meshes])]).astype(np.int64)
    xyz = np.concatenate([m.xyz for m in meshes])
    conn = np.concatenate([m.conn.astype(np.int64) + voff[p] for p, m in enumerate(meshes)]).astype(np.int32)
    maps, adr = [], []
    for p, nm in enumerate(nums):
        loc = np.arange(nm.nDofs)
        gm = np.where(loc < nm.nInc, uoff[p] + loc, NU + eoff[p] + (loc - nm.nInc))
        maps.append(gm)
        adr.append(gm[nm.adr[0]])
    adr = np.concatenate(adr).astype(np.int32)
    k = np.concatenate(ks)
    um = UnionMesh(2, int(voff[-1]), int(conn.size // 3), 3, xyz, conn)
    S = A.LinearSystemHIP(device)
    S.upload_mesh(um)
    f = S.add_field(kind, adr)
    S.set_sizes(NU, NU + int(eoff[-1]))
    S.add_form(A.DIFFUSION, [f], 2, [A.Fn(A.FN_ELEM, data=k)])
    S.add_form(A.SOURCE, [f], 2, [A.Fn(A.FN_CONST, -1.0)])
    S.build_csr()
    S.set_solution(np.zeros(NU + int(eoff[-1])))
    b = Batch(kind, [N_ for N_, _ in probs], [s for _, s in probs], S, uoff.astype(np.int32), nInc, maps)
    if keep_meshes:
        b.meshes, b.coeffs = meshes, ks
    b.setup_s = time.perf_counter() - t0
    return b


def union_layout(problems):
    """Host side of build_union_batch (no device work): per problem its mesh,
    numbering and block coefficient; the union mesh (V vertices, 3 V
    coordinates xyz, conn); the unknown
    rows (rows[p]..rows[p+1] for problem p, in draw order) and per problem the
    union DOF of each of its local DOFs (maps); per element kind the union DOF
    map of the field over every element of the union mesh (adr), the other
    degree's elements mapped to the extra essential DOF `dummy`."""
    meshes, nums, ks, kinds = [], [], [], []
    for N_, kind, s in problems:
        m = A.Mesh.square(N_)
        meshes.append(m)
        nums.append(A.Numbering(m, [kind], [1]))
        ks.append(m.coeff_blocks(8, 0.1, 10.0, s))
        kinds.append(kind)
    nInc = [nm.nInc for nm in nums]
    nEss = [nm.nDofs - nm.nInc for nm in nums]
    NU = int(sum(nInc))
    uoff = np.concatenate([[0], np.cumsum(nInc)]).astype(np.int64)
    eoff = np.concatenate([[0], np.cumsum(nEss)]).astype(np.int64)
    voff = np.concatenate([[0], np.cumsum([m.V for m in meshes])]).astype(np.int64)
    dummy = NU + int(eoff[-1])  # the essential DOF of the other degree's elements
    xyz = np.concatenate([m.xyz for m in meshes])
    conn = np.concatenate([m.conn.astype(np.int64) + voff[p] for p, m in enumerate(meshes)]).astype(np.int32)
    maps, adr = [], {A.TRI_P1: [], A.TRI_P2: []}
    for p, nm in enumerate(nums):
        loc = np.arange(nm.nDofs)
        gm = np.where(loc < nm.nInc, uoff[p] + loc, NU + eoff[p] + (loc - nm.nInc))
        maps.append(gm)
        for kind, nl in ((A.TRI_P1, 3), (A.TRI_P2, 6)):
            adr[kind].append(gm[nm.adr[0]] if kinds[p] == kind else np.full(meshes[p].E * nl, dummy))
    adr = {kind: np.concatenate(v).astype(np.int32) for kind, v in adr.items()}
    return dict(meshes=meshes, nums=nums, coeffs=ks, kinds=kinds, nInc=nInc, NU=NU, rows=uoff.astype(np.int32),
                dummy=dummy, V=int(voff[-1]), xyz=xyz, conn=conn, maps=maps, adr=adr, coeff=np.concatenate(ks))


def build_union_batch(problems, device=0, keep_meshes=False) -> Batch:
    """One device system holding every problem of the draw, P1 and P2
    together: the disjoint union mesh carries a P1 and a P2 field; problem p's
    unknowns are rows [rows[p], rows[p+1]) in draw order, numbered by its own
    field.  On an element of the other degree a field's map points at one
    extra essential DOF (value 0), so that element adds no row, column or
    gather incidence of that field.  One assembly and one batched solve then
    cover the whole draw."""
    t0 = time.perf_counter()
    U = union_layout(problems)
    um = UnionMesh(2, U["V"], int(U["conn"].size // 3), 3, U["xyz"], U["conn"])
    S = A.LinearSystemHIP(device)
    S.upload_mesh(um)
    S.set_sizes(U["NU"], U["dummy"] + 1)
    for kind in (A.TRI_P1, A.TRI_P2):
        f = S.add_field(kind, U["adr"][kind])
        S.add_form(A.DIFFUSION, [f], 2, [A.Fn(A.FN_ELEM, data=U["coeff"])])
        S.add_form(A.SOURCE, [f], 2, [A.Fn(A.FN_CONST, -1.0)])
    S.build_csr()
    S.set_solution(np.zeros(U["dummy"] + 1))
    b = Batch(0, [N_ for N_, _, _ in problems], [s for _, _, s in problems], S, U["rows"], U["nInc"], U["maps"])
    if keep_meshes:
        b.meshes, b.coeffs = U["meshes"], U["coeffs"]
    b.setup_s = time.perf_counter() - t0
    return b


def solve_batch(b: Batch, precision=A.PREC_FP32_IR, rtol=1e-8, inner_rtol=1e-2, max_iter=100000, autotune=True,
                check=True, pc=A.PC_JACOBI, profile=False):
    """Assemble every problem (zero + all element kernels + scatter) and solve
    them all; returns timings and per-problem statistics.  autotune: the
    assembly kernel variant (fe_hip_autotune_assembly) and, for the Jacobi
    method, the iteration's launch shape are timed and the fastest kept; for
    the multigrid method, the coarse correction's over-correction factor
    (amg_oc_x100 in AMG_OC_CHOICES) is chosen by timing one solve with each.
    profile: events around the dominant kernel's launches (kernel_ms,
    kernel_launches)."""
    L, S = b.S.L, b.S
    rows = np.ascontiguousarray(b.rows, np.int32)
    nprob = len(b.nInc)
    tuned = None
    if autotune:
        ak, sl, gx, ams = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        S._c(L.fe_hip_autotune_assembly(S.ctx, precision, C.byref(ak), C.byref(sl), C.byref(gx), C.byref(ams)),
             "fe_hip_autotune_assembly")
        tuned = dict(asm_kernel=ak.value, scalar_lanes=sl.value, gather_xcd=gx.value, asm_ms=ams.value)
    S.set_profiling(profile)
    S.setToZero()
    S.assemble(3, precision)
    asm_ms = S.kernel_stats(1)[0]
    if autotune and pc == A.PC_JACOBI:
        lanes, block, ms = C.c_int(), C.c_int(), C.c_double()
        S._c(L.fe_hip_autotune_batched(S.ctx, nprob, N.ptr(rows, C.c_int), precision, C.byref(lanes),
                                       C.byref(block), C.byref(ms)), "fe_hip_autotune_batched")
        tuned.update(lanes=lanes.value, block=block.value, ms_per_iter=ms.value)
    o = N.FeHipSolverOpts()
    L.fe_hip_default_opts(C.byref(o))
    o.precision, o.rtol, o.atol, o.inner_rtol, o.max_iter = precision, rtol, 1e-300, inner_rtol, max_iter
    o.pc = pc
    st = (N.FeHipBatchStats * nprob)()
    ms = C.c_double()
    if autotune and pc == A.PC_AMG:
        def timed():
            S._c(L.fe_hip_solve_batched(S.ctx, nprob, N.ptr(rows, C.c_int), C.byref(o), st, C.byref(ms)),
                 "fe_hip_solve_batched")
            return ms.value if all(x.converged for x in st) else None
        trial = {}
        for oc in AMG_OC_CHOICES:
            S.set_tuning("amg_oc_x100", oc)
            t = timed()
            if t is not None:
                trial[oc] = t
        oc = min(trial, key=trial.get) if trial else AMG_OC_CHOICES[-1]
        S.set_tuning("amg_oc_x100", oc)
        tuned.update(amg_oc_x100=oc, amg_oc_ms={k: round(v, 1) for k, v in trial.items()})
    S._c(L.fe_hip_solve_batched(S.ctx, nprob, N.ptr(rows, C.c_int), C.byref(o), st, C.byref(ms)),
         "fe_hip_solve_batched")
    its = np.array([s.iterations for s in st])
    rel = np.array([s.relres_fp64 for s in st])
    conv = np.array([bool(s.converged) for s in st])
    kms, kl = S.kernel_stats(0)
    return dict(asm_ms=asm_ms, solve_ms=ms.value, iterations=its, relres=rel, converged=conv, tuned=tuned,
                n=int(b.rows[-1]), nnz=S.nnz, kernel_ms=kms, kernel_launches=kl)
