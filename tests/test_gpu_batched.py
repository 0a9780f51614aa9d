"""Batched independent problems (fe_batch.hip): one disjoint-mesh system per
element degree or one for both degrees, per-problem Jacobi- or multigrid-PCG
in shared launches.  Each problem's
block of the union system must equal the problem assembled alone (CSR
bit-exact, values to 1e-12), and each problem's solution must match its own
single-system solve.  On the full 256-problem draw of the batched workload,
every problem's block equals the oracle's assembly of that problem
(or_assemble_diffusion, feSysElm.cpp:16-27,520-589) and every problem's
solution meets rel-res 1e-8 on the oracle's system (or_spmv)."""
import numpy as np
import pytest

from tests.cases import relerr

pytestmark = pytest.mark.gpu


def _alone(N_, kind, seed):
    from feng_amd.api import FN_CONST, FN_ELEM, Fn, Mesh
    from feng_amd.problems import diffusion_system
    m = Mesh.square(N_)
    k = m.coeff_blocks(8, 0.1, 10.0, seed)
    S, num = diffusion_system(m, kind, 2, Fn(FN_ELEM, data=k), Fn(FN_CONST, -1.0))
    return S, num


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_individual_systems(kind, precision):
    from feng_amd.api import PREC_FP64
    from feng_amd.batched import build_batch, solve_batch
    probs = [(9, kind, 11), (17, kind, 12), (12, kind, 13), (24, kind, 14)]
    b = build_batch(probs, kind)
    res = solve_batch(b, precision=precision, rtol=1e-10, autotune=(precision == 1))
    assert res["converged"].all() and (res["relres"] <= 1e-10).all()
    ia, ja = b.S.get_csr()
    val, rhs = b.S.get_system()
    b.S.correctSolution()
    u = b.S.get_solution()
    for p, (N_, _, seed) in enumerate(probs):
        S1, num = _alone(N_, kind, seed)
        S1.setToZero()
        S1.assemble(3, PREC_FP64)
        ia1, ja1 = S1.get_csr()
        v1, r1 = S1.get_system()
        r0, r1_ = b.rows[p], b.rows[p + 1]
        assert r1_ - r0 == num.nInc
        blk = slice(ia[r0], ia[r1_])
        assert np.array_equal(ia[r0:r1_ + 1] - ia[r0], ia1)
        assert np.array_equal(ja[blk] - r0, ja1)
        assert relerr(val[blk], v1) <= 1e-12 and relerr(rhs[r0:r1_], r1) <= 1e-12
        r = S1.solve(rtol=1e-10)
        S1.correctSolution()
        u1 = S1.get_solution()[:num.nInc]
        up = u[b.dof_maps[p][:num.nInc]]
        assert relerr(up, u1) <= 1e-6
        if precision == 0:
            # batching costs no convergence: no more iterations than the
            # problem solved alone (same tolerance, other restart rule)
            assert int(res["iterations"][p]) <= 1.1 * r.iterations + 5, (res["iterations"][p], r.iterations)
        S1.close()
    b.S.close()


@pytest.mark.parametrize("kind", [1, 2])
def test_batched_amg_matches_individual_systems(kind):
    """AMG-PCG on the union system: every problem meets its own FP64 rel-res
    target and matches its single-system solve."""
    from feng_amd.api import PC_AMG, PREC_FP64
    from feng_amd.batched import build_batch, solve_batch
    probs = [(9, kind, 21), (17, kind, 22), (12, kind, 23), (24, kind, 24), (31, kind, 25)]
    b = build_batch(probs, kind)
    res = solve_batch(b, precision=PREC_FP64, rtol=1e-10, pc=PC_AMG)
    assert res["converged"].all() and (res["relres"] <= 1e-10).all(), res["relres"]
    b.S.correctSolution()
    u = b.S.get_solution()
    for p, (N_, _, seed) in enumerate(probs):
        S1, num = _alone(N_, kind, seed)
        S1.setToZero()
        S1.assemble(3, PREC_FP64)
        S1.solve(rtol=1e-10)
        S1.correctSolution()
        u1 = S1.get_solution()[:num.nInc]
        assert relerr(u[b.dof_maps[p][:num.nInc]], u1) <= 1e-6
        S1.close()
    b.S.close()


@pytest.mark.parametrize("kind", [1, 2])
def test_assembly_autotuner_keeps_a_variant_with_the_same_system(kind):
    """fe_hip_autotune_assembly times the row gather (two slice orders), the
    colored thread-per-element and lane-per-row sweeps and the colored MFMA
    sweep, and keeps the fastest in the tuning; every variant assembles the
    same system (values and RHS to 1e-13), so the choice changes time only."""
    import ctypes as C
    from feng_amd.api import PREC_FP64
    from feng_amd.batched import build_batch
    print("assembly autotune", kind, flush=True)
    probs = [(20, kind, 41), (33, kind, 42), (27, kind, 43)]
    b = build_batch(probs, kind)
    S = b.S
    systems = []
    for ak, sl in ((2, 0), (1, 0), (0, 0), (0, 1)):
        S.set_tuning("asm_kernel", ak)
        S.set_tuning("scalar_lanes", sl)
        S.setToZero()
        S.assemble(3, PREC_FP64)
        systems.append(S.get_system())
    v0, r0 = systems[0]
    for v, r in systems[1:]:
        assert np.linalg.norm(v - v0) <= 1e-13 * np.linalg.norm(v0) and np.linalg.norm(r - r0) <= 1e-13 * np.linalg.norm(r0)
    ak, sl, gx, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    S._c(S.L.fe_hip_autotune_assembly(S.ctx, PREC_FP64, C.byref(ak), C.byref(sl), C.byref(gx), C.byref(ms)),
         "fe_hip_autotune_assembly")
    print("autotuned assembly: asm_kernel", ak.value, "scalar_lanes", sl.value, "gather_xcd", gx.value,
          f"{ms.value:.3f} ms", flush=True)
    assert (S.get_tuning("asm_kernel"), S.get_tuning("scalar_lanes"), S.get_tuning("gather_xcd")) == \
        (ak.value, sl.value, gx.value)
    assert 0 < ms.value < 1e3
    S.setToZero()
    S.assemble(3, PREC_FP64)
    v, r = S.get_system()
    assert np.linalg.norm(v - v0) <= 1e-13 * np.linalg.norm(v0)
    S.close()


def test_batched_amg_graph_replays_equal_direct_launches():
    """The per-problem multigrid PCG replays its chunks of 16 iterations from a
    captured graph (batch_graph 1, the default): solutions, iterations and
    residuals equal direct launches (batch_graph 0) bit for bit, over many
    replays, with eager host work between them and a recapture each time
    problems drop out of the active tiles."""
    from feng_amd.api import PC_AMG, PREC_FP64
    from feng_amd.batched import build_batch, solve_batch
    print("batched AMG graph", flush=True)
    probs = [(9, 2, 31), (40, 2, 32), (17, 2, 33), (64, 2, 34), (25, 2, 35), (48, 2, 36)]
    b = build_batch(probs, 2)
    out = {}
    for g in (1, 0, 1):
        b.S.set_tuning("batch_graph", g)
        b.S.set_solution(np.zeros(b.S.nDofs))
        res = solve_batch(b, precision=PREC_FP64, rtol=1e-10, pc=PC_AMG, autotune=False)
        assert res["converged"].all()
        b.S.correctSolution()
        out.setdefault(g, []).append((b.S.get_solution().copy(), res["iterations"].copy(), res["relres"].copy()))
    assert len(set(out[0][0][1])) > 1  # problems stop at different iterations: the active set shrinks
    for x, it, rr in out[1]:
        assert np.array_equal(x, out[0][0][0]) and np.array_equal(it, out[0][0][1]) and np.array_equal(rr, out[0][0][2])
    b.S.close()


def test_batched_amg_tile_skipping_equals_every_tile():
    """The V-cycle's level launches skip the tiles whose problems have all
    stopped (batch_skip 1, the default): each row only couples to rows of its
    own problem on every level, so solutions, iterations and residuals equal
    the launches over every tile (batch_skip 0) bit for bit, on one degree
    and on the union system of both degrees.  The same holds with the row
    passes that store the gathered vectors of every level first
    (amg_pre_rows 1) and without them (amg_pre_rows 0), with four and with
    two lanes per row."""
    from feng_amd.api import PC_AMG, PREC_FP64
    from feng_amd.batched import build_batch, build_union_batch, solve_batch
    print("batched AMG tile skipping", flush=True)
    probs = [(9, 2, 41), (40, 1, 42), (17, 2, 43), (64, 2, 44), (25, 1, 45), (48, 2, 46), (33, 1, 47)]
    for b in (build_batch(probs, 2), build_union_batch(probs)):
        out = {}
        for skip, pre in ((1, 1), (0, 0), (1, 0), (0, 1)):
            b.S.set_tuning("batch_skip", skip)
            b.S.set_tuning("amg_pre_rows", pre)
            b.S.set_solution(np.zeros(b.S.nDofs))
            res = solve_batch(b, precision=PREC_FP64, rtol=1e-10, pc=PC_AMG, autotune=False)
            assert res["converged"].all()
            b.S.correctSolution()
            out[skip, pre] = (b.S.get_solution().copy(), res["iterations"].copy(), res["relres"].copy())
        assert len(set(out[0, 0][1])) > 1  # problems stop at different iterations
        for key in out:
            for a_, b_ in zip(out[key], out[0, 0]):
                assert np.array_equal(a_, b_), key
        # two lanes per row (the other summation order): row passes on and off
        b.S.set_tuning("amg_lanes", 2)
        two = []
        for pre in (1, 0):
            b.S.set_tuning("amg_pre_rows", pre)
            b.S.set_solution(np.zeros(b.S.nDofs))
            res = solve_batch(b, precision=PREC_FP64, rtol=1e-10, pc=PC_AMG, autotune=False)
            assert res["converged"].all()
            b.S.correctSolution()
            two.append((b.S.get_solution().copy(), res["iterations"].copy(), res["relres"].copy()))
        for a_, b_ in zip(two[0], two[1]):
            assert np.array_equal(a_, b_)
        b.S.close()


@pytest.mark.parametrize("method", ["jacobi", "amg"])
def test_union_batch_of_both_degrees_matches_problems_alone(method):
    """One system for P1 and P2 problems together (build_union_batch): every
    problem's block equals the problem assembled alone (CSR bit-exact, values
    and RHS to 1e-12) and its solution, by the batched Jacobi- or
    multigrid-PCG, its single-system solve to 1e-6."""
    from feng_amd.api import PC_AMG, PC_JACOBI, PREC_FP64
    from feng_amd.batched import build_union_batch, solve_batch
    print(f"union batch {method}", flush=True)
    pc = PC_AMG if method == "amg" else PC_JACOBI
    probs = [(9, 1, 51), (17, 2, 52), (12, 2, 53), (24, 1, 54), (31, 2, 55), (20, 1, 56)]
    b = build_union_batch(probs)
    res = solve_batch(b, precision=PREC_FP64, rtol=1e-10, pc=pc)
    assert res["converged"].all() and (res["relres"] <= 1e-10).all(), res["relres"]
    ia, ja = b.S.get_csr()
    val, rhs = b.S.get_system()
    b.S.correctSolution()
    u = b.S.get_solution()
    for p, (N_, kind, seed) in enumerate(probs):
        S1, num = _alone(N_, kind, seed)
        S1.setToZero()
        S1.assemble(3, PREC_FP64)
        ia1, ja1 = S1.get_csr()
        v1, r1 = S1.get_system()
        r0, r1_ = b.rows[p], b.rows[p + 1]
        assert r1_ - r0 == num.nInc
        blk = slice(ia[r0], ia[r1_])
        assert np.array_equal(ia[r0:r1_ + 1] - ia[r0], ia1)
        assert np.array_equal(ja[blk] - r0, ja1)
        assert relerr(val[blk], v1) <= 1e-12 and relerr(rhs[r0:r1_], r1) <= 1e-12
        S1.solve(rtol=1e-10)
        S1.correctSolution()
        u1 = S1.get_solution()[:num.nInc]
        assert relerr(u[b.dof_maps[p][:num.nInc]], u1) <= 1e-6
        S1.close()
    b.S.close()


def _check_blocks_vs_oracle(oracle, b, kinds, res, label):
    """Every problem of batch b (element kind kinds[p]): CSR block = the
    oracle's pattern, values and RHS = the oracle's assembly to 1e-12, the
    solution's rel-res on the oracle's system <= 1.01e-8, P1 solutions
    non-negative (the split-square P1 stiffness is an M-matrix, so the
    discrete maximum principle holds for a non-negative load)."""
    from concurrent.futures import ThreadPoolExecutor

    from feng_amd.api import Numbering
    assert res["converged"].all() and (res["relres"] <= 1e-8).all(), res["relres"].max()
    ia, ja = b.S.get_csr()
    val, rhs = b.S.get_system()
    b.S.correctSolution()
    u = b.S.get_solution()

    def check(p):
        m, k, kind = b.meshes[p], b.coeffs[p], kinds[p]
        num = Numbering(m, [kind], [1])
        nl = num.adr[0].size // m.E
        ia0, ja0 = oracle.csr(num.nInc, [(num.adr[0], nl, num.adr[0], nl)])
        val0, rhs0 = oracle.assemble_diffusion(kind, m.xyz, m.conn, num.adr[0], num.nInc, ia0, ja0, 2,
                                               (4, 0.0, 0.0, k), (0, -1.0, 0.0, None), np.zeros(num.nDofs))
        r0, r1 = int(b.rows[p]), int(b.rows[p + 1])
        blk = slice(ia[r0], ia[r1])
        ok_csr = np.array_equal(ia[r0:r1 + 1] - ia[r0], ia0) and np.array_equal(ja[blk] - r0, ja0)
        e_val, e_rhs = relerr(val[blk], val0), relerr(rhs[r0:r1], rhs0)
        up = u[b.dof_maps[p][:num.nInc]]
        rr = np.linalg.norm(rhs0 - oracle.spmv(ia0, ja0, val0, up)) / np.linalg.norm(rhs0)
        return ok_csr, e_val, e_rhs, rr, up.min()

    with ThreadPoolExecutor(8) as ex:
        out = list(ex.map(check, range(len(b.nInc))))
    for p, (ok_csr, e_val, e_rhs, rr, umin) in enumerate(out):
        assert ok_csr, p
        assert e_val <= 1e-12 and e_rhs <= 1e-12, (p, e_val, e_rhs)
        assert rr <= 1.01e-8, (p, rr)
        if kinds[p] == 1:
            assert umin >= -1e-12, (p, umin)
    print(f"{label}: blocks and solutions of {len(out)} problems checked against the oracle, "
          f"max rel-res on the oracle's systems {max(o[3] for o in out):.2e}", flush=True)
    return len(out)


@pytest.mark.timeout(900)
def test_batched_256_problem_draw_vs_oracle(oracle):
    """draw_problems(256, 1), one system per degree: the FP32-assembly,
    FP32-inner + FP64-refined batched Jacobi-PCG brings every problem to FP64
    rel-res <= 1e-8, and every problem's block and solution pass the oracle
    checks (_check_blocks_vs_oracle)."""
    from feng_amd.api import PREC_FP32_IR
    from feng_amd.batched import build_batch, draw_problems, solve_batch
    probs = draw_problems(256, 1)
    checked = 0
    for kind in (1, 2):
        b = build_batch(probs, kind, keep_meshes=True)
        print(f"kind {kind}: {len(b.nInc)} problems, n = {int(b.rows[-1])}", flush=True)
        res = solve_batch(b, precision=PREC_FP32_IR, rtol=1e-8)
        checked += _check_blocks_vs_oracle(oracle, b, [kind] * len(b.nInc), res, f"kind {kind}")
        b.S.close()
    assert checked == 256


@pytest.mark.timeout(900)
def test_union_256_problem_draw_amg_vs_oracle(oracle):
    """draw_problems(256, 2) as ONE system of both degrees (the batched
    workload's layout), solved by the batched multigrid PCG with the tiles of
    stopped problems skipped: every problem reaches FP64 rel-res <= 1e-8 and
    its block and solution pass the oracle checks."""
    from feng_amd.api import PC_AMG, PREC_FP64
    from feng_amd.batched import build_union_batch, draw_problems, solve_batch
    probs = draw_problems(256, 2)
    b = build_union_batch(probs, keep_meshes=True)
    print(f"union: {len(b.nInc)} problems, n = {int(b.rows[-1])}", flush=True)
    res = solve_batch(b, precision=PREC_FP64, rtol=1e-8, pc=PC_AMG)
    print(f"union AMG: iterations {res['iterations'].min()}..{res['iterations'].max()}, "
          f"solve {res['solve_ms']:.0f} ms, tuned {res['tuned']}", flush=True)
    assert _check_blocks_vs_oracle(oracle, b, [k for _, k, _ in probs], res, "union") == 256
    b.S.close()
