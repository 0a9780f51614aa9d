"""Batched independent problems (fe_batch.hip): one disjoint-mesh system per
element degree, per-problem Jacobi-PCG in shared launches.  Each problem's
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
        res = solve_batch(b, precision=PREC_FP64, rtol=1e-10, pc=PC_AMG)
        assert res["converged"].all()
        b.S.correctSolution()
        out.setdefault(g, []).append((b.S.get_solution().copy(), res["iterations"].copy(), res["relres"].copy()))
    assert len(set(out[0][0][1])) > 1  # problems stop at different iterations: the active set shrinks
    for x, it, rr in out[1]:
        assert np.array_equal(x, out[0][0][0]) and np.array_equal(it, out[0][0][1]) and np.array_equal(rr, out[0][0][2])
    b.S.close()


@pytest.mark.timeout(900)
def test_batched_256_problem_draw_vs_oracle(oracle):
    """draw_problems(256, 1), both degrees: the FP32-assembly, FP32-inner +
    FP64-refined batched Jacobi-PCG (the batched configuration) brings every
    problem to FP64 rel-res <= 1e-8; every problem's CSR block is the
    oracle's pattern, its values and RHS the oracle's assembly to 1e-12, and
    its solution has rel-res <= 1.01e-8 on the oracle's system; P1 solutions
    are non-negative (the split-square P1 stiffness is an M-matrix, so the
    discrete maximum principle holds for a non-negative load)."""
    from concurrent.futures import ThreadPoolExecutor

    from feng_amd.api import PREC_FP32_IR, Numbering
    from feng_amd.batched import build_batch, draw_problems, solve_batch
    probs = draw_problems(256, 1)
    checked = 0
    for kind in (1, 2):
        b = build_batch(probs, kind, keep_meshes=True)
        print(f"kind {kind}: {len(b.nInc)} problems, n = {int(b.rows[-1])}", flush=True)
        res = solve_batch(b, precision=PREC_FP32_IR, rtol=1e-8)
        assert res["converged"].all() and (res["relres"] <= 1e-8).all(), res["relres"].max()
        ia, ja = b.S.get_csr()
        val, rhs = b.S.get_system()
        b.S.correctSolution()
        u = b.S.get_solution()

        def check(p):
            m, k = b.meshes[p], b.coeffs[p]
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
            if kind == 1:
                assert umin >= -1e-12, (p, umin)
        checked += len(out)
        print(f"kind {kind}: blocks and solutions of {len(out)} problems checked against the oracle, "
              f"max rel-res on the oracle's systems {max(o[3] for o in out):.2e}", flush=True)
        b.S.close()
    assert checked == 256
