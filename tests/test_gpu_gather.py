"""Row-gather assembly (fe_gather.hip, asm_kernel = 2, the default) and the
FP32 values of FE_PREC_FP32_IR against the FP64 oracle.

- Gather vs the oracle's assembly (feSysElm.cpp:520-589 diffusion, :16-27
  source, the colored Pardiso scatter feLinearSystemMklPardiso.cpp:501-749):
  values and RHS <= 1e-12 relative on P1 triangles, P2 triangles with a
  1e4-contrast block coefficient, P1 tets with a GRF coefficient; against
  the colored per-color sweep (asm_kernel = 0) at rounding level; bitwise
  repeatable.  The transient mass term (feSysElm.cpp:426-458) too.
- FP32 values: the default keeps the FP32 rounding of the FP64 values
  (exactly float(val64)); FP32 element rows (asm_fp32_math = 1) and the
  colored FP32 MFMA sweep stay within 1e-6 of the FP64 oracle's values
  (relative to max |val|); the FP32-inner + FP64-refinement solve stays within
  the FP64 CG solution's own accuracy.  The errors are printed per case.
- A matrix assembled by a gather sweep and a colored FP64-only sweep (a
  MASS form at another quadrature degree) still gets FP32 values equal to
  the FP32 rounding of the FP64 values.
"""
import math

import numpy as np
import pytest

from tests.cases import relerr

pytestmark = pytest.mark.gpu

BOUND_FP32_VALUES = 1e-6  # max |val32 - val64_oracle| / max |val64_oracle|


def _fn(t):
    from feng_amd.api import Fn
    kind, c0, c1, data = t
    return Fn(kind, c0, c1, data)


def _case(name):
    from feng_amd.api import Mesh
    if name == "p1_64":
        m = Mesh.square(64)
        return m, 1, 2, (0, 1.0, 0.0, None), (1, -2 * math.pi ** 2, 0.0, None)
    if name == "p2_blocks_48":
        m = Mesh.square(48, 0.2, 1)
        return m, 2, 2, (4, 0.0, 0.0, m.coeff_blocks(16, 1e-2, 1e2, 9)), (0, -1.0, 0.0, None)
    m = Mesh.cube(12)
    return m, 4, 1, (4, 0.0, 0.0, m.coeff_grf(0.1, 1.5, 256, 4)), (0, -1.0, 0.0, None)


def _oracle(oracle, m, num, kind, deg, k, s, u=None):
    nl = num.adr[0].size // m.E
    ia, ja = oracle.csr(num.nInc, [(num.adr[0], nl, num.adr[0], nl)])
    val, rhs = oracle.assemble_diffusion(kind, m.xyz, m.conn, num.adr[0], num.nInc, ia, ja, deg, k, s,
                                         np.zeros(num.nDofs) if u is None else u)
    return ia, ja, val, rhs


CASES = ["p1_64", "p2_blocks_48", "tet_grf_12"]


@pytest.mark.parametrize("name", CASES)
def test_gather_matches_oracle_and_colored_sweep(oracle, name):
    from feng_amd.problems import diffusion_system
    print("gather vs oracle", name, flush=True)
    m, kind, deg, k, s = _case(name)
    S, num = diffusion_system(m, kind, deg, _fn(k), _fn(s))
    assert S.get_tuning("asm_kernel") == 2
    rng = np.random.default_rng(3)
    u = S.get_solution()
    u[:num.nInc] = rng.standard_normal(num.nInc)  # a nonzero state: the residual's gradient term
    S.set_solution(u)
    ia0, ja0, val0, rhs0 = _oracle(oracle, m, num, kind, deg, k, s, u)
    S.setToZero()
    S.assemble(3)
    val, rhs = S.get_system()
    assert relerr(val, val0) <= 1e-12 and relerr(rhs, rhs0) <= 1e-12
    S.setToZero()
    S.assemble(3)
    v2, r2 = S.get_system()
    assert np.array_equal(val, v2) and np.array_equal(rhs, r2)  # bitwise repeatable
    S.set_tuning("asm_kernel", 0)  # the colored per-color sweep
    S.setToZero()
    S.assemble(3)
    vc, rc = S.get_system()
    assert relerr(val, vc) <= 1e-14 and relerr(rhs, rc) <= 1e-14
    # matrix only and residual only
    S.set_tuning("asm_kernel", 2)
    S.setToZero()
    S.assemble(1)
    vm, rm = S.get_system()
    assert np.array_equal(vm, val) and not rm.any()
    S.setToZero()
    S.assemble(2)
    vr, rr = S.get_system()
    assert not vr.any() and np.array_equal(rr, rhs)
    S.close()


def test_gather_stores_after_a_zeroing_and_adds_otherwise():
    """The first matrix sweep after setToZero stores its sums without reading
    the values; every other assembly adds to what is there: a second assembly
    without zeroing doubles the values, and values set from the host are
    added to."""
    from feng_amd.problems import diffusion_system
    print("gather store / add", flush=True)
    m, kind, deg, k, s = _case("tet_grf_12")
    S, num = diffusion_system(m, kind, deg, _fn(k), _fn(s))
    S.setToZero()
    S.assemble(3)
    v1, r1 = S.get_system()
    S.assemble(1)
    v2, _ = S.get_system()
    assert np.array_equal(v2, 2.0 * v1)
    base = np.linspace(-1.0, 1.0, v1.size)
    S.set_system(base, r1)
    S.assemble(1)
    v3, _ = S.get_system()
    assert np.array_equal(v3, base + v1)
    S.setToZero()
    S.assemble(1)
    v4, _ = S.get_system()
    assert np.array_equal(v4, v1)
    S.close()


def test_gather_transient_mass_matches_colored_sweep():
    from feng_amd.api import FN_CONST, FN_ELEM, Fn, Mesh
    from feng_amd.problems import heat_system
    print("gather transient", flush=True)
    m = Mesh.cube(10)
    kk = m.coeff_grf(0.1, 1.5, 128, 2)
    S, num, _ = heat_system(m, 4, 1, Fn(FN_ELEM, data=kk), Fn(FN_CONST, 1.0), Fn(FN_CONST, -1.0))
    rng = np.random.default_rng(5)
    u = S.get_solution()
    u[:num.nInc] = rng.standard_normal(num.nInc)
    S.set_solution(u)
    S.set_bdf([1.5e3, -2e3, 0.5e3])
    S.set_past_solution(1, rng.standard_normal(num.nDofs))
    S.set_past_solution(2, rng.standard_normal(num.nDofs))
    out = {}
    for kernel in (2, 0):
        S.set_tuning("asm_kernel", kernel)
        S.setToZero()
        S.assemble(3)
        out[kernel] = S.get_system()
    assert relerr(out[2][0], out[0][0]) <= 1e-14 and relerr(out[2][1], out[0][1]) <= 1e-14
    S.close()


@pytest.mark.parametrize("name", CASES)
def test_fp32_values_and_solution_vs_fp64_oracle(oracle, name):
    from feng_amd.api import PREC_FP32_IR
    from feng_amd.problems import diffusion_system
    print("fp32 values", name, flush=True)
    m, kind, deg, k, s = _case(name)
    S, num = diffusion_system(m, kind, deg, _fn(k), _fn(s))
    ia0, ja0, val0, rhs0 = _oracle(oracle, m, num, kind, deg, k, s)
    scale = np.abs(val0).max()
    errs = {}
    # default: FP32 rounding of the FP64 gather values
    S.setToZero()
    S.assemble(3, PREC_FP32_IR)
    val, rhs = S.get_system()
    v32 = S.get_system32()
    assert np.array_equal(v32, val.astype(np.float32))
    errs["rounded"] = np.abs(v32.astype(np.float64) - val0).max() / scale
    # FP32 element rows summed in FP32
    S.set_tuning("asm_fp32_math", 1)
    S.setToZero()
    S.assemble(3, PREC_FP32_IR)
    v32m = S.get_system32()
    errs["fp32_rows"] = np.abs(v32m.astype(np.float64) - val0).max() / scale
    S.set_tuning("asm_fp32_math", 0)
    # the colored FP32 MFMA sweep
    S.set_tuning("asm_kernel", 1)
    S.setToZero()
    S.assemble(3, PREC_FP32_IR)
    v32c = S.get_system32()
    errs["mfma_fp32"] = np.abs(v32c.astype(np.float64) - val0).max() / scale
    if kind != 1:  # FP32 arithmetic rounds differently from the FP64 values: the MFMA values are kept
        assert not np.array_equal(v32c, val.astype(np.float32))
    S.set_tuning("asm_kernel", 2)
    print(name, "FP32 value error vs FP64 oracle:", {k_: f"{v:.2e}" for k_, v in errs.items()}, flush=True)
    assert errs["rounded"] <= 2.0 ** -24 * 1.01
    assert errs["fp32_rows"] <= BOUND_FP32_VALUES and errs["mfma_fp32"] <= BOUND_FP32_VALUES
    # FP32 inner solves + FP64 refinement against the oracle's FP64 CG
    S.setToZero()
    S.assemble(3, PREC_FP32_IR)
    rtol = 1e-8 if kind == 4 else 1e-10
    r, du = S.solve(precision=PREC_FP32_IR, rtol=rtol, want_du=True)
    assert r.converged and r.relres <= rtol
    res = rhs0 - oracle.spmv(ia0, ja0, val0, du)
    assert np.linalg.norm(res) <= 1.01 * rtol * np.linalg.norm(rhs0)
    x0, _, _ = oracle.cg_jacobi(ia0, ja0, val0, rhs0, rtol * 0.1, 40000)
    err_x = np.abs(du - x0).max() / np.abs(x0).max()
    print(name, f"FP32-refined solution vs FP64 oracle CG: {err_x:.2e}", flush=True)
    assert err_x <= 1e-5
    S.close()


def test_fp32_values_with_a_colored_fp64_sweep():
    """DIFFUSION (P1, degree 2, gather) plus MASS (degree 3, colored FP64
    sweep): the FP32 copy must still be the FP32 rounding of every value."""
    from feng_amd import api as A
    from feng_amd.api import FN_CONST, PREC_FP32_IR, Fn, Mesh
    print("fp32 values with a colored sweep", flush=True)
    m = Mesh.square(32, 0.2, 4)
    num = A.Numbering(m, [1], [1])
    S = A.LinearSystemHIP(0)
    S.upload_mesh(m)
    f = S.add_field(1, num.adr[0])
    S.set_sizes(num.nInc, num.nDofs)
    S.add_form(A.DIFFUSION, [f], 2, [Fn(FN_CONST, 1.0)])
    S.add_form(A.MASS, [f], 3, [Fn(FN_CONST, 2.5)])
    S.add_form(A.SOURCE, [f], 2, [Fn(FN_CONST, -1.0)])
    S.build_csr()
    S.set_solution(np.zeros(num.nDofs))
    for kernel in (2, 1, 0):
        S.set_tuning("asm_kernel", kernel)
        S.setToZero()
        S.assemble(3, PREC_FP32_IR)
        val, _ = S.get_system()
        assert np.array_equal(S.get_system32(), val.astype(np.float32)), kernel
    S.close()
