"""GPU parity of the fused Taylor-Hood Navier-Stokes kernel (libfe_hip.so)
against the reference fixtures and the CPU oracle."""
import ctypes as C

import numpy as np
import pytest

from tests.cases import NS_CASES, load, product_mesh, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", NS_CASES)
def test_ns_csr_and_assembly_match_reference(name):
    from feng_amd.problems import navier_stokes_system
    print("case", name, flush=True)
    g = load(name)
    m = product_mesh(g)
    S, num = navier_stokes_system(m, float(g["extra"]), g["deg"])
    assert (num.nInc, num.nDofs) == (g["n"], g["ndofs"])
    u0 = S.get_solution()
    assert relerr(u0, g["u0"]) <= 1e-13
    ia, ja = S.get_csr()
    assert np.array_equal(ia, g["ia"]) and np.array_equal(ja, g["ja"])
    S.assemble(3)
    val, rhs = S.get_system()
    assert relerr(val, g["val"]) <= 1e-12
    assert relerr(rhs, g["rhs"]) <= 1e-12
    S.close()


def test_ns_jittered_32_vs_oracle(oracle):
    """32x32 jittered Taylor-Hood mesh, nonzero interior velocity (a Newton
    state): device values vs the oracle's restatement."""
    from feng_amd.api import Mesh
    from feng_amd.problems import navier_stokes_system
    from tests.oracle_bind import P
    m = Mesh.square(32, 0.2, 0)
    S, num = navier_stokes_system(m, 40.0, 5)
    rng = np.random.default_rng(3)
    u = S.get_solution()
    u[: num.nInc] = rng.standard_normal(num.nInc) * 0.3
    S.set_solution(u)
    ia, ja = S.get_csr()
    uU, pP = num.adr
    forms = [(uU, 12, uU, 12), (pP, 3, uU, 12),
             (uU, 12, np.hstack([uU.reshape(-1, 12), pP.reshape(-1, 3)]).ravel(), 15)]
    ia0, ja0 = oracle.csr(num.nInc, forms)
    assert np.array_equal(ia, ia0) and np.array_equal(ja, ja0)
    S.assemble(3)
    val, rhs = S.get_system()
    val0 = np.zeros(ja0.size)
    rhs0 = np.zeros(num.nInc)
    rc = oracle.L.or_assemble_ns(m.V, P(m.xyz, C.c_double), m.E, P(m.conn, C.c_int), P(uU, C.c_int),
                                 P(pP, C.c_int), num.nInc, P(ia0, C.c_int), P(ja0, C.c_int), 5, 1.0 / 40.0,
                                 P(u, C.c_double), None, P(val0, C.c_double), P(rhs0, C.c_double))
    assert rc == 0
    assert relerr(val, val0) <= 1e-12 and relerr(rhs, rhs0) <= 1e-12
    S.close()


def test_ns_jacobian_on_the_matrix_cores(oracle):
    """Matrix-only sweeps of the Navier-Stokes stack (convection, Newtonian
    stress, divergence) take the MFMA form A = sum_q B_q^T D_q B_q
    (asmNsMfma): on the 32x32 jittered mesh at a random Newton state its
    Jacobian equals the scalar fused sweep (ns_mfma 0) to 1e-13 and the
    oracle's to 1e-12."""
    from feng_amd.api import Mesh
    from feng_amd.problems import navier_stokes_system
    from tests.oracle_bind import P
    print("NS Jacobian MFMA", flush=True)
    m = Mesh.square(32, 0.2, 0)
    S, num = navier_stokes_system(m, 40.0, 5)
    rng = np.random.default_rng(7)
    u = S.get_solution()
    u[: num.nInc] = rng.standard_normal(num.nInc) * 0.3
    S.set_solution(u)
    ia, ja = S.get_csr()
    vals = {}
    for mf in (1, 0):
        S.set_tuning("ns_mfma", mf)
        S.setToZero()
        S.assembleMatrices()
        vals[mf] = S.get_system()[0]
    uU, pP = num.adr
    val0 = np.zeros(ja.size)
    rhs0 = np.zeros(num.nInc)
    rc = oracle.L.or_assemble_ns(m.V, P(m.xyz, C.c_double), m.E, P(m.conn, C.c_int), P(uU, C.c_int),
                                 P(pP, C.c_int), num.nInc, P(ia, C.c_int), P(ja, C.c_int), 5, 1.0 / 40.0,
                                 P(u, C.c_double), None, P(val0, C.c_double), P(rhs0, C.c_double))
    assert rc == 0
    print("MFMA vs scalar", relerr(vals[1], vals[0]), "vs oracle", relerr(vals[1], val0), flush=True)
    assert relerr(vals[1], vals[0]) <= 1e-13
    assert relerr(vals[1], val0) <= 1e-12
    assert not np.array_equal(vals[1], vals[0])  # a different summation: the MFMA kernel ran
    S.close()
