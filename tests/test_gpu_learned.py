"""Learned IC(0)-pattern preconditioner on the GPU (fe_learned.hip) against the
CPU oracle (oracle/fe_oracle_mpnn.py, pinned by finite differences, FSAI
optimality and exact cases in tests/test_oracle_mpnn.py).

Tolerances: the device stores activations in bfloat16 exactly where the oracle
rounds them, but accumulates in FP32 in another order, so an occasional
activation lands one bfloat16 step (2^-8 relative) away; norm-wise relative
errors of 1e-2 cover that.  The factor, row products and loss are FP32 on the
device (FP64 accumulation of the row products and the loss sums).
"""
import numpy as np
import pytest

from oracle import fe_oracle_mpnn as mp

pytestmark = pytest.mark.gpu


def nrel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _system(N=10, seed=5):
    from feng_amd.learned import cube_instance
    return cube_instance(N, seed, nfeat=64)


def _theta(seed=7):
    rng = np.random.default_rng(seed)
    d = mp.unpack(mp.init_params(seed))
    d["P"] *= 20.0
    d["Q"] *= 20.0
    d["Hs"] = rng.normal(0, 0.05, d["Hs"].shape)
    d["hb"] = np.array([0.1, 0.02, -0.03, 0.3, 0.2])
    return mp.pack(d).astype(np.float32)


def test_init_matches_oracle():
    m, S, num = _system(4, 1)
    S.mpnn_init(123)
    assert np.array_equal(S.mpnn_get_params(), mp.init_params(123).astype(np.float32))
    assert S.L.fe_hip_mpnn_param_count() == mp.param_count()
    S.close()


def test_forward_loss_and_gradient_match_oracle(oracle):
    m, S, num = _system(10, 5)
    th = _theta()
    S.mpnn_set_params(th)
    loss, grad = S.mpnn_loss_grad()
    ia, ja = S.get_csr()
    val, _ = S.get_system()
    lu = S.get_factor()
    # the learned path starts from three parallel IC(0) sweeps, close to the
    # exact IC(0) factor of the oracle
    lu_o, rc = oracle.ilu0(ia, ja, val)
    assert rc == 0 and nrel(lu, lu_o) < 2e-2
    g = mp.Graph(ia, ja, val, lu)
    L_o, grad_o, f = mp.loss_and_grad(g, th.astype(np.float64))
    n = num.nInc
    low = g.lower
    assert nrel(S.mpnn_tensor("w0")[low], g.w0[low]) < 1e-5
    assert nrel(S.mpnn_tensor("X").reshape(n, 16), f["X"]) < 1e-6
    for name in ("h0", "m1", "h1", "m2", "h2", "m3", "h3"):
        assert nrel(S.mpnn_tensor(name).reshape(n, 32), f[name]) < 1e-2, name
    for name, shape in (("p", (n, 8)), ("q", (n, 8)), ("heads", (n, 5)), ("B", (n,)), ("d", (n,))):
        assert nrel(S.mpnn_tensor(name).reshape(shape), f[name]) < 1e-2, name
    assert nrel(S.mpnn_tensor("g")[low], f["g"][low]) < 1e-2
    assert abs(loss - L_o) <= 1e-3 * abs(L_o) + 1e-6
    cos = float(grad @ grad_o / (np.linalg.norm(grad) * np.linalg.norm(grad_o)))
    assert cos > 0.999 and nrel(grad, grad_o) < 3e-2
    o = 0
    for name, shape in mp.LAYOUT:
        c = int(np.prod(shape))
        go = grad_o[o:o + c]
        if np.linalg.norm(go) > 1e-3 * np.linalg.norm(grad_o):
            assert nrel(grad[o:o + c], go) < 6e-2, name
        o += c
    S.close()


def test_row_products_match_fsai_condition(oracle):
    # with the net's heads zero, B and w are the IC(0) row products
    m, S, num = _system(8, 3)
    d = mp.unpack(mp.init_params(3))
    d["P"][:] = 0.0
    S.mpnn_set_params(mp.pack(d).astype(np.float32))
    S.mpnn_loss_grad(want_grad=False)
    ia, ja = S.get_csr()
    val, _ = S.get_system()
    g = mp.Graph(ia, ja, val, S.get_factor())
    B0, w0 = mp.row_products(g, g.g0)
    assert nrel(S.mpnn_tensor("B"), B0) < 1e-6
    assert nrel(S.mpnn_tensor("w")[g.lower], w0[g.lower]) < 1e-5
    S.close()


def test_training_lowers_the_loss_and_the_iterations():
    from feng_amd.api import PC_IC0, PC_LEARNED
    from feng_amd.learned import train
    theta, hist = train(N=10, instances=60, seed0=300, lr=3e-3, nfeat=64, log_every=0)
    assert np.mean(hist[-10:]) < np.mean(hist[:3])
    # held-out instance
    m, S, num = _system(12, 777)
    r_ic = S.solve(pc=PC_IC0, rtol=1e-8)
    S.mpnn_set_params(theta)
    r_l = S.solve(pc=PC_LEARNED, rtol=1e-8)
    assert r_l.converged and r_l.relres <= 1e-8
    assert r_l.iterations < r_ic.iterations, (r_l.iterations, r_ic.iterations)
    S.close()


def test_learned_pcg_converges_in_fp64_and_fp32():
    from feng_amd.api import PC_IC0, PC_LEARNED, PREC_FP32_IR
    m, S, num = _system(12, 9)
    d = mp.unpack(mp.init_params(3))
    d["P"][:] = 0.0
    S.mpnn_set_params(mp.pack(d).astype(np.float32))
    r0 = S.solve(pc=PC_LEARNED, rtol=1e-8)
    assert r0.converged and r0.relres <= 1e-8
    S.mpnn_set_params(_theta(4))
    r1 = S.solve(pc=PC_LEARNED, rtol=1e-8)
    assert r1.converged and r1.relres <= 1e-8
    S.setToZero()
    S.assemble(3, PREC_FP32_IR)
    r2 = S.solve(pc=PC_LEARNED, precision=PREC_FP32_IR, rtol=1e-8)
    assert r2.converged and r2.relres <= 1e-8
    r3 = S.solve(pc=PC_IC0, rtol=1e-8)
    assert r3.converged
    S.close()


def test_learned_without_weights_is_an_error():
    from feng_amd.api import PC_LEARNED, FeError
    m, S, num = _system(4, 1)
    with pytest.raises(FeError, match="no learned preconditioner weights"):
        S.solve(pc=PC_LEARNED)
    S.close()
