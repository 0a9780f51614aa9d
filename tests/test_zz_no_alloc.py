"""fe_hip_norm_noalloc (the allocation-free variant of fe_hip_norm) against
fe_hip_norm through the C ABI.

Both entry points run the same element and folding kernels in the same
order, so their values must be bitwise equal.  Two calls in a row, the second
on a smaller mesh (a second context on the same GPU, so the second call finds
the first call's contents in the shared scratch), then the first context
again; an exact-solution table beyond the scratch is rejected before any
launch.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(N, kind, seed):
    from feng_amd.api import FN_CONST, Fn, Mesh
    from feng_amd.problems import diffusion_system
    m = Mesh.square(N, 0.2, seed)
    S, num = diffusion_system(m, kind, 2, Fn(FN_CONST, 1.0), Fn(FN_CONST, -1.0))
    rng = np.random.default_rng(seed)
    S.set_solution(rng.standard_normal(num.nDofs))
    return m, S, num


def _both(S, *args):
    a = S.norm(*args)
    b = S.norm(*args, noalloc=True)
    return a, b


def test_noalloc_norm_equals_norm_on_two_calls():
    from feng_amd.api import (FN_ELEM, FN_SINSIN, NORM_H1_ERROR, NORM_L2_ERROR, NORM_SEMI_H1, Fn)
    print("noalloc norm: 32^2 P2, then 12^2 P1", flush=True)
    m1, S1, _ = _system(32, 2, 1)
    ex = np.random.default_rng(7).standard_normal(m1.E)
    a, b = _both(S1, NORM_L2_ERROR, 0, 4, [Fn(FN_ELEM, data=ex)])
    assert a == b and a > 0.0
    a1, b1 = _both(S1, NORM_SEMI_H1, 0, 4)
    assert a1 == b1 and a1 > 0.0
    # second call: a smaller mesh on another context of the same GPU
    m2, S2, _ = _system(12, 1, 2)
    c, d = _both(S2, NORM_L2_ERROR, 0, 2, [Fn(FN_SINSIN, 1.0)])
    assert c == d and c > 0.0
    grad = [Fn(FN_SINSIN, 1.0), Fn(FN_SINSIN, 0.5), Fn(FN_SINSIN, 0.25)]
    e, f = _both(S2, NORM_H1_ERROR, 0, 3, grad)
    assert e == f and e > 0.0
    # the first context again, after the smaller call overwrote the scratch
    assert S1.norm(NORM_L2_ERROR, 0, 4, [Fn(FN_ELEM, data=ex)], noalloc=True) == a
    assert S1.norm(NORM_SEMI_H1, 0, 4, noalloc=True) == a1
    S2.close()
    S1.close()


def test_noalloc_norm_rejects_tables_beyond_its_scratch():
    from feng_amd.api import FN_QP, NORM_L2_ERROR, FeError, Fn
    print("noalloc norm: oversized exact table", flush=True)
    m, S, _ = _system(128, 1, 3)
    nq = 33  # the degree-12 triangle rule
    big = np.zeros(m.E * nq)
    assert big.size > 1 << 20
    with pytest.raises(FeError, match="exceed"):
        S.norm(NORM_L2_ERROR, 0, 12, [Fn(FN_QP, data=big)], noalloc=True)
    # the allocating entry point takes the same table
    assert math.isfinite(S.norm(NORM_L2_ERROR, 0, 12, [Fn(FN_QP, data=big)]))
    S.close()
