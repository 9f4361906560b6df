"""The parity cases are well posed: perturbing the trust-region start by 1e-5 px moves the
oracle's converged displacement by far less than the 0.02 px parity tolerance. (A case that
fails this is chaotic under the trust region, and no two implementations could be held to
0.02 px on it.)"""
import numpy as np
import pytest

import oracle as O
from helpers import MANY_NODE_CASES, PARITY_CASES, small_stack, truth_start


@pytest.mark.parametrize("name,shape,nD,scan,gen", PARITY_CASES, ids=[c[0] for c in PARITY_CASES])
def test_parity_case_is_well_posed(name, shape, nD, scan, gen):
    st = small_stack(shape, nD=nD, seed=3, **gen)
    o = O.Oracle().set_reference(st["ref_w"], nD, **scan)
    rng = np.random.default_rng(0)
    n = nD * o.n_dim
    for f in range(len(st["frames"])):
        x_ref, _ = o.align(st["frames_w"][f], rel_delta=1e-4, max_iter=50, skip_repeat=1)
        for _ in range(3):
            x, _ = o.align(st["frames_w"][f], x0=rng.normal(0, 1e-5, n), rel_delta=1e-4,
                           max_iter=50, skip_repeat=1)
            assert np.max(np.abs(x - x_ref)) < 0.005


@pytest.mark.parametrize("name,shape,nD,scan,gen", MANY_NODE_CASES, ids=[c[0] for c in MANY_NODE_CASES])
def test_many_node_case_is_well_posed(name, shape, nD, scan, gen):
    st = small_stack(shape, nD=nD, seed=3, **gen)
    o = O.Oracle().set_reference(st["ref_w"], nD, **scan)
    rng = np.random.default_rng(0)
    n = nD * o.n_dim
    for f in range(len(st["frames"])):
        x0 = truth_start(st, f, o.n_dim)
        x_ref, _ = o.align(st["frames_w"][f], x0=x0, rel_delta=1e-4, max_iter=50, skip_repeat=1)
        for _ in range(3):
            x, _ = o.align(st["frames_w"][f], x0=x0 + rng.normal(0, 1e-5, n), rel_delta=1e-4,
                           max_iter=50, skip_repeat=1)
            assert np.max(np.abs(x - x_ref)) < 0.005
