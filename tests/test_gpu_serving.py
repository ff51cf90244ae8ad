"""Serving loop: solves of different degrees and sizes advanced together
(one batched launch per LSRK stage and degree) give the results they give
alone (AdvectionOperation + LSRK(5,4) with the same configuration), also
when the memory budget holds only some of them at a time."""
import numpy as np
import pytest
import torch

from advection_miniapp_amd import (AdvectionOperation,
                                   LowStorageRungeKuttaIntegrator)
from advection_miniapp_amd import _lib as L
from advection_miniapp_amd.serving import Server, request_config, \
    serve_request

pytestmark = pytest.mark.gpu


def requests():
    rng = np.random.default_rng(21)
    out = []
    for k, N in ((2, 8), (3, 8), (3, 16), (4, 8), (5, 8), (6, 8), (4, 3),
                 (1, 8), (7, 3)):
        a = rng.standard_normal(3)
        out.append(serve_request(k, N, a / np.linalg.norm(a),
                                 rng.uniform(0.2, 0.8, 3),
                                 rng.uniform(0.05, 0.2), final_time=0.05,
                                 courant=0.3))
    return out


def solo(r):
    cfg, dt, n = request_config(r)
    op = AdvectionOperation(cfg)
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    r1, r2 = op.initialize_dof_vector(), op.initialize_dof_vector()
    integ = LowStorageRungeKuttaIntegrator()
    for s in range(n):
        integ.perform_time_step(op, s * dt, dt, u, r1, r2)
    torch.cuda.synchronize()
    return u


def run_server(budget, reqs):
    srv = Server(budget)
    ids = [srv.submit(r) for r in reqs]
    max_active, calls, finished = 0, 0, 0
    while True:
        # a few steps per call: the call returns early at a completion
        a, q, f = srv.step(max_steps=7)
        finished += f
        max_active = max(max_active, a + f)
        calls += 1
        if a == 0 and q == 0:
            break
        assert calls < 10000
    assert finished == len(reqs)
    out = []
    for i in ids:
        assert srv.info(i)["state"] == L.DG_REQUEST_DONE
        out.append(srv.take_result(i))
    assert srv.memory() == 0
    return out, max_active


def test_batched_results_equal_solo_results():
    reqs = requests()
    out, max_active = run_server(1 << 34, reqs)
    assert max_active == len(reqs)
    for r, u in zip(reqs, out):
        ref = solo(r)
        err = ((u - ref).abs().max() / ref.abs().max()).item()
        assert err < 1e-14, (r.degree, r.cells_per_direction, err)


def test_budget_limits_concurrency_and_keeps_results():
    reqs = requests()
    # room for about two of the larger solves at a time
    cfg, _, _ = request_config(reqs[2])
    big = 3 * 8 * 16 ** 3 * 4 ** 3 + 16 ** 3 * 216 + 40000
    out, max_active = run_server(2 * big, reqs)
    assert 1 <= max_active < len(reqs)
    for r, u in zip(reqs, out):
        ref = solo(r)
        err = ((u - ref).abs().max() / ref.abs().max()).item()
        assert err < 1e-14, (r.degree, r.cells_per_direction, err)


def test_oversized_request_is_rejected():
    srv = Server(1 << 20)
    r = serve_request(4, 16, [1, 0, 0], [0.5] * 3, 0.1)
    with pytest.raises(L.DgError):
        srv.submit(r)
