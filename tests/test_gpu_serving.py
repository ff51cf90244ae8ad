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


def config5_stream(n_req, final_time, seed=5):
    """Requests of the serving workload: Poisson arrivals (mean 200 per
    second), degree uniform in 2..6, cells per direction from {8, 16, 24,
    32, 48}, unit speed uniform on the sphere, Gaussian with a random
    centre and width in [0.05, 0.2]; a short end time bounds the test."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, n_req))
    reqs = []
    for _ in range(n_req):
        a = rng.standard_normal(3)
        reqs.append(serve_request(int(rng.integers(2, 7)),
                                  int(rng.choice([8, 16, 24, 32, 48])),
                                  a / np.linalg.norm(a),
                                  rng.uniform(0.2, 0.8, 3),
                                  rng.uniform(0.05, 0.2),
                                  final_time=final_time, courant=0.3))
    return arrivals, reqs


def oracle_lsrk(r):
    """The same solve on the CPU oracle: projection and LSRK(5,4) steps."""
    import oracle
    from helpers import Cfg
    cfg, dt, n = request_config(r)
    c = Cfg()
    for name, _ in oracle.OrProblem._fields_:
        v = getattr(cfg, name)
        setattr(c, name, list(v) if not isinstance(v, (int, float)) else v)
    o = c.oracle_handle()
    u = o.project(0.0)
    r1, r2 = o.zeros(), o.zeros()
    for s in range(n):
        o.lsrk_time_step(1, s * dt, dt, u, r1, r2)
    return u


def test_config5_stream_with_queueing():
    """A seeded stream of the serving workload at its own sizes (up to
    48^3 cells at degree 6, 37.9 M DoF), submitted on a virtual clock while
    earlier requests run, under a budget that holds only part of it: every
    result equals the solve run alone to 1e-14, the small ones equal the
    oracle's LSRK run to 1e-12, and the memory in use returns to zero."""
    arrivals, reqs = config5_stream(14, final_time=0.01)
    assert max(r.cells_per_direction for r in reqs) == 48
    assert {24, 32}.issubset({r.cells_per_direction for r in reqs})
    # 1.2 GiB: the degree-6 48^3 solve (0.85 GiB of vectors) and a
    # degree-4 48^3 one (0.31 GiB) do not fit together with the rest
    srv = Server(int(1.2 * (1 << 30)))
    # virtual clock: one server step call advances it by 2 ms
    ids, nxt, clock, calls, max_active, queued_seen = {}, 0, 0.0, 0, 0, 0
    while True:
        while nxt < len(reqs) and arrivals[nxt] <= clock:
            ids[srv.submit(reqs[nxt])] = nxt
            nxt += 1
        a, q, f = srv.step(max_steps=3)
        max_active = max(max_active, a + f)
        queued_seen = max(queued_seen, q)
        clock += 0.002
        calls += 1
        if a == 0 and q == 0 and nxt == len(reqs):
            break
        assert calls < 20000
    assert queued_seen > 0 and max_active < len(reqs)
    results = {}
    for rid, j in ids.items():
        assert srv.info(rid)["state"] == L.DG_REQUEST_DONE
        results[j] = srv.take_result(rid)
    assert srv.memory() == 0
    for j, r in enumerate(reqs):
        ref = solo(r)
        u = results[j]
        err = ((u - ref).abs().max() / ref.abs().max()).item()
        assert err < 1e-14, (j, r.degree, r.cells_per_direction, err)
        if r.cells_per_direction <= 8:
            o = oracle_lsrk(r)
            d = np.max(np.abs(u.cpu().numpy() - o)) / np.max(np.abs(o))
            assert d < 1e-12, (j, d)
        del ref
