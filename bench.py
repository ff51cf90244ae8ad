"""Headline benchmark on one MI355X.

Metric: DoF/s and fraction of the HBM roofline of the 3D, degree-4, FP64 DG
operator apply; and the time of the mixed-precision FDM-preconditioned IRK
solve to a 1e-10 relative residual against the all-FP64 solve.

Apply workload (the metric's configuration): 64^3 cells on [0,1]^3 with
interior vertices perturbed by up to 10% of h (seed 42), degree 4 on
Gauss-Lobatto nodes (32,768,000 DoF), periodic, divergence-free velocity
(sin2πy cos2πz, sin2πz cos2πx, sin2πx cos2πy), upwind flux, Gaussian pulse
initial data of width 0.1. One step = one operator apply
dst = M^-1(-L src) over the whole mesh, src/dst resident in HBM.

IRK workload: 32^3 cells on [0,1]^3, degree 4, the same velocity field,
Radau IIA with 3 stages (12,288,000 unknowns), dt = 5 h_min / 2 (Courant
number 5, advection_solver_irk_real.cc:2352), flexible GMRES(30) in FP64 to
a 1e-10 relative residual from a zero initial guess, cellwise FDM
preconditioner: three algorithms (FDM alone, inner GMRES of 4 or 8 steps
around the FDM), each in FP64 and with the preconditioner in FP32; the
speedup is the fastest all-FP64 solve over the fastest mixed one, with the
same-algorithm ratios beside it.

Also by default: the apply workload at degrees 1..7 with the autotuned
tile shapes, configuration 1 (2D run, L2 error), configuration 3
(SIPG, degrees 3/5/7, FP64 and FP32), the Burgers add-on of configuration 4,
the serving stream (configuration 5) against serial solves, and the CPU
baseline (the oracle on the host cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-irk]
                    [--no-configs] [--no-serving] [--extra]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6
# one 1D sweep over n^d values costs 2 n^(d+1) flop; the apply runs 9 cell
# sweeps and 6 face interpolations/integrations of 2 (d-1)-dimensional
# sweeps each, element-centric
APPLY_FLOP_PER_DOF = (9 * 2 * 5 ** 4 + 6 * 3 * 2 * 5 ** 3) / 5 ** 3


def apply_workload(refine=6, degree=4, dim=3):
    return dict(dim=dim, degree=degree, refine=refine, perturbation=0.1,
                seed=42, velocity_kind=1, flux_alpha=1.0, factor_skew=0.0,
                solution_kind=3, solution_param=[0.5, 0.5, 0.5, 0.1],
                periodic=[1, 1, 1], lo=[0, 0, 0], hi=[1, 1, 1])


def irk_workload(refine=5):
    return dict(dim=3, degree=4, refine=refine, perturbation=0.0,
                velocity_kind=1, flux_alpha=1.0, factor_skew=0.5,
                irk_stages=3, solution_kind=3,
                solution_param=[0.5, 0.5, 0.5, 0.1],
                periodic=[1, 1, 1], lo=[0, 0, 0], hi=[1, 1, 1])


def _oracle_apply_rate(threads, refine, target_seconds, min_reps):
    import oracle
    from tests.helpers import Cfg
    oracle.load().or_set_threads(threads)
    o = Cfg(**apply_workload(refine=refine)).oracle_handle()
    u = o.project(0.0)
    reps, t0 = 0, time.perf_counter()
    while True:
        o.apply(u, 0.0)
        reps += 1
        el = time.perf_counter() - t0
        if (el > target_seconds and reps >= min_reps) or reps >= 2000:
            break
    n = o.n_dofs
    del o, u
    return n * reps / el, reps, el, n


def cpu_baseline(target_seconds=10.0):
    """The oracle's apply (restated reference algorithm, OpenMP over cells)
    on the host cores of this GPU's box: all of this process's cores, at
    most the 16 a one-GPU box allots, on the full apply workload (64^3
    cells); and one core on a 32^3-cell sample of the same workload."""
    sys.path.insert(0, ROOT)
    cores = min(len(os.sched_getaffinity(0)), 16)
    v_all, reps, el, n = _oracle_apply_rate(cores, 6, target_seconds, 2)
    v_one, reps1, el1, n1 = _oracle_apply_rate(1, 5, target_seconds, 1)
    return dict(value=v_all, unit="DoF/s", cores=cores, kind="port",
                sample=f"{reps} applies of the full workload (64^3 cells, "
                       f"degree 4, {n} DoF) in {el:.1f} s on {cores} cores; "
                       "restated reference algorithm (oracle/oracle.cpp, "
                       "OpenMP over cells), not deal.II",
                one_core=dict(value=v_one, unit="DoF/s", cores=1,
                              sample=f"{reps1} applies on 32^3 cells of the "
                                     f"same workload ({n1} DoF) in "
                                     f"{el1:.1f} s"))


def measured_traffic():
    """HBM bytes per apply launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated),
    or None when no summary exists."""
    path = os.path.join(ROOT, "profiles", "apply_3d_k4_64cube_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def uniform_apply_workload(refine=6, degree=4):
    """The apply workload's mesh size and degree on axis-aligned cells with
    a constant speed (1, 0.5, 0.25): the operator kernel without geometry
    tables, for profiling against the tabulated one."""
    return dict(apply_workload(refine=refine, degree=degree),
                perturbation=0.0, velocity_kind=0, velocity=[1.0, 0.5, 0.25])


def time_apply(args):
    import torch
    from advection_miniapp_amd import AdvectionOperation, make_config
    wl = (uniform_apply_workload if args.apply_geometry == "uniform"
          else apply_workload)
    op = AdvectionOperation(make_config(**wl(refine=args.refine)))
    src = op.initialize_dof_vector()
    op.project_initial(src, 0.0)
    dst = op.initialize_dof_vector()
    for _ in range(max(1, args.warmup)):
        op.apply(src, dst, 0.0)
    torch.cuda.synchronize()
    # the library launches on torch's current stream (op._stream), so
    # events recorded there bracket exactly the K operator kernels
    stream = torch.cuda.current_stream()
    assert stream.cuda_stream == op._stream
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        op.apply(src, dst, 0.0)
    stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    op.tile = op.tile_shape()
    return op, ms


def time_irk(args):
    import torch
    from advection_miniapp_amd import AdvectionOperation, make_config
    from advection_miniapp_amd import _lib as L
    from advection_miniapp_amd.operators import solver_control
    cfg = irk_workload(refine=args.irk_refine)
    op = AdvectionOperation(make_config(**cfg))
    s = 3
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    h = 1.0 / 2 ** args.irk_refine
    dt = 5.0 * h / 2.0
    op.set_time(0.0, dt)
    rhs = op.initialize_dof_vector(blocks=s)
    op.compute_rhs(rhs, u)
    rn = rhs.norm().item()
    # each algorithm (FDM-preconditioned FGMRES, and FGMRES around an inner
    # GMRES of 4 or 8 steps with the FDM) in FP64 and with the
    # preconditioner (and the inner solver) in FP32
    variants = {
        "fp64": dict(prec_precision=L.DG_FP64),
        "fp32_fdm": dict(prec_precision=L.DG_FP32),
        "fp64_inner4": dict(prec_precision=L.DG_FP64, inner_iterations=4),
        "fp32_inner4": dict(prec_precision=L.DG_FP32, inner_iterations=4),
        "fp64_inner8": dict(prec_precision=L.DG_FP64, inner_iterations=8),
        "fp32_inner8": dict(prec_precision=L.DG_FP32, inner_iterations=8),
    }
    if args.irk_variants:
        names = args.irk_variants.split(",")
        unknown = [n for n in names if n not in variants]
        if unknown:
            raise SystemExit(f"unknown IRK variants: {unknown}")
        variants = {n: variants[n] for n in names}
    out, sol = {}, {}
    for name, kw in variants.items():
        times, z = [], None
        for rep in range(1 + args.irk_reps):
            z = op.initialize_dof_vector(blocks=s)
            ctl = solver_control(max_iterations=400, rel_tolerance=1e-10,
                                 restart=30, **kw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            op.irk_solve(z, rhs, ctl)
            torch.cuda.synchronize()
            if rep > 0:      # the first solve builds tables and buffers
                times.append(time.perf_counter() - t0)
        res = op.initialize_dof_vector(blocks=s)
        op.vmult_irk(res, z)
        res -= rhs
        times.sort()
        out[name] = dict(ms=1e3 * times[len(times) // 2],
                         iterations=ctl.iterations,
                         rel_residual=res.norm().item() / rn)
        sol[name] = z
    if "fp64" in sol:
        ref = sol["fp64"]
        for name in variants:
            out[name]["rel_diff_vs_fp64"] = \
                ((sol[name] - ref).norm() / ref.norm()).item()
    # headline: the fastest all-FP64 solve against the fastest solve with
    # an FP32 preconditioner; beside it the ratio of each algorithm run in
    # both precisions
    head = dict(workload=f"3D IRK Radau IIA(3), degree 4, "
                         f"{2 ** args.irk_refine}^3 cells, Courant 5, "
                         "FGMRES(30) FP64 to 1e-10 from zero",
                unknowns=op.n_dofs * s)
    fp64_names = [n for n in variants if n.startswith("fp64")]
    fp32_names = [n for n in variants if n.startswith("fp32")]
    if fp64_names and fp32_names:
        best64 = min(fp64_names, key=lambda n: out[n]["ms"])
        best32 = min(fp32_names, key=lambda n: out[n]["ms"])
        head.update(fp64_best=best64, mixed=best32,
                    speedup=out[best64]["ms"] / out[best32]["ms"])
    head["speedup_same_algorithm"] = {
        alg: out[a]["ms"] / out[b]["ms"] for alg, a, b in
        (("fdm", "fp64", "fp32_fdm"),
         ("inner4", "fp64_inner4", "fp32_inner4"),
         ("inner8", "fp64_inner8", "fp32_inner8"))
        if a in out and b in out}
    return dict(head, **out)


def serving_requests(n_req, final_time, seed=2024, rate=200.0):
    """The serving workload: Poisson arrivals (mean `rate` per second),
    degree uniform in 2..6, cells per direction from {8, 16, 24, 32, 48},
    unit velocity uniform on the sphere, Gaussian with random centre and
    width uniform in [0.05, 0.2], end time `final_time`."""
    import numpy as np
    from advection_miniapp_amd.serving import serve_request
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
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


def serving_cpu_baseline(reqs, target_seconds=10.0):
    """The oracle (restated reference algorithm, OpenMP over cells) on the
    serving stream, on this box's host cores (at most the 16 of one GPU):
    in arrival order, each request is set up and advanced by one LSRK(5,4)
    step until about `target_seconds` of stepping time; the rate counts the
    DoF of the 5 stage evaluations per step, as the GPU figures do."""
    sys.path.insert(0, ROOT)
    import oracle
    from advection_miniapp_amd.serving import request_config
    from tests.helpers import Cfg
    cores = min(len(os.sched_getaffinity(0)), 16)
    oracle.load().or_set_threads(cores)
    work, el, used = 0.0, 0.0, 0
    for r in reqs:
        cfg, dt, _ = request_config(r)
        c = Cfg()
        for name, _ in oracle.OrProblem._fields_:
            v = getattr(cfg, name)
            setattr(c, name, list(v) if not isinstance(v, (int, float)) else v)
        o = c.oracle_handle()
        u = o.project(0.0)
        r1, r2 = o.zeros(), o.zeros()
        t0 = time.perf_counter()
        o.lsrk_time_step(1, 0.0, dt, u, r1, r2)
        el += time.perf_counter() - t0
        work += 5.0 * o.n_dofs
        used += 1
        del o, u, r1, r2
        if el > target_seconds:
            break
    return dict(value=work / el, unit="DoF/s", cores=cores, kind="port",
                sample=f"one LSRK(5,4) step of each of the first {used} "
                       f"requests in arrival order ({el:.1f} s of "
                       "stepping; set-up not timed)")


def time_serving(args):
    """Continuous batching (Server) against serial one-at-a-time solves of
    the same arrival stream; latency = completion - arrival. With the CPU
    leg on, the oracle's rate on the same stream is timed beside them."""
    import numpy as np
    import torch
    from advection_miniapp_amd import (AdvectionOperation,
                                       LowStorageRungeKuttaIntegrator)
    from advection_miniapp_amd import _lib as L
    from advection_miniapp_amd.serving import Server, request_config
    arrivals, reqs = serving_requests(args.serve_requests,
                                      args.serve_final_time)
    work = []
    for r in reqs:
        cfg, dt, n = request_config(r)
        nd = (r.cells_per_direction * (r.degree + 1)) ** 3
        work.append(nd * n)
    # batched server in wall-clock time
    srv = Server(int(args.serve_budget_gb * (1 << 30)))
    ids, done = {}, {}
    nxt, t0 = 0, time.perf_counter()
    while len(done) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and arrivals[nxt] <= now:
            ids[srv.submit(reqs[nxt])] = nxt
            nxt += 1
        if len(ids) == len(done):
            time.sleep(max(0.0, arrivals[nxt] - now))
            continue
        _, _, fin = srv.step(max_steps=args.serve_steps_per_call)
        now = time.perf_counter() - t0
        if fin:
            for rid, j in ids.items():
                if j not in done and \
                        srv.info(rid)["state"] == L.DG_REQUEST_DONE:
                    done[j] = now
                    srv.take_result(rid)
    lat_b = np.array([done[j] - arrivals[j] for j in range(len(reqs))])
    span_b = max(done.values()) - arrivals[0]
    del srv
    # serial: one solve at a time, in arrival order
    integ = LowStorageRungeKuttaIntegrator()
    finish, lat_s = 0.0, []
    for j, r in enumerate(reqs):
        # setup (mesh, tables, vectors) counts as in the server's admission
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        cfg, dt, n = request_config(r)
        op = AdvectionOperation(cfg)
        u = op.initialize_dof_vector()
        r1, r2 = op.initialize_dof_vector(), op.initialize_dof_vector()
        op.project_initial(u, 0.0)
        for s in range(n):
            integ.perform_time_step(op, s * dt, dt, u, r1, r2)
        torch.cuda.synchronize()
        dur = time.perf_counter() - t1
        finish = max(finish, arrivals[j]) + dur
        lat_s.append(finish - arrivals[j])
        del op, u, r1, r2
    lat_s = np.array(lat_s)
    span_s = finish - arrivals[0]
    tot = float(sum(work))
    extra = {}
    if not args.no_cpu:
        extra["cpu_baseline"] = serving_cpu_baseline(reqs)
    return dict(
        **extra,
        workload=f"{len(reqs)} requests, Poisson 200/s (seed 2024), degree "
                 "2..6, 8..48 cells per direction, T = "
                 f"{args.serve_final_time}, LSRK(5,4) at Courant 0.3",
        dof_steps=tot,
        batched=dict(p50_s=float(np.percentile(lat_b, 50)),
                     p99_s=float(np.percentile(lat_b, 99)),
                     makespan_s=span_b, dof_per_s=5 * tot / span_b),
        serial=dict(p50_s=float(np.percentile(lat_s, 50)),
                    p99_s=float(np.percentile(lat_s, 99)),
                    makespan_s=span_s, dof_per_s=5 * tot / span_s))


def time_one_apply(cfg_kw, steps=20, src=None):
    """Average apply time (HIP events on the launch stream) of one
    operator configuration; returns (operator, src, dst, ms)."""
    import torch
    from advection_miniapp_amd import AdvectionOperation, make_config
    op = AdvectionOperation(make_config(**cfg_kw))
    if src is None:
        src = op.initialize_dof_vector()
        op.project_initial(src, 0.0)
    dst = op.initialize_dof_vector()
    for _ in range(3):
        op.apply(src, dst, 0.0)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        op.apply(src, dst, 0.0)
    stop.record(stream)
    torch.cuda.synchronize()
    return op, src, dst, start.elapsed_time(stop) / steps


def _events():
    import torch
    return (torch.cuda.Event(enable_timing=True),
            torch.cuda.Event(enable_timing=True))


def time_apply_by_degree():
    """The configuration-2 apply (perturbed cells, cellular speed, FP64)
    at degrees 1..7: 64^3 cells up to degree 4, 32^3 above."""
    import torch
    out = {}
    for k in range(1, 8):
        r = 6 if k <= 4 else 5
        op, src, dst, ms = time_one_apply(apply_workload(refine=r, degree=k))
        n = op.n_dofs
        out[f"k{k}"] = dict(n_dofs=n, ms=ms, dof_per_s=n / (ms * 1e-3),
                            roofline_frac=16 * n / (ms * 1e-3) / 8e12,
                            tile=op.tile_shape())
        del op, src, dst
        torch.cuda.empty_cache()
    return out


def time_config2_run(n_steps=200, courant=0.3):
    """Configuration 2 as a run: 200 LSRK(5,4) time steps at Courant 0.3
    with fused stages (dg_lsrk_time_step, three registers); per stage 32
    B/DoF of vector traffic (read r and u, write u and r')."""
    import torch
    from advection_miniapp_amd import (AdvectionOperation, make_config,
                                       explicit_time_step)
    from advection_miniapp_amd import _lib as L
    op = AdvectionOperation(make_config(**apply_workload()))
    # |a| <= sqrt(3) for the cellular field
    dt = explicit_time_step(op.min_vertex_distance, 3 ** 0.5, 4, courant,
                            1e9)
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    ri, ki = op.initialize_dof_vector(), op.initialize_dof_vector()
    m0 = op.compute_mass_and_energy(u)
    op._lsrk_time_step(L.DG_LSRK_5_4, 0.0, dt, u, ri, ki)
    torch.cuda.synchronize()
    start, stop = _events()
    start.record()
    for s in range(1, n_steps):
        op._lsrk_time_step(L.DG_LSRK_5_4, s * dt, dt, u, ri, ki)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    steps = n_steps - 1
    stage_ms = ms / (5 * steps)
    m1 = op.compute_mass_and_energy(u)
    n = op.n_dofs
    return dict(
        workload="configuration 2 run: 200 LSRK(5,4) steps, Courant 0.3, "
                 "fused stages (199 timed after one warm-up step)",
        time_step=dt, ms_per_step=ms / steps, ms_per_stage=stage_ms,
        dof_per_s_per_stage=n / (stage_ms * 1e-3),
        stage_gb_per_s=32 * n / (stage_ms * 1e-3) / 1e9,
        stage_roofline_frac=32 * n / (stage_ms * 1e-3) / 8e12,
        run_seconds_200_steps=ms / steps * n_steps / 1e3,
        mass_drift=abs(m1[0] - m0[0]) / abs(m0[0]))


def time_inverse_mass(steps=20):
    """Standalone cellwise inverse mass matrix on configuration 2 (in
    place, 16 B/DoF)."""
    import torch
    from advection_miniapp_amd import AdvectionOperation, make_config
    op = AdvectionOperation(make_config(**apply_workload()))
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    for _ in range(3):
        op.apply_inverse_mass_matrix(u)
    torch.cuda.synchronize()
    start, stop = _events()
    start.record()
    for _ in range(steps):
        op.apply_inverse_mass_matrix(u)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    n = op.n_dofs
    return dict(ms=ms, dof_per_s=n / (ms * 1e-3),
                gb_per_s=16 * n / (ms * 1e-3) / 1e9)


def time_config1_run():
    """Configuration 1 end to end on the GPU: 2D, degree 3, 32^2 cells on
    [-1,1]^2, velocity (1, 0.15), u0 = sin(pi x) sin(pi y), LSRK(5,4) to
    t = 2 at Courant 0.4 (421 steps), relative L2 error at t = 2. Times
    (projection and all steps) of a warm run with one call per step and of
    a warm run with all steps in one call."""
    import math
    import torch
    from advection_miniapp_amd import (AdvectionOperation, make_config,
                                       explicit_time_step)
    from advection_miniapp_amd import _lib as L
    op = AdvectionOperation(make_config(
        dim=2, degree=3, refine=5, lo=[-1, -1, 0], hi=[1, 1, 1],
        velocity=[1.0, 0.15, 0.0], solution_kind=2,
        solution_param=[math.pi, 0.0], periodic=[1, 1, 1]))
    dt = explicit_time_step(op.min_vertex_distance, math.hypot(1.0, 0.15),
                            3, 0.4, 2.0)
    n = int(round(2.0 / dt))
    u = op.initialize_dof_vector()
    ri, ki = op.initialize_dof_vector(), op.initialize_dof_vector()
    start, stop = _events()
    # two runs with one call per step; the first loads the kernels' code
    # objects and builds the tables, the second is timed
    for _ in range(2):
        torch.cuda.synchronize()
        start.record()
        op.project_initial(u, 0.0)
        for s in range(n):
            op._lsrk_time_step(L.DG_LSRK_5_4, s * dt, dt, u, ri, ki)
        stop.record()
        torch.cuda.synchronize()
    err, nrm = op.l2_error(u, 2.0)
    seconds = start.elapsed_time(stop) / 1e3
    # the same run with the steps in one call: the operator is autonomous,
    # so the first call captures a HIP graph of 8 steps, times stream
    # launches against its replays on its first 24 steps and keeps the
    # faster form; the second run replays the cached graph
    ug = op.initialize_dof_vector()
    times = []
    for _ in range(2):
        torch.cuda.synchronize()
        start.record()
        op.project_initial(ug, 0.0)
        op._lsrk_time_steps(L.DG_LSRK_5_4, 0.0, dt, n, ug, ri, ki)
        stop.record()
        torch.cuda.synchronize()
        times.append(start.elapsed_time(stop) / 1e3)
    return dict(steps=n, seconds=seconds, rel_l2_error=err / nrm,
                seconds_steps_call=times[1],
                seconds_steps_call_first=times[0],
                steps_call_launch=op.step_launch(),
                steps_call_equals_eager=bool(torch.equal(u, ug)))


def time_cell_solver(refine=5, reps=20):
    """The batched cell solver of the block-Jacobi preconditioner
    (advection_solver_irk_real.cc:2079-2126) on the IRK workload: one
    application against the FDM alone, and FGMRES(30) solves with each."""
    import torch
    from advection_miniapp_amd import AdvectionOperation, make_config
    from advection_miniapp_amd import _lib as L
    from advection_miniapp_amd.operators import solver_control
    op = AdvectionOperation(make_config(**irk_workload(refine=refine)))
    s = 3
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    dt = 5.0 / 2 ** refine / 2.0
    op.set_time(0.0, dt)
    rhs = op.initialize_dof_vector(blocks=s)
    op.compute_rhs(rhs, u)
    out = {}
    x = op.initialize_dof_vector(blocks=s)
    for name, f in (("fdm", op.precondition_block_jacobi_irk),
                    ("cell_solver", op.precondition_cell_solver_irk)):
        f(x, rhs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f(x, rhs)
        torch.cuda.synchronize()
        out[f"{name}_apply_ms"] = 1e3 * (time.perf_counter() - t0) / reps
    for name, kw in (("fp64", dict(prec_precision=L.DG_FP64)),
                     ("fp32_inner4", dict(prec_precision=L.DG_FP32,
                                          inner_iterations=4))):
        for cs in (False, True):
            times = []
            for rep in range(3):
                z = op.initialize_dof_vector(blocks=s)
                ctl = solver_control(max_iterations=400, rel_tolerance=1e-10,
                                     restart=30, cell_solver=cs, **kw)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                op.irk_solve(z, rhs, ctl)
                torch.cuda.synchronize()
                if rep > 0:
                    times.append(time.perf_counter() - t0)
            out[f"{name}_{'cell' if cs else 'fdm'}"] = dict(
                ms=1e3 * min(times), iterations=ctl.iterations)
    return out


def time_implicit_parts(steps=20):
    """The single-stage implicit operator (M/dt + L) on the apply workload
    (advection_solver_implicit_real.cc:876-1058), and Radau IIA(3) time
    steps with the initial-guess projection on the IRK workload
    (advection_solver_irk_real.cc:2484-2779): per-step time and outer
    iterations once the projection has four previous stage solutions."""
    import torch
    from advection_miniapp_amd import AdvectionOperation, make_config
    from advection_miniapp_amd import _lib as L
    from advection_miniapp_amd.operators import (ImplicitRungeKuttaIntegrator,
                                                 solver_control)
    out = {}
    op = AdvectionOperation(make_config(**dict(apply_workload(),
                                               factor_skew=0.5)))
    src = op.initialize_dof_vector()
    op.project_initial(src, 0.0)
    dst = op.initialize_dof_vector()
    op.set_time(0.0, 1e-3)
    for _ in range(3):
        op.vmult(dst, src)
    start, stop = _events()
    torch.cuda.synchronize()
    start.record()
    for _ in range(steps):
        op.vmult(dst, src)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    out["implicit_vmult"] = dict(n_dofs=op.n_dofs, ms=ms,
                                 dof_per_s=op.n_dofs / (ms * 1e-3))
    del op, src, dst
    torch.cuda.empty_cache()
    refine = 5
    op = AdvectionOperation(make_config(**irk_workload(refine)))
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    integ = ImplicitRungeKuttaIntegrator(op)
    dt = 5.0 / 2 ** refine / 2.0
    rows = []
    for step in range(8):
        ctl = solver_control(max_iterations=400, rel_tolerance=1e-10,
                             restart=30, prec_precision=L.DG_FP32,
                             inner_iterations=8)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        integ.perform_time_step(step * dt, dt, u, ctl)
        torch.cuda.synchronize()
        rows.append((1e3 * (time.perf_counter() - t0), ctl.iterations))
    out["irk_steps"] = dict(
        workload="Radau IIA(3) steps, 32^3 cells, degree 4, Courant 5, "
                 "FP32 inner GMRES(8), 1e-10",
        ms_and_iterations_per_step=rows)
    return out


def time_config3():
    """Configuration 3: advection + SIPG diffusion, 32^3 cells, degrees
    3/5/7, a = (1, 0.5, 0.25), nu = 1e-3, penalty (k+1)^2, product of
    sines; FP64 and FP32 operator applies and the FP32 result's error
    against the FP64 one (the FP64 apply equals the oracle to 1e-12 in the
    parity tests)."""
    import math
    import torch
    out = {}
    for k in (3, 5, 7):
        base = dict(dim=3, degree=k, refine=5, velocity=[1.0, 0.5, 0.25],
                    operator_kind=1, viscosity=1e-3,
                    penalty_factor=(k + 1) ** 2, solution_kind=2,
                    solution_param=[2 * math.pi, 1e-3])
        op, src, dst64, ms64 = time_one_apply(base)
        n = op.n_dofs
        tile64 = op.tile_shape()
        del op
        op32, _, dst32, ms32 = time_one_apply(dict(base, precision=1),
                                              src=src.float())
        tile32 = op32.tile_shape()
        del op32
        err = ((dst32.double() - dst64).abs().max() /
               dst64.abs().max()).item()
        out[f"k{k}"] = dict(
            n_dofs=n, fp64_ms=ms64, fp64_dof_per_s=n / (ms64 * 1e-3),
            fp32_ms=ms32, fp32_dof_per_s=n / (ms32 * 1e-3),
            fp32_rel_err_vs_fp64=err,
            tile_fp64=tile64, tile_fp32=tile32)
        del src, dst64, dst32
        torch.cuda.empty_cache()
    return out


def time_burgers():
    """The Burgers add-on of configuration 4: degree 3, 48^3 cells on
    [-1,1]^3 (3 x 3 x 3 coarse cells, 4 refinements), 3 components,
    u0 = (-sin pi x, 0, 0), over-integration QGauss(6): one right-hand
    side per step."""
    bk = dict(dim=3, degree=3, n_components=3, operator_kind=2,
              n_coarse=[3, 3, 3], refine=4, lo=[-1, -1, -1], hi=[1, 1, 1],
              solution_kind=6)
    op, _, _, ms = time_one_apply(bk)
    return dict(workload="3D Burgers, degree 3, 48^3 cells, 3 components, "
                         "over-integration 6^3",
                unknowns=op.vector_size, ms=ms,
                unknowns_per_s=op.vector_size / (ms * 1e-3),
                gb_per_s=16 * op.vector_size / (ms * 1e-3) / 1e9)


def time_sipg_perturbed():
    """Configuration 3's operator (degree 3, 32^3 cells, a = (1, 0.5,
    0.25), nu = 1e-3) on cells perturbed by 10% of h: advection kernel in
    integrate form + general-geometry SIPG kernel + inverse mass (FP64),
    against the fused apply on the axis-aligned mesh."""
    import math
    import torch
    base = dict(dim=3, degree=3, refine=5, velocity=[1.0, 0.5, 0.25],
                operator_kind=1, viscosity=1e-3, penalty_factor=16,
                solution_kind=2, solution_param=[2 * math.pi, 1e-3])
    out = {}
    for name, pert in (("perturbed", 0.1), ("axis_aligned", 0.0)):
        op, src, dst, ms = time_one_apply(dict(base, perturbation=pert))
        out[name] = dict(n_dofs=op.n_dofs, ms=ms,
                         dof_per_s=op.n_dofs / (ms * 1e-3))
        del op, src, dst
        torch.cuda.empty_cache()
    return out


def time_extra():
    """Further measurements: the configuration-2 run, the standalone
    inverse mass, the batched cell solver, the implicit operator parts and
    SIPG on perturbed cells."""
    out = {}
    out["sipg_perturbed"] = time_sipg_perturbed()
    out["config2_run"] = time_config2_run()
    out["inverse_mass"] = time_inverse_mass()
    out["cell_solver"] = time_cell_solver()
    out["implicit"] = time_implicit_parts()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--refine", type=int, default=6)
    ap.add_argument("--apply-geometry", choices=["config2", "uniform"],
                    default="config2",
                    help="uniform: time the apply on the same mesh size "
                         "with axis-aligned cells and a constant speed "
                         "(no geometry tables), for profiling")
    ap.add_argument("--irk-refine", type=int, default=5)
    ap.add_argument("--irk-reps", type=int, default=3)
    ap.add_argument("--irk-variants", default=None,
                    help="comma-separated subset of the IRK variants "
                         "(fp64, fp32_fdm, fp64_inner4, fp32_inner4, "
                         "fp64_inner8, fp32_inner8), for profiling one solve; "
                         "the ratios need both variants of a pair")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-irk", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the configuration 1, 3 and Burgers legs")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving workload (about 20 s)")
    ap.add_argument("--extra", action="store_true",
                    help="also time the apply by degree, the configuration-2 "
                         "run, the inverse mass, the cell solver and the "
                         "implicit operator parts")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-final-time", type=float, default=0.5)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-steps-per-call", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run plain copies of one vector (for the "
                         "HBM counter calibration of rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): the library runs one "
              "operator per GPU and has no multi-GPU decomposition",
              file=sys.stderr)
        sys.exit(2)

    op, ms = time_apply(args)
    n = op.n_dofs
    if args.copy_calibration:
        # a plain device copy of the same size as one vector (read n*8 B,
        # write n*8 B): the counter reading of this kernel calibrates the
        # FETCH_SIZE / WRITE_SIZE readings of the operator kernel
        import torch
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
    bytes_per_apply = 16.0 * n
    achieved = bytes_per_apply / (ms * 1e-3) / 1e9
    traffic = measured_traffic()
    out = {
        "metric": "DoF/s & %HBM peak, 3D p=4 DG apply; mixed vs FP64 IRK "
                  "time at 1e-10 res",
        "value": n / (ms * 1e-3),
        "unit": "DoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D advection apply, degree 4, 64^3 cells "
                               "perturbed 10% (seed 42), cellular velocity, "
                               "FP64" if args.apply_geometry == "config2" else
                               "3D advection apply, degree 4, 64^3 "
                               "axis-aligned cells, constant velocity (no "
                               "geometry tables), FP64",
                   "n_dofs": n, "refine": args.refine},
        "roofline": {"bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "fp64_tflops": APPLY_FLOP_PER_DOF * n / (ms * 1e-3) / 1e12,
        "tile": op.tile,
    }
    del op
    if not args.no_irk:
        out["irk"] = time_irk(args)
    if not args.no_configs:
        out["apply_by_degree"] = time_apply_by_degree()
        out["config1"] = time_config1_run()
        out["config3_sipg"] = time_config3()
        out["config4_burgers"] = time_burgers()
    if not args.no_serving:
        out["serving"] = time_serving(args)
    if args.extra:
        out["extra"] = time_extra()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
